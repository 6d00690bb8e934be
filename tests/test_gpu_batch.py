"""Several events on one model: the plain path (one event after another)
gives each event exactly the single-event result, and the batched path
(all events per launch from cost-sorted work lists, replayed from a HIP
graph; one block per item or persistent blocks on atomic queues; marches
shared between events) gives it to rounding; the oracle checks both at 1e-9
(spectra) / 1e-8 (seismograms)."""
import numpy as np
import pytest

from conftest import rel_l2
from yspec_amd.synth import batch_events, synth_deck

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from yspec_amd import api as a
    a.lib()
    return a


def small_events():
    ev = batch_events(nev=5, seed=3, nr=3, fmax_range=(3.0, 6.0), tout=40.0, dt=20.0,
                      lmax_per_mhz=6.0)
    ev[1].src_depth = 650.0   # deep: minors start in the core at low l
    ev[2].src_depth = 1.5     # shallow crustal source
    ev[3].lmin = 2            # no radial pairs in this event
    ev[4].grav = 1            # Cowling: spheroidal solves in the per-lane kernel
    ev[0].grav = 0            # no gravity: per-lane spheroidal and radial solvers
    for p in ev:
        p.f21, p.f22 = 0.6 * p.f2, 0.8 * p.f2
    return ev


@pytest.fixture(scope="module")
def single(api):
    deck = synth_deck(0)
    out = []
    for p in small_events():
        g = api.YSpec(deck, p)
        spec = g.spectra()
        out.append((spec, g.seismograms()))
        g.close()
    return out


def _check_events(g, single, tol=0.0):
    """tol = 0: bit-identical; else relative L2"""
    for e, (spec, seis) in enumerate(single):
        got = g.event_spectra(e)
        for a, b in zip(got, spec):
            assert (np.array_equal(a, b) if tol == 0.0 else rel_l2(a, b) <= tol), e
        s = g.event_seismograms(e)
        assert (np.array_equal(s, seis) if tol == 0.0 else rel_l2(s, seis) <= tol), e


def _snapshot(g, n):
    return [(g.event_spectra(e), g.event_seismograms(e)) for e in range(n)]


def test_plain_and_batched_equal_single(api, single):
    """The plain path gives every event its single-event result bit for bit.
    The batched path shares the b-vector rows (bit-identical) and, for the
    events that share marches, takes the secular determinant from the
    transposed minor row (rounding-level differences); its graph replays and
    launch shapes agree bit for bit with each other."""
    g = api.YSpec(synth_deck(0))
    g.set_events(small_events())
    g.forward_events(0)
    _check_events(g, single)
    g.forward_events(1)
    _check_events(g, single, tol=1e-12)
    first = _snapshot(g, len(single))
    for rep in range(3):  # graph replays, with eager work in between
        g.forward_events(0)
        _check_events(g, single)
        g.forward_events(1)
        for (sa, ka), (sb, kb) in zip(_snapshot(g, len(single)), first):
            assert all(np.array_equal(x, y) for x, y in zip(sa, sb)) and np.array_equal(ka, kb)
    g.set_batch_tuning(persist=1, chunk=2, blocks_per_cu=2)
    for rep in range(2):
        g.forward_events(1)
        for (sa, ka), (sb, kb) in zip(_snapshot(g, len(single)), first):
            assert all(np.array_equal(x, y) for x, y in zip(sa, sb)) and np.array_equal(ka, kb)


def test_radial_lanes_knob(api, single):
    """ysx_set_lanes: one lane or eight per l = 0 radial solve, set before
    or after the events (the work lists are rebuilt), on the plain and the
    batched path: every event agrees with its single-event result to
    rounding; other values are refused."""
    g = api.YSpec(synth_deck(0))
    g.set_lanes(1)
    g.set_events(small_events())
    for n in (1, 8, 0, 1):
        g.set_lanes(n)
        for mode in (0, 1):
            g.forward_events(mode)
            _check_events(g, single, tol=1e-12)
    with pytest.raises(api.YSpecError):
        g.set_lanes(4)
    g.close()


def test_tuner_applies_its_fastest_setting(api, single):
    """yspec_amd.tune: times the given launch shapes and both lane counts per
    radial solve on the loaded events, reports every candidate, and leaves
    the fastest applied; the results under it agree with the single-event
    ones to rounding."""
    from yspec_amd.tune import tune

    g = api.YSpec(synth_deck(0))
    g.set_events(small_events())
    shapes = [(0, 1, 4), (1, 2, 2)]
    out = tune(g, reps=1, candidates=shapes)
    best = out["best"]
    assert [(c["persist"], c["chunk"], c["blocks_per_cu"]) for c in out["candidates"]] == shapes
    assert sorted(r["rad_lanes"] for r in out["rad_lanes"]) == [1, 8]
    assert all(c["ms"] > 0.0 for c in out["candidates"] + out["rad_lanes"])
    times = [c["ms"] for c in out["candidates"]] + [r["ms"] for r in out["rad_lanes"]]
    assert best["ms"] == min(times)
    assert (best["persist"], best["chunk"], best["blocks_per_cu"]) in shapes
    assert best["rad_lanes"] in (0, 1, 8)
    g.forward_events(1)
    _check_events(g, single, tol=1e-12)
    g.close()


def test_events_vs_oracle(api, oracle_mod, single):
    deck = synth_deck(0)
    for e, p in enumerate(small_events()):
        o = oracle_mod.Oracle(deck, p)
        ospec = o.spectra(nthreads=16)
        spec, seis = single[e]
        for a, b in zip(spec, ospec):
            assert rel_l2(a, b) <= 1e-9, e
        assert rel_l2(seis, o.seismograms(*ospec)) <= 1e-8, e


def test_bad_event_is_named(api):
    ev = small_events()
    ev[2].f2 = 30.0  # above the 25 mHz Nyquist frequency of dt = 20 s
    g = api.YSpec(synth_deck(0))
    with pytest.raises(api.YSpecError, match="event 2"):
        g.set_events(ev)


def test_batched_columns_match_rows(api, single):
    """ysx_set_records(ctx, 1) before the events are set: the batched path
    integrates the b-vector system as the reference's four source-side
    columns; spectra and seismograms match the default receiver-side row
    march (the single-event results) to rounding."""
    g = api.YSpec(synth_deck(0))
    g.set_records(True)
    g.set_events(small_events())
    g.forward_events(1)
    for e, (spec, seis) in enumerate(single):
        for a, b in zip(g.event_spectra(e), spec):
            assert rel_l2(a, b) <= 1e-11, e
        assert rel_l2(g.event_seismograms(e), seis) <= 1e-11, e
    g.close()


def test_shared_marches_edge_cases(api, oracle_mod):
    """Events sharing the receiver-side row marches: two sources in one knot
    interval, two identical sources, a source at the receiver (zero-length
    march) and different frequency ranges; batched results against the
    single-event ones (1e-12) and the oracle (1e-9 / 1e-8)."""
    deck = synth_deck(0)
    ev = batch_events(nev=5, seed=7, nr=2, fmax_range=(4.0, 6.0), tout=40.0, dt=20.0,
                      lmax_per_mhz=5.0)
    depths = (300.0, 301.0, 120.0, 120.0, 0.0)
    for p, d in zip(ev, depths):
        p.src_depth = d
        p.f21, p.f22 = 0.6 * p.f2, 0.8 * p.f2
    g = api.YSpec(deck)
    g.set_events(ev)
    g.forward_events(1)
    for e, p in enumerate(ev):
        h = api.YSpec(deck, p)
        spec, seis = h.spectra(), h.seismograms()
        h.close()
        for a, b in zip(g.event_spectra(e), spec):
            assert rel_l2(a, b) <= 1e-12, (e, p.src_depth)
        assert rel_l2(g.event_seismograms(e), seis) <= 1e-12, (e, p.src_depth)
        o = oracle_mod.Oracle(deck, p)
        os_ = o.spectra(nthreads=16)
        for a, b in zip(g.event_spectra(e), os_):
            assert rel_l2(a, b) <= 1e-9, (e, p.src_depth)
        assert rel_l2(g.event_seismograms(e), o.seismograms(*os_)) <= 1e-8, (e, p.src_depth)
    g.close()


def test_64_event_workload(api, oracle_mod):
    """The 64-event workload the batched benchmark times (depth U[0,700] km,
    fmax U[10,40] mHz, lmax = 15 fmax, 16 receivers each, 180 min at 1 s):
    the batched path under each launch shape the tuner may pick (one block
    per item; persistent blocks) gives every event its plain-path spectra
    and seismograms to 1e-12, and the three smallest events match the
    oracle at 1e-9 (spectra) / 1e-8 (seismograms)."""
    ev = batch_events()
    deck = synth_deck(0)
    g = api.YSpec(deck)
    g.set_events(ev)
    g.forward_events(0)
    plain = [(g.event_spectra(e), g.event_seismograms(e)) for e in range(len(ev))]
    for shape in ((0, 1, 4), (1, 1, 4), (1, 4, 8)):
        g.set_batch_tuning(*shape)
        g.forward_events(1)
        worst = 0.0
        for e, (spec, seis) in enumerate(plain):
            for a, b in zip(g.event_spectra(e), spec):
                worst = max(worst, rel_l2(a, b))
            worst = max(worst, rel_l2(g.event_seismograms(e), seis))
        print("64 events, batched shape", shape, f"worst {worst:.2e}", flush=True)
        assert worst <= 1e-12, shape
    npairs = [g.info[e]["npair"] for e in range(len(ev))]
    for e in np.argsort(npairs)[:3]:
        o = oracle_mod.Oracle(deck, ev[e])
        os_ = o.spectra(nthreads=16)
        es = max(rel_l2(a, b) for a, b in zip(plain[e][0], os_))
        ek = rel_l2(plain[e][1], o.seismograms(*os_))
        print(f"event {e}: {npairs[e]} pairs, oracle spectra {es:.2e} seis {ek:.2e}",
              flush=True)
        assert es <= 1e-9 and ek <= 1e-8, e
    g.close()


def test_fluid_core_receiver_below_source_is_rejected(api):
    """Self-gravitation, a fluid inner core and the receiver below the
    source: the reference's sph_solver_fic starts solid minors in the fluid
    inner core on that branch (src/module_solver.f90:1323-1331) and gives no
    result to match, so the run is refused with YSX_ERR_UNSUPPORTED (4)
    rather than computed; the same deck with the receiver above the source,
    or l = 0 only, is accepted."""
    from make_golden import fic_deck
    from yspec_amd.synth import workload

    deck = fic_deck(0)
    g = api.YSpec(deck)
    below = workload("cfg1", lmax=8, nr=1, src_depth=20.0, rec_depth=150.0)
    with pytest.raises(api.YSpecError) as ei:
        g.set_run(below)
    assert ei.value.code == 4 and "fluid inner core" in ei.value.msg
    ev = small_events()[:2]
    ev[1].grav, ev[1].src_depth, ev[1].rec_depth = 2, 20.0, 150.0
    with pytest.raises(api.YSpecError, match="event 1"):
        g.set_events(ev)
    g.set_run(workload("cfg1", lmax=8, nr=1, src_depth=150.0, rec_depth=20.0))
    g.set_run(workload("cfg1", lmin=0, lmax=0, nr=1, src_depth=20.0, rec_depth=150.0))
    g.close()
