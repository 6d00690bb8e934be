"""The gradient of an L2 waveform misfit with respect to the deck's rho, vp,
vs and Qmu at its knots: ysx_model_vjp (the discrete adjoint of every
per-pair solve, starting solution and source vector) followed by the host
adjoint of the deck -> table transform, against central differences
(h = 1e-6 relative) on seeded knots:
  - of the reference program's own misfit (tests/golden/fd_reference.npz,
    written by tests/golden/make_golden.py from the reference modules),
  - of GPU forwards, at the lmax = 300, 20 mHz, 50 km, 50-receiver gradient
    workload and at smaller runs with and without attenuation.
The data are seismograms of the same deck with 2% Gaussian bumps in vp and
vs.  The reference has no gradient of its own."""
import os

import numpy as np
import pytest

from yspec_amd.inputs import Deck, Params
from yspec_amd.synth import bump_deck as bumped, synth_deck, workload

pytestmark = pytest.mark.gpu
COLS = {1: "rho", 2: "vp", 3: "vs", 5: "Qmu"}


@pytest.fixture(scope="module")
def api():
    from yspec_amd import api as a
    a.lib()
    return a


def seeded_entries(deck, n, seed):
    """n (knot, column) entries with nonzero values, spread over the deck"""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        k, c = int(rng.integers(0, deck.nknot)), int(rng.choice(list(COLS)))
        if deck.cols[k, c] != 0.0 and (k, c) not in out:
            out.append((k, c))
    return out


def check_gradient(api, deck, p, entries):
    g = api.YSpec(bumped(deck), p)
    obs = g.forward()
    g.close()
    g = api.YSpec(deck, p)
    nout = g.nout
    misfit, dbar = g.misfit_gradient(obs)
    print(f"misfit {misfit:.4e}, gradient kernels {g.grad_timings()[-1]:.2f} ms", flush=True)

    def J(cols):
        h = api.YSpec(Deck(deck.title, deck.ians, deck.tref, deck.nknot, deck.nic, deck.noc,
                           cols), p)
        r = h.forward()[:nout] - obs[:nout]
        h.close()
        return 0.5 * float(np.sum(r * r))

    got, fd, val = [], [], []
    for k, c in entries:
        v = deck.cols[k, c]
        h = 1e-6 * abs(v)
        cp, cm = deck.cols.copy(), deck.cols.copy()
        cp[k, c] += h
        cm[k, c] -= h
        fd.append((J(cp) - J(cm)) / (2 * h))
        got.append(dbar[k, c])
        val.append(v)
    got, fd, val = np.array(got), np.array(fd), np.array(val)
    # errors on the scale of the largest d misfit / d log(parameter)
    scale = np.max(np.abs(fd * val))
    err = np.abs(got - fd) * np.abs(val) / scale
    for (k, c), a, b, e in zip(entries, got, fd, err):
        print(f"knot {k:3d} {COLS[c]:4s} gradient {a: .6e} fd {b: .6e} scaled error {e:.1e}",
              flush=True)
    rel = float(np.linalg.norm((got - fd) * val) / np.linalg.norm(fd * val))
    print(f"relative L2 over the entries {rel:.2e}", flush=True)
    g.close()
    return rel, err


def test_gradient_vs_central_differences(api):
    """Attenuation on, all degrees from 0 (radial solves included), a 100 km
    source: 20 seeded entries"""
    deck = synth_deck(0)
    p = workload("cfg1", lmin=0, lmax=40, nr=3, ats=1, src_depth=100.0)
    rel, err = check_gradient(api, deck, p, seeded_entries(deck, 20, 17))
    assert rel <= 1e-6 and np.all(err <= 1e-6)


def test_gradient_deep_source_no_attenuation(api):
    """Attenuation off, a 600 km source, degrees 0..30"""
    deck = synth_deck(0)
    p = workload("cfg1", lmin=0, lmax=30, nr=2, ats=0, src_depth=600.0, f2=6.0, tout=40.0,
                 dt=20.0, f21=4.0, f22=5.5)
    rel, err = check_gradient(api, deck, p, seeded_entries(deck, 16, 23))
    assert rel <= 1e-6 and np.all(err <= 1e-6)


def test_gradient_unsupported_runs(api):
    """Runs outside the adjoint's scope are refused (code 4), not computed"""
    from make_golden import ocean_deck

    deck = synth_deck(0)
    for d, p in ((deck, workload("cfg1", lmax=8, nr=1, rec_depth=10.0)),
                 (deck, workload("cfg1", lmax=8, nr=1, grav=1)),
                 (ocean_deck(0), workload("cfg1", lmax=8, nr=1, rec_depth=4.0))):
        g = api.YSpec(d, p)
        g.forward()
        g.seismograms_vjp(np.ones((g.nt, 3, p.nr)))
        g.spectra_vjp(np.zeros(0, np.int32), np.zeros(0, np.int32))
        with pytest.raises(api.YSpecError) as ei:
            g.model_vjp()
        assert ei.value.code == 4
        g.close()


def _fixture_run(z):
    deck = Deck("fd", int(z["deck_head"][3]), float(z["deck_tref"]), int(z["deck_head"][0]),
                int(z["deck_head"][1]), int(z["deck_head"][2]), z["deck_cols"])
    ip, fp = z["ip"], z["fp"]
    p = Params(prefix="yspec.out", model_file="fd", ats=int(ip[2]), grav=int(ip[3]),
               out=int(ip[4]), cor=int(ip[5]), lmin=int(ip[0]), lmax=int(ip[1]), f1=fp[0],
               f2=fp[1], tout=fp[2], dt=fp[3], f11=fp[4], f12=fp[5], f21=fp[6], f22=fp[7],
               src_depth=fp[8], src_lat=fp[9], src_lon=fp[10], mm=fp[11:17].copy(),
               rec_depth=fp[17], lat=z["lat"], lon=z["lon"])
    return deck, p


def test_gradient_vs_reference_differences(api):
    """ysx_misfit_grad against central differences of the reference
    program's misfit (lmax 40, 6 mHz, 50 km source, attenuation, 4
    receivers; obs from the reference on the bumped deck): the misfit to
    1e-6 and 20 seeded entries to 1e-5 (relative L2, and each entry on the
    scale of the largest d misfit / d log(parameter))"""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "fd_reference.npz"))
    deck, p = _fixture_run(z)
    g = api.YSpec(deck, p)
    obs = z["obs"]
    assert g.nout == obs.shape[0]
    misfit, grad = g.misfit_grad(obs)
    g.close()
    assert abs(misfit - float(z["misfit"])) <= 1e-6 * float(z["misfit"]), (misfit, z["misfit"])
    ent = z["entries"]
    got = np.array([grad[k, q] for k, q in ent])
    fd = z["fd"]
    val = np.array([deck.cols[k, (1, 2, 3, 5)[q]] for k, q in ent])
    scale = np.max(np.abs(fd * val))
    err = np.abs(got - fd) * np.abs(val) / scale
    rel = float(np.linalg.norm((got - fd) * val) / np.linalg.norm(fd * val))
    for (k, q), a, b, e in zip(ent, got, fd, err):
        print(f"knot {k:3d} {api.KNOT_PARAMS[q]:4s} gradient {a: .6e} reference fd {b: .6e} "
              f"scaled error {e:.1e}", flush=True)
    print(f"misfit {misfit:.10e} reference {float(z['misfit']):.10e}; relative L2 {rel:.2e}",
          flush=True)
    assert rel <= 1e-5 and np.all(err <= 1e-5)


def test_gradient_workload_vs_central_differences(api):
    """The gradient workload itself (lmax 300, 20 mHz, 120 min at 2 s, 50 km
    source, 50 receivers, attenuation, 98,728 pairs): ysx_misfit_grad against
    central differences of GPU forwards at 20 seeded (knot, parameter)
    entries, to 1e-5"""
    deck = synth_deck(0)
    p = workload("cfg4")
    g = api.YSpec(bumped(deck), p)
    obs = g.forward()
    g.close()
    g = api.YSpec(deck, p)
    nout = g.nout
    misfit, grad = g.misfit_grad(obs)
    print(f"pairs {g.npair}, misfit {misfit:.6e}, gradient kernels {g.grad_timings()[-1]:.1f} ms",
          flush=True)
    g.close()
    knots = np.stack([deck.cols[:, c] for c in (1, 2, 3, 5)], axis=1)
    rng = np.random.default_rng(20)
    ent = []
    while len(ent) < 20:
        k, q = int(rng.integers(0, deck.nknot)), int(rng.integers(0, 4))
        if knots[k, q] != 0.0 and (k, q) not in ent:
            ent.append((k, q))
    h = api.YSpec(deck, p)

    def J(kn):
        h.set_model(api.deck_with_knots(deck, kn))
        h.set_run(p)
        r = h.forward()[:nout] - obs[:nout]
        return 0.5 * float(np.sum(r * r))

    got, fd, val = [], [], []
    for k, q in ent:
        v = knots[k, q]
        kp, km = knots.copy(), knots.copy()
        kp[k, q] = v * (1 + 1e-6)
        km[k, q] = v * (1 - 1e-6)
        fd.append((J(kp) - J(km)) / (kp[k, q] - km[k, q]))
        got.append(grad[k, q])
        val.append(v)
    h.close()
    got, fd, val = map(np.array, (got, fd, val))
    scale = np.max(np.abs(fd * val))
    err = np.abs(got - fd) * np.abs(val) / scale
    rel = float(np.linalg.norm((got - fd) * val) / np.linalg.norm(fd * val))
    for (k, q), a, b, e in zip(ent, got, fd, err):
        print(f"knot {k:3d} {api.KNOT_PARAMS[q]:4s} gradient {a: .6e} fd {b: .6e} "
              f"scaled error {e:.1e}", flush=True)
    print(f"relative L2 over the entries {rel:.2e}", flush=True)
    assert rel <= 1e-5 and np.all(err <= 1e-5)


def test_forward_stored_states_match_the_sweep_march(api):
    """ysx_misfit_grad's single-event forward stores the minor march's state
    before every step, and the minor sweep starts from those states instead
    of marching again; misfit_gradient (a plain forward, then the entry
    points one by one) leaves the sweep to march them itself.  On the
    gradient workload (lmax 300, 20 mHz, 50 receivers, 98,728 pairs) the
    two gradients agree to 1e-12 of the largest entry, twice in a row (the
    second time with the observed data packed beforehand by pack_obs)."""
    deck = synth_deck(0)
    p = workload("cfg4")
    g = api.YSpec(bumped(deck), p)
    obs = g.forward()
    g.set_model(deck)
    g.set_run(p)
    packed = g.pack_obs(obs)
    with pytest.raises(ValueError):
        g.misfit_grad(packed[:-1])
    for trial in range(2):
        misfit, grad = g.misfit_grad(obs if trial == 0 else packed)
        m2, db = g.misfit_gradient(obs)
        grad2 = api.knot_gradient(db)
        d = float(np.max(np.abs(grad2 - grad)))
        print(f"trial {trial}: misfit {misfit:.9e} / {m2:.9e}, max |stored - marched| "
              f"{d:.2e} of {np.max(np.abs(grad)):.2e}", flush=True)
        assert abs(m2 - misfit) <= 1e-12 * misfit
        assert d <= 1e-12 * np.max(np.abs(grad))
    g.close()


def test_stored_states_with_few_sweep_items(api):
    """The stored-state sweeps cut the minor items into three cost-sorted
    groups (the first eighth, up to half, the rest); with seven items (lmax
    24, 7 frequencies) the first group is empty.  The gradient still equals
    the one whose sweeps march the states themselves, to 1e-12 of the
    largest entry."""
    deck = synth_deck(0)
    p = workload("cfg1", lmax=24, nr=2, ats=1, src_depth=50.0, f2=1.2, f21=0.9, f22=1.1)
    g = api.YSpec(bumped(deck), p)
    assert g.nw * ((p.lmax - p.lmin + 1 + 63) // 64) < 8
    obs = g.forward()
    g.set_model(deck)
    g.set_run(p)
    misfit, grad = g.misfit_grad(obs)
    m2, db = g.misfit_gradient(obs)
    grad2 = api.knot_gradient(db)
    d = float(np.max(np.abs(grad2 - grad)))
    print(f"nw {g.nw}: misfit {misfit:.9e} / {m2:.9e}, max |stored - marched| {d:.2e} of "
          f"{np.max(np.abs(grad)):.2e}", flush=True)
    assert np.max(np.abs(grad)) > 0.0
    assert abs(m2 - misfit) <= 1e-12 * misfit
    assert d <= 1e-12 * np.max(np.abs(grad))
    g.close()


def test_autograd_knot_gradient(api):
    """KnotSeismograms (torch.autograd) against ysx_misfit_grad: the same
    misfit through torch, backward through the library's adjoints"""
    import torch

    from yspec_amd.autograd import knot_seismograms

    deck = synth_deck(0)
    p = workload("cfg1", lmin=0, lmax=24, nr=2, ats=1, src_depth=50.0)
    g = api.YSpec(bumped(deck), p)
    obs = g.forward()
    g.set_model(deck)
    g.set_run(p)
    nout = g.nout
    misfit, grad = g.misfit_grad(obs)
    knots = torch.tensor(np.stack([deck.cols[:, c] for c in (1, 2, 3, 5)], axis=1),
                         dtype=torch.float64, requires_grad=True)
    seis = knot_seismograms(knots, g, deck, p)
    loss = 0.5 * ((seis[:nout] - torch.from_numpy(obs[:nout])) ** 2).sum()
    loss.backward()
    g.close()
    tg = knots.grad.numpy()
    lv = float(loss.detach())
    print(f"misfit {misfit:.6e} torch {lv:.6e}", flush=True)
    assert abs(lv - misfit) <= 1e-12 * misfit
    assert np.max(np.abs(tg - grad)) <= 1e-12 * np.max(np.abs(grad))


def test_batched_events_gradient_is_the_sum(api):
    """Several events on one context (one batched forward launch, each
    adjoint sweep one persistent launch over every event's pairs): the
    misfit and the knot gradient are the sums of the single-event ones
    (events with different depths, frequency grids, receivers, all from
    degree 0)"""
    deck = synth_deck(0)
    events = [workload("cfg1", seed=s, lmin=0, lmax=20, nr=2, ats=1, src_depth=d, f2=f,
                       f21=0.8 * f, f22=0.9 * f)
              for s, d, f in ((1, 50.0, 5.0), (2, 300.0, 6.0), (3, 650.0, 4.0))]
    obs = []
    for p in events:
        h = api.YSpec(bumped(deck), p)
        obs.append(h.forward())
        h.close()
    total, gsum = 0.0, np.zeros((deck.nknot, 4))
    for p, o in zip(events, obs):
        h = api.YSpec(deck, p)
        m, gr = h.misfit_grad(o)
        h.close()
        total += m
        gsum += gr
    g = api.YSpec(deck)
    g.set_events(events)
    mb, gb = g.misfit_grad(obs)
    g.close()
    print(f"misfit batched {mb:.12e} summed {total:.12e}; gradient max difference "
          f"{np.max(np.abs(gb - gsum)) / np.max(np.abs(gsum)):.2e} of its largest entry",
          flush=True)
    assert abs(mb - total) <= 1e-10 * total
    assert np.max(np.abs(gb - gsum)) <= 1e-9 * np.max(np.abs(gsum))
