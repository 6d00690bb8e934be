"""GPU path against the oracle beyond the golden sizes, and size-independent
properties at the full lmax = 600, fmax = 40 mHz workload.

* the whole lmax = 150, fmax = 10 mHz workload (7,852 pairs, 10 receivers):
  spectra <= 1e-9, seismograms <= 1e-8 relative L2 against the oracle;
* the full 787,310-pair workload: spectra and seismograms against the
  oracle's (16 host threads); per-pair solver outputs of a spread sample of
  (l, w), highest l and w included, against the oracle's;
  linearity of the spectra in the moment tensor; run-to-run identity."""
import numpy as np
import pytest

from conftest import ysph_rows, rel_l2
from yspec_amd.synth import synth_deck, workload

pytestmark = pytest.mark.gpu
KEYS = ("svt", "ytor", "svs", "ysph", "sr0", "yrad", "uu", "vv", "ww")


@pytest.fixture(scope="module")
def api():
    from yspec_amd import api as a
    a.lib()
    return a


def test_cfg1_full_vs_oracle(api, oracle_mod):
    deck, p = synth_deck(0), workload("cfg1")
    g = api.YSpec(deck, p)
    o = oracle_mod.Oracle(deck, p)
    assert (g.nw, g.nt) == (o.nw, o.nt)
    gs = g.spectra()
    os_ = o.spectra(nthreads=16)
    for a, b, key in zip(gs, os_, ("ur", "ut", "up")):
        e = rel_l2(a, b)
        print("cfg1", key, f"{e:.2e}", flush=True)
        assert e <= 1e-9
    e = rel_l2(g.seismograms(), o.seismograms(*os_))
    print("cfg1 seis", f"{e:.2e}", flush=True)
    assert e <= 1e-8


@pytest.fixture(scope="module")
def cfg2(api):
    deck, p = synth_deck(0), workload("cfg2")
    g = api.YSpec(deck, p)
    return deck, p, g


def test_cfg2_pairs_vs_oracle(cfg2, oracle_mod):
    deck, p, g = cfg2
    g.spectra(copy=False)
    rng = np.random.default_rng(11)
    ls = np.concatenate([[0, 1, 2, p.lmax, p.lmax, 300], rng.integers(0, p.lmax + 1, 60)])
    # j counts frequencies from 1, as the reference's loop does
    js = np.concatenate([[1, g.nw, g.nw // 2, 1, g.nw, g.nw],
                         rng.integers(1, g.nw + 1, 60)])
    rec = g.pair_records(ls, js)
    o = oracle_mod.Oracle(deck, p)
    want = {k: [] for k in KEYS}
    for l, j in zip(ls, js):
        r = o.pair(int(l), int(j))
        for k in KEYS:
            want[k].append(r[k])
    for k in KEYS:
        a, b = rec[k], np.array(want[k])
        if k == "ysph":  # rows the default b-vector path fills
            a, b = a[:, :ysph_rows(p), :], b[:, :ysph_rows(p), :]
        e = rel_l2(a, b)
        print("cfg2 pairs", k, f"{e:.2e}", flush=True)
        assert e <= 1e-9, k


@pytest.mark.timeout(600)
def test_cfg2_full_vs_oracle(cfg2, oracle_mod):
    """The whole lmax = 600, fmax = 40 mHz workload (787,310 pairs, 200
    receivers) against the oracle run on 16 host threads: spectra <= 1e-9,
    seismograms <= 1e-8 relative L2."""
    import time

    deck, p, g = cfg2
    gs = g.spectra()
    seis = g.seismograms()
    o = oracle_mod.Oracle(deck, p)
    t0 = time.perf_counter()
    os_ = o.spectra(nthreads=16)
    print(f"cfg2 oracle spectra: {time.perf_counter() - t0:.1f} s on 16 threads", flush=True)
    for a, b, key in zip(gs, os_, ("ur", "ut", "up")):
        e = rel_l2(a, b)
        print("cfg2 full", key, f"{e:.2e}", flush=True)
        assert e <= 1e-9, key
    e = rel_l2(seis, o.seismograms(*os_))
    print("cfg2 full seis", f"{e:.2e}", flush=True)
    assert e <= 1e-8


def test_cfg2_pairs_vs_reference(cfg2):
    """Per-pair records of the lmax = 600, fmax = 40 mHz workload against the
    reference's own at l = 0, 1, 2, 150, 300, 450, 600 (a spread of each
    degree's 1310 frequencies; tests/golden/cfg2_records.npz from
    oracle/_ref/ref_dump)."""
    from refdump import PAIR_FIELDS
    from test_oracle_golden import _cfg2_fixture

    deck, p, g = cfg2
    _, q, z = _cfg2_fixture()
    assert np.array_equal(q.int_block(), p.int_block())
    g.spectra(copy=False)
    rec = g.pair_records(z["pair_l"], z["pair_j"])
    n, at = ysph_rows(p), 0
    for k, shape in PAIR_FIELDS:
        cnt = int(np.prod(shape))
        want = z["rec"][:, at:at + cnt].reshape((len(z["rec"]),) + shape, order="F")
        at += cnt
        a = rec[k]
        if k == "ysph":  # rows the default b-vector path fills
            a, want = a[:, :n, :], want[:, :n, :]
        e = rel_l2(a, want)
        print("cfg2 pairs vs reference", k, f"{e:.2e}", flush=True)
        assert e <= 1e-9, k


def test_cfg2_rows_match_columns(cfg2, api):
    """The receiver-side row march (transposed steps) against the
    reference's source-side column march on the same pairs: the same
    product of step matrices, so rounding-level differences only."""
    deck, p, g = cfg2
    ur, ut, up = g.spectra()
    h = api.YSpec(deck)
    h.set_records(True)
    h.set_run(p)
    vr, vt, vp = h.spectra()
    for a, b, key in zip((ur, ut, up), (vr, vt, vp), ("ur", "ut", "up")):
        e = rel_l2(a, b)
        print("cfg2 rows/columns", key, f"{e:.2e}", flush=True)
        assert e <= 1e-11, key
    rng = np.random.default_rng(5)
    ls = np.concatenate([[1, 2, 300, p.lmax], rng.integers(1, p.lmax + 1, 60)])
    js = np.concatenate([[1, g.nw, g.nw // 2, g.nw], rng.integers(1, g.nw + 1, 60)])
    ra, rb = g.pair_records(ls, js)["ysph"], h.pair_records(ls, js)["ysph"]
    n = ysph_rows(p)
    e = rel_l2(ra[:, :n, :], rb[:, :n, :])
    print("cfg2 rows/columns ysph", f"{e:.2e}", flush=True)
    assert e <= 1e-11
    assert not np.any(ra[:, n:, :])  # rows the row march does not produce are zero
    h.close()


def test_cfg2_linear_in_moment_tensor(cfg2, api):
    deck, p, g = cfg2
    rng = np.random.default_rng(3)
    m1, m2 = rng.uniform(-1e20, 1e20, 6), rng.uniform(-1e20, 1e20, 6)
    out = []
    for mm in (m1, m2, 2.0 * m1 - 3.0 * m2):
        p.mm = mm
        g.set_run(p)
        out.append(g.spectra())
    for c in range(3):
        e = rel_l2(out[2][c], 2.0 * out[0][c] - 3.0 * out[1][c])
        print("cfg2 linearity", c, f"{e:.2e}", flush=True)
        assert e < 1e-12


def test_cfg2_repeatable(cfg2):
    deck, p, g = cfg2
    a = g.forward()
    b = g.forward()
    assert np.array_equal(a, b)
    assert np.all(np.isfinite(a))


def test_cfg2_pair_steps(cfg2, api):
    """Per-pair step counts: they add up to the run's totals; l = 0 pairs
    take only radial steps; every l >= 1 pair at one w takes the b-vector
    steps that the reference's step rule gives from the source up to the
    receiver (yint, src/module_int.f90:1903-1907: 1 + floor(dr / drr)
    steps per knot interval, drr = 1 / (4 Re w / max(alpha_i, alpha_i+1)),
    steps, src/module_int.f90:2121-2138), restated here in numpy."""
    deck, p, g = cfg2
    g.spectra(copy=False)
    nl = p.lmax - p.lmin + 1
    ls = np.repeat(np.arange(p.lmin, p.lmax + 1), g.nw)
    js = np.tile(np.arange(1, g.nw + 1), nl)
    c = g.pair_steps(ls, js).reshape(nl, g.nw, 5)
    assert np.array_equal(c.reshape(-1, 5).sum(0, dtype=np.int64), g.step_counts())
    assert (c[0, :, :4] == 0).all() and (c[0, :, 4] > 0).all()
    assert (c[1:, :, 4] == 0).all() and (c[1:, :, 3] > 0).all()
    tab = g.model_table()
    r, alpha = tab[0], tab[api.TABLE_FIELDS.index("alpha")]
    nk = len(r)
    assert g.rr >= g.rs
    want = np.zeros(g.nw, np.int64)
    for jj in range(g.nw):
        rew = g.wmin + jj * g.dw
        for i in range(g.is_, g.ir + 1):  # 1-based knot intervals
            a = g.rs if i == g.is_ else r[i - 1]
            b = g.rr if i == g.ir else r[i]
            if a == b:
                continue
            tmp = max(alpha[i - 1], alpha[i]) if i < nk else alpha[i - 1]
            drr = 1.0 / (4.0 * rew / tmp)
            want[jj] += 1 + int(np.floor(abs(b - a) / drr))
    assert np.array_equal(c[1:, :, 3], np.broadcast_to(want, (nl - 1, g.nw)))


def _refine(deck, factor):
    """The deck with every region resampled `factor` times finer (linear in
    each column between the original knots; discontinuities kept)."""
    from yspec_amd.inputs import Deck
    c = deck.cols
    cuts = [0] + [i + 1 for i in range(len(c) - 1) if c[i + 1, 0] == c[i, 0]] + [len(c)]
    rows, nic, noc = [], None, None
    for a, b in zip(cuts[:-1], cuts[1:]):
        seg = c[a:b]
        t = np.linspace(0.0, len(seg) - 1.0, factor * (len(seg) - 1) + 1)
        rows.append(np.stack([np.interp(t, np.arange(len(seg)), seg[:, k])
                              for k in range(c.shape[1])], axis=1))
        n = sum(len(r) for r in rows)
        if b == deck.nic:
            nic = n
        if b == deck.noc:
            noc = n
    cols = np.concatenate(rows)
    return Deck("refined", deck.ians, deck.tref, len(cols), nic, noc, cols)


def test_large_deck_vs_oracle(api, oracle_mod):
    """A deck of about 790 knots: its model table (29 doubles per knot) is
    larger than the LDS of a block, so the kernels that stage the table in
    LDS read it from memory instead; results as for any deck."""
    deck = _refine(synth_deck(0), 4)
    assert deck.nknot * 29 * 8 > 160 * 1024
    p = workload("cfg1", lmax=20, nr=2, ats=1, src_depth=300.0)
    g = api.YSpec(deck, p)
    o = oracle_mod.Oracle(deck, p)
    gs, os_ = g.spectra(), o.spectra(nthreads=16)
    for a, b in zip(gs, os_):
        assert rel_l2(a, b) <= 1e-9
    assert rel_l2(g.seismograms(), o.seismograms(*os_)) <= 1e-8


# receiver exactly at the source depth (a zero-length b-vector and minor
# continuation), a surface source and receiver, and a receiver below the
# source in the mantle; both b-vector paths
@pytest.mark.parametrize("src,rec", [(100.0, 100.0), (0.0, 0.0), (35.0, 410.0)],
                         ids=["rec_at_source", "surface_both", "rec_deeper"])
@pytest.mark.parametrize("full", [False, True], ids=["rows", "columns"])
def test_geometry_vs_oracle(api, oracle_mod, src, rec, full):
    deck = synth_deck(0)
    p = workload("cfg1", lmax=40, nr=3, src_depth=src, rec_depth=rec)
    g = api.YSpec(deck)
    g.set_records(full)
    g.set_run(p)
    o = oracle_mod.Oracle(deck, p)
    gs = g.spectra()
    os_ = o.spectra(nthreads=16)
    for a, b, key in zip(gs, os_, ("ur", "ut", "up")):
        e = rel_l2(a, b)
        print(src, rec, "full" if full else "rows", key, f"{e:.2e}", flush=True)
        assert e <= 1e-9, key
    e = rel_l2(g.seismograms(), o.seismograms(*os_))
    print(src, rec, "seis", f"{e:.2e}", flush=True)
    assert e <= 1e-8
    g.close()


def test_long_series_single_pass_ifft(api, oracle_mod):
    """nt = 131072, past the 65536 limit of the two-pass transform: the
    single-pass inverse FFT (k_seis) against the oracle's ifft and
    post-processing applied to the GPU's own spectra, and the whole forward
    against the oracle."""
    import dataclasses

    from conftest import load_case

    deck, p, _ = load_case("cfg1_small")
    p = dataclasses.replace(p, lmax=10, f1=0.1, f2=1.0, tout=1500.0, dt=1.0, f11=0.2,
                            f12=0.3, f21=0.7, f22=0.9)
    g = api.YSpec(deck, p)
    assert g.nt == 131072
    spec = g.spectra()
    seis = g.seismograms()
    o = oracle_mod.Oracle(deck, p)
    e_ifft = rel_l2(seis, o.seismograms(*spec))
    os_ = o.spectra(nthreads=16)
    e_spec = max(rel_l2(a, b) for a, b in zip(spec, os_))
    e_seis = rel_l2(seis, o.seismograms(*os_))
    print(f"long series: ifft {e_ifft:.2e}, spectra {e_spec:.2e}, seismograms {e_seis:.2e}",
          flush=True)
    assert e_ifft <= 1e-12 and e_spec <= 1e-9 and e_seis <= 1e-8
    g.close()
