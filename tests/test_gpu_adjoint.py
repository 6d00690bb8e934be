"""Adjoints of the linear back end (ysx_seismograms_vjp, ysx_spectra_vjp)
through the C-ABI.  The reference has no adjoint, so these are pinned by
what the mathematics fixes:

- the dot-product identity <x_bar, J x> = <J^T x_bar, x> with J the
  forward's own discrete map (seismograms from the spectra, spectra from
  the coefficients uu, vv, ww), x the forward's values and x_bar random;
  complex pairs enter as sum Re(conj(a) b);
- Euler's identity for the coefficient assembly, which is linear in the
  source vectors for fixed solutions and linear in the solutions for fixed
  source vectors: per pair, <c_bar, c> = <sv_bar, sv> = <y_bar, y>.

A plausible mistake (a missing conjugate, a wrong sign or factor on one
coefficient plane, a dropped Hermitian-fill term or filter weight) breaks
one of them: every case has a general moment tensor, so no coefficient plane
that reaches the spectra is zero."""
import dataclasses

import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu

# no_filter: all filter corners zero; receivers_37: more than two 16-receiver
# tiles; long_series: nt = 131072, past the two-pass FFT limit (the
# single-pass global-memory transforms on both sides)
CASES = ["cfg1_small", "aniso_acc_cor", "cow_ocean_vel", "epicentre_ragged", "radial_only",
         "one_degree", "prem_example", "no_filter", "receivers_37", "long_series"]
SRC = ("svt", "svs", "sr0")
SOL = ("ytor", "ysph", "yrad")
TOL = 1e-11


@pytest.fixture(scope="module")
def api():
    from yspec_amd import api as a
    a.lib()
    return a


def _case(name):
    if name == "no_filter":
        deck, p, _ = load_case("prem_noocean_disp")
        return deck, dataclasses.replace(p, f11=0.0, f12=0.0, f21=0.0, f22=0.0)
    if name == "receivers_37":
        deck, p, _ = load_case("cfg1_small")
        rng = np.random.default_rng(37)
        lat = rng.uniform(-80.0, 80.0, 37)
        lon = rng.uniform(-180.0, 180.0, 37)
        return deck, dataclasses.replace(p, lat=lat, lon=lon)
    if name == "long_series":
        deck, p, _ = load_case("cfg1_small")
        return deck, dataclasses.replace(p, lmax=10, f1=0.1, f2=1.0, tout=1500.0, dt=1.0,
                                         f11=0.2, f12=0.3, f21=0.7, f22=0.9)
    deck, p, _ = load_case(name)
    return deck, p


def _forward(api, name):
    deck, p = _case(name)
    g = api.YSpec(deck)
    g.set_run(p)
    spec = g.spectra()
    seis = g.seismograms()
    nl = p.lmax - p.lmin + 1
    ls = np.repeat(np.arange(p.lmin, p.lmax + 1), g.nw)
    js = np.tile(np.arange(1, g.nw + 1), nl)
    return g, p, spec, seis, ls, js


def _rdot(a, b):
    return float(np.sum((np.conj(a) * b).real))


def _coefs(rec):
    return np.concatenate([rec["uu"], rec["vv"], rec["ww"]], axis=1)


@pytest.mark.parametrize("name", CASES)
def test_seismograms_vjp_dot_product(name, api):
    g, p, spec, seis, _, _ = _forward(api, name)
    rng = np.random.default_rng(1)
    seis_bar = rng.standard_normal(seis.shape)
    spec_bar = g.seismograms_vjp(seis_bar)
    lhs = float(np.sum(seis_bar * seis))
    rhs = sum(_rdot(b, s) for b, s in zip(spec_bar, spec))
    scale = np.linalg.norm(seis_bar) * np.linalg.norm(seis)
    print(name, f"<sb, J s> = {lhs:.15e}  <J^T sb, s> = {rhs:.15e}  "
          f"rel {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= TOL * scale
    # bins past nt/2 never reach the series
    for b in spec_bar:
        assert np.all(b[np.arange(g.nw) + g.i1 - 1 > g.nt // 2] == 0)


@pytest.mark.parametrize("name", CASES)
def test_spectra_vjp_dot_product(name, api):
    g, p, spec, _, ls, js = _forward(api, name)
    rng = np.random.default_rng(2)
    spec_bar = [rng.standard_normal(s.shape) + 1j * rng.standard_normal(s.shape) for s in spec]
    rb = g.spectra_vjp(ls, js, spec_bar)
    c = _coefs(g.pair_records(ls, js))
    cb = _coefs(rb)
    lhs = sum(_rdot(b, s) for b, s in zip(spec_bar, spec))
    rhs = _rdot(cb, c)
    scale = max(sum(np.linalg.norm(b) * np.linalg.norm(s) for b, s in zip(spec_bar, spec)),
                float(np.sum(np.abs(cb) * np.abs(c))))
    print(name, f"<sb, L c> = {lhs:.15e}  <L^T sb, c> = {rhs:.15e}  "
          f"rel {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= TOL * scale
    # ww0 never reaches the spectra
    assert np.all(rb["ww"][:, 2] == 0)


@pytest.mark.parametrize("name", CASES)
def test_assembly_vjp_euler(name, api):
    g, p, spec, seis, ls, js = _forward(api, name)
    rng = np.random.default_rng(3)
    g.seismograms_vjp(rng.standard_normal(seis.shape))
    rb = g.spectra_vjp(ls, js)  # from the device-resident spectra adjoint
    rec = g.pair_records(ls, js)
    n = len(ls)

    def per_pair(keys, a, b):
        return sum(np.sum((np.conj(a[k].reshape(n, -1)) * b[k].reshape(n, -1)).real, axis=1)
                   for k in keys)

    cc = np.sum((np.conj(_coefs(rb)) * _coefs(rec)).real, axis=1)
    scale = np.sum(np.abs(_coefs(rb)) * np.abs(_coefs(rec)), axis=1) + 1e-300
    es = np.abs(per_pair(SRC, rb, rec) - cc) / scale
    ey = np.abs(per_pair(SOL, rb, rec) - cc) / scale
    print(name, f"pairs {n}  source side {es.max():.2e}  solution side {ey.max():.2e}",
          flush=True)
    assert es.max() <= 1e-12 and ey.max() <= 1e-12
    # rows 3..5 of ysph and source-vector columns 3 and 6 do not enter
    assert np.all(rb["ysph"][:, 3:, :] == 0)
    assert np.all(rb["svs"][:, :, [2, 5]] == 0)


def test_chain_dot_product(api):
    """seismograms_vjp followed by spectra_vjp on the device: <sb, seis> =
    sum over pairs of <c_bar, c>"""
    g, p, spec, seis, ls, js = _forward(api, "cfg1_small")
    rng = np.random.default_rng(4)
    sb = rng.standard_normal(seis.shape)
    g.seismograms_vjp(sb)
    cb = _coefs(g.spectra_vjp(ls, js))
    c = _coefs(g.pair_records(ls, js))
    lhs = float(np.sum(sb * seis))
    rhs = _rdot(cb, c)
    scale = max(np.linalg.norm(sb) * np.linalg.norm(seis), float(np.sum(np.abs(cb) * np.abs(c))))
    assert abs(lhs - rhs) <= TOL * scale


def test_vjp_errors(api):
    deck, p = _case("cfg1_small")
    g = api.YSpec(deck)
    g.set_run(p)
    with pytest.raises(api.YSpecError):
        g.spectra_vjp([p.lmin], [1])  # no forward yet
    g.spectra()
    with pytest.raises(api.YSpecError):
        g.spectra_vjp([p.lmin], [1])  # no spectra adjoint on the device yet
    with pytest.raises(api.YSpecError):
        g.spectra_vjp([p.lmax + 1], [1], [np.zeros((g.nw, p.nr), complex)] * 3)
    with pytest.raises(ValueError):
        g.seismograms_vjp(np.zeros((g.nt + 1, 3, p.nr)))


@pytest.mark.parametrize("name", ["cfg1_small", "aniso_acc_cor", "prem_example", "radial_only",
                                  "epicentre_ragged"])
def test_moment_tensor_gradient_vs_forwards(name, api):
    """The seismograms are linear in the moment tensor, so the gradient of
    sum(seis_bar * seis) is <seis_bar, seis(e_i)> / M0 with seis(e_i) the
    forward for the i-th unit tensor of size M0: six independent forwards
    pin the whole chain (seismogram, receiver-sum, assembly and source
    adjoints)."""
    deck, p = _case(name)
    g = api.YSpec(deck)
    g.set_run(p)
    seis = g.forward()
    rng = np.random.default_rng(5)
    sb = rng.standard_normal(seis.shape)
    grad = g.misfit_gradient_mt(sb)
    m0 = 1e19
    want = np.zeros(6)
    for i in range(6):
        mm = np.zeros(6)
        mm[i] = m0
        gi = api.YSpec(deck)
        gi.set_run(dataclasses.replace(p, mm=mm))
        want[i] = float(np.sum(sb * gi.forward())) / m0
    # linearity of the forward itself: seis = sum_i M_i seis(e_i)
    err = np.abs(grad - want) / np.linalg.norm(want)
    print(name, "gradient", grad, "forwards", want, f"max rel {err.max():.2e}", flush=True)
    assert err.max() <= 1e-10
    # <seis_bar, seis> = sum_i M_i dF/dM_i
    assert abs(float(np.sum(sb * seis)) - float(np.dot(p.mm, grad))) <= 1e-10 * (
        np.linalg.norm(sb) * np.linalg.norm(seis))


def test_moment_tensor_vjp_needs_record_adjoints(api):
    deck, p = _case("cfg1_small")
    g = api.YSpec(deck)
    g.set_run(p)
    g.forward()
    with pytest.raises(api.YSpecError):
        g.moment_tensor_vjp()
    # record adjoints belong to the forward whose records they were built
    # from: a new forward invalidates them
    g.seismograms_vjp(np.ones((g.nt, 3, p.nr)))
    g.spectra_vjp([2], [1])
    g.moment_tensor_vjp()
    g.forward()
    with pytest.raises(api.YSpecError, match="no record adjoints"):
        g.moment_tensor_vjp()
    with pytest.raises(api.YSpecError, match="no record adjoints"):
        g.source_model_vjp()


def test_autograd_misfit_vs_central_differences(api):
    """L2 misfit through the torch.autograd wrapper against central finite
    differences (h = 1e-6 relative) of the GPU forward, per tensor component"""
    import torch

    from yspec_amd.autograd import seismograms

    deck, p = _case("prem_example")
    g = api.YSpec(deck)
    m0 = np.asarray(p.mm, dtype=np.float64)
    rng = np.random.default_rng(6)
    obs_mm = m0 * (1.0 + 0.2 * rng.standard_normal(6))
    g.set_run(dataclasses.replace(p, mm=obs_mm))
    obs = torch.from_numpy(np.ascontiguousarray(g.forward()))

    def misfit(m):
        return 0.5 * torch.sum((seismograms(m, g, p) - obs) ** 2)

    mm = torch.tensor(m0, requires_grad=True)
    misfit(mm).backward()
    grad = mm.grad.numpy()
    fd = np.zeros(6)
    for i in range(6):
        h = 1e-6 * np.abs(m0).max()
        mp, mn = m0.copy(), m0.copy()
        mp[i] += h
        mn[i] -= h
        with torch.no_grad():
            fd[i] = (misfit(torch.tensor(mp)) - misfit(torch.tensor(mn))).item() / (2 * h)
    err = np.abs(grad - fd) / np.linalg.norm(fd)
    print("autograd", grad, "central differences", fd, f"max rel {err.max():.2e}", flush=True)
    assert err.max() <= 1e-6


@pytest.mark.parametrize("name", ["prem_example", "cfg1_small"])
def test_source_model_gradient_vs_central_differences(name, api):
    """Source-vector term of the model gradient: with the record adjoints W
    of a base run held fixed, F(deck) = sum over pairs of Re<W, sv(deck)>
    (sv the source vectors of a GPU forward on that deck) against central
    differences (h = 1e-6 relative) on rho, vpv, vsv, qmu at the knots around
    the source, through source_model_vjp and model_table_vjp."""
    deck, p = _case(name)
    g = api.YSpec(deck)
    g.set_run(p)
    spec = g.spectra()
    n_l = p.lmax - p.lmin + 1
    ls = np.repeat(np.arange(p.lmin, p.lmax + 1), g.nw)
    js = np.tile(np.arange(1, g.nw + 1), n_l)
    rng = np.random.default_rng(8)
    rb = g.spectra_vjp(ls, js, [rng.standard_normal(s.shape) + 1j * rng.standard_normal(s.shape)
                                for s in spec])
    deck_bar = api.model_table_vjp(deck, g.source_model_vjp())

    def F(cols):
        gi = api.YSpec(dataclasses.replace(deck, cols=cols))
        gi.set_run(p)
        gi.spectra(copy=False)
        rec = gi.pair_records(ls, js)
        return sum(float(np.sum((np.conj(rb[k]) * rec[k]).real)) for k in SRC)

    r = deck.cols[:, 0]
    rsrc = r[-1] - 1000.0 * p.src_depth
    lo = int(np.searchsorted(r, rsrc)) - 1   # knot interval of the source (0-based low end)
    knots = [k for k in (lo - 1, lo, lo + 1, lo + 2) if 0 <= k < deck.nknot and r[k] != r[k + 1 if k + 1 < deck.nknot else k]]
    got, want = [], []
    for k in knots:
        for c in (1, 2, 3, 5):      # rho, vpv, vsv, qmu
            v = deck.cols[k, c]
            if v == 0.0 or (c == 3 and v <= 0.0):
                continue
            h = 1e-6 * abs(v)
            cp, cm = deck.cols.copy(), deck.cols.copy()
            cp[k, c] += h
            cm[k, c] -= h
            want.append((F(cp) - F(cm)) / (2 * h))
            got.append(deck_bar[k, c])
    got, want = np.array(got), np.array(want)
    err = np.abs(got - want) / np.abs(want).max()
    print(name, "knots", knots, "max rel", f"{err.max():.2e}", "entries", len(got), flush=True)
    assert np.abs(want).max() > 0
    assert err.max() <= 1e-6


@pytest.mark.parametrize("mode", [0, 1], ids=["plain", "batched"])
def test_events_gradients_match_single_events(mode, api):
    """Several events in one context: per-event moment-tensor gradients and
    the summed source-vector model term, after the plain (mode 0) or batched
    (mode 1) forward, against one-event contexts: bitwise-level after the
    plain forward, to rounding after the batched one (its shared
    receiver-side marches reorder the same arithmetic)."""
    from yspec_amd.synth import batch_events, synth_deck

    ev = batch_events(nev=4, seed=21, nr=3, fmax_range=(3.0, 5.0), tout=40.0, dt=20.0,
                      lmax_per_mhz=6.0)
    ev[1].src_depth = 650.0
    ev[2].src_depth = 1.5
    ev[3].grav = 0          # per-lane spheroidal and radial solvers
    deck = synth_deck(0)
    g = api.YSpec(deck)
    g.set_events(ev)
    g.forward_events(mode)
    rng = np.random.default_rng(9)
    sbs = [rng.standard_normal(g.event_seismograms(e).shape) for e in range(len(ev))]
    G = g.events_gradient_mt(sbs)
    tb = g.source_model_vjp()
    want_tb = np.zeros_like(tb)
    for e, p in enumerate(ev):
        gi = api.YSpec(deck)
        gi.set_run(p)
        gi.forward()
        want = gi.misfit_gradient_mt(sbs[e])
        want_tb += gi.source_model_vjp()
        err = np.abs(G[e] - want).max() / np.abs(want).max()
        print("mode", mode, "event", e, f"moment-tensor gradient rel {err:.2e}", flush=True)
        assert err <= (1e-13 if mode == 0 else 1e-10)
    err = np.abs(tb - want_tb).max() / np.abs(want_tb).max()
    print("mode", mode, f"source model term rel {err:.2e}", flush=True)
    assert err <= (1e-13 if mode == 0 else 1e-10)


def test_events_reject_host_copies(api):
    from yspec_amd.synth import batch_events, synth_deck

    ev = batch_events(nev=2, seed=22, nr=2, fmax_range=(3.0, 4.0), tout=40.0, dt=20.0,
                      lmax_per_mhz=6.0)
    g = api.YSpec(synth_deck(0))
    g.set_events(ev)
    g.forward_events(0)
    with pytest.raises(api.YSpecError):
        g.spectra_vjp([ev[0].lmin], [1])
