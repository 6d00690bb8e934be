"""The gradient of an L2 waveform misfit with respect to the deck's rho, vp,
vs and Qmu at its knots: ysx_model_vjp (the discrete adjoint of every
per-pair solve, starting solution and source vector) followed by the host
adjoint of the deck -> table transform, against central differences
(h = 1e-6 relative) of GPU forwards on seeded knots.

The data are seismograms of the same deck with 2% Gaussian bumps in vp and
vs.  The reference has no gradient; the check is against the GPU forward's
own finite differences (the forward matches the reference to 1e-9)."""
import numpy as np
import pytest

from yspec_amd.inputs import Deck
from yspec_amd.synth import synth_deck, workload

pytestmark = pytest.mark.gpu
COLS = {1: "rho", 2: "vp", 3: "vs", 5: "Qmu"}


@pytest.fixture(scope="module")
def api():
    from yspec_amd import api as a
    a.lib()
    return a


def bumped(deck, amp=0.02, at=0.8, width=0.05):
    """2% Gaussian bumps in vp and vs centred at radius `at` (fraction of
    the Earth's radius)"""
    cols = deck.cols.copy()
    x = cols[:, 0] / 6371000.0
    for c in (2, 3, 6, 7):
        cols[:, c] *= 1.0 + amp * np.exp(-((x - at) / width) ** 2)
    return Deck(deck.title, deck.ians, deck.tref, deck.nknot, deck.nic, deck.noc, cols)


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
