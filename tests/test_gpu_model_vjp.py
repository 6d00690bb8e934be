"""The deck -> model-table adjoint on the device (ysx_model_vjp_deck,
k_model_vjp) against the host function ysx_model_table_vjp, which
tests/test_model_vjp.py pins with central differences of the transform:
bit for bit, on the synthetic deck and on the golden decks (the reference's
PREM example, a transversely isotropic deck, an ocean, a fluid inner core),
and on one deck directly against central differences of
ysx_model_table_host (h = 1e-6 relative)."""
import dataclasses

import numpy as np
import pytest

from conftest import load_case
from yspec_amd.synth import synth_deck

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from yspec_amd import api as a
    a.lib()
    return a


def _decks():
    out = [("synth", synth_deck(0))]
    for name in ("prem_example", "aniso_acc_cor", "ocean_vel", "fic_grav2"):
        deck, _, _ = load_case(name)
        out.append((name, deck))
    return out


@pytest.mark.parametrize("name,deck", _decks(), ids=[n for n, _ in _decks()])
def test_device_deck_adjoint_matches_host(name, deck, api):
    g = api.YSpec(deck)
    rng = np.random.default_rng(5)
    for trial in range(2):
        tbar = rng.standard_normal((len(api.TABLE_FIELDS), deck.nknot))
        tbar[0] = 0.0  # radii are not differentiated
        dev = g.model_vjp_deck(tbar)
        host = api.model_table_vjp(deck, tbar)
        print(f"{name} trial {trial}: max |device - host| {np.max(np.abs(dev - host)):.1e}",
              flush=True)
        assert np.array_equal(dev, host), name
    g.close()


def test_device_deck_adjoint_vs_central_differences(api):
    deck = synth_deck(1)
    g = api.YSpec(deck)
    rng = np.random.default_rng(9)
    tbar = rng.standard_normal((len(api.TABLE_FIELDS), deck.nknot))
    tbar[0] = 0.0
    db = g.model_vjp_deck(tbar)
    g.close()
    fd, ad = [], []
    while len(ad) < 20:
        i, c = int(rng.integers(0, deck.nknot)), int(rng.choice([1, 2, 3, 5]))
        v = deck.cols[i, c]
        if v == 0.0:
            continue
        h = 1e-6 * abs(v)
        out = []
        for s in (h, -h):
            cc = deck.cols.copy()
            cc[i, c] = v + s
            out.append(api.model_table_host(dataclasses.replace(deck, cols=cc)))
        fd.append(np.sum(tbar * (out[0] - out[1])) / (2.0 * h))
        ad.append(db[i, c])
    fd, ad = np.array(fd), np.array(ad)
    err = float(np.linalg.norm(fd - ad) / np.linalg.norm(ad))
    print(f"device deck adjoint vs central differences: {err:.2e}", flush=True)
    assert err <= 1e-5
