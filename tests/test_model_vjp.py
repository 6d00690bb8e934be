"""Adjoint of the deck -> model-table transform (ysx_model_table_vjp)
against central finite differences of the transform itself
(ysx_model_table_host), on the synthetic isotropic deck and on a
transversely isotropic deck of the golden cases: rho, vp, vs, Qmu (and
vph, vsh, eta) at random knots, h = 1e-6 relative, zero-valued parameters
(fluid vs and Qmu) excluded, and on the anisotropic deck the fluid knots."""
import dataclasses

import numpy as np
import pytest

from conftest import load_case
from yspec_amd import api
from yspec_amd.synth import synth_deck


def _fd_check(deck, cols, seed, n=40, h=1e-6, solid_only=False):
    rng = np.random.default_rng(seed)
    t0 = api.model_table_host(deck)
    tbar = rng.standard_normal(t0.shape)
    tbar[0] = 0.0  # radii are not differentiated
    db = api.model_table_vjp(deck, tbar)
    fd, ad = [], []
    for _ in range(n):
        i, c = int(rng.integers(0, deck.nknot)), int(rng.choice(cols))
        v = deck.cols[i, c]
        if v == 0.0 or (solid_only and deck.cols[i, 3] == 0.0):
            continue
        step = h * abs(v)
        out = []
        for s in (step, -step):
            cc = deck.cols.copy()
            cc[i, c] = v + s
            out.append(api.model_table_host(dataclasses.replace(deck, cols=cc)))
        fd.append(np.sum(tbar * (out[0] - out[1])) / (2.0 * step))
        ad.append(db[i, c])
    fd, ad = np.array(fd), np.array(ad)
    assert len(ad) >= n // 2
    err = np.linalg.norm(fd - ad) / np.linalg.norm(ad)
    assert err <= 1e-5, err
    return err


def _lib():
    try:
        api.lib()
    except (RuntimeError, OSError) as e:
        pytest.skip(str(e))


def test_vjp_isotropic_deck():
    _lib()
    _fd_check(synth_deck(0), cols=(1, 2, 3, 5), seed=1)


def test_vjp_anisotropic_deck():
    _lib()
    deck, _, _ = load_case("aniso_acc_cor")
    assert deck.ians == 1
    # in a fluid, vpv moved without vph makes mu = (C - A) / 15 negative for
    # one sign of the step (beta undefined): solid knots only
    _fd_check(deck, cols=(1, 2, 3, 5, 6, 7, 8), seed=2, solid_only=True)


def test_vjp_is_linear_in_table_bar():
    _lib()
    deck = synth_deck(0)
    rng = np.random.default_rng(3)
    shape = api.model_table_host(deck).shape
    a, b = rng.standard_normal(shape), rng.standard_normal(shape)
    va, vb = api.model_table_vjp(deck, a), api.model_table_vjp(deck, b)
    vab = api.model_table_vjp(deck, 2.0 * a - 3.0 * b)
    assert np.allclose(vab, 2.0 * va - 3.0 * vb, rtol=1e-12, atol=1e-12 * np.abs(vab).max())
