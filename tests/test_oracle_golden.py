"""The C restatement (oracle/) against FP64 dumps of the reference itself
(tests/golden/case_*.npz, written by tests/golden/make_golden.py from the
unmodified reference modules)."""
import numpy as np
import pytest

from conftest import case_names, load_case, rel_l2

CASES = case_names()


@pytest.mark.parametrize("name", CASES)
def test_model_and_geometry_bit_exact(name, oracle_mod):
    deck, p, z = load_case(name)
    o = oracle_mod.Oracle(deck, p)
    # the model transform is pure FP64 arithmetic in the same order: identical
    np.testing.assert_array_equal(o.model_arrays(), z["model"])
    assert (o.nw, o.nt, o.i1, o.i2, o.is_, o.ir) == tuple(int(v) for v in z["run_i"][:6])
    g = o.geometry()
    for a, key in zip(g, ("delta", "azep", "azst", "xa", "xpa", "xca")):
        np.testing.assert_array_equal(a, z[key])


@pytest.mark.parametrize("name", CASES)
def test_pair_records(name, oracle_mod):
    deck, p, z = load_case(name)
    o = oracle_mod.Oracle(deck, p)
    worst = 0.0
    for i, (l, j) in enumerate(zip(z["pair_l"], z["pair_j"])):
        q = o.pair(int(l), int(j))
        assert q["starts"][0] == z["tor_start"][i] and q["starts"][1] == z["sph_start"][i]
        for key in ("svt", "ytor", "svs", "ysph", "sr0", "yrad", "uu", "vv", "ww"):
            worst = max(worst, rel_l2(q[key], z["pair_" + key][i]))
    assert worst < 1e-12, worst


@pytest.mark.parametrize("name", CASES)
def test_spectra_and_seismograms(name, oracle_mod):
    deck, p, z = load_case(name)
    o = oracle_mod.Oracle(deck, p)
    ur, ut, up = o.spectra(nthreads=4)
    for a, key in zip((ur, ut, up), ("ur", "ut", "up")):
        assert rel_l2(a, z[key]) < 1e-12
    s = o.seismograms(ur, ut, up)
    assert rel_l2(s, z["seis"]) < 1e-12
