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


def _cfg2_fixture():
    import os
    from conftest import GOLDEN
    from yspec_amd.synth import synth_deck, workload

    z = dict(np.load(os.path.join(GOLDEN, "cfg2_records.npz"), allow_pickle=False))
    deck, p = synth_deck(0), workload("cfg2")
    # the fixture was written from this very workload
    np.testing.assert_array_equal(z["deck_cols"], deck.cols)
    np.testing.assert_array_equal(z["ip"], p.int_block())
    np.testing.assert_array_equal(z["fp"], p.float_block())
    return deck, p, z


def test_cfg2_records_l600_40mhz(oracle_mod):
    """The oracle where the headline throughput is measured: per-pair
    records of the lmax = 600, fmax = 40 mHz workload (attenuation,
    self-gravitation, 600 km source) at l = 0, 1, 2, 150, 300, 450, 600
    against the reference's own (tests/golden/cfg2_records.npz), start knots
    exactly."""
    from refdump import PAIR_FIELDS

    deck, p, z = _cfg2_fixture()
    o = oracle_mod.Oracle(deck, p)
    worst = 0.0
    for i, (l, j) in enumerate(zip(z["pair_l"], z["pair_j"])):
        q = o.pair(int(l), int(j))
        assert q["starts"][0] == z["pair_tor_start"][i]
        assert q["starts"][1] == z["pair_sph_start"][i]
        got = np.concatenate([np.asarray(q[k]).reshape(-1, order="F") for k, _ in PAIR_FIELDS])
        worst = max(worst, rel_l2(got, z["rec"][i]))
    assert sorted(set(z["pair_l"].tolist())) == [0, 1, 2, 150, 300, 450, 600]
    assert worst < 1e-12, worst
