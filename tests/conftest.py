"""Shared test set-up.  GPU tests are marked ``gpu``; everything else runs
on the CPU (the oracle, the host logic and the C-ABI symbol table)."""
import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle"), GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)

from yspec_amd.inputs import Deck, Params  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def case_names():
    return sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, "case_*.npz")))


def load_case(name):
    """(deck, params, fixture dict) of a golden case"""
    z = dict(np.load(os.path.join(GOLDEN, f"case_{name}.npz"), allow_pickle=False))
    nk, nic, noc, ians = (int(v) for v in z["deck_head"])
    deck = Deck(name, ians, float(z["deck_tref"]), nk, nic, noc, z["deck_cols"])
    ip, fp = z["ip"], z["fp"]
    p = Params(prefix="yspec.out", model_file="model.deck", ats=int(ip[2]), grav=int(ip[3]),
               out=int(ip[4]), cor=int(ip[5]), lmin=int(ip[0]), lmax=int(ip[1]),
               f1=fp[0], f2=fp[1], tout=fp[2], dt=fp[3], f11=fp[4], f12=fp[5], f21=fp[6],
               f22=fp[7], src_depth=fp[8], src_lat=fp[9], src_lon=fp[10], mm=fp[11:17].copy(),
               rec_depth=fp[17], lat=z["lat"], lon=z["lon"])
    return deck, p, z


def ysph_rows(p):
    """Rows of the pair records' ysph (n, 6, 4) that the default b-vector path
    (receiver-side rows) fills: 0..1, and 0..2 with the acceleration
    potential/tilt correction; all six need ysx_set_records(ctx, 1)."""
    return 3 if (p.out == 2 and p.cor != 0) else 2


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    try:
        pyoracle.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    return pyoracle
