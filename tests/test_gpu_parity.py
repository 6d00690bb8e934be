"""GPU path (libysx.so through its C-ABI) against the reference's own FP64
results (tests/golden) and against the oracle.  Bars: relative L2 <= 1e-9 on
the per-pair solver outputs and the (l, w) spectra, <= 1e-8 on the filtered
seismograms."""
import numpy as np
import pytest

from conftest import case_names, load_case, rel_l2, ysph_rows

pytestmark = pytest.mark.gpu
CASES = case_names()
KEYS = ("svt", "ytor", "svs", "ysph", "sr0", "yrad", "uu", "vv", "ww")


@pytest.fixture(scope="module")
def api():
    from yspec_amd import api as a
    a.lib()
    return a


def _cut(rec_k, want_k, k, full, p):
    if k == "ysph" and not full:
        n = ysph_rows(p)
        return rec_k[:, :n, :], want_k[:, :n, :]
    return rec_k, want_k


# full=False: the default b-vector path (receiver-side rows, transposed
# steps); full=True: the reference's four source-side columns
@pytest.mark.parametrize("full", [False, True], ids=["rows", "columns"])
@pytest.mark.parametrize("name", CASES)
def test_gpu_vs_reference(name, full, api):
    deck, p, z = load_case(name)
    g = api.YSpec(deck)
    g.set_records(full)
    g.set_run(p)
    ur, ut, up = g.spectra()
    rec = g.pair_records(z["pair_l"], z["pair_j"])
    errs = {k: rel_l2(*_cut(rec[k], z["pair_" + k], k, full, p)) for k in KEYS}
    print(name, "full" if full else "rows", {k: f"{v:.2e}" for k, v in errs.items()},
          flush=True)
    for k, v in errs.items():
        assert v <= 1e-9, (k, v)
    for a, key in zip((ur, ut, up), ("ur", "ut", "up")):
        e = rel_l2(a, z[key])
        print(name, key, f"{e:.2e}", flush=True)
        assert e <= 1e-9
    s = g.seismograms()
    e = rel_l2(s, z["seis"])
    print(name, "seis", f"{e:.2e}", flush=True)
    assert e <= 1e-8
