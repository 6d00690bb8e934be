"""The allocation-free model table (ysx_model_table_noalloc) against the
allocating entry points (ysx_set_model, then ysx_model_table), bit for bit,
on two calls in a row that share the device scratch, the second with a
smaller deck; and its rejection of a deck too large for the scratch.

The caller's device buffer comes from hipMalloc of the HIP runtime that
libysx.so itself loaded (torch's GPU side cannot initialise once another
runtime build holds the process)."""
import ctypes

import numpy as np
import pytest

from conftest import load_case
from yspec_amd import api
from yspec_amd.inputs import Deck
from yspec_amd.synth import synth_deck

H2D, D2H = 1, 2  # hipMemcpyKind


def _hip():
    api.lib()  # loads the runtime libysx.so links
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipSetDevice.argtypes = [ctypes.c_int]
    return h


class DeviceBuffer:
    """n doubles of device memory on device 0, filled from / read to numpy"""

    def __init__(self, hip, n, fill):
        self.hip, self.n, self.p = hip, n, ctypes.c_void_p()
        assert hip.hipSetDevice(0) == 0
        assert hip.hipMalloc(ctypes.byref(self.p), n * 8) == 0
        host = np.full(n, fill, dtype=np.float64)
        assert hip.hipMemcpy(self.p, host.ctypes.data, n * 8, H2D) == 0

    def read(self):
        host = np.empty(self.n, dtype=np.float64)
        assert self.hip.hipMemcpy(host.ctypes.data, self.p, self.n * 8, D2H) == 0
        return host

    def free(self):
        self.hip.hipFree(self.p)


def _table_of(deck):
    g = api.YSpec(deck)
    t = g.model_table()
    g.close()
    return t


@pytest.mark.gpu
def test_noalloc_model_table_matches_set_model():
    hip = _hip()
    big, _, _ = load_case("prem_example")  # 220 knots, transversely isotropic
    small = synth_deck(0)  # 200 knots, isotropic
    assert small.nknot < big.nknot
    g = api.YSpec(small)
    before = g.model_table()
    nf = len(api.TABLE_FIELDS)
    for deck in (big, small):
        want = _table_of(deck)
        buf = DeviceBuffer(hip, nf * deck.nknot, np.nan)
        g.model_table_noalloc(deck, buf.p.value)
        got = buf.read().reshape(nf, deck.nknot)
        buf.free()
        np.testing.assert_array_equal(got, want)
    # the context's own model is untouched
    np.testing.assert_array_equal(g.model_table(), before)
    g.close()


@pytest.mark.gpu
def test_noalloc_rejects_a_deck_beyond_the_scratch():
    hip = _hip()
    n = 6000  # one layer: 13 x 3 x 6001 work doubles alone exceed 2 MiB
    cols = np.zeros((n, 9))
    cols[:, 0] = np.linspace(0.0, 6371000.0, n)
    cols[:, 1:4] = (3000.0, 8000.0, 4000.0)
    cols[:, 4:6] = (57823.0, 300.0)
    deck = Deck("long", 0, 1.0, n, 1, 2, cols)
    g = api.YSpec(synth_deck(0))
    buf = DeviceBuffer(hip, len(api.TABLE_FIELDS) * n, 0.0)
    with pytest.raises(api.YSpecError) as ei:
        g.model_table_noalloc(deck, buf.p.value)
    buf.free()
    assert ei.value.code == 1 and "bytes of device scratch" in ei.value.msg
    g.close()
