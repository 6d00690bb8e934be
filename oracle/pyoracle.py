"""ctypes wrapper of the C restatement (oracle/liboracle.so).

Test infrastructure only: imported by tests/, the package smoke check and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def _dp(a):
    return a.ctypes.data_as(D)


def _ip(a):
    return a.ctypes.data_as(I)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        L.ysxo_create.restype = C.c_void_p
        L.ysxo_destroy.argtypes = [C.c_void_p]
        L.ysxo_error.restype = C.c_char_p
        L.ysxo_error.argtypes = [C.c_void_p]
        L.ysxo_set_model.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_double, D]
        L.ysxo_model_arrays.argtypes = [C.c_void_p, D]
        L.ysxo_model_flags.argtypes = [C.c_void_p, I]
        L.ysxo_norms.argtypes = [C.c_void_p, D]
        L.ysxo_set_run.argtypes = [C.c_void_p, I, D, D, D]
        L.ysxo_run_info.argtypes = [C.c_void_p, I, D]
        L.ysxo_geometry.argtypes = [C.c_void_p, D, D, D, D, D, D]
        L.ysxo_pair.argtypes = [C.c_void_p, C.c_int, C.c_int, D, I]
        L.ysxo_spectra.argtypes = [C.c_void_p, C.c_int, C.c_int, D, D, D, C.c_int]
        L.ysxo_seismograms.argtypes = [C.c_void_p, D, D, D, D]
        L.ysxo_fourrow.argtypes = [D, C.c_int, C.c_int]
        L.ysxo_legendre.argtypes = [C.c_double, C.c_int, C.c_int, D, D, D]
        L.ysxo_delaz.argtypes = [C.c_double] * 4 + [D]
        L.ysxo_fcal.argtypes = [C.c_double] * 4 + [I, D]
        L.ysxo_bfs.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, D]
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    pass


PAIR_FIELDS = (("svt", (5, 2)), ("ytor", (2, 2)), ("svs", (5, 6)),
               ("ysph", (6, 4)), ("sr0", (2,)), ("yrad", (2, 2)),
               ("uu", (5,)), ("vv", (5,)), ("ww", (5,)))


class Oracle:
    """One model + one run, mirroring the reference driver's state."""

    def __init__(self, deck, params=None):
        self.L = lib()
        self.h = self.L.ysxo_create()
        cols = np.ascontiguousarray(deck.cols, dtype=np.float64)
        self._check(self.L.ysxo_set_model(self.h, deck.nknot, deck.nic, deck.noc,
                                          deck.ians, deck.tref, _dp(cols)))
        self.nknot = deck.nknot
        if params is not None:
            self.set_run(params)

    def _check(self, rc):
        if rc:
            raise OracleError(self.L.ysxo_error(self.h).decode())

    def __del__(self):
        try:
            self.L.ysxo_destroy(self.h)
        except Exception:
            pass

    def set_run(self, p):
        ip = p.int_block()
        fp = p.float_block()
        lat = np.ascontiguousarray(p.lat, dtype=np.float64)
        lon = np.ascontiguousarray(p.lon, dtype=np.float64)
        self._check(self.L.ysxo_set_run(self.h, _ip(ip), _dp(fp), _dp(lat), _dp(lon)))
        self.p = p
        io = np.zeros(7, np.int32)
        fo = np.zeros(12)
        self._check(self.L.ysxo_run_info(self.h, _ip(io), _dp(fo)))
        self.nw, self.nt, self.i1, self.i2, self.is_, self.ir, self.nout = map(int, io)
        (self.rs, self.rr, self.dt, self.tout, self.ep, self.df, self.wmin,
         self.dw, self.f11, self.f12, self.f21, self.f22) = map(float, fo)

    def model_arrays(self):
        out = np.zeros((32, self.nknot))
        self._check(self.L.ysxo_model_arrays(self.h, _dp(out)))
        return out

    def model_flags(self):
        out = np.zeros(8, np.int32)
        self.L.ysxo_model_flags(self.h, _ip(out))
        return out

    def norms(self):
        out = np.zeros(9)
        self.L.ysxo_norms(self.h, _dp(out))
        return out

    def geometry(self):
        nr, L1 = self.p.nr, self.p.lmax + 1
        d = [np.zeros(nr) for _ in range(3)]
        t = [np.zeros((L1, 3, nr), order="F") for _ in range(3)]
        self._check(self.L.ysxo_geometry(self.h, *[_dp(a) for a in d],
                                         *[_dp(a) for a in t]))
        return d + t

    def pair(self, l, j):
        rec = np.zeros(89, np.complex128)
        st = np.zeros(2, np.int32)
        self._check(self.L.ysxo_pair(self.h, l, j, rec.ctypes.data_as(D),
                                     _ip(st)))
        out, n = {}, 0
        for name, shape in PAIR_FIELDS:
            cnt = int(np.prod(shape))
            out[name] = rec[n:n + cnt].reshape(shape, order="F")
            n += cnt
        out["starts"] = st
        return out

    def spectra(self, l0=None, l1=None, nthreads=1):
        l0 = self.p.lmin if l0 is None else l0
        l1 = self.p.lmax if l1 is None else l1
        shp = (self.nw, self.p.nr)
        ur, ut, up = (np.zeros(shp, np.complex128, order="F") for _ in range(3))
        self._check(self.L.ysxo_spectra(self.h, l0, l1, *[a.ctypes.data_as(D) for a in (ur, ut, up)], nthreads))
        return ur, ut, up

    def seismograms(self, ur, ut, up):
        """(nt, 3, nr) real"""
        arrs = [np.asfortranarray(a, dtype=np.complex128) for a in (ur, ut, up)]
        seis = np.zeros((self.nt, 3, self.p.nr), order="F")
        self._check(self.L.ysxo_seismograms(self.h, *[a.ctypes.data_as(D) for a in arrs], _dp(seis)))
        return seis


def fourrow(data, isign=1):
    a = np.ascontiguousarray(data, dtype=np.complex128).copy()
    if lib().ysxo_fourrow(a.ctypes.data_as(D), len(a), isign):
        raise OracleError("n must be a power of 2")
    return a


def legendre(theta, l, m):
    x, xp, xc = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().ysxo_legendre(theta, l, m, _dp(x), _dp(xp), _dp(xc))
    return x[:m + 1], xp[:m + 1], xc[:m + 1]


def delaz(eplat, eplon, stlat, stlon):
    out = np.zeros(3)
    lib().ysxo_delaz(eplat, eplon, stlat, stlon, _dp(out))
    return out


def fcal(f1, f2, dt, tout):
    io = np.zeros(3, np.int32)
    fo = np.zeros(4)
    if lib().ysxo_fcal(f1, f2, dt, tout, _ip(io), _dp(fo)):
        raise OracleError("fmax above Nyquist")
    return tuple(int(v) for v in io), tuple(float(v) for v in fo)


def bfs(l, xsq, eps=1e-9):
    out = np.zeros(2)
    lib().ysxo_bfs(l, xsq.real, xsq.imag, eps, _dp(out))
    return complex(out[0], out[1])
