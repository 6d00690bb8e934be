"""Python host over the C-ABI of libysx.so (include/ysx.h).

Mirrors the reference driver's flow (src/yspec.f90): a model is loaded,
a run is set, the (l, w) spectra are computed and summed over receivers,
then the filtered inverse FFT gives the seismograms.  All compute is on
the GPU; there is no CPU fallback -- loading fails loudly when the HIP
library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libysx.so")
_LIB = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)

EXPORTS = ("ysx_create", "ysx_destroy", "ysx_last_error", "ysx_set_model", "ysx_set_run",
           "ysx_run_info", "ysx_spectra", "ysx_seismograms", "ysx_forward",
           "ysx_set_events", "ysx_event_info", "ysx_forward_events", "ysx_set_batch_tuning",
           "ysx_set_records",
           "ysx_event_spectra", "ysx_event_seismograms", "ysx_model_table",
           "ysx_model_table_host", "ysx_model_table_vjp",
           "ysx_pair_records", "ysx_pair_steps", "ysx_timings", "ysx_step_counts",
           "ysx_seismograms_vjp", "ysx_spectra_vjp", "ysx_moment_tensor_vjp",
           "ysx_vjp_timings", "ysx_source_model_vjp", "ysx_model_table_noalloc",
           "ysx_model_vjp", "ysx_grad_timings", "ysx_misfit_grad", "ysx_grad_memory",
           "ysx_model_vjp_deck", "ysx_set_lanes")

PAIR_FIELDS = (("svt", (5, 2)), ("ytor", (2, 2)), ("svs", (5, 6)),
               ("ysph", (6, 4)), ("sr0", (2,)), ("yrad", (2, 2)),
               ("uu", (5,)), ("vv", (5,)), ("ww", (5,)))


def _split_records(rec):
    """(n, 89) complex records -> dict of the fields of PAIR_FIELDS"""
    out, n = {}, 0
    for name, shape in PAIR_FIELDS:
        cnt = int(np.prod(shape))
        out[name] = rec[:, n:n + cnt].reshape((len(rec),) + shape, order="F")
        n += cnt
    return out


class YSpecError(RuntimeError):
    """A nonzero return code of the C-ABI, with its message."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


def lib():
    """Load libysx.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run python -m yspec_amd.build")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.ysx_create.argtypes = [C.POINTER(vp), C.c_int]
        L.ysx_destroy.argtypes = [vp]
        L.ysx_destroy.restype = None
        L.ysx_last_error.argtypes = [vp]
        L.ysx_last_error.restype = C.c_char_p
        L.ysx_set_model.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, D]
        L.ysx_set_run.argtypes = [vp, I, D, D, D]
        L.ysx_run_info.argtypes = [vp, I, D]
        L.ysx_spectra.argtypes = [vp, D, D, D]
        L.ysx_seismograms.argtypes = [vp, D]
        L.ysx_forward.argtypes = [vp, D]
        L.ysx_set_events.argtypes = [vp, C.c_int, I, D, D, D]
        L.ysx_event_info.argtypes = [vp, C.c_int, I, D]
        L.ysx_forward_events.argtypes = [vp, C.c_int, D]
        L.ysx_set_batch_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.ysx_set_records.argtypes = [vp, C.c_int]
        L.ysx_event_spectra.argtypes = [vp, C.c_int, D, D, D]
        L.ysx_event_seismograms.argtypes = [vp, C.c_int, D]
        L.ysx_model_table.argtypes = [vp, D]
        L.ysx_model_table_host.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, D, D]
        L.ysx_model_table_vjp.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, D, D, D]
        L.ysx_pair_records.argtypes = [vp, C.c_int, I, I, D]
        L.ysx_pair_steps.argtypes = [vp, C.c_int, I, I, I]
        L.ysx_timings.argtypes = [vp, D]
        L.ysx_step_counts.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.ysx_seismograms_vjp.argtypes = [vp, D, D, D, D]
        L.ysx_spectra_vjp.argtypes = [vp, D, D, D, C.c_int, I, I, D]
        L.ysx_moment_tensor_vjp.argtypes = [vp, D]
        L.ysx_vjp_timings.argtypes = [vp, D]
        L.ysx_source_model_vjp.argtypes = [vp, D]
        L.ysx_model_vjp.argtypes = [vp, D]
        L.ysx_grad_timings.argtypes = [vp, D]
        L.ysx_misfit_grad.argtypes = [vp, D, C.c_int, D, D]
        L.ysx_grad_memory.argtypes = [vp, D]
        L.ysx_model_vjp_deck.argtypes = [vp, D, D]
        L.ysx_set_lanes.argtypes = [vp, C.c_int]
        L.ysx_model_table_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_double, D, C.c_void_p]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(D)


def _ip(a):
    return a.ctypes.data_as(I)


# names of the 29 fields of the model table (include/ysx.h, ysx_model_table)
TABLE_FIELDS = ("r", "rho", "rho_cs", "acon", "acon_cs", "ccon", "ccon_cs", "fcon", "fcon_cs",
                "lcon", "lcon_cs", "ncon", "ncon_cs", "kappa", "kappa_cs", "mu", "mu_cs",
                "qkappa", "qkappa_cs", "qmu", "qmu_cs", "xlam", "xlam_cs", "xa2", "xa2_cs",
                "grav", "grav_cs", "alpha", "beta")


def model_table_host(deck):
    """(29, nknot) model table computed on the host (no GPU needed)."""
    cols = np.ascontiguousarray(deck.cols, dtype=np.float64)
    out = np.zeros((len(TABLE_FIELDS), deck.nknot))
    rc = lib().ysx_model_table_host(deck.nknot, deck.nic, deck.noc, deck.ians, deck.tref,
                                    _dp(cols), _dp(out))
    if rc:
        raise YSpecError(rc, "bad deck")
    return out


# deck columns of a knot's rho, vp, vs and Qmu (vp and vs: both polarisations)
KNOT_PARAMS = ("rho", "vp", "vs", "Qmu")
_KNOT_COLS = ((1,), (2, 6), (3, 7), (5,))


def knot_gradient(deck_bar):
    """(nknot, 4) gradient with respect to each knot's rho, vp, vs, Qmu from
    the (nknot, 9) deck-column gradient (vp and vs move vpv and vph, vsv and
    vsh together), as ysx_misfit_grad folds it."""
    return np.stack([sum(deck_bar[:, c] for c in cs) for cs in _KNOT_COLS], axis=1)


def deck_with_knots(deck, knots):
    """A copy of `deck` whose rho, vp (vpv = vph), vs (vsv = vsh) and Qmu
    columns are the (nknot, 4) array knots."""
    from .inputs import Deck

    cols = np.array(deck.cols, dtype=np.float64)
    for q, cs in enumerate(_KNOT_COLS):
        for c in cs:
            cols[:, c] = knots[:, q]
    return Deck(deck.title, deck.ians, deck.tref, deck.nknot, deck.nic, deck.noc, cols)


def model_table_vjp(deck, table_bar):
    """(nknot, 9) deck_bar = J^T table_bar for the deck -> table transform
    (table_bar: (29, nknot), the layout of model_table_host)."""
    cols = np.ascontiguousarray(deck.cols, dtype=np.float64)
    tb = np.ascontiguousarray(table_bar, dtype=np.float64)
    if tb.shape != (len(TABLE_FIELDS), deck.nknot):
        raise ValueError("table_bar must be (29, nknot)")
    out = np.zeros((deck.nknot, 9))
    rc = lib().ysx_model_table_vjp(deck.nknot, deck.nic, deck.noc, deck.ians, deck.tref,
                                   _dp(cols), _dp(tb), _dp(out))
    if rc:
        raise YSpecError(rc, "bad deck")
    return out


class YSpec:
    """One GPU context: a deck model and a run (parameter block)."""

    def __init__(self, deck, params=None, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        rc = self.L.ysx_create(C.byref(self.h), device)
        if rc:
            msg = self.L.ysx_last_error(self.h).decode() if self.h else ""
            raise YSpecError(rc, msg or "ysx_create failed")
        self.set_model(deck)
        if params is not None:
            self.set_run(params)

    def set_model(self, deck):
        """A new deck model; the run must be set again afterwards."""
        cols = np.ascontiguousarray(deck.cols, dtype=np.float64)
        self._check(self.L.ysx_set_model(self.h, deck.nknot, deck.nic, deck.noc, deck.ians,
                                         deck.tref, _dp(cols)))
        self.deck = deck

    def _check(self, rc):
        if rc:
            raise YSpecError(rc, self.L.ysx_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.ysx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_run(self, p):
        ip = p.int_block()
        fp = p.float_block()
        lat = np.ascontiguousarray(p.lat, dtype=np.float64)
        lon = np.ascontiguousarray(p.lon, dtype=np.float64)
        self._check(self.L.ysx_set_run(self.h, _ip(ip), _dp(fp), _dp(lat), _dp(lon)))
        self.p = p
        self.events = [p]
        io = np.zeros(8, np.int32)
        fo = np.zeros(8)
        self._check(self.L.ysx_run_info(self.h, _ip(io), _dp(fo)))
        self.nw, self.nt, self.i1, self.i2, self.is_, self.ir, self.nout, self.npair = map(int, io)
        (self.rs, self.rr, self.dt, self.tout, self.ep, self.df, self.wmin,
         self.dw) = map(float, fo)
        self.info = [dict(zip(("nw", "nt", "i1", "i2", "is", "ir", "nout", "npair"),
                              map(int, io)))]

    def set_records(self, full):
        """full=True: ysph of the pair records holds all six rows (the
        b-vector system integrated as the reference's source-side columns);
        False (default): the rows the seismograms use.  Re-applies the
        current run, if any."""
        self._check(self.L.ysx_set_records(self.h, 1 if full else 0))
        if getattr(self, "p", None) is not None and not getattr(self, "events", None):
            self.set_run(self.p)

    def spectra(self, copy=True):
        """ur, ut, up: (nw, nr) complex (Fortran order), or None if copy=False."""
        if not copy:
            self._check(self.L.ysx_spectra(self.h, None, None, None))
            return None
        shp = (self.nw, self.p.nr)
        out = [np.zeros(shp, np.complex128, order="F") for _ in range(3)]
        self._check(self.L.ysx_spectra(self.h, *[a.ctypes.data_as(D) for a in out]))
        return tuple(out)

    def seismograms(self, copy=True):
        """(nt, 3, nr) real: Z, N, E in SI units, rows t = i*dt."""
        if not copy:
            self._check(self.L.ysx_seismograms(self.h, None))
            return None
        seis = np.zeros((self.nt, 3, self.p.nr), order="F")
        self._check(self.L.ysx_seismograms(self.h, _dp(seis)))
        return seis

    def forward(self, copy=True):
        if not copy:
            self._check(self.L.ysx_forward(self.h, None))
            return None
        seis = np.zeros((self.nt, 3, self.p.nr), order="F")
        self._check(self.L.ysx_forward(self.h, _dp(seis)))
        return seis

    def pair_steps(self, ls, js):
        """(n, 5) step counts: toroidal, solid minors, fluid minors,
        b-vector, radial."""
        ls = np.ascontiguousarray(ls, dtype=np.int32)
        js = np.ascontiguousarray(js, dtype=np.int32)
        out = np.zeros((len(ls), 5), np.int32)
        self._check(self.L.ysx_pair_steps(self.h, len(ls), _ip(ls), _ip(js), _ip(out)))
        return out

    def pair_records(self, ls, js):
        ls = np.ascontiguousarray(ls, dtype=np.int32)
        js = np.ascontiguousarray(js, dtype=np.int32)
        rec = np.zeros((len(ls), 89), np.complex128)
        self._check(self.L.ysx_pair_records(self.h, len(ls), _ip(ls), _ip(js),
                                            rec.ctypes.data_as(D)))
        return _split_records(rec)

    def seismograms_vjp(self, seis_bar):
        """Adjoint of seismograms(): (nt, 3, nr) real seis_bar -> ur_bar,
        ut_bar, up_bar, (nw, nr) complex each (dF/dRe + i dF/dIm)."""
        sb = np.asfortranarray(seis_bar, dtype=np.float64)
        if sb.shape != (self.nt, 3, self.p.nr):
            raise ValueError("seis_bar must be (nt, 3, nr)")
        shp = (self.nw, self.p.nr)
        out = [np.zeros(shp, np.complex128, order="F") for _ in range(3)]
        self._check(self.L.ysx_seismograms_vjp(self.h, _dp(sb),
                                               *[a.ctypes.data_as(D) for a in out]))
        return tuple(out)

    def spectra_vjp(self, ls, js, spec_bar=None):
        """Adjoints of the records of pairs (ls, js) (the fields of
        pair_records) from the spectra adjoints spec_bar = (ur_bar, ut_bar,
        up_bar), or from the device-resident result of seismograms_vjp."""
        ls = np.ascontiguousarray(ls, dtype=np.int32)
        js = np.ascontiguousarray(js, dtype=np.int32)
        rec = np.zeros((len(ls), 89), np.complex128)
        if spec_bar is None:
            ins = [None, None, None]
        else:
            shp = (self.nw, self.p.nr)
            arrs = [np.asfortranarray(a, dtype=np.complex128) for a in spec_bar]
            if any(a.shape != shp for a in arrs):
                raise ValueError("spec_bar: three (nw, nr) arrays")
            ins = [a.ctypes.data_as(D) for a in arrs]
        self._check(self.L.ysx_spectra_vjp(self.h, *ins, len(ls), _ip(ls), _ip(js),
                                           rec.ctypes.data_as(D)))
        return _split_records(rec)

    def moment_tensor_vjp(self):
        """Gradient with respect to Mrr, Mrt, Mrp, Mtt, Mtp, Mpp (per N m) from
        the record adjoints of the last spectra_vjp: (6,) for one event,
        (nev, 6) for several."""
        nev = len(self.events)
        out = np.zeros(6 * nev)
        self._check(self.L.ysx_moment_tensor_vjp(self.h, _dp(out)))
        return out if nev == 1 else out.reshape(nev, 6)

    def source_model_vjp(self):
        """(29, nknot) table adjoint of the source-vector term, from the record
        adjoints of the last spectra_vjp (input of model_table_vjp)."""
        out = np.zeros((len(TABLE_FIELDS), self.deck.nknot))
        self._check(self.L.ysx_source_model_vjp(self.h, _dp(out)))
        return out

    def vjp_timings(self):
        """kernel ms of the last seismograms_vjp, spectra_vjp, and
        moment_tensor_vjp or source_model_vjp (whichever ran last)"""
        out = np.zeros(3)
        self._check(self.L.ysx_vjp_timings(self.h, _dp(out)))
        return out

    def events_gradient_mt(self, seis_bars):
        """(nev, 6): d/dM_e of sum_e sum(seis_bar_e * seismograms_e) after
        forward_events; seis_bars: one (nt_e, 3, nr_e) array per event."""
        flat = np.concatenate([np.asfortranarray(b, dtype=np.float64).ravel(order="F")
                               for b in seis_bars])
        self._check(self.L.ysx_seismograms_vjp(self.h, _dp(flat), None, None, None))
        z = np.zeros(0, np.int32)
        self._check(self.L.ysx_spectra_vjp(self.h, None, None, None, 0, _ip(z), _ip(z), None))
        return self.moment_tensor_vjp().reshape(len(self.events), 6)

    def misfit_gradient_mt(self, seis_bar):
        """d/dM of sum(seis_bar * seismograms) after a forward: the three
        adjoint stages on the device, no per-pair copies to the host."""
        self.seismograms_vjp(seis_bar)
        self.spectra_vjp(np.zeros(0, np.int32), np.zeros(0, np.int32))
        return self.moment_tensor_vjp()

    def model_vjp(self):
        """(29, nknot) table adjoint through the per-pair solves, starting
        solutions and source vectors, from the record adjoints of the last
        spectra_vjp (input of model_table_vjp)."""
        out = np.zeros((len(TABLE_FIELDS), self.deck.nknot))
        self._check(self.L.ysx_model_vjp(self.h, _dp(out)))
        return out

    def grad_memory(self):
        """device bytes of the model gradient: per-step states, step scratch
        with the radial tape, per-pair adjoints (ysx_grad_memory)"""
        out = np.zeros(3)
        self._check(self.L.ysx_grad_memory(self.h, _dp(out)))
        return out

    def grad_timings(self):
        """kernel ms of the last model_vjp: b-vector, minor and toroidal
        sweeps, starting solutions, radial solves, source vectors, sum"""
        out = np.zeros(7)
        self._check(self.L.ysx_grad_timings(self.h, _dp(out)))
        return out

    def misfit_gradient(self, obs, nout=None):
        """L2 misfit 0.5 * sum((seis - obs)^2) over the first nout samples
        (the samples the drop-in program writes) and its gradient with
        respect to the deck columns, (nknot, 9) in the deck's layout, after
        one forward: seismograms_vjp, spectra_vjp, model_vjp and the deck ->
        table adjoint (model_vjp_deck), each on the device."""
        nout = self.nout if nout is None else nout
        seis = self.forward()
        res = np.zeros_like(seis)
        res[:nout] = seis[:nout] - obs[:nout]
        misfit = 0.5 * float(np.sum(res * res))
        self.seismograms_vjp(res)
        self.spectra_vjp(np.zeros(0, np.int32), np.zeros(0, np.int32))
        tb = self.model_vjp()
        return misfit, self.model_vjp_deck(tb)

    def _obs_sizes(self, nout=None):
        out = []
        for e, p in enumerate(self.events):
            own = self.nout if len(self.events) == 1 else self.info[e]["nout"]
            out.append((own if nout is None else nout) * 3 * p.nr)
        return out

    def pack_obs(self, obs, nout=None):
        """The observed seismograms in ysx_misfit_grad's layout: per event,
        the first nout samples of every trace, time fastest, events one
        after another, as one float64 array (one copy).  obs: the (>= n, 3,
        nr) seismograms of the run, or a list with one per event.  A caller
        that reuses the same data (an inversion's iterations) packs them
        once and hands the packed array to misfit_grad."""
        if not isinstance(obs, (list, tuple)):
            obs = [obs]
        if len(obs) != len(self.events):
            raise ValueError(f"obs must hold one array per event ({len(self.events)})")
        sizes = self._obs_sizes(nout)
        flat = np.empty(sum(sizes))
        off = 0
        for e, (o, p, m) in enumerate(zip(obs, self.events, sizes)):
            n = m // (3 * p.nr)
            a = np.asarray(o, dtype=np.float64)
            if a.ndim != 3 or a.shape[0] < n or a.shape[1:] != (3, p.nr):
                raise ValueError(f"event {e}: obs must hold (nout, 3, nr) = ({n}, 3, {p.nr}) "
                                 "samples")
            flat[off:off + m].reshape((n, 3, p.nr), order="F")[...] = a[:n]
            off += m
        return flat

    def misfit_grad(self, obs, nout=None):
        """ysx_misfit_grad: one forward (the batched launch when several
        events are set), the misfit 0.5 * sum((seis - obs)^2) over the first
        nout samples of every trace (default: each event's own output
        length, the samples the drop-in program writes) and its (nknot, 4)
        gradient with respect to each knot's rho, vp, vs and Qmu (vp and vs
        moving both polarisations), summed over the events.  obs: the
        (>= n, 3, nr) seismograms of the run, a list with one per event, or
        their pack_obs array."""
        if isinstance(obs, np.ndarray) and obs.ndim == 1:
            flat = obs  # already in the library's layout (pack_obs)
            if flat.dtype != np.float64 or not flat.flags.c_contiguous or \
                    flat.size != sum(self._obs_sizes(nout)):
                raise ValueError("a packed obs must be pack_obs's float64 array for this nout")
        else:
            flat = self.pack_obs(obs, nout)
        misfit = np.zeros(1)
        grad = np.zeros((self.deck.nknot, 4))
        self._check(self.L.ysx_misfit_grad(self.h, _dp(flat), 0 if nout is None else nout,
                                           _dp(misfit), _dp(grad)))
        return float(misfit[0]), grad

    def model_vjp_deck(self, table_bar):
        """ysx_model_vjp_deck: (nknot, 9) deck_bar = J^T table_bar for this
        context's deck, computed on the device (table_bar: (29, nknot))."""
        tb = np.ascontiguousarray(table_bar, dtype=np.float64)
        if tb.shape != (len(TABLE_FIELDS), self.deck.nknot):
            raise ValueError("table_bar must be (29, nknot)")
        out = np.zeros((self.deck.nknot, 9))
        self._check(self.L.ysx_model_vjp_deck(self.h, _dp(tb), _dp(out)))
        return out

    def model_table(self):
        """(29, nknot) model table the kernels read (built on the device)."""
        out = np.zeros((len(TABLE_FIELDS), self.deck.nknot))
        self._check(self.L.ysx_model_table(self.h, _dp(out)))
        return out

    def model_table_noalloc(self, deck, d_table):
        """The model table of `deck` written into the caller's device buffer
        d_table (address of F_COUNT * nknot doubles on this context's device)
        by ysx_model_table_noalloc; the context's own model is untouched."""
        cols = np.ascontiguousarray(deck.cols, dtype=np.float64)
        self._check(self.L.ysx_model_table_noalloc(self.h, deck.nknot, deck.nic, deck.noc,
                                                   deck.ians, deck.tref, _dp(cols),
                                                   C.c_void_p(int(d_table))))

    def timings(self):
        """ms: source, toroidal, minors, b-vector, radial, assemble, l-sum, IFFT"""
        t = np.zeros(8)
        self._check(self.L.ysx_timings(self.h, _dp(t)))
        return t

    def step_counts(self):
        """exact Cash-Karp step totals of the last spectra (all events):
        toroidal, solid minors, fluid minors, b-vector (one per step of the
        march, which the four columns or the receiver-side rows take
        together), radial"""
        out = np.zeros(5, np.int64)
        self._check(self.L.ysx_step_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong))))
        return out

    # ---- several events on one model
    def set_events(self, events):
        """events: a list of Params (each with its own receivers)."""
        ip = np.concatenate([p.int_block() for p in events]).astype(np.int32)
        fp = np.concatenate([p.float_block() for p in events])
        lat = np.ascontiguousarray(np.concatenate([p.lat for p in events]), dtype=np.float64)
        lon = np.ascontiguousarray(np.concatenate([p.lon for p in events]), dtype=np.float64)
        self._check(self.L.ysx_set_events(self.h, len(events), _ip(ip), _dp(fp), _dp(lat),
                                          _dp(lon)))
        self.events = list(events)
        self.p = events[0]
        self.info = []
        for e in range(len(events)):
            io = np.zeros(8, np.int32)
            fo = np.zeros(8)
            self._check(self.L.ysx_event_info(self.h, e, _ip(io), _dp(fo)))
            self.info.append(dict(zip(("nw", "nt", "i1", "i2", "is", "ir", "nout", "npair"),
                                      map(int, io))))
        self.npair = sum(d["npair"] for d in self.info)
        self.nw, self.nt = self.info[0]["nw"], self.info[0]["nt"]

    def set_batch_tuning(self, persist=0, chunk=1, blocks_per_cu=4):
        self._check(self.L.ysx_set_batch_tuning(self.h, persist, chunk, blocks_per_cu))

    def set_lanes(self, rad_lanes=0):
        """ysx_set_lanes: lanes per l = 0 radial solve (0 default, 1 or 8)."""
        self._check(self.L.ysx_set_lanes(self.h, rad_lanes))

    def forward_events(self, mode=1):
        """mode 0: one event after another (plain); 1: batched.  Returns ms."""
        self._check(self.L.ysx_forward_events(self.h, mode, None))
        return float(self.timings()[7])

    def event_spectra(self, e):
        d = self.info[e]
        shp = (d["nw"], self.events[e].nr)
        out = [np.zeros(shp, np.complex128, order="F") for _ in range(3)]
        self._check(self.L.ysx_event_spectra(self.h, e, *[a.ctypes.data_as(D) for a in out]))
        return tuple(out)

    def event_seismograms(self, e, out=None):
        """event e's (nt, 3, nr) seismograms of the last forward (a device
        to host copy; into `out` when given, a Fortran-ordered float64 array
        of that shape)."""
        shape = (self.info[e]["nt"], 3, self.events[e].nr)
        if out is None:
            out = np.zeros(shape, order="F")
        elif out.shape != shape or out.dtype != np.float64 or not out.flags.f_contiguous:
            raise ValueError(f"out must be a Fortran-ordered float64 array of shape {shape}")
        self._check(self.L.ysx_event_seismograms(self.h, e, _dp(out)))
        return out
