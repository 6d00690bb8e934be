"""Seeded synthetic inputs of the benchmark workloads.

* ``synth_deck``: a 200-knot isotropic deck (ians=0, tref = 1 s) of radius
  6371 km: solid inner core to 1221.5 km (30 knots), fluid outer core to
  3480 km (60 knots, vs = 0, Qmu = 0), mantle to 6347 km (104 knots) and a
  24 km crust (6 knots); smooth polynomial profiles per region (rho 13.0 ->
  2.6 g/cc, vs 3.6 in the inner core, 7.2 -> 3.2 from the mantle to the
  crust, Qmu 84..600, Qkappa 57823) with a seeded smooth perturbation of at
  most 1 % on rho, vp and vs.  Radii and the duplicated knots that mark the
  ICB, CMB and Moho are exact.
* ``workload``: the parameter blocks of the named workloads, with seeded
  moment tensors (uniform in +-1e20 N m) and receivers uniform on the sphere.
* ``bump_deck``: the data deck of the misfit-gradient workload (2% Gaussian
  bumps in vp and vs).
* ``batch_events`` / ``inversion_events``: the seeded multi-event workloads
  (64 events of uneven frequency grids; 32 events at lmax 600, 40 mHz).
"""
from __future__ import annotations

import numpy as np

from .inputs import Deck, Params

R_EARTH = 6371000.0
REGIONS = (  # name, r0 (m), r1 (m), knots, (rho, vp, vs) at r0 and r1, qmu
    ("inner core", 0.0, 1221500.0, 30, (13.0, 11.2, 3.6), (12.76, 11.02, 3.5), 84.6),
    ("outer core", 1221500.0, 3480000.0, 60, (12.16, 10.36, 0.0), (9.90, 8.06, 0.0), 0.0),
    ("mantle", 3480000.0, 6347000.0, 104, (5.57, 13.72, 7.2), (3.38, 8.10, 4.50), 312.0),
    ("crust", 6347000.0, 6371000.0, 6, (2.90, 6.80, 3.90), (2.60, 5.80, 3.20), 600.0),
)
QKAPPA = 57823.0


def synth_deck(seed: int = 0, knots: int = 200, perturb: float = 0.01) -> Deck:
    if knots != 200:
        raise ValueError("the workload deck has 200 knots")
    rng = np.random.default_rng(seed)
    # smooth perturbation: a few low-order sines in normalised radius,
    # scaled so that its maximum magnitude is `perturb`
    xs = np.linspace(0.0, 1.0, 2001)
    pert = []
    for _ in range(3):
        a = rng.uniform(-1.0, 1.0, 4)
        f = sum(a[k] * np.sin((k + 1) * np.pi * xs + rng.uniform(0, np.pi)) for k in range(4))
        f = f / np.max(np.abs(f)) * perturb * rng.uniform(0.5, 1.0)
        pert.append(f)
    rows = []
    for name, r0, r1, n, top0, top1, qmu in REGIONS:
        rr = np.linspace(r0, r1, n)
        t = (rr - r0) / (r1 - r0)
        # quadratic blend between the end values, slightly concave
        blend = t + 0.15 * t * (1.0 - t)
        for i, r in enumerate(rr):
            x = r / R_EARTH
            vals = [top0[q] + (top1[q] - top0[q]) * blend[i] for q in range(3)]
            vals = [v * (1.0 + np.interp(x, xs, pert[q])) for q, v in enumerate(vals)]
            rho, vp, vs = vals
            if name == "mantle":
                qm = qmu + (143.0 - qmu) * t[i]
            else:
                qm = qmu
            rows.append([r, rho * 1000.0, vp * 1000.0, vs * 1000.0, QKAPPA, qm])
    cols = np.zeros((len(rows), 9))
    cols[:, :6] = np.array(rows)
    cols[:, 6] = cols[:, 2]
    cols[:, 7] = cols[:, 3]
    cols[:, 8] = 1.0
    # 12+ significant digits survive a text round trip; keep the in-memory
    # deck identical to what a file would hold
    cols = np.array([[float(f"{v:.13g}") for v in row] for row in cols])
    nic = REGIONS[0][3]
    noc = nic + REGIONS[1][3]
    return Deck("synth.200", 0, 1.0, len(rows), nic, noc, cols)


def bump_deck(deck: Deck, amp: float = 0.02, at: float = 0.8, width: float = 0.05) -> Deck:
    """The deck with Gaussian bumps of relative amplitude amp in vp and vs
    (both polarisations) centred at radius `at` (fraction of 6371 km), width
    `width`: the data deck of the misfit-gradient workload."""
    cols = deck.cols.copy()
    x = cols[:, 0] / R_EARTH
    for c in (2, 3, 6, 7):
        cols[:, c] *= 1.0 + amp * np.exp(-((x - at) / width) ** 2)
    return Deck(deck.title, deck.ians, deck.tref, deck.nknot, deck.nic, deck.noc, cols)


def _receivers(rng, n):
    u = rng.uniform(0.0, 1.0, (n, 2))
    lat = np.degrees(np.arcsin(2.0 * u[:, 0] - 1.0))
    lon = 360.0 * u[:, 1] - 180.0
    return lat, lon


WORKLOADS = {
    # name: ats, lmax, fmin, fmax, tout (min), dt (s), filters, depth (km), nr
    "cfg1": (0, 150, 0.1, 10.0, 60.0, 10.0, (0.2, 0.5, 8.0, 9.5), 100.0, 10),
    "cfg2": (1, 600, 0.1, 40.0, 360.0, 1.0, (0.2, 0.5, 35.0, 38.0), 600.0, 200),
    # the gradient workload: attenuation on, so that Qmu enters the forward
    "cfg4": (1, 300, 0.1, 20.0, 120.0, 2.0, (0.2, 0.5, 17.0, 19.0), 50.0, 50),
}


def workload(name: str, seed: int = 0, lmax: int | None = None, nr: int | None = None,
             model_file: str = "synth.200", **over) -> Params:
    """Parameter block of a named workload; ``lmax``/``nr`` shrink it for
    tests.  Extra keyword arguments override fields of :class:`Params`."""
    ats, L, f1, f2, tout, dt, flt, depth, n = WORKLOADS[name]
    rng = np.random.default_rng(1000 + seed)
    mm = rng.uniform(-1e20, 1e20, 6)
    slat, slon = _receivers(rng, 1)
    nr = n if nr is None else nr
    lat, lon = _receivers(rng, nr)
    p = Params(prefix="yspec.out", model_file=model_file, ats=ats, grav=2, out=0, cor=0,
               lmin=0, lmax=L if lmax is None else lmax, f1=f1, f2=f2, tout=tout, dt=dt,
               f11=flt[0], f12=flt[1], f21=flt[2], f22=flt[3], src_depth=depth,
               src_lat=float(slat[0]), src_lon=float(slon[0]), mm=mm, rec_depth=0.0,
               lat=lat, lon=lon)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def batch_events(nev: int = 64, seed: int = 0, nr: int = 16, fmax_range=(10.0, 40.0),
                 depth_range=(0.0, 700.0), tout: float = 180.0, dt: float = 1.0,
                 lmax_per_mhz: float = 15.0, model_file: str = "synth.200") -> list:
    """Seeded multi-event workload: per event a source depth uniform in
    ``depth_range`` km, fmax uniform in ``fmax_range`` mHz with lmax =
    ``lmax_per_mhz`` * fmax, ``nr`` receivers at the surface, a moment tensor
    uniform in +-1e20 N m; 180 min at 1 s, attenuation and full
    self-gravitation; cosine filters 0.2/0.5 mHz and 0.85/0.95 fmax."""
    rng = np.random.default_rng(5000 + seed)
    out = []
    for _ in range(nev):
        fmax = float(rng.uniform(*fmax_range))
        depth = float(rng.uniform(*depth_range))
        mm = rng.uniform(-1e20, 1e20, 6)
        slat, slon = _receivers(rng, 1)
        lat, lon = _receivers(rng, nr)
        out.append(Params(prefix="yspec.out", model_file=model_file, ats=1, grav=2, out=0,
                          cor=0, lmin=0, lmax=int(lmax_per_mhz * fmax), f1=0.1, f2=fmax,
                          tout=tout, dt=dt, f11=0.2, f12=0.5, f21=0.85 * fmax,
                          f22=0.95 * fmax, src_depth=depth, src_lat=float(slat[0]),
                          src_lon=float(slon[0]), mm=mm, rec_depth=0.0, lat=lat, lon=lon))
    return out


def inversion_events(nev: int = 32, seed: int = 0, nr: int = 100, depth_range=(0.0, 700.0),
                     model_file: str = "synth.200") -> list:
    """Seeded batched inversion workload: per event a source depth uniform in
    ``depth_range`` km, ``nr`` receivers at the surface and a moment tensor
    uniform in +-1e20 N m; every event lmax = 600, fmax = 40 mHz, 180 min at
    1 s, filters 0.2/0.5/35/38 mHz, attenuation and full self-gravitation."""
    rng = np.random.default_rng(7000 + seed)
    out = []
    for _ in range(nev):
        depth = float(rng.uniform(*depth_range))
        mm = rng.uniform(-1e20, 1e20, 6)
        slat, slon = _receivers(rng, 1)
        lat, lon = _receivers(rng, nr)
        out.append(Params(prefix="yspec.out", model_file=model_file, ats=1, grav=2, out=0,
                          cor=0, lmin=0, lmax=600, f1=0.1, f2=40.0, tout=180.0, dt=1.0,
                          f11=0.2, f12=0.5, f21=35.0, f22=38.0, src_depth=depth,
                          src_lat=float(slat[0]), src_lon=float(slon[0]), mm=mm,
                          rec_depth=0.0, lat=lat, lon=lon))
    return out
