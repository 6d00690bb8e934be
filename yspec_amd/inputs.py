"""Readers for the reference's two input formats.

* the ``yspec`` parameter file, read the way ``src/yspec.f90:79-139`` reads it
  with the skip-count helpers of ``src/module_util.f90:301-352``: skip a fixed
  number of records (lines), then read one value with Fortran list-directed
  input (blank records are passed over, values end at a blank, comma or
  slash, strings may be quoted);
* the minos "deck" model format, read the way ``src/module_model.f90:194-222``
  reads it (title, ``ians tref``, ``nknot nic noc``, then one knot per line
  with 6 columns for ``ians=0`` or 9 for ``ians=1``).

Errors carry the reference's own ``stop`` messages where it has one.
"""
from __future__ import annotations

import dataclasses
import re
from typing import List

import numpy as np


class YSpecInputError(RuntimeError):
    """Raised where the reference would ``stop`` while reading its inputs."""


_TOKEN = re.compile(r"""\s*(?:'((?:[^']|'')*)'|"((?:[^"]|"")*)"|([^\s,/]+)|(/))\s*,?""")


class _Records:
    """Line reader with Fortran list-directed semantics for value reads."""

    def __init__(self, text: str, name: str):
        self.lines = text.splitlines()
        self.pos = 0
        self.name = name

    def skip(self, n: int) -> None:
        # read(io,*) with no items consumes exactly one record each
        if self.pos + n > len(self.lines):
            raise YSpecInputError(f"{self.name}: end of file while skipping records")
        self.pos += n

    def values(self, count: int) -> List[str]:
        out: List[str] = []
        while len(out) < count:
            if self.pos >= len(self.lines):
                raise YSpecInputError(f"{self.name}: end of file reading {count} value(s)")
            line = self.lines[self.pos]
            self.pos += 1
            i = 0
            while i < len(line) and len(out) < count:
                m = _TOKEN.match(line, i)
                if not m or m.end() == i:
                    break
                i = m.end()
                if m.group(4):  # slash ends the read
                    return out + [None] * (count - len(out))
                tok = m.group(1) if m.group(1) is not None else (
                    m.group(2) if m.group(2) is not None else m.group(3))
                out.append(tok)
        return out

    def read_string(self, skip: int) -> str:
        self.skip(skip)
        return self.values(1)[0]

    def read_int(self, skip: int) -> int:
        self.skip(skip)
        tok = self.values(1)[0]
        try:
            return int(tok)
        except (TypeError, ValueError) as e:
            raise YSpecInputError(f"{self.name}: bad integer {tok!r}") from e

    def read_float(self, skip: int) -> float:
        self.skip(skip)
        tok = self.values(1)[0]
        try:
            return float(tok.replace("d", "e").replace("D", "e"))
        except (AttributeError, ValueError) as e:
            raise YSpecInputError(f"{self.name}: bad real {tok!r}") from e


@dataclasses.dataclass
class Deck:
    """A deck model as read from file; ``cols`` is (nknot, 9) in SI units:
    r, rho, vpv, vsv, qkappa, qmu, vph, vsh, eta (ians=0 decks fill the last
    three with vpv, vsv, 1 -- the reference's own completion,
    src/module_model.f90:214-221)."""
    title: str
    ians: int
    tref: float
    nknot: int
    nic: int
    noc: int
    cols: np.ndarray

    def to_text(self, digits: int = 13) -> str:
        """Deck-format text; ``digits`` significant digits per value."""
        lines = [f" {self.title}", f" {self.ians} {self.tref!r} 1 2",
                 f" {self.nknot} {self.nic} {self.noc}"]
        ncol = 9 if self.ians == 1 else 6
        for row in self.cols:
            lines.append(" ".join(f"{v:.{digits}g}" for v in row[:ncol]))
        return "\n".join(lines) + "\n"


def read_deck(path: str) -> Deck:
    """src/module_model.f90:194-222"""
    try:
        with open(path) as f:
            text = f.read()
    except OSError as e:
        raise YSpecInputError(" problem opening model file") from e
    rec = _Records(text, path)
    title = rec.values(1)[0]
    v = rec.values(2)
    ians, tref = int(v[0]), float(v[1])
    v = rec.values(3)
    nknot, nic, noc = int(v[0]), int(v[1]), int(v[2])
    ncol = 9 if ians == 1 else 6
    cols = np.zeros((nknot, 9))
    for i in range(nknot):
        vals = rec.values(ncol)
        cols[i, :ncol] = [float(t) for t in vals]
    if ians != 1:
        cols[:, 6] = cols[:, 2]
        cols[:, 7] = cols[:, 3]
        cols[:, 8] = 1.0
    return Deck(title, ians, tref, nknot, nic, noc, cols)


@dataclasses.dataclass
class Params:
    """The parameter file's contents, unnormalised (mHz, minutes, s, km, deg,
    N m), in the order src/yspec.f90:88-137 reads them."""
    prefix: str
    model_file: str
    ats: int
    grav: int
    out: int
    cor: int
    lmin: int
    lmax: int
    f1: float
    f2: float
    tout: float
    dt: float
    f11: float
    f12: float
    f21: float
    f22: float
    src_depth: float
    src_lat: float
    src_lon: float
    mm: np.ndarray
    rec_depth: float
    lat: np.ndarray
    lon: np.ndarray

    @property
    def nr(self) -> int:
        return int(len(self.lat))

    def int_block(self) -> np.ndarray:
        """lmin lmax ats grav out cor nr (int32), the C-ABI's run integers"""
        return np.array([self.lmin, self.lmax, self.ats, self.grav, self.out,
                         self.cor, self.nr], dtype=np.int32)

    def float_block(self) -> np.ndarray:
        """f1 f2 tout dt f11 f12 f21 f22 src_depth src_lat src_lon mm[6]
        rec_depth, the C-ABI's run reals"""
        return np.array([self.f1, self.f2, self.tout, self.dt, self.f11,
                         self.f12, self.f21, self.f22, self.src_depth,
                         self.src_lat, self.src_lon, *self.mm,
                         self.rec_depth], dtype=np.float64)

    def to_text(self) -> str:
        """A parameter file the reference reads back to these values."""
        def blk(comment, value):
            return f"\n# {comment}\n  {value}\n"
        s = ("# This file contains the parameters needed to run the\n# program yspec\n\n"
             f"# prefix for output files\n  {self.prefix}\n")
        s += blk("Earth model", f"'{self.model_file}'")
        s += blk("attenuation switch: 1 = on, 0 = off", self.ats)
        s += blk("gravitation: 0 = none, 1 = cowling, 2 = self", self.grav)
        s += blk("output: 0 = displacement, 1 = velocity, 2 = acceleration", self.out)
        s += blk("potential and tilt corrections: 0 = no, 1 = yes", self.cor)
        s += blk("lmin", self.lmin)
        s += blk("lmax", self.lmax)
        for name, v in (("fmin (mHz)", self.f1), ("fmax (mHz)", self.f2),
                        ("length of time series (min)", self.tout),
                        ("time step (sec)", self.dt), ("f11 filter (mHz)", self.f11),
                        ("f12 filter (mHz)", self.f12), ("f21 filter (mHz)", self.f21),
                        ("f22 filter (mHz)", self.f22),
                        ("source depth (km)", self.src_depth),
                        ("source latitude (deg)", self.src_lat),
                        ("source longitude (deg)", self.src_lon)):
            s += blk(name, repr(float(v)))
        for name, v in zip(("M_{r,r}", "M_{r,theta}", "M_{r,phi}", "M_{theta,theta}",
                            "M_{theta,phi}", "M_{phi,phi}"), self.mm):
            s += blk(name + " (Nm)", repr(float(v)))
        s += blk("receiver depth", repr(float(self.rec_depth)))
        s += blk("number of receivers", self.nr)
        s += "\n# receiver latitudes and longitudes\n"
        for a, b in zip(self.lat, self.lon):
            s += f"  {float(a)!r}  {float(b)!r}\n"
        return s


def read_params(path: str) -> Params:
    """src/yspec.f90:79-137 with src/module_util.f90:301-352"""
    try:
        with open(path) as f:
            text = f.read()
    except OSError as e:
        raise YSpecInputError("can't find parameter file") from e
    rec = _Records(text, path)
    prefix = rec.read_string(4)
    model = rec.read_string(2)
    ats = rec.read_int(2)
    grav = rec.read_int(2)
    out = rec.read_int(2)
    cor = rec.read_int(2)
    lmin = rec.read_int(2)
    lmax = rec.read_int(2)
    f1 = rec.read_float(2)
    f2 = rec.read_float(2)
    tout = rec.read_float(2)
    dt = rec.read_float(2)
    f11 = rec.read_float(2)
    f12 = rec.read_float(2)
    f21 = rec.read_float(2)
    f22 = rec.read_float(2)
    sdep = rec.read_float(2)
    slat = rec.read_float(2)
    slon = rec.read_float(2)
    mm = np.array([rec.read_float(2) for _ in range(6)])
    rdep = rec.read_float(2)
    nr = rec.read_int(2)
    rec.skip(2)
    lat = np.zeros(nr)
    lon = np.zeros(nr)
    for i in range(nr):
        v = rec.values(2)
        lat[i], lon[i] = float(v[0]), float(v[1])
    return Params(prefix, model, ats, grav, out, cor, lmin, lmax, f1, f2, tout,
                  dt, f11, f12, f21, f22, sdep, slat, slon, mm, rdep, lat, lon)
