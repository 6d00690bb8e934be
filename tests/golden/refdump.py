"""Reader for the FP64 dumps written by oracle/ref_dump (the reference's own
modules driven by the project's dump driver).  Layout: see the header of
oracle/ref_dump.f90."""
from __future__ import annotations

import os

import numpy as np

PAIR_FIELDS = (("svt", (5, 2)), ("ytor", (2, 2)), ("svs", (5, 6)),
               ("ysph", (6, 4)), ("sr0", (2,)), ("yrad", (2, 2)),
               ("uu", (5,)), ("vv", (5,)), ("ww", (5,)))
MODEL_NAMES = ("r", "rho", "rho_cs", "acon", "acon_cs", "ccon", "ccon_cs",
               "fcon", "fcon_cs", "lcon", "lcon_cs", "ncon", "ncon_cs",
               "kappa", "kappa_cs", "mu", "mu_cs", "qkappa", "qkappa_cs",
               "qmu", "qmu_cs", "rat", "rat_cs", "xlam", "xlam_cs", "xa2",
               "xa2_cs", "grav", "grav_cs", "lambda", "beta", "alpha")


def read_dump(d: str) -> dict:
    out = {}
    with open(os.path.join(d, "header.bin"), "rb") as f:
        flags = np.frombuffer(f.read(32), "<i4")
        norms = np.frombuffer(f.read(72), "<f8")
        run = np.frombuffer(f.read(64), "<i4")
        runf = np.frombuffer(f.read(176), "<f8")
    keys = ("nknot", "nic", "noc", "nsl", "ifanis", "ocean", "fic", "nlayer")
    out.update({k: int(v) for k, v in zip(keys, flags)})
    keys = ("tref", "r_norm", "rho_norm", "vel_norm", "acl_norm", "fre_norm",
            "t_norm", "moment_norm", "pibigg")
    out.update({k: float(v) for k, v in zip(keys, norms)})
    keys = ("lmin", "lmax", "nw", "nt", "i1", "i2", "is", "ir", "nr", "ats",
            "grav", "out", "cor")
    out.update({k: int(v) for k, v in zip(keys, run)})
    keys = ("rs", "rr", "dt", "tout", "ep", "df", "wmin", "dw", "f11", "f12",
            "f21", "f22", "f1", "f2", "src_depth", "rec_depth")
    out.update({k: float(v) for k, v in zip(keys, runf[:16])})
    out["mm"] = runf[16:22].copy()
    nk, nr, L1 = out["nknot"], out["nr"], out["lmax"] + 1
    m = np.fromfile(os.path.join(d, "model.bin"), "<f8", count=32 * nk)
    out["model"] = {n: m[i * nk:(i + 1) * nk] for i, n in enumerate(MODEL_NAMES)}
    out["model_stack"] = m.reshape(32, nk)
    g = np.fromfile(os.path.join(d, "geom.bin"), "<f8")
    out["delta"], out["azep"], out["azst"] = g[:nr], g[nr:2 * nr], g[2 * nr:3 * nr]
    n3 = L1 * 3 * nr
    for i, name in enumerate(("xa", "xpa", "xca")):
        out[name] = g[3 * nr + i * n3:3 * nr + (i + 1) * n3].reshape((L1, 3, nr), order="F")
    recdt = np.dtype([("ints", "<i4", 4), ("c", "<c16", 89)])
    pr = np.fromfile(os.path.join(d, "pairs.bin"), recdt)
    pairs = {"l": pr["ints"][:, 0].copy(), "j": pr["ints"][:, 1].copy(),
             "tor_start": pr["ints"][:, 2].copy(), "sph_start": pr["ints"][:, 3].copy()}
    n = 0
    for name, shape in PAIR_FIELDS:
        cnt = int(np.prod(shape))
        pairs[name] = pr["c"][:, n:n + cnt].reshape((len(pr),) + shape, order="F")
        n += cnt
    out["pairs"] = pairs
    nw = out["nw"]
    s = np.fromfile(os.path.join(d, "spec.bin"), "<c16")
    out["spec"] = [s[i * nw * nr:(i + 1) * nw * nr].reshape((nw, nr), order="F")
                   for i in range(3)]
    with open(os.path.join(d, "seis.bin"), "rb") as f:
        nt, nout = np.frombuffer(f.read(8), "<i4")
        z = np.frombuffer(f.read(), "<c16")
    out["nt_seis"], out["nout"] = int(nt), int(nout)
    out["seis"] = [z[i * nt * nr:(i + 1) * nt * nr].reshape((nt, nr), order="F")
                   for i in range(3)]
    return out
