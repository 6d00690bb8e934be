"""Build libysx.so (HIP, gfx950) and the Fortran drop-in host in-tree.

    python -m yspec_amd.build            # library + Fortran host
    python -m yspec_amd.build --lib      # library only

The library is compiled with hipcc for gfx950 only; packed FP32 VALU
instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libysx.so")
FDIR = os.path.join(HERE, "fortran")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("YSX_FC", "/opt/rocm/bin/amdflang")

SOURCES = ["ysx_gpu.hip", "ysx_host.cpp"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_lib(extra=None):
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *HIP_FLAGS, *(extra or []), "-I", os.path.join(ROOT, "include"),
               "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"),
                   "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    if not os.path.exists(FC):
        print("amdflang not found; Fortran host not built", flush=True)
        return None
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    exe = os.path.join(FDIR, "yspec")
    srcs = [os.path.join(FDIR, "ysx_bind.f90"), os.path.join(FDIR, "yspec_host.f90")]
    _run([FC, "-O2", "-J", bdir, *srcs, "-o", exe, "-L", HERE, "-lysx",
          "-Wl,-rpath,$ORIGIN/.."])
    return exe


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_lib()
    if "--lib" not in argv and os.path.exists(os.path.join(FDIR, "yspec_host.f90")):
        build_fortran()


if __name__ == "__main__":
    main()
