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
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libysx.so")
FDIR = os.path.join(HERE, "fortran")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("YSX_FC", "/opt/rocm/bin/amdflang")

# (source, object, extra flags): the march kernels dominate compile time, so
# ysx_k_march.hip is compiled once per kernel and work-distribution mode and
# all objects compile in parallel.
SOURCES = [("ysx_k_march.hip", f"ysx_k_march_{k}{p}.o", [f"-DYSX_MARCH={k}", f"-DYSX_PERSIST={p}"])
           for k in range(6) for p in range(2)]
# The receiver-sum MFMAs accumulate in VGPRs: with the default AGPR form the
# loop-carried accumulators are copied between AGPRs and VGPRs every step.
SOURCES += [("ysx_k_misc.hip", "ysx_k_misc.o", ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]),
            ("ysx_k_ngc.hip", "ysx_k_ngc.o", []),
            ("ysx_k_adjoint.hip", "ysx_k_adjoint.o", []),
            ("ysx_api.hip", "ysx_api.o", []), ("ysx_host.cpp", "ysx_host.o", [])]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def _compile(entry, extra):
    src, obj, defs = entry
    obj = os.path.join(HERE, "_build", obj)
    inc = ["-I", os.path.join(ROOT, "include"), *defs]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", *inc, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC, *HIP_FLAGS, *(extra or []), *inc, "-c", os.path.join(CSRC, src), "-o", obj]
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return obj, r


def build_lib(extra=None):
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda s: _compile(s, extra), SOURCES))
    objs = []
    failed = []
    for obj, r in results:
        if r.stdout:
            print(r.stdout, end="", flush=True)
        if r.stderr:
            print(r.stderr, end="", file=sys.stderr, flush=True)
        if r.returncode != 0:
            failed.append(os.path.basename(obj))
        objs.append(obj)
    if failed:
        raise RuntimeError("compilation failed: " + ", ".join(failed))
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
