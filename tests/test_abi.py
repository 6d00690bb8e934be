"""The C-ABI library and the Fortran host: symbols, bindings, linkage.

No compute call is made here (this runs without a GPU): the shared library
is loaded and every entry point that include/ysx.h declares must resolve;
the Python and Fortran bindings must name exactly those entry points."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ysx.h")
LIB = os.path.join(ROOT, "yspec_amd", "libysx.so")
FDIR = os.path.join(ROOT, "yspec_amd", "fortran")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ysx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = header_symbols()
    for s in ("ysx_create", "ysx_destroy", "ysx_last_error", "ysx_set_model", "ysx_set_run",
              "ysx_run_info", "ysx_spectra", "ysx_seismograms", "ysx_forward"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libysx.so is not built (python -m yspec_amd.build)")
    lib = ctypes.CDLL(LIB)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, f"declared in ysx.h but not exported: {missing}"


def test_library_exports_only_declared_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines()
                       if re.search(r"\sT\sysx_", ln)})
    assert exported == header_symbols()


def test_python_binding_matches_header():
    from yspec_amd import api
    assert sorted(api.EXPORTS) == header_symbols()


def test_fortran_binding_names_are_declared():
    text = open(os.path.join(FDIR, "ysx_bind.f90")).read().lower()
    names = set(re.findall(r"bind\(c,\s*name='(ysx_[a-z0-9_]+)'\)", text))
    assert names and names <= set(header_symbols())
    # the drop-in host needs at least these
    assert {"ysx_create", "ysx_set_model", "ysx_set_run", "ysx_run_info",
            "ysx_forward", "ysx_destroy", "ysx_last_error"} <= names


def test_fortran_host_links_the_library():
    exe = os.path.join(FDIR, "yspec")
    if not os.path.exists(exe):
        pytest.fail("Fortran host is not built (python -m yspec_amd.build)")
    out = subprocess.run(["ldd", exe], capture_output=True, text=True, check=True).stdout
    line = [ln for ln in out.splitlines() if "libysx.so" in ln]
    assert line and "not found" not in line[0]
    assert os.path.realpath(line[0].split("=>")[1].split()[0]) == os.path.realpath(LIB)


def test_host_rejects_missing_parameter_file(tmp_path):
    """the reference's first stop, before any GPU work"""
    exe = os.path.join(FDIR, "yspec")
    r = subprocess.run([exe, str(tmp_path / "nope.in")], capture_output=True, text=True,
                       cwd=tmp_path)
    # a Fortran STOP with a message, exit status 0, as the reference's own
    assert "can't find parameter file" in (r.stdout + r.stderr)
    assert os.listdir(tmp_path) == []
