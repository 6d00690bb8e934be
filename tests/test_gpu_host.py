"""The drop-in program: `yspec <parameter file>` through the Fortran host
(yspec_amd/fortran/yspec) and libysx.so, against the text files the
reference program wrote for the same inputs (the golden fixtures' `ascii`
entries, oracle/_ref/yspec run on the same parameter file and deck).

Output lines are t, ur, ut, up in format 5e15.6 (src/yspec.f90:497-503).
Values must agree to the printed precision: a 6-digit rendering may differ
by one unit in the last digit where the value sits on a rounding boundary,
since the spectra agree to ~1e-15 but not bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, case_names, load_case

EXE = os.path.join(ROOT, "yspec_amd", "fortran", "yspec")


def _quantum(v):
    """one unit in the 6th significant digit of an E15.6 rendering"""
    a = np.abs(v)
    e = np.floor(np.log10(np.where(a > 0, a, 1.0))) + 1
    return np.where(a > 0, 10.0 ** (e - 6), 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", case_names())
def test_host_program_matches_reference_output(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built: python -m yspec_amd.build")
    deck, p, z = load_case(name)
    p.prefix = "out"
    (tmp_path / "model.deck").write_text(deck.to_text(digits=17))
    (tmp_path / "in.txt").write_text(p.to_text())
    print(f"case {name}: yspec in.txt", flush=True)
    r = subprocess.run([EXE, "in.txt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = [str(s) for s in z["ascii"]]
    assert len(ref) == p.nr
    bumped = 0
    for k in range(1, p.nr + 1):
        got = (tmp_path / f"out.{k}").read_text()
        gl, rl = got.splitlines(), ref[k - 1].splitlines()
        assert len(gl) == len(rl) == int(z["nout"])
        for a, b in zip(gl, rl):
            assert len(a) == len(b) == 60
            if a == b:
                continue
            va = np.array(a.split(), dtype=float)
            vb = np.array(b.split(), dtype=float)
            assert va[0] == vb[0]
            assert np.all(np.abs(va - vb) <= 1.0001 * _quantum(vb)), (k, a, b)
            bumped += 1
    total = p.nr * int(z["nout"])
    assert bumped <= max(2, total // 100), f"{bumped} of {total} lines differ in the last digit"


@pytest.mark.gpu
def test_host_program_stops_like_the_reference(tmp_path):
    """Input the reference stops on (a source in the ocean,
    src/yspec.f90:210-215): the drop-in program prints the same text and,
    like the reference program, ends with exit status 0 and no output files."""
    from make_golden import ocean_deck
    from yspec_amd.synth import workload

    ref_exe = os.path.join(ROOT, "oracle", "_ref", "yspec")
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built: python -m yspec_amd.build")
    p = workload("cfg1", lmax=4, nr=1, src_depth=1.0, model_file="model.deck")
    p.prefix = "out"
    (tmp_path / "model.deck").write_text(ocean_deck(0).to_text(digits=17))
    (tmp_path / "in.txt").write_text(p.to_text())
    r = subprocess.run([EXE, "in.txt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "source in the ocean" in r.stdout + r.stderr
    assert not (tmp_path / "out.1").exists()
    if os.path.exists(ref_exe):  # the reference program, built from its sources
        q = subprocess.run([ref_exe, "in.txt"], cwd=tmp_path, capture_output=True, text=True,
                           timeout=100)
        assert q.returncode == 0 and "source in the ocean" in q.stdout + q.stderr
