"""The numerics of the discrete adjoint (yspec_amd/csrc/ysx_adjoint.hpp) on
the CPU: tests/adjoint_check.cpp is compiled with hipcc for the host and run
on the model table of the seeded 200-knot deck.  It checks identities the
mathematics fixes -- bilinear forms against each system's own right-hand
side, transposed systems against the forward ones, one Cash-Karp step
against its transpose -- and every adjoint (per-lane base adjoints, base
adjoints to table entries, the stage pairings of a step) against central
differences."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from yspec_amd import api
from yspec_amd.synth import synth_deck

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("adjoint_check")
    exe = d / "adjoint_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang",
                    "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", os.path.join(ROOT, "yspec_amd", "csrc"),
                    os.path.join(ROOT, "tests", "adjoint_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    deck = synth_deck(0)
    api.model_table_host(deck).astype("<f8").tofile(d / "table.bin")
    out = subprocess.run([str(exe), str(d / "table.bin"), str(deck.nknot)], check=True,
                         capture_output=True, text=True, timeout=300).stdout
    return {k: float(v) for k, v in (line.split() for line in out.splitlines())}


IDENTITIES = ["bilinear_solid_minor", "bilinear_bvec", "streamed_solid_minor", "streamed_bvec",
              "streamed_fluid_minor", "streamed_tor", "transpose_solid_minor",
              "transpose_bvec", "transpose_fluid_minor", "transpose_tor",
              "step_transpose_solid_minor", "step_transpose_bvec",
              "step_transpose_fluid_minor", "step_transpose_tor", "rad_segments"]
DIFFERENCES = ["lane_ob_solid_minor", "lane_ob_bvec", "lane_ob_fluid_minor", "lane_ob_tor",
               "base_adj_solid", "base_adj_fluid", "base_adj_tor",
               "step_gradient_solid_minor", "step_gradient_bvec",
               "step_gradient_fluid_minor", "step_gradient_tor", "rad_coef_adj",
               "start_tangent_solid", "start_tangent_fluid", "start_tangent_tor",
               "start_tangent_rad"]


@pytest.mark.parametrize("name", IDENTITIES)
def test_identity(results, name):
    assert results[name] <= 1e-13, results[name]


# the starting solutions read the Bessel-function ratio from a continued
# fraction or series stopped at a 1e-9 tolerance (bfs, src/module_function.f90:
# 497-581); differences of the forward carry that noise divided by the step,
# so they use a 4-point stencil with steps of 1e-3 relative, where the
# solid start's noise floor is ~5e-6 (1e-4 at 1e-6 steps: the tangent itself
# follows the iterations exactly)
BOUND = {"start_tangent_solid": 2e-5}


@pytest.mark.parametrize("name", DIFFERENCES)
def test_against_central_differences(results, name):
    # central differences with h = 1e-6 relative: truncation ~1e-8 or below
    assert results[name] <= BOUND.get(name, 1e-6), results[name]
