"""The spatial ordering's 3-D Hilbert index (fpsample_amd/csrc/fpsx_curve.hpp).

The index only decides which points share a wave of the pruned kernels (the
picks never depend on it), so it must be a permutation of the grid; unit
steps between consecutive indices are what make its blocks compact.  The
header is compiled for the host with g++ and checked on twelve aligned 128^3
sub-cubes (each must be one contiguous, aligned run of indices with unit
steps) and for injectivity on 4 M random grid points.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_hilbert_index_is_a_unit_step_bijection(tmp_path):
    exe = tmp_path / "curve_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fpsample_amd", "csrc"),
                    os.path.join(ROOT, "tests", "curve_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout + out.stderr
