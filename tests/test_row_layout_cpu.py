"""The range-interleaved row order of long two-step runs, checked on the host:
lbm_rowlayout.h (the same code the kernels use) compiled with g++ against
the HIP headers; every (lx, rows) pair up to 2100 x 128 maps rows to a
permutation of slots with each offset's rows consecutive in range order."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "lattice_boltzmann_parallel_solver_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def test_row_slots_are_a_permutation(tmp_path):
    if shutil.which("g++") is None or not os.path.exists(os.path.join(ROCM, "include", "hip", "hip_runtime.h")):
        pytest.skip("g++ or the HIP headers are missing")
    exe = str(tmp_path / "row_layout_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"),
                    "-I", CSRC, "-o", exe, os.path.join(HERE, "row_layout_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
