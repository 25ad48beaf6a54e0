"""The drop-in boundary compiles: hip_accelerator.cpp against this package's
restatement of the reference's Accelerator interface (always), and against the
reference's own src/compute_backend.hpp when those sources are present (the
integration build), so the two declarations cannot drift apart unnoticed."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "sparse-dft_amd", "csrc", "host", "hip_accelerator.cpp")
REFERENCE_SRC = "/root/reference/src"
FLAGS = ["-std=c++17", "-fsyntax-only", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def _compile(extra):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    r = subprocess.run(["g++", *FLAGS, *extra, SRC], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]


def test_compiles_against_restated_interface():
    _compile([])


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE_SRC, "compute_backend.hpp")),
                    reason="reference sources not present")
def test_compiles_against_reference_header():
    _compile(["-DSDFT_HIP_WITH_REFERENCE_HEADER", f"-I{REFERENCE_SRC}"])
