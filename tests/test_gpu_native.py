"""The `hip` Accelerator behind the reference's interface, exercised by a
native C++ smoke program (tests/native/hip_accelerator_smoke.cpp) in the
style of the reference's own native smokes."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_hip_accelerator_native_smoke():
    exe = os.path.join(HERE, "native", "hip_accelerator_smoke")
    assert os.path.exists(exe), "build first (__graft_entry__.build())"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, (p.returncode, p.stdout, p.stderr)
    out = json.loads(p.stdout.strip().splitlines()[-1])
    assert out["passed"] and out["gemm_scaled_error"] < 2e-11 and out["fft_scaled_error"] < 2e-11


def test_hip_accelerator_fp32_search_precision():
    exe = os.path.join(HERE, "native", "hip_accelerator_smoke")
    env = dict(os.environ, SDFT_HIP_SEARCH_PRECISION="fp32")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=240, env=env)
    # FP32 search arithmetic must fail the FP64 gate (exit 3), never silently pass
    assert p.returncode == 3, (p.returncode, p.stdout, p.stderr)
    err = json.loads(p.stdout.strip().splitlines()[-1])["error"]
    assert 1e-9 < err < 2e-5
