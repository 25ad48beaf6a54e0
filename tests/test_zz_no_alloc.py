"""The allocation-free operator application (sdft_hip_operator_apply_noalloc)
against the allocating device-pointer application (sdft_hip_operator_apply
with device_pointers = 1), both through the C ABI.

A block that fits one staging pass of the scratch array runs the same kernels
with the same GEMM plans as the allocating entry point, so the outputs agree
bit for bit. A block the scratch array splits into several passes runs its
GEMMs on other column counts (another split-K summation order) and agrees to
within 1e-13 of the largest output entry, the FP64 bar of the operator tests.
Device buffers of the test come from the HIP runtime the library links."""
import ctypes as ct

import numpy as np
import pytest

import pyoracle
from sdft_amd import native
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20  # SDFT_HIP_SCRATCH_BYTES
H2D, D2H = 1, 2


class _Runtime:
    """hipMalloc / hipMemcpy / hipFree of the runtime libsdft_hip.so loaded."""

    def __init__(self):
        native.lib()
        rt = ct.CDLL("libamdhip64.so.7")
        rt.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        rt.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        rt.hipFree.argtypes = [ct.c_void_p]
        self.rt = rt

    def alloc(self, nbytes):
        p = ct.c_void_p()
        assert self.rt.hipMalloc(ct.byref(p), ct.c_size_t(nbytes)) == 0
        return p.value

    def upload(self, ptr, a):
        a = np.asfortranarray(a)
        assert self.rt.hipMemcpy(ct.c_void_p(ptr), a.ctypes.data_as(ct.c_void_p), ct.c_size_t(a.nbytes), H2D) == 0

    def download(self, ptr, shape):
        out = np.empty(shape, dtype=np.complex128, order="F")
        assert self.rt.hipMemcpy(out.ctypes.data_as(ct.c_void_p), ct.c_void_p(ptr), ct.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self, *ptrs):
        for p in ptrs:
            assert self.rt.hipFree(ct.c_void_p(p)) == 0


@pytest.fixture(scope="module")
def hip_rt():
    return _Runtime()


def _both(op, rt, x, which, precision):
    """(allocating, allocation-free) outputs of `which` on x: dicts h/s."""
    n, m = x.shape
    nbytes = 16 * n * m
    bufs = [rt.alloc(nbytes) for _ in range(5)]
    xd, h0, s0, h1, s1 = bufs
    try:
        rt.upload(xd, x)
        op.apply_device(xd, m, h1, s1, which, precision, allocate=False)
        free = {k: rt.download(p, (n, m)) for k, p in (("h", h1), ("s", s1)) if which in (k, "pair")}
        op.apply_device(xd, m, h0, s0, which, precision, allocate=True)
        owned = {k: rt.download(p, (n, m)) for k, p in (("h", h0), ("s", s0)) if which in (k, "pair")}
    finally:
        rt.free(*bufs)
    return owned, free


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_apply(hip_ctx, hip_rt, si8, precision):
    """Si8 (one pass): two calls in a row, the second on a smaller block, each
    bitwise equal to the allocating entry point for H, S and the pair; the
    allocation-free calls leave the operator's device memory unchanged."""
    op = native.Operator(hip_ctx, si8)
    try:
        held = op.device_bytes()
        x40 = syn.initial_psi(si8, 40, 0, 7)
        x9 = syn.initial_psi(si8, 9, 0, 8)
        for x in (x40, x9):
            for which in ("h", "s", "pair"):
                # the allocation-free call first, on an operator that has not
                # grown its own staging for this block
                before = op.device_bytes()
                n, m = x.shape
                bufs = [hip_rt.alloc(16 * n * m) for _ in range(3)]
                try:
                    hip_rt.upload(bufs[0], x)
                    op.apply_device(bufs[0], m, bufs[1], bufs[2], which, precision, allocate=False)
                finally:
                    hip_rt.free(*bufs)
                assert op.device_bytes() == before
                owned, free = _both(op, hip_rt, x, which, precision)
                for k in owned:
                    assert np.array_equal(_bits(owned[k]), _bits(free[k])), (which, k, x.shape)
                    assert np.isfinite(free[k]).all()
        assert op.device_bytes() > held  # the allocating entry point grew its staging
    finally:
        op.close()


def test_noalloc_multi_pass_si64(hip_ctx, hip_rt):
    """Si64 (72^3 grid, 512 projectors): a 64-column pair application needs
    more staging than the scratch array holds, so it goes in several passes;
    it agrees with the allocating entry point within 1e-13 relative (FP64)."""
    prob = syn.config_problem(1, enumerate_fn=pyoracle.enumerate_g)
    op = native.Operator(hip_ctx, prob)
    try:
        per_column = 2 * op.sticks * op.shape[2] * 16 + 48 * op.np
        assert 64 * per_column > SCRATCH_BYTES  # the case really takes several passes
        x = syn.initial_psi(prob, 64, 1, 3)
        owned, free = _both(op, hip_rt, x, "pair", 0)
        for k in ("h", "s"):
            scale = np.abs(owned[k]).max()
            assert np.abs(owned[k] - free[k]).max() <= 1e-13 * scale, k
    finally:
        op.close()


def test_noalloc_rejects_invalid_calls(hip_ctx, hip_rt, si8):
    """Null buffers and FP32 with the real-space projector search enabled are
    invalid arguments; FP64 on that operator still runs."""
    op = native.Operator(hip_ctx, si8)
    n = si8.n
    bufs = [hip_rt.alloc(16 * n * 4) for _ in range(3)]
    try:
        hip_rt.upload(bufs[0], syn.initial_psi(si8, 4, 0, 9))
        with pytest.raises(ValueError):
            op.apply_device(bufs[0], 4, 0, bufs[2], "pair", 0, allocate=False)
        with pytest.raises(ValueError):
            op.apply_device(bufs[0], 0, bufs[1], bufs[2], "pair", 0, allocate=False)
        op.real_space_search(1.3)
        with pytest.raises(ValueError):
            op.apply_device(bufs[0], 4, bufs[1], bufs[2], "pair", 1, allocate=False)
        op.apply_device(bufs[0], 4, bufs[1], bufs[2], "pair", 0, allocate=False)
        assert np.isfinite(hip_rt.download(bufs[1], (n, 4))).all()
    finally:
        hip_rt.free(*bufs)
        op.close()
