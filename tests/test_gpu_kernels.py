"""GPU parity of the HIP kernels through the C ABI: complex GEMM, the local
Hamiltonian FFT pipeline, the PAW chain and the device-resident operator,
against the oracle and the reference's known-answer tests."""
import numpy as np
import pytest

import pyoracle
from sdft_amd import native
from sdft_amd import synthetic as syn
from test_oracle import smoke_problem, stencil_case

pytestmark = pytest.mark.gpu

FP64_TOL = 2e-11  # tests/native/compute_backend_smoke.cpp:93-94 (FP64)
FP32_TOL = 2e-5   # same file, raw FP32 arithmetic


def naive_gemm(aop, bop, a, b):
    A = a.conj().T if aop else a
    B = b.conj().T if bop else b
    return A @ B


# (100, 70, 2600): K split eight ways with a ragged last split; (4097, 1, 1031):
# the tall plans (128 x 32 tiles for normal A), split along K
@pytest.mark.parametrize("mnk", [(23, 19, 37), (127, 89, 367), (4097, 1, 1031), (64, 64, 16), (1, 1, 1),
                                 (100, 70, 2600)])
@pytest.mark.parametrize("precision", [0, 1])
def test_zgemm_all_ops_padded(hip_ctx, mnk, precision):
    m, n, k = mnk
    tol = FP64_TOL if precision == 0 else FP32_TOL
    for aop in (0, 1):
        for bop in (0, 1):
            ar, ac = (m, k) if aop == 0 else (k, m)
            br, bc = (k, n) if bop == 0 else (n, k)
            lda, ldb, ldc = ar + 3, br + 5, m + 7
            i, j = np.meshgrid(np.arange(lda), np.arange(ac), indexing="ij")
            a = (np.sin(.13 * (i + 2) * (j + 1)) + 1j * np.cos(.17 * (i + 4) * (j + 2))).ravel(order="F")
            i, j = np.meshgrid(np.arange(ldb), np.arange(bc), indexing="ij")
            b = (np.cos(.21 * (i + 1) * (j + 3)) + 1j * np.sin(.31 * (i + 3) * (j + 2))).ravel(order="F")
            A = a.reshape(lda, ac, order="F")[:ar]
            B = b.reshape(ldb, bc, order="F")[:br]
            ref = naive_gemm(aop, bop, A, B)
            scale = max(1.0, np.abs(ref).max())
            sentinel = 9876. - 5432.j
            c = np.full(ldc * n, sentinel, dtype=np.complex128)
            C = c.reshape(ldc, n, order="F")
            C[:m] = np.nan
            hip_ctx.zgemm_strided(aop, bop, m, n, k, a, lda, b, ldb, c, ldc, add=False, precision=precision)
            C = c.reshape(ldc, n, order="F")
            assert np.all(np.isfinite(C[:m])), "beta=0 must not read the output"
            assert np.abs(C[:m] - ref).max() / scale < tol
            assert np.all(C[m:] == sentinel), "GEMM overwrote output stride padding"
            C[:m] = (.01 * (np.arange(m)[:, None] + 1)) - .02j * (np.arange(n)[None, :] + 1)
            prev = C[:m].copy()
            hip_ctx.zgemm_strided(aop, bop, m, n, k, a, lda, b, ldb, c, ldc, add=True, precision=precision)
            C = c.reshape(ldc, n, order="F")
            assert np.abs(C[:m] - (ref + prev)).max() / scale < tol


def test_zgemm_immutable_identity(hip_ctx):
    """Same host storage, new identity when the contents change
    (tests/native/compute_backend_smoke.cpp:60-83)."""
    rng = np.random.default_rng(5)
    a = np.asfortranarray(rng.standard_normal((43, 67)) + 1j * rng.standard_normal((43, 67)))
    key = 101
    for rep in range(4):
        if rep == 2:
            a[:] = rng.standard_normal((43, 67)) + 1j * rng.standard_normal((43, 67))
            key = 202
        b = rng.standard_normal((67, 17)) + 1j * rng.standard_normal((67, 17))
        out = hip_ctx.zgemm(0, 0, a, b, left_identity=key)
        assert np.abs(out - a @ b).max() < 1e-11 * max(1, np.abs(a @ b).max())


def test_zgemm_triple_right(hip_ctx):
    rng = np.random.default_rng(9)
    a = [rng.standard_normal((300, 40)) + 1j * rng.standard_normal((300, 40)) for _ in range(3)]
    b = rng.standard_normal((40, 12)) + 1j * rng.standard_normal((40, 12))
    outs = hip_ctx.zgemm_triple_right(a, b)
    for ai, o in zip(a, outs):
        assert np.abs(o - ai @ b).max() < FP64_TOL * np.abs(ai @ b).max()


@pytest.mark.parametrize("shape", [(8, 15, 14), (27, 25, 16), (4, 6, 7), (32, 32, 32)])
def test_local_hamiltonian_stencil(hip_ctx, shape):
    """Closed-form stencil of tests/native/fft_backend_smoke.cpp:35-40, limit 2e-11."""
    idx, kin, pot, inp, expected = stencil_case(shape)
    scale = max(1.0, np.abs(expected).max())
    got = hip_ctx.local_hamiltonian(shape, idx, kin, pot, inp, precision=0)
    assert np.abs(got - expected).max() / scale < 2e-11
    got32 = hip_ctx.local_hamiltonian(shape, idx, kin, pot, inp, precision=1)
    assert np.abs(got32 - expected).max() / scale < 2e-6


def test_local_hamiltonian_identity_cache(hip_ctx):
    """A nonzero token keeps the potential resident; a new token (or zero)
    must reload it even at the same host address
    (tests/native/fft_operator_cache_smoke.cpp:31-50)."""
    shape = (8, 15, 14)
    idx, kin, pot, inp, expected = stencil_case(shape)
    got = hip_ctx.local_hamiltonian(shape, idx, kin, pot, inp, operator_identity=77)
    pot2 = pot * 1.5
    stale = hip_ctx.local_hamiltonian(shape, idx, kin, pot2, inp, operator_identity=77)
    assert np.abs(stale - got).max() < 1e-13  # promised immutable: cached values used
    fresh = hip_ctx.local_hamiltonian(shape, idx, kin, pot2, inp, operator_identity=78)
    fresh0 = hip_ctx.local_hamiltonian(shape, idx, kin, pot2, inp, operator_identity=0)
    ref2 = hip_ctx.local_hamiltonian(shape, idx, kin, pot2, inp, operator_identity=0)
    assert np.abs(fresh - fresh0).max() < 1e-13 and np.abs(fresh0 - ref2).max() == 0
    assert np.abs(fresh - got).max() > 1e-3


def test_paw_pair_against_dense(hip_ctx):
    rng = np.random.default_rng(2)
    npw, npj, m = 500, 24, 7
    b = rng.standard_normal((npj, npw)) + 1j * rng.standard_normal((npj, npw))
    d = rng.standard_normal((npj, npj)) + 1j * rng.standard_normal((npj, npj))
    d = d + d.conj().T
    q = rng.standard_normal((npj, npj))
    q = q + q.T
    x = rng.standard_normal((npw, m)) + 1j * rng.standard_normal((npw, m))
    for rep in range(3):  # static hits after the first call
        xr = x * (1 + .1 * rep)
        h, s = hip_ctx.paw_pair(11, b, d, q, xr, projector_identity=12)
        eh = b.conj().T @ d @ b @ xr
        es = b.conj().T @ q @ b @ xr
        assert np.abs(h - eh).max() < FP64_TOL * np.abs(eh).max()
        assert np.abs(s - es).max() < FP64_TOL * np.abs(es).max()


def test_onsite_blocks_bitwise_against_oracle(hip_ctx, si8):
    """FP64 onsite D/Q products follow the compensated scalar order of
    src/paw_compensated_product.hpp:41-56: bit-identical to the oracle."""
    rng = np.random.default_rng(4)
    npj = si8.np_total
    p = rng.standard_normal((npj, 9)) + 1j * rng.standard_normal((npj, 9))
    d = [0.5 * (x + x.conj().T) for x in si8.d_blocks]
    q = [0.5 * (x + x.conj().T) for x in si8.q_blocks]
    w = hip_ctx.onsite_blocks(si8.atom_rows, d, q, p, precision=0)
    o = 0
    for r, db, qb in zip(si8.atom_rows, d, q):
        ed = pyoracle.compensated_product(0, db, p[o:o + r])
        eq = pyoracle.compensated_product(0, qb, p[o:o + r])
        bits = lambda a: np.ascontiguousarray(a).view(np.uint64)  # noqa: E731
        assert np.array_equal(bits(w[o:o + r, :9]), bits(ed))
        assert np.array_equal(bits(w[o:o + r, 9:]), bits(eq))
        o += r


@pytest.fixture(scope="module")
def si8_device(hip_ctx, si8):
    return native.Operator(hip_ctx, si8)


def test_device_operator_setup_matches_oracle(si8_device, si8_oracle):
    info = si8_device.info()
    assert si8_device.shape == si8_oracle.shape
    assert np.array_equal(info["indices"], si8_oracle.indices)
    assert np.abs(info["local_real"] - si8_oracle.local_real).max() < 1e-12
    assert np.array_equal(info["kinetic"], si8_oracle.kinetic)
    assert np.abs(info["h_diag"] - si8_oracle.h_diag).max() < 1e-12
    assert np.abs(info["s_diag"] - si8_oracle.s_diag).max() < 1e-12


@pytest.mark.parametrize("which", ["h", "s", "pair"])
def test_device_operator_apply_si8(si8_device, si8_oracle, si8, which):
    x = syn.initial_psi(si8, 13, 0)
    ref = si8_oracle.apply(x, which)
    got = si8_device.apply(x, which, precision=0)
    got32 = si8_device.apply(x, which, precision=1)
    refs = ref if which == "pair" else (ref,)
    gots = got if which == "pair" else (got,)
    gots32 = got32 if which == "pair" else (got32,)
    for r, g, g32 in zip(refs, gots, gots32):
        scale = max(1.0, np.abs(r).max())
        assert np.abs(g - r).max() / scale < 1e-12
        assert np.abs(g32 - r).max() / scale < 2e-5


def test_kernel_clock_counts_launches_and_work(si8_device, si8):
    """The event clock sees exactly the GEMM and local-H launches of one pair
    apply and credits each with its flops / algorithmic bytes."""
    ctx = si8_device.ctx
    x = syn.initial_psi(si8, 13, 0)
    ctx.kernel_clock(True)
    try:
        si8_device.apply(x, "pair", precision=1)
        st = ctx.kernel_clock_stats()
    finally:
        ctx.kernel_clock(False)
    n, npr, m, count = si8.n, si8.np_total, 13, int(np.prod(si8_device.shape))
    # P = B X, H += B^H (DP), S = B^H (QP) + (1 + reg) X: three FP32-arithmetic GEMMs
    assert st["gemm_fp32"]["launches"] == 3
    assert st["gemm_fp32"]["work"] == pytest.approx(3 * 8.0 * n * npr * m)
    assert st["local_fp32"]["launches"] == 1
    assert st["local_fp32"]["work"] == pytest.approx(m * (2 * n * 8 + 4 * count * 8) + count * 4 + n * 8)
    assert st["gemm_fp64"]["launches"] == 0 and st["local_fp64"]["launches"] == 0
    assert all(v["ms"] > 0 for v in st.values() if v["launches"])


def test_device_operator_smoke_geometry(hip_ctx):
    prob, hd, sd = smoke_problem()
    op = native.Operator(hip_ctx, prob, shape=(0, 0, 0))
    assert op.count == 225
    eye = np.eye(prob.n, dtype=np.complex128)
    hy, sy = op.apply(eye, "pair")
    assert np.abs(hy - hd).max() <= 1e-12
    assert np.abs(sy - sd).max() <= 1e-12


def test_device_operator_si64_columns(hip_ctx):
    """Full-size Si64 (npw 23847, 72^3, 512 projectors): device FP64 against
    the oracle on a few columns, FP32 within search tolerance."""
    prob = syn.config_problem(1, enumerate_fn=pyoracle.enumerate_g)
    orc = pyoracle.Operator(prob)
    dev = native.Operator(hip_ctx, prob)
    assert dev.shape == (72, 72, 72) and dev.np == 512
    x = syn.initial_psi(prob, 3, 1)
    rh, rs = orc.apply(x, "pair")
    gh, gs = dev.apply(x, "pair", precision=0)
    assert np.abs(gh - rh).max() / np.abs(rh).max() < 1e-12
    assert np.abs(gs - rs).max() / np.abs(rs).max() < 1e-12
    fh, fs = dev.apply(x, "pair", precision=1)
    assert np.abs(fh - rh).max() / np.abs(rh).max() < 2e-5
