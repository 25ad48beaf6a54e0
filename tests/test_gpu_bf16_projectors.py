"""The refinement's correction operator (H - theta_j S) x_j through the C ABI
in FP64, in FP32 and with the projector GEMMs on BF16 MFMA (FP32
accumulation), against the oracle's pair application; the BF16 form also
against the same products restated on the host with BF16-rounded operands,
which pins its lane layout and conversions; and certified solves whose
correction solves use the BF16 projector products."""
import numpy as np
import pytest

from sdft_amd import native
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round to BF16 (nearest even) through FP32, as the device converts."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64).reshape(np.shape(a))


def cbf16(z):
    return bf16(z.real) + 1j * bf16(z.imag)


@pytest.mark.parametrize("columns", [1, 7, 45])
def test_shifted_operator_precisions(si8, si8_oracle, hip_ctx, columns):
    dev = native.Operator(hip_ctx, si8)
    try:
        x = syn.initial_psi(si8, max(columns, 2), 0)[:, :columns]
        rh, rs = si8_oracle.apply(x, "pair")
        theta = np.linspace(-0.4, 0.9, columns)
        ref = rh - rs * theta
        scale = np.abs(ref).max()
        y64 = dev.apply_shifted(x, theta, 0)
        assert np.abs(y64 - ref).max() / scale < 1e-12
        y32 = dev.apply_shifted(x, theta, 1)
        assert np.abs(y32 - ref).max() / scale < 2e-5
        y16 = dev.apply_shifted(x, theta, 2)
        # the same products with BF16-rounded projectors, block and onsite
        # output (B P and B^H W), FP64 elsewhere
        b, d, q = si8.all_b(), si8.block_d(), si8.block_q()
        exact_nl = b.conj().T @ ((d @ (b @ x)) - (q @ (b @ x)) * theta)
        bq = cbf16(b)
        p = bq @ cbf16(x)
        w = cbf16(d @ p - (q @ p) * theta)
        emulated = ref - exact_nl + bq.conj().T @ w
        assert np.abs(y16 - emulated).max() / scale < 1e-4
        err = np.abs(y16 - ref).max() / scale
        assert 1e-6 < err < 2e-2, err
    finally:
        dev.close()


def test_davidson_bf16_projectors_certify(si8, hip_ctx):
    """FP32 search with the refinement whose correction solves use BF16
    projector products: certified on fresh FP64 applications, eigenvalues
    within 1e-10 Ha of the FP64 path."""
    import scipy.linalg as sl
    h, s = syn.dense_operator(si8)
    w = sl.eigh(h, s, eigvals_only=True, subset_by_index=[0, 15])
    dev = native.Operator(hip_ctx, si8)
    try:
        init = syn.initial_psi(si8, 32, 0)
        r = dev.davidson(16, 256, 1e-12, init, search_precision=1, refinement=1, search_handoff=1.0,
                         bf16_projectors=1)
        assert r["converged"] and r["refinement_steps"] > 0
        assert np.abs(r["eigenvalues"] - w).max() < 1e-10
        x = r["eigenvectors"]
        assert np.linalg.norm(h @ x - (s @ x) * r["eigenvalues"], axis=0).max() <= 1.01e-12
    finally:
        dev.close()
