"""GPU parity of the Rayleigh-Ritz eigensolver (one-launch Householder
tridiagonalisation + divide and conquer + back-transformation) against
LAPACK (numpy.linalg.eigh), on random, clustered, degenerate and structured
Hermitian matrices, at the orders the block Davidson uses."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_hermitian(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return (a + a.conj().T) / 2


def check_pairs(a, w, v, k, ref):
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(w - ref).max() <= 1e-12 * scale * max(1, np.sqrt(a.shape[0]) / 8)
    if k:
        res = np.abs(a @ v - v * w[:k]).max()
        assert res <= 1e-11 * scale
        assert np.abs(v.conj().T @ v - np.eye(k)).max() <= 1e-11


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 160, 333, 640, 960])
def test_random_hermitian(hip_ctx, n):
    a = random_hermitian(n, n)
    k = min(n, 320)
    w, v, fused = hip_ctx.hermitian_eigen(a, k)
    assert fused
    check_pairs(a, w, v, k, np.linalg.eigvalsh(a))


def test_degenerate_and_clustered_spectrum(hip_ctx):
    """Exact multiplicities and 1e-10 splittings (the shells of a symmetric
    crystal): eigenvalues to rounding, an orthonormal basis of each cluster."""
    n = 480
    rng = np.random.default_rng(7)
    base = np.repeat(np.linspace(-3, 5, 96), 5)
    base[::7] += 1e-10 * np.arange(base[::7].size)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    a = (q * base) @ q.conj().T
    a = (a + a.conj().T) / 2
    w, v, fused = hip_ctx.hermitian_eigen(a, 320)
    assert fused
    check_pairs(a, w, v, 320, np.sort(base))


def test_structured_inputs(hip_ctx):
    """Diagonal input (every reflector is the identity, tau = 0) and a
    tridiagonal input with complex off-diagonal (order-1 complex reflectors)."""
    d = np.linspace(2, -1, 50)
    w, v, _ = hip_ctx.hermitian_eigen(np.diag(d).astype(complex), 50)
    check_pairs(np.diag(d).astype(complex), w, v, 50, np.sort(d))
    rng = np.random.default_rng(3)
    off = rng.standard_normal(79) + 1j * rng.standard_normal(79)
    t = np.diag(rng.standard_normal(80)).astype(complex) + np.diag(off, -1) + np.diag(off.conj(), 1)
    w, v, _ = hip_ctx.hermitian_eigen(t, 80)
    check_pairs(t, w, v, 80, np.linalg.eigvalsh(t))


def test_repeated_calls_reuse_state(hip_ctx):
    """The grid-barrier counter and the published vectors are reset per call."""
    for n in (300, 200, 300):
        a = random_hermitian(n, 100 + n)
        w, _, _ = hip_ctx.hermitian_eigen(a, 0)
        assert np.abs(w - np.linalg.eigvalsh(a)).max() <= 1e-12 * max(1, np.abs(w).max()) * 4
