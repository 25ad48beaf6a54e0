"""GPU parity of the Rayleigh-Ritz eigensolver (one-launch Householder
tridiagonalisation, tridiagonal eigenvalues by multisection and eigenvectors
by inverse iteration, back-transformation) against LAPACK (numpy.linalg.eigh),
on random, clustered, degenerate and structured Hermitian matrices, at the
orders the block Davidson uses."""
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


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 129, 160, 333, 640, 960, 1100])
def test_random_hermitian(hip_ctx, n):
    a = random_hermitian(n, n)
    k = min(n, 320)
    w, v, fused = hip_ctx.hermitian_eigen(a, k)
    assert fused == 1
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
    assert fused == 1
    check_pairs(a, w, v, 320, np.sort(base))


def _from_spectrum(lam, seed):
    n = len(lam)
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    a = (q * lam) @ q.conj().T
    return (a + a.conj().T) / 2


@pytest.mark.parametrize("case", ["12-fold", "split-1e-9", "split-1e-6", "6-fold-960"])
def test_inverse_iteration_clusters(hip_ctx, case):
    """Spectra whose lowest eigenvalues come in groups: exact 12-fold and
    6-fold multiplicities (one inverse-iteration cluster each, Gram-Schmidt
    inside it) and triplets split by 1e-9 or 1e-6 (inside or outside the
    clustering gap of 1e-7 |T|). The inverse-iteration path must accept them
    (no divide-and-conquer fallback) with LAPACK-level residuals and an
    orthonormal basis."""
    rng = np.random.default_rng(11)
    if case == "12-fold":
        lam = np.concatenate([np.repeat(rng.uniform(-1, 1, 20), 12), rng.uniform(1, 15, 400)])
    elif case == "6-fold-960":
        lam = np.concatenate([np.repeat(rng.uniform(-1, 1, 40), 6), rng.uniform(1, 15, 720)])
    else:
        split = 1e-9 if case == "split-1e-9" else 1e-6
        lam = np.concatenate([np.repeat(rng.uniform(-1, 1, 60), 3) + split * rng.standard_normal(180),
                              rng.uniform(1, 15, 300)])
    lam = np.sort(lam)
    a = _from_spectrum(lam, 5)
    k = 320 if len(lam) >= 640 else 240
    w, v, fused = hip_ctx.hermitian_eigen(a, k)
    assert fused == 1
    check_pairs(a, w, v, k, np.linalg.eigvalsh(a))


def test_glued_wilkinson_tridiagonal(hip_ctx):
    """Twenty copies of Wilkinson's W21 glued by 1e-14 (pairs of eigenvalues
    equal to many digits, twenty-fold near-multiplicities across copies): the
    classic hard case for inverse iteration."""
    w21 = np.abs(np.arange(-10, 11)).astype(float)
    d = np.tile(w21, 20)
    off = np.ones(d.size - 1)
    off[20::21] = 1e-14
    t = (np.diag(d) + np.diag(off, 1) + np.diag(off, -1)).astype(complex)
    w, v, fused = hip_ctx.hermitian_eigen(t, 200)
    assert fused == 1
    check_pairs(t, w, v, 200, np.linalg.eigvalsh(t))


def test_structured_inputs(hip_ctx):
    """Diagonal input (every column already reduced), a tridiagonal input with
    complex off-diagonal (order-1 complex reflectors) and a real tridiagonal
    input of order 300 (every column already reduced; the reduction's
    reflectors are then sign flips, which the back-transformation must keep
    finite)."""
    d = np.linspace(2, -1, 50)
    w, v, _ = hip_ctx.hermitian_eigen(np.diag(d).astype(complex), 50)
    check_pairs(np.diag(d).astype(complex), w, v, 50, np.sort(d))
    rng = np.random.default_rng(3)
    off = rng.standard_normal(79) + 1j * rng.standard_normal(79)
    t = np.diag(rng.standard_normal(80)).astype(complex) + np.diag(off, -1) + np.diag(off.conj(), 1)
    w, v, _ = hip_ctx.hermitian_eigen(t, 80)
    check_pairs(t, w, v, 80, np.linalg.eigvalsh(t))
    d = rng.standard_normal(300)
    off = rng.standard_normal(299)
    t = (np.diag(d) + np.diag(off, 1) + np.diag(off, -1)).astype(complex)
    w, v, path = hip_ctx.hermitian_eigen(t, 100)
    assert path == 1
    check_pairs(t, w, v, 100, np.linalg.eigvalsh(t))


def test_repeated_calls_reuse_state(hip_ctx):
    """The grid-barrier counter and the published vectors are reset per call."""
    for n in (300, 200, 300):
        a = random_hermitian(n, 100 + n)
        w, _, _ = hip_ctx.hermitian_eigen(a, 0)
        assert np.abs(w - np.linalg.eigvalsh(a)).max() <= 1e-12 * max(1, np.abs(w).max()) * 4


def test_tridiagonal_stage_matches_host_core(hip_ctx):
    """The device tridiagonal stage (multisection + inverse iteration) against
    the host build of the same arithmetic (tests/native/tridiag_core_check.cpp):
    eigenvalues to the bisection tolerance, vectors spanning the same
    eigenspaces (equal up to sign outside clusters)."""
    import ctypes as ct
    import pathlib
    lib = ct.CDLL(str(pathlib.Path(__file__).parent / "native" / "libtridiag_core_check.so"))
    rng = np.random.default_rng(21)
    w21 = np.abs(np.arange(-10, 11)).astype(float)
    glued_d = np.tile(w21, 10)
    glued_e = np.ones(glued_d.size - 1)
    glued_e[20::21] = 1e-14
    for d, e, k in ((rng.standard_normal(400), rng.standard_normal(399), 130), (glued_d, glued_e, 100)):
        n = d.size
        w, z, status = hip_ctx.tridiagonal_eigen(d, e, k)
        assert status == 0
        wh = np.zeros(n)
        p = lambda a: a.ctypes.data_as(ct.c_void_p)  # noqa: E731
        lib.tridiag_core_values(n, p(d), p(e), n, p(wh))
        zh = np.zeros((k, n))
        assert lib.tridiag_core_vectors(n, p(d), p(e), p(wh), k, ct.c_double(1e-7), p(zh)) > 0
        zh = zh.T
        t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        tn = np.abs(wh).max()
        assert np.abs(w - wh).max() <= 2e-15 * tn
        # same eigenspaces: projector difference of the k-dimensional spans
        assert np.abs(z @ (z.T @ zh) - zh).max() <= 1e-10
        assert np.abs(t @ z - z * w[:k]).max() <= 1e-14 * tn
