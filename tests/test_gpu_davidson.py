"""Device-resident block Davidson through the C ABI against a dense FP64
generalized eigensolve and a fresh residual certificate on the CPU
(docs/THEORY.md §9: ||H x - e S x|| <= 1e-12 Ha with x^H S x = 1)."""
import numpy as np
import pytest
import scipy.linalg as sl

from sdft_amd import native
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def si8_dense(si8):
    h, s = syn.dense_operator(si8)
    w = sl.eigh(h, s, eigvals_only=True, subset_by_index=[0, 31])
    return h, s, w


@pytest.fixture(scope="module")
def si8_dev(hip_ctx, si8):
    return native.Operator(hip_ctx, si8)


def certify(h, s, r, roots):
    x = r["eigenvectors"]
    res = np.linalg.norm(h @ x - (s @ x) * r["eigenvalues"], axis=0)
    gram = x.conj().T @ s @ x
    return res, np.abs(gram - np.eye(roots)).max()


@pytest.mark.parametrize("search_precision,refinement,handoff", [(0, 0, 1e-8), (1, 0, 1e-8), (1, 1, 1e-8),
                                                                   (1, 1, 3e-3), (1, 1, 1.0)])
def test_davidson_si8_certified(si8, si8_dev, si8_dense, search_precision, refinement, handoff):
    """FP64 path, FP32 search with the reference's FP64 Davidson tail, and FP32
    search with the Jacobi-Davidson refinement (hand-off at 1e-8, 3e-3 and 1, the
    last handing off after the first Rayleigh-Ritz steps)."""
    h, s, w = si8_dense
    init = syn.initial_psi(si8, 32, 0)
    r = si8_dev.davidson(16, 256, 1e-12, init, search_precision=search_precision, refinement=refinement,
                         search_handoff=handoff)
    if refinement:
        assert r["refinement_steps"] > 0
    assert r["converged"], r
    assert np.abs(r["eigenvalues"] - w[:16]).max() < 1e-10
    res, gerr = certify(h, s, r, 16)
    assert res.max() <= 1.01e-12
    assert gerr < 1e-10
    assert np.abs(res - r["residual_norms"]).max() < 1e-13
    if search_precision:
        assert r["handoff_iteration"] > 0 and r["h_columns_search"] > 0


def test_davidson_si8_matches_reference_golden(si8, si8_dev):
    """Eigenvalues of the reference's own compiled Davidson on the same
    operator and warm start (tests/golden/make_golden.py)."""
    import os
    golden = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_golden.npz"))
    init = syn.initial_psi(si8, 32, 0)
    for precision in (0, 1):
        r = si8_dev.davidson(16, 256, 1e-12, init, search_precision=precision)
        assert r["converged"]
        assert np.abs(r["eigenvalues"] - golden["si8_davidson16_eigenvalues"]).max() < 1e-10
    r = si8_dev.davidson(8, 256, 1e-12)
    assert np.abs(r["eigenvalues"] - golden["si8_cold8_eigenvalues"]).max() < 1e-10


def test_davidson_cold_start_si8(si8_dev, si8_dense, si8):
    h, s, w = si8_dense
    r = si8_dev.davidson(8, 256, 1e-12)
    assert r["converged"]
    assert np.abs(r["eigenvalues"] - w[:8]).max() < 1e-10
    res, _ = certify(h, s, r, 8)
    assert res.max() <= 1.01e-12


def test_davidson_iteration_limited_reports_fresh_residuals(si8_dev, si8_dense):
    """An iteration-limited solve must not claim convergence and must report
    its fresh FP64 residuals (tests/native/plane_wave_operator_smoke.cpp:253-267)."""
    h, s, w = si8_dense
    r = si8_dev.davidson(8, 1, 1e-12)
    assert not r["converged"]
    res, gerr = certify(h, s, r, 8)
    assert np.abs(res - r["residual_norms"]).max() < 1e-13
    assert gerr < 1e-12


@pytest.mark.parametrize("search_precision,band_group_size", [(0, 0), (1, 0), (0, 8)])
def test_davidson_device_blocks_equal_host_buffers(si8, si8_dev, search_precision, band_group_size):
    """device_pointers = 1 (initial block and eigenvectors resident in HBM)
    runs the same kernels as the host-buffer call: the same eigenpairs bit
    for bit, warm and cold, and the device block holds the eigenvectors."""
    init = syn.initial_psi(si8, 16, 0)[:, :16]
    kw = dict(search_precision=search_precision, refinement=search_precision, search_handoff=1.0,
              band_group_size=band_group_size)
    host = si8_dev.davidson(16, 256, 1e-12, init, **kw)
    blocks = [native.DeviceBlock(si8.n, 16) for _ in range(2)]
    try:
        blocks[0].upload(init)
        dev = si8_dev.davidson(16, 256, 1e-12, device_blocks=(blocks[0].ptr, blocks[1].ptr), **kw)
        assert dev["eigenvectors"] is None and dev["converged"] and host["converged"]
        assert np.array_equal(dev["eigenvalues"], host["eigenvalues"])
        assert np.array_equal(blocks[1].download(), host["eigenvectors"])
        assert np.array_equal(blocks[0].download(), np.asfortranarray(init))  # the initial block is only read
        if not band_group_size:
            cold_host = si8_dev.davidson(8, 256, 1e-12, search_precision=search_precision)
            cold = si8_dev.davidson(8, 256, 1e-12, device_blocks=(0, blocks[1].ptr), search_precision=search_precision)
            assert np.array_equal(cold["eigenvalues"], cold_host["eigenvalues"])
            assert np.array_equal(blocks[1].download()[:, :8], cold_host["eigenvectors"])
    finally:
        for b in blocks:
            b.close()


def test_davidson_device_blocks_need_destination(si8_dev):
    with pytest.raises(ValueError):
        si8_dev.davidson(8, device_blocks=(0, 0))
    with pytest.raises(ValueError):
        si8_dev.davidson(8, initial=np.zeros((si8_dev.n, 8), complex), device_blocks=(0, 1))


def test_davidson_invalid_arguments(si8_dev):
    with pytest.raises(ValueError):
        si8_dev.davidson(0)
    with pytest.raises(ValueError):
        si8_dev.davidson(4, subspace_factor=3)


def _reference_seeds(n, roots, tol, initial, order):
    """src/pw_core.cpp:1027-1112 seeding, restated with the host recipe's
    sequential sums (splitmix64 in exact integer arithmetic)."""
    import math
    mask = (1 << 64) - 1
    state = 0x6A09E667F3BCC909

    def uniform():
        nonlocal state
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        return float(z >> 11) * 2.0 ** -53 - 0.5

    def broad():
        b = [complex(uniform(), uniform()) for _ in range(n)]
        nn = 0.0
        for z in b:
            nn += z.real * z.real + z.imag * z.imag
        nn = math.sqrt(nn)
        return [complex(z.real / nn, z.imag / nn) for z in b]

    pert = min(max(100.0 * tol, 1e-10), 1e-6)
    warm = initial is not None and 2 * roots < n
    cols = []
    if warm:
        for j in range(roots):
            src = initial[:, j]
            nn = 0.0
            for z in src:
                nn += z.real * z.real + z.imag * z.imag
            s = pert * math.sqrt(nn)
            b = broad()
            cols.append([complex(src[i].real + s * b[i].real, src[i].imag + s * b[i].imag) for i in range(n)])
    for j in range(min(n, 2 * roots)):
        b = broad()
        col = [complex(0.05 * z.real, 0.05 * z.imag) for z in b]
        col[order[j]] += 1.0
        cols.append(col)
    return np.array(cols, dtype=np.complex128).T


@pytest.mark.parametrize("warm", [True, False])
def test_device_seeds_equal_reference_recipe(si8, si8_dev, warm):
    """The device seed block (generated in parallel from the splitmix64
    stream) equals the reference's sequential recipe bit for bit."""
    roots = 16
    init = syn.initial_psi(si8, roots, 0) if warm else None
    dev = si8_dev.davidson_seeds(roots, 1e-12, init)
    info = si8_dev.info()
    order = np.argsort(info["h_diag"] / info["s_diag"], kind="stable")
    ref = _reference_seeds(si8_dev.n, roots, 1e-12, init, order)
    assert dev.shape == ref.shape
    assert np.array_equal(dev, ref)
