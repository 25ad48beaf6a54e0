"""The tridiagonal eigensolver's arithmetic (sparse-dft_amd/csrc/hip/tridiag_core.hpp,
shared by the device kernels) built for the host and driven as the kernels
drive it (tests/native/tridiag_core_check.cpp), checked against LAPACK through
scipy: eigenvalues against dstebz bisection and the dense eigensolver, vectors
by residual and orthogonality after the symmetric orthonormalisation the
device applies, on the matrix families that break naive inverse iteration."""
import ctypes as ct
import pathlib
import shutil
import subprocess

import numpy as np
import pytest
import scipy.linalg as sl

LIB = pathlib.Path(__file__).parent / "native" / "libtridiag_core_check.so"
GAP = 1e-7  # kTridiagClusterGap


@pytest.fixture(scope="module")
def core(tmp_path_factory):
    path = LIB
    if not path.exists():
        # built by `make -C sparse-dft_amd`; compile a private copy otherwise
        cxx = shutil.which("g++")
        if cxx is None:
            pytest.fail(f"{LIB} not built and no g++ to build it")
        path = tmp_path_factory.mktemp("tridiag") / LIB.name
        subprocess.run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(path),
                        str(LIB.parent / "tridiag_core_check.cpp")], check=True)
    lib = ct.CDLL(str(path))
    lib.tridiag_core_values.restype = ct.c_int
    lib.tridiag_core_vectors.restype = ct.c_int
    return lib


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def solve(core, d, e, k):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    n = d.size
    w = np.zeros(n)
    rounds = core.tridiag_core_values(n, _p(d), _p(e), n, _p(w))
    z = np.zeros((k, n))
    largest = core.tridiag_core_vectors(n, _p(d), _p(e), _p(w), k, ct.c_double(GAP), _p(z))
    return w, z.T.copy(), rounds, largest


def lowdin(z):
    g = z.T @ z
    e = g - np.eye(len(g))
    return z @ (np.eye(len(g)) - e / 2 + 3 * e @ e / 8), np.abs(e).max()


def check(core, d, e, k, max_cluster=None):
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    tn = np.abs(np.linalg.eigvalsh(t)).max()
    w, z, rounds, largest = solve(core, d, e, k)
    assert largest > 0, "an inverse-iteration iterate left finite range"
    if max_cluster is not None:
        assert largest >= max_cluster
    assert rounds <= 10
    # bisection to the same tolerance as LAPACK dstebz (abstol = 0)
    bis = sl.eigvalsh_tridiagonal(d, e, lapack_driver="stebz")
    assert np.abs(w - bis).max() <= 4e-15 * tn
    assert np.abs(w - np.linalg.eigvalsh(t)).max() <= 1e-13 * tn
    zl, defect = lowdin(z)
    assert defect <= 1e-6  # the device falls back to divide and conquer beyond this
    assert np.abs(t @ zl - zl * w[:k]).max() <= 1e-14 * tn
    assert np.abs(zl.T @ zl - np.eye(k)).max() <= 1e-14


def test_random(core):
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 50, 320):
        check(core, rng.standard_normal(n), rng.standard_normal(max(n - 1, 0)), max(1, n // 3))


def test_wilkinson_pairs(core):
    """W21+ and W101+: pairs of eigenvalues equal to many digits."""
    for m in (10, 50):
        d = np.abs(np.arange(-m, m + 1)).astype(float)
        check(core, d, np.ones(2 * m), 2 * m + 1, max_cluster=2)


def test_glued_wilkinson(core):
    """Twenty W21 glued by 1e-14: twenty-fold near-multiplicities."""
    d = np.tile(np.abs(np.arange(-10, 11)).astype(float), 20)
    e = np.ones(d.size - 1)
    e[20::21] = 1e-14
    check(core, d, e, 200, max_cluster=20)


@pytest.mark.parametrize("case", ["6-fold", "12-fold", "split-1e-9", "split-1e-6"])
def test_degenerate_spectra(core, case):
    """Tridiagonals of Hermitian matrices with prescribed multiplicities
    (LAPACK zhetrd of Q diag(lam) Q^H), the shell structure of a symmetric
    crystal's Ritz values."""
    rng = np.random.default_rng(7)
    if case == "6-fold":
        lam = np.concatenate([np.repeat(rng.uniform(-1, 1, 20), 6), rng.uniform(1, 15, 200)])
    elif case == "12-fold":
        lam = np.concatenate([np.repeat(rng.uniform(-1, 1, 10), 12), rng.uniform(1, 15, 200)])
    else:
        split = 1e-9 if case == "split-1e-9" else 1e-6
        lam = np.concatenate([np.repeat(rng.uniform(-1, 1, 40), 3) + split * rng.standard_normal(120),
                              rng.uniform(1, 15, 200)])
    n = lam.size
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    a = (q * np.sort(lam)) @ q.conj().T
    _, d, e, _, info = sl.lapack.zhetrd((a + a.conj().T) / 2, lower=1)
    assert info == 0
    check(core, d.real.copy(), e.real.copy(), n // 2,
          max_cluster={"6-fold": 6, "12-fold": 12}.get(case, 1))


def test_zero_matrix(core):
    """T = 0: one cluster of every eigenvalue, an orthonormal basis."""
    w, z, _, largest = solve(core, np.zeros(10), np.zeros(9), 10)
    assert largest == 10 and np.abs(w).max() <= 1e-300  # the pivot guard's scale
    assert np.abs(z.T @ z - np.eye(10)).max() <= 1e-14
