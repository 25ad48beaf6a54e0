"""Certificate path on the device: known-answer strict polish and root-manifold
steps (restating tests/native/strict_residual_polish_smoke.cpp:1-70 of the
reference), a deliberately wrong search operator that the fresh FP64
residual must reject (tests/native/original_residual_smoke.cpp:36-63), and
injected failures (expired reduction barrier, nonfinite application) that
every kernel survives.

The known-answer operator is diagonal and exact in FP64: no potential, zero
kinetic energy, projector rows that are unit vectors, and diagonal per-atom D
and Q, so H = diag(S_jj E_j) and S = diag(1 + 0.07 j) as in the reference
test, with no eigensolver defining the answer."""
import types

import numpy as np
import pytest
import scipy.linalg as sl

import pyoracle
from sdft_amd import native
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu

N, ROOTS = 24, 6


def _diag_problem():
    metric = 1.0 + 0.07 * np.arange(N)
    energy = np.array([-0.8 + 0.01 * (i // 2) if i < ROOTS else 1.0 + 0.4 * i for i in range(N)])
    hkl = np.array([(i % 3 - 1, (i // 3) % 3 - 1, i // 9 - 1) for i in range(N)], dtype=np.int32)
    atoms, rows = 4, N // 4
    amps, dblk, qblk = [], [], []
    for a in range(atoms):
        b = np.zeros((rows, N), dtype=np.complex128)
        for r in range(rows):
            b[r, a * rows + r] = 1.0
        amps.append(b)
        sl_ = slice(a * rows, (a + 1) * rows)
        dblk.append(np.diag(metric[sl_] * energy[sl_]).astype(np.complex128))
        qblk.append(np.diag(metric[sl_] - 1.0).astype(np.complex128))
    prob = types.SimpleNamespace(n=N, hkl=hkl, kinetic_ev=np.zeros(N), v_hkl=np.zeros((0, 3), dtype=np.int32),
                                 v_values=np.zeros(0, dtype=np.complex128), atom_rows=[rows] * atoms,
                                 amplitudes=amps, d_blocks=dblk, q_blocks=qblk, regularization=0.0,
                                 shape=(32, 32, 32))
    return prob, np.diag(metric * energy).astype(np.complex128), np.diag(metric).astype(np.complex128)


@pytest.fixture(scope="module")
def diag_op(hip_ctx):
    prob, h, s = _diag_problem()
    op = native.Operator(hip_ctx, prob)
    x = np.zeros((N, 3), dtype=np.complex128)
    x[:, 0] = 1.0
    x[3, 1] = 2.0 - 1.0j
    x[:, 2] = np.arange(N) * (1 + 0.5j)
    hx, sx = op.apply(x, "pair")
    # the device operator is the diagonal pencil to FP64 rounding (the Gauss
    # form of the complex projector products rounds the imaginary parts)
    assert np.abs(hx - h @ x).max() <= 1e-15 * np.abs(h @ x).max()
    assert np.abs(sx - s @ x).max() <= 1e-15 * np.abs(s @ x).max()
    yield op, h, s
    op.close()


def _state(h, s, x):
    sx = s @ x
    e = np.real(np.einsum("ij,ij->j", x.conj(), h @ x))
    return sx, e, h @ x - sx * e


def test_polish_known_answer(diag_op):
    op, h, s = diag_op
    x = np.zeros((N, ROOTS), dtype=np.complex128)
    for j in range(ROOTS):
        x[j, j] = 1 / np.sqrt(s[j, j].real)
        x[ROOTS + j, j] = 3e-10 - 2e-10j
        x[:, j] /= np.sqrt(np.real(x[:, j].conj() @ s @ x[:, j]))
    sx, e, r = _state(h, s, x)
    direction = -(r / np.diag(s).real[:, None])
    ok, xp, ep, rp = op.strict_step("polish", x, direction, tolerance=5e-14)
    assert ok
    fresh = np.linalg.norm(h @ xp - (s @ xp) * ep, axis=0)
    assert fresh.max() <= 5e-14 and rp.max() <= 5e-14
    assert np.abs(xp.conj().T @ s @ xp - np.eye(ROOTS)).max() <= 1e-12
    for j in range(ROOTS):
        exact = np.zeros(N, dtype=np.complex128)
        exact[j] = 1 / np.sqrt(s[j, j].real)
        assert np.linalg.norm(xp[:, j] - exact) <= 1e-13
        assert abs(ep[j] - h[j, j].real / s[j, j].real) <= 1e-14
    # exact vectors stay untouched (tolerance met: no step)
    ok2, xs, _, _ = op.strict_step("polish", xp, direction, tolerance=5e-14)
    assert not ok2 and np.abs(xs - xp).max() <= 1e-15
    # a residual from contamination with the wrong direction sign: no step
    xc = xp.copy()
    xc[ROOTS, 0] += 1e-8
    _, ec, rc = _state(h, s, xc)
    wrong_sign = rc / np.diag(s).real[:, None]
    ok3, xr, _, _ = op.strict_step("polish", xc, wrong_sign, tolerance=5e-14)
    assert not ok3 and np.abs(xr - xc / np.sqrt(np.real(np.einsum("ij,ij->j", xc.conj(), s @ xc)))).max() <= 1e-15
    bad = wrong_sign.copy()
    bad[0, 0] = np.nan
    with pytest.raises(ValueError):
        op.strict_step("polish", xc, bad, tolerance=5e-14)


def test_root_manifold_known_answer(diag_op):
    op, h, s = diag_op
    solved = np.zeros((N, ROOTS), dtype=np.complex128)
    for j in range(ROOTS):
        solved[j, j] = 1 / np.sqrt(s[j, j].real)
    rot = np.eye(ROOTS, dtype=np.complex128)
    angle = 2e-8
    rot[0, 0] = rot[2, 2] = np.cos(angle)
    rot[0, 2] = rot[2, 0] = 1j * np.sin(angle)
    internal = solved @ rot
    _, _, r = _state(h, s, internal)
    assert np.linalg.norm(r, axis=0).max() >= 1e-11  # a tangent step cannot repair an internal rotation
    ok, xm, em, rm = op.strict_step("manifold", internal, tolerance=5e-14)
    assert ok
    assert np.linalg.norm(h @ xm - (s @ xm) * em, axis=0).max() <= 5e-14
    assert np.abs(xm.conj().T @ s @ xm - np.eye(ROOTS)).max() <= 1e-12
    ok2, xm2, _, _ = op.strict_step("manifold", xm, tolerance=5e-14)
    assert not ok2 and np.abs(xm2 - xm).max() <= 1e-15


def test_wrong_search_operator_rejected(hip_ctx, si8):
    """A real-space search operator with a deliberately tiny projector sphere
    (0.3 bohr: most of the PAW coupling dropped) converges in search
    arithmetic to the wrong vectors; the fresh FP64 residual at the hand-off
    rejects them and the solve still certifies the original pencil."""
    h, s = syn.dense_operator(si8)
    ref = sl.eigh(h, s, eigvals_only=True, subset_by_index=[0, 15])
    op = native.Operator(hip_ctx, si8)
    init = syn.initial_psi(si8, 16, 0)
    good = op.davidson(16, 256, 1e-12, init, search_precision=1)
    assert good["converged"] and good["search_rejected"] == 0
    lat = si8.cell.lattice
    frac = si8.cell.positions @ np.linalg.inv(lat)
    op.real_space_search(1.0, lat, frac, np.full(len(si8.atom_rows), 0.3))
    try:
        r = op.davidson(16, 256, 1e-12, init, search_precision=1)
    finally:
        op.real_space_search(0)
    assert r["converged"] and r["search_rejected"] >= 1
    assert np.abs(r["eigenvalues"] - ref).max() < 1e-10
    x = r["eigenvectors"]
    assert np.linalg.norm(h @ x - (s @ x) * r["eigenvalues"], axis=0).max() <= 1.01e-12
    op.close()


def test_barrier_expiry_falls_back(hip_ctx, si8):
    """Forced expiry of the one-launch reduction's grid barrier (raised from
    the host, so every workgroup leaves at its first wait): the Rayleigh-Ritz
    solves are redone by zheevd and the solve still certifies."""
    a = np.random.default_rng(7).standard_normal((96, 96)) + 1j * np.random.default_rng(8).standard_normal((96, 96))
    a = a + a.conj().T
    hip_ctx.inject_fault("barrier_expiry", 1)
    try:
        w, v, path = hip_ctx.hermitian_eigen(a, 8)
    finally:
        hip_ctx.inject_fault("barrier_expiry", 0)
    assert path == 0
    assert np.abs(w - np.linalg.eigvalsh(a)).max() < 1e-12 * np.abs(w).max()
    assert np.linalg.norm(a @ v - v * w[:8], axis=0).max() < 1e-11 * np.abs(w).max()
    op = native.Operator(hip_ctx, si8)
    init = syn.initial_psi(si8, 16, 0)
    base = op.davidson(16, 256, 1e-12, init)
    hip_ctx.inject_fault("barrier_expiry", 3)
    try:
        r = op.davidson(16, 256, 1e-12, init)
    finally:
        hip_ctx.inject_fault("barrier_expiry", 0)
    assert r["converged"] and r["rr_barrier_fallbacks"] == 3 and base["rr_barrier_fallbacks"] == 0
    assert np.abs(r["eigenvalues"] - base["eigenvalues"]).max() < 1e-10
    op.close()


def test_nonfinite_apply_fails_the_solve_alone(hip_ctx, si8):
    op = native.Operator(hip_ctx, si8)
    init = syn.initial_psi(si8, 16, 0)
    hip_ctx.inject_fault("nonfinite_apply", 1)
    try:
        with pytest.raises(native.HipError, match="nonfinite"):
            op.davidson(16, 256, 1e-12, init)
    finally:
        hip_ctx.inject_fault("nonfinite_apply", 0)
    r = op.davidson(16, 256, 1e-12, init)  # the operator and context stay usable
    assert r["converged"] and r["residual_norms"].max() <= 1e-12
    op.close()
