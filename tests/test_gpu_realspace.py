"""GPU parity of the real-space projector search
(src/plane_wave_operator.cpp:104-212) against the oracle restatement
(oracle/pyrealspace.py), the reference's own real-space tests
(tests/native/sparsity_operator_smoke.cpp:107-133) and the certified solve."""
import numpy as np
import pytest

import pyoracle
import pyrealspace
from realspace_model import model
from sdft_amd import native
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def cell_geometry(prob):
    lat = np.asarray(prob.cell.lattice, dtype=np.float64)
    frac = np.asarray(prob.cell.positions) @ np.linalg.inv(lat)
    radius = np.array([sp.rc for sp in prob.cell.species])
    return lat, frac, radius


def _dev_block_error(dev, dense, which, precision=0):
    n = dense.shape[0]
    return np.abs(dev.apply_real_space(np.eye(n, dtype=np.complex128), which, precision) - dense).max()


def test_full_sphere_equals_dense_operator(hip_ctx):
    prob, hd, sd, geo = model(radius=3.0 * np.sqrt(3.0))
    dev = native.Operator(hip_ctx, prob)
    points, retained = dev.real_space_search(1.0, *geo)
    assert points == 2 * dev.count and retained == 1.0
    assert _dev_block_error(dev, hd, "h") <= 1e-12
    assert _dev_block_error(dev, sd, "s") <= 1e-12
    # FP64 (exact) applications stay on the reciprocal projectors
    n = hd.shape[0]
    hy, sy = dev.apply(np.eye(n, dtype=np.complex128), "pair", 0)
    assert np.abs(hy - hd).max() <= 1e-12 and np.abs(sy - sd).max() <= 1e-12


def test_radius_sweep_matches_oracle_and_is_monotone(hip_ctx):
    errors = []
    for r in (0.8, 1.4, 2.2, 3.2):
        prob, hd, sd, geo = model(radius=r)
        dev = native.Operator(hip_ctx, prob)
        points, retained = dev.real_space_search(1.0, *geo)
        rs = pyrealspace.RealSpaceSearch(pyoracle.Operator(prob), prob, factor=1.0, geometry=geo)
        nat, per_atom = dev.real_space_points(2)
        ref_nat, ref_counts = rs.natural_points()
        assert np.array_equal(per_atom, ref_counts)
        off = 0
        for a in range(2):
            assert np.array_equal(np.sort(nat[off:off + per_atom[a]]), ref_nat[off:off + per_atom[a]])
            off += per_atom[a]
        assert abs(retained - rs.retained) < 1e-12
        n = hd.shape[0]
        x = np.eye(n, dtype=np.complex128)
        for which in ("h", "s"):
            ref = rs.apply(x, which)
            assert np.abs(dev.apply_real_space(x, which, 0) - ref).max() <= 1e-12
            assert np.abs(dev.apply_real_space(x, which, 1) - ref).max() <= 2e-6 * max(1.0, np.abs(ref).max())
        errors.append(_dev_block_error(dev, hd, "h"))
    assert all(b <= a * (1 + 1e-9) for a, b in zip(errors, errors[1:]))
    assert errors[-1] < 0.1 * errors[0]


def _compare_selection(dev, rs, atoms, points):
    nat, per_atom = dev.real_space_points(atoms)
    ref_nat, ref_counts = rs.natural_points()
    off_d = off_r = 0
    for a in range(atoms):
        got = set(nat[off_d:off_d + per_atom[a]].tolist())
        exp = set(ref_nat[off_r:off_r + ref_counts[a]].tolist())
        # the device compares squared norms; only threshold ties may differ
        assert len(got ^ exp) <= max(2, 1e-4 * len(exp))
        off_d += per_atom[a]
        off_r += ref_counts[a]


@pytest.mark.parametrize("factor", [1.0, 1.5])
def test_si8_threshold_and_sphere_match_oracle(hip_ctx, si8, si8_oracle, factor):
    dev = native.Operator(hip_ctx, si8)
    x = np.asfortranarray(syn.initial_psi(si8, 6, 0))
    for geo in (None, cell_geometry(si8)):
        points, retained = dev.real_space_search(factor, *(geo or (None, None, None)))
        rs = pyrealspace.RealSpaceSearch(si8_oracle, si8, factor=factor, geometry=geo)
        atoms = len(si8.atom_rows)
        if geo is None:
            _compare_selection(dev, rs, atoms, points)
        else:
            nat, per_atom = dev.real_space_points(atoms)
            ref_nat, ref_counts = rs.natural_points()
            assert np.array_equal(per_atom, ref_counts)
            assert np.array_equal(np.sort(nat), np.sort(ref_nat))
        if points != rs.total_points:
            continue  # a threshold tie changed one point: the selection check above covers it
        for which in ("h", "s"):
            ref = rs.apply(x, which)
            scale = np.abs(ref).max()
            assert np.abs(dev.apply_real_space(x, which, 0) - ref).max() <= 1e-11 * scale
            assert np.abs(dev.apply_real_space(x, which, 1) - ref).max() <= 1e-5 * scale


def test_search_operator_solve_is_certified_exactly(hip_ctx, si8):
    # src/plane_wave_operator.cpp:678-681: real-space projectors accelerate the
    # search only; the certificate uses FP64 reciprocal projectors
    dev = native.Operator(hip_ctx, si8)
    init = syn.initial_psi(si8, 16, 0)
    exact = dev.davidson(16, 256, 1e-12, init, search_precision=0)
    dev.real_space_search(1.3, *cell_geometry(si8))
    r = dev.davidson(16, 256, 1e-12, init, search_precision=1)
    assert r["converged"] and r["h_columns_search"] > 0
    assert np.abs(r["eigenvalues"] - exact["eigenvalues"]).max() < 1e-10
    h, s = syn.dense_operator(si8)
    x = r["eigenvectors"]
    assert np.linalg.norm(h @ x - (s @ x) * r["eigenvalues"], axis=0).max() <= 1.01e-12
    dev.real_space_search(0)
    assert dev.real_space_search(0)[0] == 0


def test_invalid_factor_and_disabled_apply(hip_ctx, si8):
    dev = native.Operator(hip_ctx, si8)
    with pytest.raises(ValueError):
        dev.real_space_search(2.5)
    with pytest.raises(ValueError):
        dev.apply_real_space(np.zeros((si8.n, 1), dtype=np.complex128), "h")
