"""CPU pins of the real-space projector search oracle (oracle/pyrealspace.py)
with the reference's own real-space tests
(tests/native/sparsity_operator_smoke.cpp:107-133) and the exact limit."""
import numpy as np
import pytest

import pyoracle
import pyrealspace
from realspace_model import model


def _block_error(rs, dense, which):
    n = dense.shape[0]
    return np.abs(rs.apply(np.eye(n, dtype=np.complex128), which) - dense).max()


def test_full_sphere_equals_dense_operator():
    # real-full: radius 3 sqrt(3) at factor 1 covers the cell
    prob, hd, sd, geo = model(radius=3.0 * np.sqrt(3.0))
    op = pyoracle.Operator(prob)
    rs = pyrealspace.RealSpaceSearch(op, prob, factor=1.0, geometry=geo)
    assert rs.total_points == 2 * op.count
    assert _block_error(rs, hd, "h") <= 1e-12
    assert _block_error(rs, sd, "s") <= 1e-12
    # the exact (reciprocal) operator is untouched
    n = hd.shape[0]
    assert np.abs(op.apply(np.eye(n), "h") - hd).max() <= 1e-12


def test_error_shrinks_monotonically_with_radius():
    # real-radius: radii 0.8 .. 3.2 at factor 1, monotone and tenfold
    errors = []
    for r in (0.8, 1.4, 2.2, 3.2):
        prob, hd, sd, geo = model(radius=r)
        op = pyoracle.Operator(prob)
        rs = pyrealspace.RealSpaceSearch(op, prob, factor=1.0, geometry=geo)
        errors.append(_block_error(rs, hd, "h"))
    assert all(b <= a * (1 + 1e-9) for a, b in zip(errors, errors[1:]))
    assert errors[-1] < 0.1 * errors[0]


def test_keep_all_points_is_exact_without_geometry():
    # every point kept: sum_r w_i(r) psi(r) = sum_G B_iG psi_G exactly
    prob, hd, sd, _ = model()
    op = pyoracle.Operator(prob)
    rs = pyrealspace.RealSpaceSearch(op, prob, keep_all=True)
    assert rs.retained == 1.0
    assert _block_error(rs, hd, "h") <= 1e-12
    assert _block_error(rs, sd, "s") <= 1e-12


def test_sphere_selection_is_the_minimum_image_ball():
    # brute force over the 27 neighbouring images
    prob, _, _, geo = model(radius=1.4)
    op = pyoracle.Operator(prob)
    rs = pyrealspace.RealSpaceSearch(op, prob, factor=1.3, geometry=geo)
    shape = np.array(op.shape)
    idx = np.stack(np.unravel_index(np.arange(op.count), tuple(shape)), axis=1) / shape
    shifts = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=np.float64)
    for a in range(2):
        diff = idx[:, None, :] - geo[1][a][None, None, :] + shifts[None, :, :]
        dist = np.sqrt(((diff @ geo[0]) ** 2).sum(axis=2)).min(axis=1)
        expect = np.nonzero(dist <= 1.3 * 1.4 + 1e-12)[0]
        margin = np.abs(dist - 1.3 * 1.4) > 1e-9
        got = rs.atoms[a]["points"]
        assert np.array_equal(np.intersect1d(got, np.nonzero(margin)[0]), np.intersect1d(expect, np.nonzero(margin)[0]))


@pytest.mark.parametrize("factor", [1.0, 1.3, 2.0])
def test_threshold_selection_rule(factor):
    # amplitude mode: kept points exceed the threshold, dropped ones do not,
    # and a larger factor keeps a superset
    prob, hd, _, _ = model()
    op = pyoracle.Operator(prob)
    rs = pyrealspace.RealSpaceSearch(op, prob, factor=factor)
    tighter = pyrealspace.RealSpaceSearch(op, prob, factor=min(2.0, factor + 0.3))
    for a in range(2):
        assert set(rs.atoms[a]["points"]) <= set(tighter.atoms[a]["points"])
    assert 0 < rs.retained <= 1.0
    assert tighter.retained >= rs.retained
    # search error decreases as the threshold falls
    assert _block_error(tighter, hd, "h") <= _block_error(rs, hd, "h") * (1 + 1e-9)


def test_si8_threshold_search_is_close_to_exact(si8, si8_oracle):
    rs = pyrealspace.RealSpaceSearch(si8_oracle, si8, factor=1.5)
    x = np.random.default_rng(3).standard_normal((si8.n, 3)) + 0j
    exact = si8_oracle.apply(x, "h")
    approx = rs.apply(x, "h")
    rel = np.abs(approx - exact).max() / np.abs(exact).max()
    assert rel < 1e-3
    assert rs.retained > 0.999
