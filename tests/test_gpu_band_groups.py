"""GPU parity of the bounded band-group solver (src/band_group_solver.cpp:57-211):
local Davidson solves of `group` roots with the completed bands excluded,
then the global Rayleigh-Ritz rotation certified on fresh FP64 H and S;
against the reference's own compiled band-group solver on the same operator
(tests/golden/make_golden.py) and the dense eigensolve."""
import os

import numpy as np
import pytest

from sdft_amd import native
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu

GOLDEN = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_golden.npz"))


@pytest.fixture(scope="module")
def si8_dev(hip_ctx, si8):
    return native.Operator(hip_ctx, si8)


@pytest.mark.parametrize("group", [8, 5])
def test_band_groups_match_reference(si8, si8_dev, group):
    init = syn.initial_psi(si8, 16, 0)
    r = si8_dev.davidson(16, 256, 1e-12, init, band_group_size=group)
    assert r["converged"]
    assert np.abs(r["eigenvalues"] - GOLDEN[f"si8_bandgroups{group}_eigenvalues"]).max() < 1e-10
    h, s = syn.dense_operator(si8)
    x = r["eigenvectors"]
    assert np.linalg.norm(h @ x - (s @ x) * r["eigenvalues"], axis=0).max() <= 1.01e-12
    assert np.abs(x.conj().T @ s @ x - np.eye(16)).max() < 1e-10


SWEEP = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "band_sweep_golden.npz"))


@pytest.mark.parametrize("group", [8, 5])
def test_swept_band_groups_match_reference(si8, si8_dev, group):
    """The swept variant (SDFT_BAND_GROUP_SWEEP=1: 3-iteration local solves over
    half the iteration budget, then the Davidson over all roots from the swept
    block) against the reference's own compiled swept solver
    (tests/golden/make_golden_sweep.py) and the dense certificate."""
    init = syn.initial_psi(si8, 16, 0)
    r = si8_dev.davidson(16, 256, 1e-12, init, band_group_size=group, band_group_sweep=1)
    assert r["converged"]
    assert r["iterations"] <= 256
    assert np.abs(r["eigenvalues"] - SWEEP[f"si8_bandsweep{group}_eigenvalues"]).max() < 1e-10
    h, s = syn.dense_operator(si8)
    x = r["eigenvectors"]
    assert np.linalg.norm(h @ x - (s @ x) * r["eigenvalues"], axis=0).max() <= 1.01e-12
    print(f"swept groups of {group}: {r['iterations']} iterations (reference "
          f"{int(SWEEP[f'si8_bandsweep{group}_iterations'])})")


def test_swept_band_groups_small_budget_solve_all_roots(si8, si8_dev):
    """A budget too small for one sweep of 3-iteration local solves (4 groups
    x 3 > 20 / 2) solves all roots with the Davidson at once, within the limit."""
    init = syn.initial_psi(si8, 16, 0)
    r = si8_dev.davidson(16, 20, 1e-12, init, band_group_size=5, band_group_sweep=1)
    assert r["iterations"] <= 20
    assert np.all(np.isfinite(r["eigenvalues"])) and np.all(np.isfinite(r["residual_norms"]))
    if r["converged"]:
        assert np.abs(r["eigenvalues"] - SWEEP["si8_bandsweep5_eigenvalues"]).max() < 1e-10


def test_band_groups_cold_start_and_invalid_size(si8, si8_dev):
    import scipy.linalg as sl
    h, s = syn.dense_operator(si8)
    ref = sl.eigh(h, s, eigvals_only=True, subset_by_index=[0, 11])
    r = si8_dev.davidson(12, 256, 1e-12, None, band_group_size=4)
    assert r["converged"] and np.abs(r["eigenvalues"] - ref).max() < 1e-10
    with pytest.raises(ValueError, match="BAND_GROUP_SIZE"):
        si8_dev.davidson(12, 256, 1e-12, None, band_group_size=65)
