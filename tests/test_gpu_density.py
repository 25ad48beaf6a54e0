"""GPU parity of the smooth density (src/density_functionals.cpp:406-460)
against the oracle's numpy-FFT restatement on the reference's own grid
max(3, 2 m + 1): Si8 bands over the full G - G' difference support, and a
few Si64 bands over a box support."""
import numpy as np
import pytest

import pyoracle
from sdft_amd import native
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def difference_support(hkl):
    d = (hkl[:, None, :] - hkl[None, :, :]).reshape(-1, 3)
    return np.unique(d, axis=0)


def test_density_si8_difference_support(hip_ctx, si8):
    dev = native.Operator(hip_ctx, si8)
    rng = np.random.default_rng(3)
    c = syn.initial_psi(si8, 16, 0)
    occ = np.full(16, 2.0)
    occ[12:] = rng.uniform(0, 2, 4)
    occ[15] = 0.0
    volume = abs(np.linalg.det(si8.cell.lattice))
    support = difference_support(np.asarray(si8.hkl))
    got = dev.density(c, occ, volume, support)
    ref = pyoracle.density_fft(si8.hkl, c, occ, volume, support)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    with pytest.raises(ValueError, match="occupation"):
        dev.density(c, np.full(16, 2.5), volume, support[:10])


def test_density_si64_box_support(hip_ctx):
    prob = syn.config_problem(1)
    dev = native.Operator(hip_ctx, prob)
    c = syn.initial_psi(prob, 6, 1)
    occ = np.array([2.0, 2.0, 2.0, 1.5, 0.5, 0.0])
    volume = abs(np.linalg.det(prob.cell.lattice))
    r = np.arange(-6, 7)
    support = np.array([[h, k, l] for h in r for k in r for l in r])
    got = dev.density(c, occ, volume, support)
    ref = pyoracle.density_fft(prob.hkl, c, occ, volume, support)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
