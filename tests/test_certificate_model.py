"""CPU check of the diagonal known-answer pencil used by the device
certificate-path tests (tests/test_gpu_certificate.py): the oracle builds
exactly H = diag(S_jj E_j), S = diag(1 + 0.07 j) from the same description."""
import numpy as np

import pyoracle
from test_gpu_certificate import N, _diag_problem


def test_oracle_diag_problem_matches():
    prob, h, s = _diag_problem()
    orc = pyoracle.Operator(prob)
    x = np.eye(N, dtype=np.complex128)[:, :5] * (1 + 1j)
    hx, sx = orc.apply(x, "pair")
    assert np.abs(hx - h @ x).max() < 1e-15 and np.abs(sx - s @ x).max() < 1e-15
