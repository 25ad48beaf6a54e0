"""Real-space projector search operator restated with numpy FFTs.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device's
real-space search. The product library never loads it.

Follows src/plane_wave_operator.cpp:104-163 (point selection and projector
values) and src/plane_wave_operator.cpp:167-212 (application):

  w_i(r) = FFT_fwd(B_i)(r) / N       (B_i placed on the grid at the basis indices)
  keep r for atom a: |r - tau_a|_min-image <= factor * radius_a   (geometry given)
                     ||w(r)||_2 > 10^(-2-5(factor-1)) max_r ||w(r)||_2   (otherwise)
  psi(r) = FFT_bwd(x) (unnormalised), p_a = W_a psi(points_a)
  c_a = D_a p_a (H) or Q_a p_a (S)
  f(r) = V(r) psi(r) (H) or 0 (S);  f(points_a) += N W_a^H c_a
  y = T x + FFT_fwd(f)[basis] / N (H) or (1 + reg) x + FFT_fwd(f)[basis] / N (S)

The grid, basis indices, kinetic diagonal and V(r) come from the oracle's
operator (pyoracle.Operator), i.e. from the reference's constructor.
"""
from __future__ import annotations

import numpy as np


def _round_half_away(v):
    """std::round: halves away from zero (numpy's round is half-to-even)."""
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def _hermitian(m):
    m = np.asarray(m, dtype=np.complex128)
    return 0.5 * (m + m.conj().T)


class RealSpaceSearch:
    """Selection and application of src/plane_wave_operator.cpp:104-212.

    op: pyoracle.Operator of `prob`; geometry: None (amplitude threshold) or
    (lattice 3x3 rows in bohr, fractional atoms x 3, radius per atom); an atom
    with radius <= 0 uses the threshold. keep_all=True keeps every grid point
    (the exact limit, src/plane_wave_operator.cpp:98-100)."""

    def __init__(self, op, prob, factor=1.3, geometry=None, keep_all=False):
        if not (1.0 <= factor <= 2.0):
            raise ValueError("SDFT_PROJECTOR_RADIUS_FACTOR must be in [1,2]")
        self.op, self.prob = op, prob
        self.shape = tuple(op.shape)
        self.count = int(op.count)
        self.indices = op.indices.astype(np.int64)
        threshold = 10.0 ** (-2.0 - 5.0 * (factor - 1.0))
        i0, i1, i2 = np.unravel_index(np.arange(self.count), self.shape)
        self.atoms = []
        kept = total = 0.0
        for a, bmat in enumerate(prob.amplitudes):
            rows = bmat.shape[0]
            full = np.zeros((self.count, rows), dtype=np.complex128)
            for i in range(rows):
                grid = np.zeros(self.count, dtype=np.complex128)
                grid[self.indices] = bmat[i]
                full[:, i] = np.fft.fftn(grid.reshape(self.shape)).ravel() / self.count
            peak = np.sqrt((np.abs(full) ** 2).sum(axis=1))
            top = peak.max()
            sphere = geometry is not None and geometry[2][a] > 0
            if keep_all:
                keep = np.ones(self.count, dtype=bool)
            elif sphere:
                lattice = np.asarray(geometry[0], dtype=np.float64)
                tau = np.asarray(geometry[1], dtype=np.float64)[a]
                f = [i0 / self.shape[0] - tau[0], i1 / self.shape[1] - tau[1], i2 / self.shape[2] - tau[2]]
                f = [v - _round_half_away(v) for v in f]
                d2 = np.zeros(self.count)
                for c in range(3):
                    u = f[0] * lattice[0, c] + f[1] * lattice[1, c] + f[2] * lattice[2, c]
                    d2 = d2 + u * u
                cut = factor * geometry[2][a]
                keep = d2 <= cut * cut
            else:
                keep = peak > threshold * top
            points = np.nonzero(keep)[0]
            total += float((peak * peak).sum())
            kept += float((peak[points] ** 2).sum())
            self.atoms.append(dict(points=points, values=full[points].T.copy(), d=_hermitian(prob.d_blocks[a]),
                                   q=_hermitian(prob.q_blocks[a])))
        self.retained = kept / total if total > 0 else 1.0
        self.total_points = int(sum(len(a["points"]) for a in self.atoms))

    def apply(self, x, which):
        """which 'h' or 's'; x: n x m complex."""
        x = np.asarray(x, dtype=np.complex128)
        if x.ndim == 1:
            x = x[:, None]
        y = np.zeros_like(x)
        overlap = which == "s"
        n_grid = self.count
        for c in range(x.shape[1]):
            work = np.zeros(n_grid, dtype=np.complex128)
            work[self.indices] = x[:, c]
            psi = (np.fft.ifftn(work.reshape(self.shape)) * n_grid).ravel()
            amps = []
            for at in self.atoms:
                p = at["values"] @ psi[at["points"]]
                amps.append((at["q"] if overlap else at["d"]) @ p)
            f = np.zeros(n_grid, dtype=np.complex128) if overlap else psi * self.op.local_real
            for at, amp in zip(self.atoms, amps):
                f[at["points"]] += (at["values"].conj().T @ amp) * n_grid
            g = np.fft.fftn(f.reshape(self.shape)).ravel()[self.indices] / n_grid
            diag = (1.0 + self.prob.regularization) if overlap else self.op.kinetic
            y[:, c] = diag * x[:, c] + g
        return y

    def natural_points(self):
        """(concatenated sorted natural indices, count per atom)."""
        return (np.concatenate([a["points"] for a in self.atoms]) if self.atoms else np.zeros(0, dtype=np.int64),
                np.array([len(a["points"]) for a in self.atoms], dtype=np.int64))
