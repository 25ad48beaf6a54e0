"""Deterministic synthetic inputs for the plane-wave H.psi / Davidson path.

Builds the operator inputs that the reference's SCF driver would hand to
``PlaneWaveOperator`` (src/plane_wave_operator.hpp:25-52): an ordered G+k
sphere, a Fourier local potential V(G) in Miller indices, and per-atom PAW
blocks (B rows, D, Q), plus seeded initial wavefunctions.

Recipes (all seeds fixed; configuration c uses base seed 20261202 + c with an
independent splitmix64 stream per quantity):

* atoms: diamond Si sites (fcc + (1/4,1/4,1/4) basis) of cubic cells, repeated
  into supercells; the slab cell places 32 seeded atoms (two species) in
  z in [0,15) bohr with minimum-image pair distance >= 3.5 bohr.
* V(G): Fourier coefficients of sum_atoms depth*exp(-|r-R|^2/(2*0.8^2)) Ha
  (depth -4 Ha, -6 Ha for species B) plus Gaussian noise whose real-space RMS
  is sigma (0.01 Ha), all low-passed to |G|^2 <= 10 bohr^-2 and Hermitian
  (V(-G) = V(G)*).  V(r) = sum_G V(G) e^{iG.r} (no 1/N), the convention of
  src/plane_wave_operator.cpp:465-475.
* projectors: B[p,G] = 4 pi / sqrt(Omega) * i^l * Y_lm(q) * beta(|q|) *
  exp(+i q.R), q = G+k (src/upf_paw_operator_v3.cpp:79-83), with real
  spherical harmonics and beta(q) = N q^l exp(-q^2 rc^2 / 2) normalised to
  int_0^inf beta^2 q^2 dq = 1.  Each l carries two channels; the second
  channel uses rc * 0.7 so the rows are linearly independent.
* D: per atom U^T diag(u) U with u ~ U[-1,1] Ha, U a seeded orthogonal matrix.
* Q: per atom U^T diag(w) U, then scaled so that the non-trivial spectrum of
  S = (1+reg) I + B^H Q B, i.e. spec(1 + G^{1/2} Q G^{1/2}) with G = B B^H,
  lies in [0.8, 1.5] (the signature computation of
  src/plane_wave_operator.cpp:702-717).
* initial psi: complex normals / (1 + |G+k|^2).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

HARTREE_EV = 27.211386245988        # src/atomic_units.hpp:13
BASE_SEED = 20261202
LOWPASS_G2 = 10.0
REGULARIZATION = 1e-12              # src/plane_wave_operator.hpp:29 default

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class SplitMix64:
    """splitmix64 stream (own implementation); uniform in [0,1) with 53 bits."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, count: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(1, count + 1, dtype=np.uint64)
            z = self.state + idx * np.uint64(0x9E3779B97F4A7C15)
            self.state = self.state + np.uint64(count) * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, count: int) -> np.ndarray:
        return (self.next_u64(count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, count: int) -> np.ndarray:
        m = (count + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:count]


def stream(config: int, quantity: int, extra: int = 0) -> SplitMix64:
    seed = (BASE_SEED + config) * 0x100000 + quantity * 0x1000 + extra
    return SplitMix64(seed)


@dataclass
class Species:
    name: str
    depth: float          # Gaussian well depth (Ha)
    rc: float             # projector radius (bohr)
    lmax: int             # channels l = 0..lmax, two radial channels each


@dataclass
class Cell:
    lattice: np.ndarray   # rows a_i (bohr)
    positions: np.ndarray  # cartesian (bohr), natoms x 3
    species: list          # Species per atom


@dataclass
class Problem:
    """One k-point's operator inputs (host arrays, FP64)."""
    cell: Cell
    ecut_ha: float
    kfrac: np.ndarray
    shape: tuple
    hkl: np.ndarray            # n x 3 int32, reference order
    kinetic_ev: np.ndarray     # n
    qcart: np.ndarray          # n x 3 (G+k cartesian)
    v_hkl: np.ndarray          # nv x 3 int32
    v_values: np.ndarray       # nv complex128
    atom_rows: np.ndarray      # natoms int64
    amplitudes: list           # per atom rows x n complex128 (row-major B)
    d_blocks: list             # per atom rows x rows complex128 (row-major)
    q_blocks: list
    regularization: float = REGULARIZATION
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.hkl.shape[0])

    @property
    def np_total(self) -> int:
        return int(self.atom_rows.sum())

    def all_b(self) -> np.ndarray:
        """np x n complex (row p = projector)."""
        return np.concatenate(self.amplitudes, axis=0)

    def block_d(self) -> np.ndarray:
        return _block_diag(self.d_blocks)

    def block_q(self) -> np.ndarray:
        return _block_diag(self.q_blocks)


def _block_diag(blocks):
    total = sum(b.shape[0] for b in blocks)
    out = np.zeros((total, total), dtype=np.complex128)
    o = 0
    for b in blocks:
        r = b.shape[0]
        out[o:o + r, o:o + r] = b
        o += r
    return out


SILICON = Species("Si", -4.0, 1.0, 1)
SLAB_A = Species("A", -4.0, 1.0, 1)
SLAB_B = Species("B", -6.0, 0.7, 2)


def diamond_cell(a: float, reps=(1, 1, 1)) -> Cell:
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=np.float64)
    frac = np.concatenate([base, base + .25])
    pos = []
    for i in range(reps[0]):
        for j in range(reps[1]):
            for k in range(reps[2]):
                for f in frac:
                    pos.append((f + np.array([i, j, k])) * a)
    lattice = np.diag([a * reps[0], a * reps[1], a * reps[2]]).astype(np.float64)
    return Cell(lattice, np.array(pos), [SILICON] * len(pos))


def slab_cell(config: int = 3) -> Cell:
    lattice = np.diag([15.0, 15.0, 30.0])
    rng = stream(config, 5)
    pos = []
    while len(pos) < 32:
        u = rng.uniform(3)
        p = np.array([u[0] * 15.0, u[1] * 15.0, u[2] * 15.0])
        ok = True
        for q in pos:
            d = p - q
            d -= np.round(d / np.diag(lattice)) * np.diag(lattice)
            if np.dot(d, d) < 3.5 ** 2:
                ok = False
                break
        if ok:
            pos.append(p)
    species = [SLAB_A] * 16 + [SLAB_B] * 16
    return Cell(lattice, np.array(pos), species)


def reciprocal(lattice: np.ndarray) -> np.ndarray:
    return 2 * np.pi * np.linalg.inv(lattice).T


def _real_harmonics(l: int, q: np.ndarray) -> np.ndarray:
    """(2l+1) x n real spherical harmonics of the directions of q."""
    r = np.linalg.norm(q, axis=1)
    safe = np.where(r > 0, r, 1.0)
    x, y, z = q[:, 0] / safe, q[:, 1] / safe, q[:, 2] / safe
    if l == 0:
        return np.full((1, q.shape[0]), 0.5 / math.sqrt(math.pi))
    if l == 1:
        c = math.sqrt(3 / (4 * math.pi))
        out = np.stack([c * y, c * z, c * x])
    elif l == 2:
        c1 = 0.5 * math.sqrt(15 / math.pi)
        c0 = 0.25 * math.sqrt(5 / math.pi)
        c2 = 0.25 * math.sqrt(15 / math.pi)
        out = np.stack([c1 * x * y, c1 * y * z, c0 * (3 * z * z - 1), c1 * x * z, c2 * (x * x - y * y)])
    else:
        raise ValueError("l > 2 not generated")
    out[:, r == 0] = 0.0
    return out


def _radial(l: int, rc: float, qn: np.ndarray) -> np.ndarray:
    # int_0^inf q^{2l+2} exp(-q^2 rc^2) dq = Gamma(l+3/2) / (2 rc^{2l+3})
    norm = math.sqrt(2 * rc ** (2 * l + 3) / math.gamma(l + 1.5))
    return norm * qn ** l * np.exp(-0.5 * qn * qn * rc * rc)


def atom_projectors(sp: Species, qcart: np.ndarray, pos: np.ndarray, volume: float) -> np.ndarray:
    qn = np.linalg.norm(qcart, axis=1)
    phase = np.exp(1j * (qcart @ pos))
    rows = []
    for l in range(sp.lmax + 1):
        ylm = _real_harmonics(l, qcart)
        for rc in (sp.rc, 0.7 * sp.rc):
            rad = _radial(l, rc, qn)
            for m in range(2 * l + 1):
                rows.append(4 * np.pi / math.sqrt(volume) * (1j ** l) * ylm[m] * rad * phase)
    return np.array(rows, dtype=np.complex128)


def projector_count(sp: Species) -> int:
    return 2 * sum(2 * l + 1 for l in range(sp.lmax + 1))


def _orthogonal(rng: SplitMix64, n: int) -> np.ndarray:
    a = rng.normal(n * n).reshape(n, n)
    q, r = np.linalg.qr(a)
    return q * np.sign(np.diag(r))


def lowpass_potential(cell: Cell, config: int, sigma: float = 0.01, perturbation: int = 0):
    """Miller-index list and complex V(G) with |G|^2 <= LOWPASS_G2."""
    b = reciprocal(cell.lattice)
    hmax = [int(math.floor(math.sqrt(LOWPASS_G2) / np.linalg.norm(b[i]) * 1.0)) + 1 for i in range(3)]
    hh, kk, ll = np.meshgrid(*[np.arange(-m, m + 1) for m in hmax], indexing="ij")
    hkl = np.stack([hh.ravel(), kk.ravel(), ll.ravel()], axis=1).astype(np.int32)
    g = hkl.astype(np.float64) @ b
    g2 = np.einsum("ij,ij->i", g, g)
    keep = g2 <= LOWPASS_G2
    hkl, g, g2 = hkl[keep], g[keep], g2[keep]
    volume = abs(np.linalg.det(cell.lattice))
    sig = 0.8
    v = np.zeros(len(hkl), dtype=np.complex128)
    for pos, sp in zip(cell.positions, cell.species):
        v += sp.depth * (2 * np.pi * sig * sig) ** 1.5 / volume * np.exp(-0.5 * sig * sig * g2) * np.exp(-1j * (g @ pos))
    # Hermitian noise: draw for the half-space representative of each +-G pair
    rng = stream(config, 1, perturbation)
    m = len(hkl)
    zr = rng.normal(m)
    zi = rng.normal(m)
    key = {tuple(x): i for i, x in enumerate(hkl.tolist())}
    noise = np.zeros(m, dtype=np.complex128)
    scale = sigma / math.sqrt(m)
    for i, t in enumerate(hkl.tolist()):
        j = key[(-t[0], -t[1], -t[2])]
        if i == j:
            noise[i] = scale * zr[i]
        elif tuple(t) > (-t[0], -t[1], -t[2]):
            c = scale * (zr[i] + 1j * zi[i]) / math.sqrt(2)
            noise[i] = c
            noise[j] = np.conj(c)
    v += noise
    # exact Hermitian symmetrisation of the atomic part as well
    vh = v.copy()
    for i, t in enumerate(hkl.tolist()):
        j = key[(-t[0], -t[1], -t[2])]
        vh[i] = 0.5 * (v[i] + np.conj(v[j]))
    return hkl, vh


def default_enumerate(lattice, cutoff_ev, kfrac, ev_per_ha=HARTREE_EV):
    """G-sphere through the product library's host enumeration."""
    from . import native
    return native.enumerate_g_vectors(lattice, cutoff_ev, kfrac, ev_per_ha)


def build_problem(cell: Cell, ecut_ha: float, kfrac, shape, config: int, enumerate_fn=None,
                  q_scale: float | None = None, perturbation: int = 0) -> Problem:
    enumerate_fn = enumerate_fn or default_enumerate
    kfrac = np.asarray(kfrac, dtype=np.float64)
    # src/scf_core.cpp:325-330: cutoff in eV plus 1e-8
    hkl, kin_ev, cart = enumerate_fn(cell.lattice, ecut_ha * HARTREE_EV + 1e-8, kfrac, HARTREE_EV)
    b = reciprocal(cell.lattice)
    qcart = cart + kfrac @ b
    v_hkl, v_values = lowpass_potential(cell, config, 0.02 if perturbation else 0.01, perturbation)
    volume = abs(np.linalg.det(cell.lattice))
    amps, ds, qs, rows = [], [], [], []
    rng_d = stream(config, 2)
    rng_q = stream(config, 3)
    for pos, sp in zip(cell.positions, cell.species):
        bmat = atom_projectors(sp, qcart, pos, volume)
        r = bmat.shape[0]
        u = _orthogonal(rng_d, r)
        ev = rng_d.uniform(r) * 2 - 1
        ds.append((u.T @ np.diag(ev) @ u).astype(np.complex128))
        u2 = _orthogonal(rng_q, r)
        w = rng_q.uniform(r) * 2 - 1
        qs.append((u2.T @ np.diag(w) @ u2).astype(np.complex128))
        amps.append(bmat)
        rows.append(r)
    prob = Problem(cell, ecut_ha, kfrac, tuple(shape), hkl.astype(np.int32), kin_ev, qcart, v_hkl, v_values,
                   np.array(rows, dtype=np.int64), amps, ds, qs)
    if q_scale is None:
        q_scale = overlap_scale(prob)
    prob.q_blocks = [q * q_scale for q in prob.q_blocks]
    prob.extra["q_scale"] = q_scale
    return prob


def overlap_scale(prob: Problem) -> float:
    """Scale s so spec(1 + G^{1/2} (s Q) G^{1/2}) lies in [0.8, 1.5]."""
    bm = prob.all_b()
    gram = bm @ bm.conj().T
    w, u = np.linalg.eigh(0.5 * (gram + gram.conj().T))
    root = (u * np.sqrt(np.maximum(w, 0))) @ u.conj().T
    q = prob.block_q()
    mu = np.linalg.eigvalsh(0.5 * (root @ q @ root + (root @ q @ root).conj().T))
    hi = max(mu.max(), 1e-30)
    lo = max(-mu.min(), 1e-30)
    return float(min(0.5 / hi, 0.2 / lo))


def initial_psi(prob: Problem, columns: int, config: int, extra: int = 0) -> np.ndarray:
    """n x columns complex, column-major order when flattened with order='F'."""
    rng = stream(config, 4, extra)
    z = (rng.normal(prob.n * columns) + 1j * rng.normal(prob.n * columns)).reshape(columns, prob.n).T
    q2 = np.einsum("ij,ij->i", prob.qcart, prob.qcart)
    return np.ascontiguousarray(z / (1.0 + q2)[:, None]).astype(np.complex128)


# ------------------------------------------------------------------ configs
CONFIGS = {
    0: dict(name="si8_gamma", cell=lambda: diamond_cell(10.26), ecut=10.0, shape=(32, 32, 32), bands=32, roots=16),
    1: dict(name="si64_gamma", cell=lambda: diamond_cell(10.26, (2, 2, 2)), ecut=15.0, shape=(72, 72, 72),
            bands=160, roots=160),
    2: dict(name="si8_mp444", cell=lambda: diamond_cell(10.26), ecut=10.0, shape=(32, 32, 32), bands=32, roots=32),
    3: dict(name="slab_2k", cell=lambda: slab_cell(3), ecut=20.0, shape=(48, 48, 96), bands=96, roots=96),
}


def mp_kpoints(n: int = 4) -> np.ndarray:
    """Monkhorst-Pack n x n x n mesh (fractional), shifted grid (2i - n - 1) / 2n."""
    s = (2 * np.arange(1, n + 1) - n - 1) / (2.0 * n)
    return np.array([[a, b, c] for a in s for b in s for c in s])


def config_problem(config: int, kfrac=(0.0, 0.0, 0.0), enumerate_fn=None, shape=None, q_scale=None,
                   cell: Cell | None = None, ecut: float | None = None, perturbation: int = 0) -> Problem:
    c = CONFIGS[config]
    cell = cell or c["cell"]()
    return build_problem(cell, ecut if ecut is not None else c["ecut"], kfrac,
                         shape if shape is not None else c["shape"], config, enumerate_fn, q_scale, perturbation)


def dense_operator(prob: Problem):
    """Dense H and S (n x n) assembled from the textbook identity
    H = T + V(G-G') + B^H D B,  S = (1+reg) I + B^H Q B
    (src/plane_wave_operator.cpp:21-23; test assembly of
    tests/native/plane_wave_operator_smoke.cpp:46-60)."""
    n = prob.n
    vmap = {tuple(h): v for h, v in zip(prob.v_hkl.tolist(), prob.v_values)}
    hkl = prob.hkl.astype(np.int64)
    diff = hkl[:, None, :] - hkl[None, :, :]
    h = np.zeros((n, n), dtype=np.complex128)
    # vectorised lookup through a dense box
    lo = prob.v_hkl.min(axis=0)
    hi = prob.v_hkl.max(axis=0)
    box = np.zeros(tuple(hi - lo + 1), dtype=np.complex128)
    for (a, b_, c), v in vmap.items():
        box[a - lo[0], b_ - lo[1], c - lo[2]] = v
    inside = np.all((diff >= lo) & (diff <= hi), axis=2)
    idx = diff - lo
    h[inside] = box[idx[..., 0][inside], idx[..., 1][inside], idx[..., 2][inside]]
    h[np.diag_indices(n)] += prob.kinetic_ev / HARTREE_EV
    bm = prob.all_b()
    h += bm.conj().T @ prob.block_d() @ bm
    s = (1 + prob.regularization) * np.eye(n, dtype=np.complex128) + bm.conj().T @ prob.block_q() @ bm
    return h, s
