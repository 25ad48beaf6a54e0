"""Two-atom model of tests/native/sparsity_operator_smoke.cpp:27-86 (the
reference's own real-space projector tests): a 6 bohr cubic cell, the basis
h^2 + k^2 + l^2 <= 9, V(G) on G = 0 and +-(1,0,0), and per atom three
Gaussian projectors (s, x, z angular factors) centred at the atom, with the
dense H and S they define."""
import numpy as np

from sdft_amd import synthetic as syn

EV = syn.HARTREE_EV


def model(radius=1.4):
    a = 6.0
    b = 2 * np.pi / a
    hkl = [(h, k, l) for h in range(-3, 4) for k in range(-3, 4) for l in range(-3, 4) if h * h + k * k + l * l <= 9]
    hkl = np.array(hkl, dtype=np.int32)
    g = b * hkl.astype(np.float64)
    g2 = (g * g).sum(axis=1)
    kin_ev = EV * 0.5 * b * b * (hkl.astype(np.float64) ** 2).sum(axis=1)
    v = {(0, 0, 0): 0.3, (1, 0, 0): 0.05 + 0.02j, (-1, 0, 0): 0.05 - 0.02j}
    n = len(hkl)
    tau = np.array([[0.1, 0.2, 0.3], [0.6, 0.55, 0.7]])
    width = [0.9, 1.1]
    d = np.array([[0.4, 0.03 + 0.02j, -0.02], [0.03 - 0.02j, -0.25, 0.01], [-0.02, 0.01, 0.3]], dtype=np.complex128)
    q = np.array([[0.05, 0.004 + 0.003j, 0], [0.004 - 0.003j, 0.03, 0], [0, 0, 0.06]], dtype=np.complex128)
    hd = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        for j in range(n):
            hd[i, j] = v.get(tuple(int(t) for t in hkl[i] - hkl[j]), 0)
    hd += np.diag(kin_ev / EV)
    sd = (1 + 1e-12) * np.eye(n, dtype=np.complex128)
    amps = []
    for at in range(2):
        phase = g @ (a * tau[at])
        env = 0.05 * np.exp(-0.5 * width[at] ** 2 * g2) * np.exp(1j * phase)
        bm = np.stack([env, g[:, 0] * env, g[:, 2] * env]).astype(np.complex128)
        amps.append(bm)
        hd += bm.conj().T @ d @ bm
        sd += bm.conj().T @ q @ bm
    prob = syn.Problem(None, 0.0, np.zeros(3), (0, 0, 0), hkl, kin_ev, g.copy(),
                       np.array(list(v.keys()), dtype=np.int32), np.array(list(v.values()), dtype=np.complex128),
                       np.array([3, 3]), amps, [d, d.copy()], [q, q.copy()])
    geometry = (a * np.eye(3), tau, np.array([radius, radius]))
    return prob, hd, sd, geometry
