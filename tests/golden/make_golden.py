"""Writes tests/golden/reference_golden.npz from the REFERENCE's own compiled
sources (oracle/_ref/libsdft_ref.so, built by oracle/build_ref.sh from
/root/reference). Run in the development container only:

    python tests/golden/make_golden.py

Fixtures (data only): ordered G-spheres (full list for Si8 Gamma, SHA-256 of
the ordered (h,k,l) int32 + kinetic FP64 bytes and counts otherwise), the
reference Davidson's eigenvalues/iteration count on the Si8 synthetic operator
(driven through the oracle operator), compensated-product vectors and
kinetic-preconditioner values (bitwise).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "sparse-dft_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import pyoracle  # noqa: E402
import pyref  # noqa: E402
from sdft_amd import synthetic as syn  # noqa: E402


def reciprocal_rows(lattice):
    out = np.zeros(9)
    lat = np.ascontiguousarray(lattice, dtype=np.float64)
    pyoracle._check(pyoracle.lib().orc_reciprocal(pyoracle._p(lat), pyoracle._p(out)))
    return out.reshape(3, 3)


def sphere_digest(hkl, kin):
    return hashlib.sha256(np.ascontiguousarray(hkl, dtype=np.int32).tobytes() +
                          np.ascontiguousarray(kin, dtype=np.float64).tobytes()).hexdigest()


def main():
    assert pyref.available(), "build oracle/_ref first (bash oracle/build_ref.sh)"
    ev = syn.HARTREE_EV
    out = {}
    cases = [("si8_gamma", syn.diamond_cell(10.26).lattice, 10.0, [(0, 0, 0)]),
             ("si8_mp444", syn.diamond_cell(10.26).lattice, 10.0, [tuple(k) for k in syn.mp_kpoints(4)]),
             ("si64_gamma", syn.diamond_cell(10.26, (2, 2, 2)).lattice, 15.0, [(0, 0, 0)]),
             ("slab_2k", syn.slab_cell(3).lattice, 20.0, [(0, 0, 0), (.5, .5, 0)])]
    for name, lat, ecut, ks in cases:
        b = reciprocal_rows(lat)
        digests, counts = [], []
        for k in ks:
            hkl, kin = pyref.enumerate_g(b, ecut * ev + 1e-8, k)
            digests.append(sphere_digest(hkl, kin))
            counts.append(len(hkl))
            if name == "si8_gamma":
                out["si8_gamma_hkl"] = hkl
                out["si8_gamma_kinetic_ev"] = kin
        out[f"{name}_kpoints"] = np.array(ks, dtype=np.float64)
        out[f"{name}_sha256"] = np.array(digests)
        out[f"{name}_counts"] = np.array(counts)
    # reference Davidson on the Si8 synthetic operator (config 0)
    prob = syn.config_problem(0, enumerate_fn=pyoracle.enumerate_g)
    op = pyoracle.Operator(prob)
    init = syn.initial_psi(prob, 32, 0)
    r = pyref.davidson(op, 16, 256, 1e-12, init)
    assert r["converged"]
    out["si8_davidson16_eigenvalues"] = r["eigenvalues"]
    out["si8_davidson16_iterations"] = np.array(r["iterations"])
    out["si8_davidson16_max_residual"] = np.array(r["residual_norms"].max())
    r8 = pyref.davidson(op, 8, 256, 1e-12, None)
    assert r8["converged"]
    out["si8_cold8_eigenvalues"] = r8["eigenvalues"]
    out["si8_cold8_iterations"] = np.array(r8["iterations"])
    # reference bounded band-group solver (src/band_group_solver.cpp), groups of 8 and 5
    init16 = syn.initial_psi(prob, 16, 0)
    for g in (8, 5):
        rg = pyref.davidson(op, 16, 256, 1e-12, init16, band_group_size=g)
        assert rg["converged"]
        out[f"si8_bandgroups{g}_eigenvalues"] = rg["eigenvalues"]
        out[f"si8_bandgroups{g}_max_residual"] = np.array(rg["residual_norms"].max())
    # compensated product vectors (src/paw_compensated_product.hpp)
    rng = np.random.default_rng(20261202)
    a = rng.standard_normal((18, 18)) + 1j * rng.standard_normal((18, 18))
    bb = rng.standard_normal((18, 7)) + 1j * rng.standard_normal((18, 7))
    bb[:, 0] = np.linalg.solve(a, np.full(18, 3e-9 + 1e-9j))  # heavy cancellation column
    out["comp_a"], out["comp_b"] = a, bb
    out["comp_normal"] = pyref.compensated_product(0, a, bb)
    out["comp_adjoint"] = pyref.compensated_product(1, a, bb)
    # kinetic preconditioner samples (src/pw_core.cpp:33-47)
    t = np.array([0., 1e-3, .05, .3, 1., 2.5, 10., 100., 1e5, 1e100])
    e = np.array([0., .01, .2, 1., 100.])
    out["kp_t"], out["kp_e"] = t, e
    out["kp_values"] = np.array([[pyref.kinetic_preconditioner(x, y) for y in e] for x in t])
    np.savez_compressed(os.path.join(HERE, "reference_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "reference_golden.npz"), {k: np.asarray(v).shape for k, v in out.items()})


if __name__ == "__main__":
    main()
