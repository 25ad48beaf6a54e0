"""Writes tests/golden/large_golden.npz: eigenvalues of the full-size
workloads from the REFERENCE's own compiled Davidson
(solve_generalized_hermitian_davidson, src/pw_core.cpp:532-1437, built by
oracle/build_ref.sh into oracle/_ref/libsdft_ref.so) driving the CPU oracle's
plane-wave operator through callbacks. Run in the development container only
(it needs /root/reference and takes tens of minutes on 8 cores):

    python tests/golden/make_golden_large.py [si64|slab0|slab1 ...]

Cases (data only; the inputs are regenerated from the seeded synthetic
recipe by the tests):
  si64   Si64 Gamma (2x2x2 diamond, Ecut 15 Ha, 72^3), 160 roots from the
         seeded warm block of the benchmark;
  slab0  two-species slab at k = (0,0,0) (48x48x96), 96 roots;
  slab1  the same slab at k = (1/2,1/2,0), 96 roots.
Each entry keeps the eigenvalues, the iteration count and the largest
certified residual the reference reported. Existing entries of other cases
are kept when only some cases are recomputed.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "sparse-dft_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import pyoracle  # noqa: E402
import pyref  # noqa: E402
from sdft_amd import synthetic as syn  # noqa: E402

OUT = os.path.join(HERE, "large_golden.npz")

# name: (config, k-point, roots, initial-block config seed)
CASES = {
    "si64": (1, (0.0, 0.0, 0.0), 160, 1),
    "slab0": (3, (0.0, 0.0, 0.0), 96, 3),
    "slab1": (3, (0.5, 0.5, 0.0), 96, 3),
}


def run_case(name):
    config, kfrac, roots, seed = CASES[name]
    prob = syn.config_problem(config, kfrac=kfrac, enumerate_fn=pyoracle.enumerate_g)
    op = pyoracle.Operator(prob)
    init = syn.initial_psi(prob, roots, seed)
    t0 = time.perf_counter()
    r = pyref.davidson(op, roots, 256, 1e-12, init)
    dt = time.perf_counter() - t0
    if not r["converged"]:
        raise RuntimeError(f"{name}: the reference Davidson did not certify")
    # fresh FP64 residuals of the reference's vectors on the oracle operator
    hx, sx = op.apply(r["eigenvectors"], "pair")
    fresh = np.linalg.norm(hx - sx * r["eigenvalues"], axis=0)
    print(f"{name}: n={prob.n} roots={roots} iterations={r['iterations']} "
          f"max residual {r['residual_norms'].max():.3e} fresh {fresh.max():.3e} ({dt:.0f} s)", flush=True)
    return {f"{name}_eigenvalues": r["eigenvalues"], f"{name}_iterations": np.array(r["iterations"]),
            f"{name}_max_residual": np.array(r["residual_norms"].max()), f"{name}_fresh_residual": np.array(fresh.max()),
            f"{name}_n": np.array(prob.n)}


def main():
    assert pyref.available(), "build oracle/_ref first (bash oracle/build_ref.sh)"
    blas = pyoracle.numpy_openblas_path()
    if blas:
        pyoracle.bind_blas(blas)
    names = sys.argv[1:] or list(CASES)
    out = dict(np.load(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        out.update(run_case(name))
        np.savez_compressed(OUT, **out)
    print("wrote", OUT, sorted(out))


if __name__ == "__main__":
    main()
