"""Writes tests/golden/band_sweep_golden.npz from the REFERENCE's own compiled
sources (oracle/_ref/libsdft_ref.so, built by oracle/build_ref.sh from
/root/reference): the swept bounded band-group solver
(src/band_group_solver.cpp with SDFT_BAND_GROUP_SWEEP=1, read by the reference
at call time) on the Si8 synthetic operator, 16 roots from the warm block,
groups of 8 and 5. Run in the development container only:

    python tests/golden/make_golden_sweep.py

Fixtures (data only): eigenvalues, the largest residual and the iteration
count of each solve.
"""
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


def main():
    prob = syn.config_problem(0, enumerate_fn=pyoracle.enumerate_g)
    op = pyoracle.Operator(prob)
    init16 = syn.initial_psi(prob, 16, 0)
    out = {}
    os.environ["SDFT_BAND_GROUP_SWEEP"] = "1"
    try:
        for g in (8, 5):
            r = pyref.davidson(op, 16, 256, 1e-12, init16, band_group_size=g)
            assert r["converged"]
            out[f"si8_bandsweep{g}_eigenvalues"] = r["eigenvalues"]
            out[f"si8_bandsweep{g}_max_residual"] = np.array(r["residual_norms"].max())
            out[f"si8_bandsweep{g}_iterations"] = np.array(r["iterations"])
    finally:
        del os.environ["SDFT_BAND_GROUP_SWEEP"]
    np.savez_compressed(os.path.join(HERE, "band_sweep_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "band_sweep_golden.npz"), {k: np.asarray(v).tolist() for k, v in out.items()
                                                                if "eigenvalues" not in k})


if __name__ == "__main__":
    main()
