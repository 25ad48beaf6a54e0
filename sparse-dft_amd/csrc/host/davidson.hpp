// Device-resident block Davidson in the S metric, restating
// solve_generalized_hermitian_davidson (src/pw_core.cpp:532-1437) with the
// reference's default controls, strict polish and root-manifold refinement
// (src/strict_residual_polish.hpp:47-197). Every plane-wave-sized block stays
// in HBM; only residual norms, Ritz values and small control scalars cross to
// the host.
#pragma once
#include <complex>
#include <vector>

#include "device.hpp"

namespace sdft_hip {

struct DavidsonParams {
  long roots = 0;
  long max_iterations = 256;
  double tolerance = 1e-12;
  long subspace_factor = 8;       // capacity = min(n, max(factor*roots, 64)) (src/pw_core.cpp:735-746)
  int search_fp32 = 0;            // 1: FP32 H/S applies and subspace GEMMs until hand-off
  double search_handoff = 1e-8;   // max residual that hands the search to FP64 (src/pw_core.cpp:1184)
  int stagnation_window = 3;      // FP32 only: hand off when max residual stops halving over this many iterations
  int handoff_restart = 1;        // FP32 only: h >= 1 thick-restarts to max(2, h) x roots Ritz vectors at the
                                  // hand-off and builds FP64 images for those; 0 re-applies FP64 to the whole
                                  // subspace
  int refinement = 0;             // FP32 only: 0 continues with the reference's Davidson iterations in FP64
                                  // after the hand-off; 1 runs block Jacobi-Davidson refinement steps (FP32
                                  // correction solves, FP64 images and Rayleigh-Ritz)
  int inner_iterations = 0;       // refinement: preconditioned CG steps per correction solve (0: default)
  long refinement_guards = 0;     // refinement: guard Ritz vectors beside the roots (0: roots / 2)
  int verbose = 0;
  int timing = 0;                 // synchronise phases and accumulate per-phase times
  long band_group_size = 0;       // > 0: bounded band-group solver (src/band_group_solver.cpp:57-211)
  int band_group_sweep = 0;       // band groups: 1 = the swept variant (SDFT_BAND_GROUP_SWEEP, :79-100, 197-208)
  // Local solve of one band group (src/band_group_solver.cpp:84-100): capacity
  // override, completed bands to keep out of the search space (S-orthonormal,
  // device n x excluded_count), and no strict polish (src/pw_core.cpp:1273).
  long capacity = 0;
  const double* excluded = nullptr;
  long excluded_count = 0;
  int band_group_local = 0;
  // host n x roots destination of the eigenvectors: copied there directly
  // (DavidsonOutput::eigenvectors stays empty); nullptr keeps them in the output
  std::complex<double>* vectors_out = nullptr;
  // 1: the initial block and vectors_out are device pointers (HBM resident,
  // n x roots complex); vectors_out is then required
  int device_pointers = 0;
  // refinement: projector GEMMs of the correction solves on BF16 MFMA (FP32 accumulation)
  int bf16_projectors = 0;
};

struct DavidsonStats {
  long iterations = 0;
  bool converged = false;
  long h_columns_search = 0, h_columns_exact = 0;
  long handoff_iteration = -1;
  double seconds_total = 0, seconds_apply = 0, seconds_gemm = 0, seconds_ritz = 0, seconds_ortho = 0;
  long cholesky_blocks = 0;  // correction blocks normalised by Cholesky QR
  long rr_cholesky = 0;      // Rayleigh-Ritz solves that needed the Cholesky reduction
  long rr_stedc = 0;         // tridiagonal solves by divide and conquer instead of inverse iteration
  long rr_barrier_fallbacks = 0;  // eigensolves redone by zheevd after the reduction's grid barrier expired
  // strict polish (src/strict_residual_polish.hpp:125-197) and root-manifold
  // (:68-123) steps whose trial was accepted / rejected on fresh FP64 images
  long polish_accepted = 0, polish_rejected = 0, manifold_accepted = 0, manifold_rejected = 0;
  // hand-offs from the search arithmetic whose first fresh FP64 residual did
  // not confirm the search residual (above max(8 tol, 10 x the search's))
  long search_rejected = 0;
  long refinement_steps = 0;  // Jacobi-Davidson refinement steps after the hand-off
  std::vector<double> max_residual_history;
};

struct DavidsonOutput {
  std::vector<double> eigenvalues, residual_norms;
  std::vector<std::complex<double>> eigenvectors;  // column-major n x roots
  DavidsonStats stats;
};

DavidsonOutput davidson(DeviceOperator& op, const DavidsonParams& params,
                        const std::complex<double>* initial /* n x roots or nullptr */);

struct StrictStepOutput {
  bool accepted = false;
  std::vector<std::complex<double>> x;  // n x roots, the block after the step
  std::vector<double> energies, residual_norms;
};
// One strict step on the block x (host, n x roots) from its fresh FP64
// certificate: kind 0 the tangent polish along `direction` (host, n x roots),
// kind 1 the root-manifold re-Ritz (src/strict_residual_polish.hpp:68-197).
StrictStepOutput strict_step(DeviceOperator& op, int kind, long roots, double tolerance, const std::complex<double>* x,
                             const std::complex<double>* direction);

// Device bytes of the Davidson buffers the operator keeps for its next solve.
std::size_t davidson_cache_bytes(DeviceOperator& op);
// Canonical order of the basis by H/S diagonal (src/pw_core.cpp:1016-1021).
std::vector<long> davidson_order(const DeviceOperator& op);
// The Davidson seed block of `roots` roots at tolerance `tol` (warm start
// from `initial_device`, n x roots on the device, or none), generated on the
// device into `block` (n x columns, ld n); returns the column count.
long davidson_seed_block(DeviceOperator& op, long roots, double tol, const double* initial_device, double* block,
                         hipStream_t s);
long davidson_seed_columns(long n, long roots, bool warm);
// The bounded band-group solver: local Davidson solves of band_group_size
// roots with the completed bands excluded, then a global roots x roots
// Rayleigh-Ritz rotation certified on fresh FP64 H and S (two sweeps).
DavidsonOutput davidson_band_groups(DeviceOperator& op, const DavidsonParams& params,
                                    const std::complex<double>* initial);

}  // namespace sdft_hip
