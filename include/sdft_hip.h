/* C ABI of the MI355X (gfx950) backend for Sparse-DFT's plane-wave H.psi /
 * S.psi application and block-Davidson step.
 *
 * Conventions: complex values are interleaved FP64 (re, im) pairs, matrices
 * are column-major with explicit leading dimensions counted in complex
 * elements, exactly the host layouts of the reference's
 * src/compute_backend.hpp:13-27 (Complex64 = std::complex<double>).
 * Every function returns a status: 0 success, 1 runtime error
 * (std::runtime_error), 2 invalid argument (std::invalid_argument), 3 overflow
 * (std::overflow_error); the message is in sdft_hip_last_error() of the
 * calling thread. A context belongs to one host thread, as the reference keeps
 * one thread_local Accelerator per thread (src/compute_backend.cpp:62-73).
 * precision: 0 = FP64 arithmetic, 1 = FP32 arithmetic (search precision).
 */
#ifndef SDFT_HIP_H
#define SDFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdft_hip_context sdft_hip_context;
typedef struct sdft_hip_operator sdft_hip_operator;
typedef struct sdft_hip_batch sdft_hip_batch;

#define SDFT_HIP_OK 0
#define SDFT_HIP_RUNTIME_ERROR 1
#define SDFT_HIP_INVALID_ARGUMENT 2
#define SDFT_HIP_OVERFLOW 3

/* Thread-local message of the last failed call ("HIP: ..."). */
const char* sdft_hip_last_error(void);

/* Capability JSON; replaces compiled_compute_capabilities()
 * (src/compute_backend.cpp:110-124) with "hip_compiled": true. No GPU needed. */
int sdft_hip_capabilities(char* buffer, size_t size);

/* Context (HIP stream + rocBLAS/rocSOLVER handle + grow-only buffers); the
 * state a `hip` Accelerator owns (src/compute_cuda.cu:82-85 for CUDA). */
int sdft_hip_context_create(int device, sdft_hip_context** out);
/* Drops the creator's reference. Every operator built on the context holds a
 * reference of its own, so the context's streams, handles and buffers are
 * released when the last of them is destroyed, in whatever order the caller
 * destroys them. */
int sdft_hip_context_destroy(sdft_hip_context* ctx);
/* Fault injection for tests of failure handling (no counterpart in the
 * reference; every kernel still exits normally): the next `count` events of
 * `kind` on this context misbehave on purpose. kind 0: the one-launch
 * tridiagonalisation of a Rayleigh-Ritz solve finds its grid barrier expired
 * (the solve is redone by rocSOLVER zheevd); kind 1: an operator application
 * that produces H x writes a NaN into that output (the Davidson solve using it
 * fails with a runtime error). count 0 clears the kind. */
#define SDFT_HIP_FAULT_BARRIER_EXPIRY 0
#define SDFT_HIP_FAULT_NONFINITE_APPLY 1
int sdft_hip_inject_fault(sdft_hip_context* ctx, int kind, long count);
/* Accelerator::device() (src/compute_backend.hpp:66) */
int sdft_hip_device_name(sdft_hip_context* ctx, char* buffer, size_t size);
int sdft_hip_synchronize(sdft_hip_context* ctx);

/* Kernel clock (measurement; no counterpart in the reference): while enabled,
 * HIP events are recorded on the context's stream around every MFMA complex
 * GEMM kernel and every local-Hamiltonian pipeline launch. Enabling or
 * disabling clears the sums. kind: 0 GEMM FP64, 1 GEMM FP32 arithmetic,
 * 2 local H FP64, 3 local H FP32, 4 GEMM BF16 arithmetic. work: flops (8 m n k per GEMM) or the
 * algorithmic bytes m (2 npw c + 4 N c) + N r + npw (4 + r) per local-H launch.
 * The stats call synchronises the stream. */
#define SDFT_HIP_CLOCK_GEMM_FP64 0
#define SDFT_HIP_CLOCK_GEMM_FP32 1
#define SDFT_HIP_CLOCK_LOCAL_FP64 2
#define SDFT_HIP_CLOCK_LOCAL_FP32 3
#define SDFT_HIP_CLOCK_GEMM_BF16 4 /* BF16-MFMA projector GEMMs of the refinement's correction solves */
int sdft_hip_kernel_clock(sdft_hip_context* ctx, int enable);
int sdft_hip_kernel_clock_stats(sdft_hip_context* ctx, int kind, long* launches, double* milliseconds, double* work);
/* Real MFMA products per complex multiply-add in the complex GEMM kernels of
 * `precision` (0 FP64, 1 FP32): 3 for the Gauss form, 4 otherwise. Measurement
 * only (the conventional complex GEMM flop count 8 m n k corresponds to 4). */
int sdft_hip_gemm_real_products(int precision, int* products);
/* Algorithmic HBM bytes summed over the same launches (GEMM: operands and
 * result once each; local H: the per-launch byte count above). */
int sdft_hip_kernel_clock_bytes(sdft_hip_context* ctx, int kind, double* bytes);

/* Accelerator::gemm / gemm_immutable_left (src/compute_backend.hpp:67-73):
 * C = op(A) op(B) (+ C if add), ops 0 = Normal, 1 = Adjoint, host pointers.
 * left_identity != 0 keeps op(A) resident on the device for that identity
 * (src/compute_backend.hpp:32-41). Never writes outside m x n of ldc. */
int sdft_hip_zgemm(sdft_hip_context* ctx, int aop, int bop, size_t m, size_t n, size_t k, const double* a, size_t lda,
                   const double* b, size_t ldb, double* c, size_t ldc, int add, uint64_t left_identity,
                   int precision);

/* Accelerator::gemm_triple_right (src/compute_backend.hpp:74-80):
 * C_i = A_i B for i = 0..2, normal operands, host pointers. */
int sdft_hip_zgemm_triple_right(sdft_hip_context* ctx, size_t m, size_t n, size_t k, const double* const a[3],
                                size_t lda, const double* b, size_t ldb, double* const c[3], size_t ldc,
                                int precision);

/* Accelerator::local_hamiltonian (src/compute_backend.hpp:81-83):
 * out[:,c] = T in[:,c] + FFT_fwd(V(r) FFT_bwd(in[:,c])) / N for `columns`
 * contiguous columns of npw values; indices into the row-major grid `shape`,
 * potential = V(r) on the grid (N values). operator_identity != 0 keeps
 * shape/indices/potential resident (src/compute_backend.hpp:44-45). */
int sdft_hip_local_hamiltonian(sdft_hip_context* ctx, const int shape[3], const uint64_t* indices, size_t npw,
                               const double* kinetic, const double* potential, const double* input, size_t columns,
                               double* output, uint64_t operator_identity, int precision);

/* Accelerator::paw_pair (src/compute_backend.hpp:88-89; CUDA chain
 * src/compute_cuda.cu:95-128): h = B^H D B in, s = B^H Q B in with B
 * (projectors x npw), D, Q (projectors x projectors, column-major). Only the
 * nonlocal parts are returned. projector_identity guards B, identity guards
 * D and Q. Returns 0 and sets *handled = 1 when computed. */
int sdft_hip_paw_pair(sdft_hip_context* ctx, uint64_t identity, size_t npw, size_t projectors, size_t columns,
                      const double* b, const double* d, const double* q, const double* input, double* h, double* s,
                      uint64_t projector_identity, int precision, int* handled);

/* G+k sphere of src/pw_core.cpp:140-171 for the real-space lattice rows
 * (bohr, reciprocal per src/scf_core.cpp:169-175), cutoff in eV. Writes at
 * most `capacity` entries; *count receives the sphere size. Host only. */
int sdft_hip_enumerate_g_vectors(const double lattice[9], double cutoff_ev, const double kfrac[3],
                                 double ev_per_hartree, long capacity, int* hkl, double* kinetic_ev,
                                 double* cartesian, long* count);

/* Row-major grid index of src/plane_wave_operator.cpp:94-97. Host only. */
int sdft_hip_grid_indices(const int shape[3], size_t n, const int* hkl, uint64_t* indices);

/* Device-resident plane-wave operator: the PlaneWaveOperator constructor of
 * src/plane_wave_operator.cpp:393-516 on the device (index map, V(r) from
 * V(G) by a device FFT, diagonals, stacked projectors, Hermitised D/Q).
 * shape[a] <= 0 selects the reference grid rule (fast_fft_length). */
typedef struct {
  size_t n;                 /* plane waves */
  const int* hkl;           /* n x 3, reference G order */
  const double* kinetic_ev; /* n */
  double ev_per_hartree;
  size_t nv;                /* Fourier potential entries */
  const int* v_hkl;         /* nv x 3 */
  const double* v_values;   /* nv complex */
  size_t atoms;
  const long* rows;         /* projectors per atom */
  const double* amplitudes; /* per atom row-major rows x n complex, concatenated */
  const double* d;          /* per atom row-major rows x rows complex, concatenated */
  const double* q;
  double regularization;
  int shape[3];
} sdft_hip_operator_desc;

int sdft_hip_operator_create(sdft_hip_context* ctx, const sdft_hip_operator_desc* desc, sdft_hip_operator** out);
int sdft_hip_operator_destroy(sdft_hip_operator* op);
/* shape, grid size, projectors, sticks; arrays may be NULL (host outputs). */
int sdft_hip_operator_info(sdft_hip_operator* op, int shape[3], long* count, long* projectors, long* sticks,
                           uint64_t* indices, double* kinetic, double* local_real, double* h_diagonal,
                           double* s_diagonal);
/* HermitianOperator::apply_block_into / apply_pair_into
 * (src/pw_core.hpp:119-130): which 0 = H, 1 = S, 2 = joint H and S.
 * device_pointers = 1: x, hy, sy are device pointers (n x columns). */
int sdft_hip_operator_apply(sdft_hip_operator* op, int which, const double* x, size_t columns, double* hy, double* sy,
                            int precision, int device_pointers);
/* y(:, j) = (H - theta_j S) x(:, j) with one back-projection of the shifted
 * onsite products (no counterpart in the reference: the operator of the
 * refinement's correction equations, exposed for tests). Host pointers, theta
 * a host array of `columns` shifts; precision 0 FP64, 1 FP32, 2 FP32 with the
 * projector GEMMs on BF16 MFMA (FP32 accumulation). */
int sdft_hip_operator_apply_shifted(sdft_hip_operator* op, const double* x, size_t columns, const double* theta,
                                    double* y, int precision);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sdft_hip_operator_apply with device_pointers = 1 that
 * allocates nothing. x, hy, sy are device pointers (n x columns); the stick
 * buffers, projections and GEMM split-K partials are staged in a statically
 * allocated SDFT_HIP_SCRATCH_BYTES array per GPU, held until the process
 * ends. Columns go in passes of the most that fit; an operator one of whose
 * columns needs more staging than the array holds is rejected
 * (SDFT_HIP_OVERFLOW), and so is precision 1 on an operator whose real-space
 * projector search is enabled. Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two contexts or threads at once) race.
 * Returns with the outputs complete, like the allocating entry point, whose
 * results it reproduces bit for bit when the block fits one pass. */
#define SDFT_HIP_SCRATCH_BYTES (128u << 20)
int sdft_hip_operator_apply_noalloc(sdft_hip_operator* op, int which, const double* x, size_t columns, double* hy,
                                    double* sy, int precision);
/* Averages `repetitions` device-resident applications of `columns` columns
 * with HIP events on the context stream (inputs resident in HBM).
 * kernel_ms[0..2]: local-H pipeline, projector GEMMs + onsite, whole apply. */
int sdft_hip_operator_benchmark(sdft_hip_operator* op, int which, size_t columns, int precision, int repetitions,
                                double* kernel_ms);

/* solve_generalized_hermitian_davidson (src/pw_core.cpp:532-1437), device
 * resident. search_precision 1 runs H/S applies and subspace products in FP32
 * until the hand-off to FP64; Rayleigh-Ritz, orthonormalisation and the
 * certificate are FP64. */
typedef struct {
  long roots;
  long max_iterations;
  double tolerance;
  long subspace_factor;
  int search_precision;
  double search_handoff;
  int stagnation_window;
  /* FP32 search hand-off: h >= 1 keeps the max(2, h) x roots lowest Ritz
   * vectors (FP64 images built for those only); 0 re-applies FP64 to the whole
   * subspace. */
  int handoff_restart;
  int verbose;
  int timing;
  /* > 0: the bounded band-group solver with groups of this many roots
   * (SDFT_BAND_GROUP_SOLVER=1 with SDFT_BAND_GROUP_SIZE, src/pw_core.cpp:564-567,
   * src/band_group_solver.cpp:57-211); 0: one block Davidson. */
  long band_group_size;
  /* FP32 search only (no counterpart in the reference, which refuses FP32
   * search): 0 continues with the reference's Davidson iterations in FP64
   * after the hand-off; 1 replaces them by block Jacobi-Davidson refinement
   * steps: FP32 preconditioned-CG correction solves against the fresh FP64
   * residual, FP64 images of the corrections and an FP64 Rayleigh-Ritz of the
   * Ritz block and the corrections. The certificate is the same. */
  int refinement;
  int inner_iterations; /* refinement: CG steps per correction solve, 0 = default */
  /* refinement: guard Ritz vectors carried beside the roots in its block
   * (1 .. roots; 0 = roots / 2) */
  long refinement_guards;
  /* 1: initial and result->eigenvectors are device pointers (n x roots
   * complex, resident in HBM; eigenvectors required), so that a caller whose
   * blocks live on the device moves nothing over PCIe; 0: host pointers, as
   * the reference's call takes them */
  int device_pointers;
  /* refinement: 1 runs the projector GEMMs of the FP32 correction solves on
   * BF16 MFMA with FP32 accumulation (the projectors as a resident BF16 copy;
   * the correction equations are solved to a few digits only), 0 on FP32 MFMA.
   * Rayleigh-Ritz, images and the certificate are unaffected. */
  int bf16_projectors;
  /* band groups: 1 runs the swept variant (SDFT_BAND_GROUP_SWEEP=1,
   * src/band_group_solver.cpp:79-100, 197-208): local solves of at most 3
   * iterations over max_iterations / 2 worth of sweeps, then the Davidson over
   * all roots from the swept block if it is not certified */
  int band_group_sweep;
} sdft_hip_davidson_params;

typedef struct {
  double* eigenvalues;    /* roots, caller-allocated */
  double* eigenvectors;   /* n x roots complex, caller-allocated or NULL */
  double* residual_norms; /* roots */
  long iterations;
  int converged;
  long h_columns_search;
  long h_columns_exact;
  long handoff_iteration;
  double seconds_total, seconds_apply, seconds_gemm, seconds_ritz, seconds_ortho;
  long cholesky_blocks; /* correction blocks normalised by Cholesky QR */
  long rr_cholesky;     /* Rayleigh-Ritz solves that took the Cholesky reduction */
  long rr_stedc;        /* Rayleigh-Ritz tridiagonal solves by divide and conquer (inverse iteration
                           failed its checks, or SDFT_HIP_TRIDIAG_EIGEN=stedc) */
  long rr_barrier_fallbacks; /* eigensolves redone by rocSOLVER zheevd after the one-launch reduction's
                                bounded grid barrier expired */
  /* strict polish (src/strict_residual_polish.hpp:125-197) and root-manifold
   * (:68-123) steps accepted / rejected on fresh FP64 images */
  long polish_accepted, polish_rejected, manifold_accepted, manifold_rejected;
  /* hand-offs from search arithmetic whose first fresh FP64 residual did not
   * confirm the search residual (the role of the reference's counter
   * davidson_coarse_original_residual_rejected) */
  long search_rejected;
  long refinement_steps; /* Jacobi-Davidson refinement steps after the hand-off */
} sdft_hip_davidson_result;

/* PlaneWaveOperator::project_into / projector_metric_into
 * (src/plane_wave_operator.cpp:375-389), host pointers: P = B X (projectors x
 * columns, column-major) and QP = Q P with the per-atom onsite metric in the
 * compensated FP64 order (dense Q for a PWO2 operator). */
int sdft_hip_operator_project(sdft_hip_operator* op, const double* x, size_t columns, double* p, int precision);
int sdft_hip_operator_projector_metric(sdft_hip_operator* op, const double* p, size_t columns, double* qp);

/* Real-space projector search (src/plane_wave_operator.cpp:104-163; the
 * reference's SDFT_PROJECTOR_SPACE=real-search with SDFT_PROJECTOR_RADIUS_FACTOR).
 * factor in [1,2] enables the approximate operator for FP32 search
 * applications (0 disables). Per atom, with geometry (lattice: 3 rows of cell
 * vectors in bohr, fractional: atoms x 3, radius: atoms, as the reference's
 * atom_geometry) and radius > 0, the grid points within factor * radius of the
 * atom (minimum image) are kept; otherwise (geometry NULL, as for snapshots)
 * those where the projector amplitude exceeds 10^(-2-5(factor-1)) of its
 * maximum. *points: retained points; *retained: kept fraction of the
 * projector power (either may be NULL). */
int sdft_hip_operator_real_space_search(sdft_hip_operator* op, double factor, const double* lattice,
                                        const double* fractional, const double* radius, size_t* points,
                                        double* retained);
/* H (which 0) or S (which 1) through the real-space projectors, host pointers;
 * precision 0 FP64, 1 FP32. */
int sdft_hip_operator_apply_real_space(sdft_hip_operator* op, int which, const double* x, size_t columns, double* y,
                                       int precision);
/* Retained points as natural grid indices, atom by atom (total entries), and
 * the count per atom. */
int sdft_hip_operator_real_space_points(sdft_hip_operator* op, long* natural, long* per_atom);

/* plane_wave_density_fft (src/density_functionals.cpp:406-460): Fourier
 * coefficients (nsupport complex) at the Miller triples support_hkl of
 * rho(r) = sum_b occupations[b] / volume |psi_b(r)|^2, psi_b the band
 * coefficients (host, n x bands complex, the operator's basis order); the
 * transforms run on the smallest device FFT grid that is alias free. */
int sdft_hip_operator_density(sdft_hip_operator* op, const double* coefficients, size_t bands,
                              const double* occupations, double volume, size_t nsupport, const int* support_hkl,
                              double* out);

/* Solver snapshots SDFTPWO1/2/3 (src/plane_wave_operator.cpp:520-653): the
 * operator arrays (indices, kinetic, H/S diagonals, V(r), B and the onsite
 * matrices in atom-block, dense or folded form) and a Davidson request
 * (roots, iterations, tolerance, warm block), as the replay harness of
 * src/solver_replay_cli.cpp consumes them. */
typedef struct {
  int format; /* 1, 2 or 3 (SDFTPWO1/2/3) */
  size_t n, count, shape[3], projectors, atoms;
  size_t roots, iterations, initial_columns;
  double regularization, tolerance;
} sdft_hip_snapshot_info;
/* Reads and validates the whole file as the reference loader does; no GPU. */
int sdft_hip_snapshot_info_read(const char* path, sdft_hip_snapshot_info* info);
/* Device operator from a snapshot; the stored warm block (n x initial_columns
 * complex) is copied to initial when not NULL. */
int sdft_hip_snapshot_load(sdft_hip_context* ctx, const char* path, double* initial, sdft_hip_operator** out);
/* Writes the operator in its snapshot form (PWO3 for atom blocks) with a
 * Davidson request; initial is n x roots complex or NULL. Refuses to
 * overwrite an existing file. */
int sdft_hip_snapshot_save(sdft_hip_operator* op, const char* path, size_t roots, size_t iterations,
                           double tolerance, const double* initial);

/* Dense Hermitian eigensolver of the Rayleigh-Ritz step (no counterpart in the
 * reference, whose Ritz solve is LAPACK zhegvd; exposed for tests): all n
 * eigenvalues ascending of the host matrix a (n x n column-major complex,
 * Hermitised on the device), eigenvectors of the lowest k into vectors
 * (n x k, may be NULL when k = 0). Orders above the one-launch reduction's
 * limit take rocSOLVER zheevd. */
int sdft_hip_hermitian_eigen(sdft_hip_context* ctx, size_t n, const double* a, size_t k, double* values,
                             double* vectors,
                             int* fused /* out, may be NULL: 1 = one-launch tridiagonalisation and inverse
                                           iteration, 2 = one-launch tridiagonalisation and divide and
                                           conquer, 0 = zheevd (also after the reduction's grid barrier
                                           expired) */);

/* Davidson seed block of the device solver (no counterpart as a function in
 * the reference; src/pw_core.cpp:1027-1112 builds it inline; exposed for
 * tests): for `roots` roots at `tolerance`, the warm columns (initial, n x
 * roots, perturbed) when 2 roots < n and initial is not NULL, then 2 roots
 * guard columns; block (host, n x (3 roots) complex at most) receives
 * *columns columns. */
int sdft_hip_davidson_seeds(sdft_hip_operator* op, size_t roots, double tolerance, const double* initial,
                            double* block, size_t* columns);

/* Symmetric tridiagonal stage of that eigensolver (no counterpart in the
 * reference; exposed for tests): all n eigenvalues of T (diagonal d, off-
 * diagonal e of length n - 1) by Sturm-count multisection, ascending to within
 * the bisection tolerance, and the eigenvectors of the lowest k by inverse
 * iteration into vectors (n x k real, column-major), before the symmetric
 * orthonormalisation. *status is 0, or 1 when an iteration did not stay
 * finite. Host pointers; n <= 2048. */
int sdft_hip_tridiagonal_eigen(sdft_hip_context* ctx, size_t n, const double* d, const double* e, size_t k,
                               double* values, double* vectors, int* status);

int sdft_hip_davidson(sdft_hip_operator* op, const sdft_hip_davidson_params* params, const double* initial,
                      sdft_hip_davidson_result* result);

/* Device memory held by an operator: its arrays, work buffers and the
 * Davidson buffers it keeps for its next solve (no counterpart in the
 * reference; the serving loop admits requests against an HBM budget). */
int sdft_hip_operator_device_bytes(sdft_hip_operator* op, size_t* bytes);

/* Batched k-point application (no counterpart as one call in the reference,
 * whose k-points are independent PlaneWaveOperator applications on OpenMP
 * teams, src/scf_core.cpp:3051-3061): H psi and S psi of `count` device
 * operators of one context that share an FFT grid shape (the k-points of one
 * crystal; uneven plane-wave counts, k-dependent projectors) for `columns`
 * columns each, as ONE persistent kernel launch over (operator, column) work
 * items, captured once in a HIP graph and replayed. The launch shape (threads,
 * xy-planes per LDS round, workgroups) is chosen at creation by timing the
 * candidates on the device, and cached per problem shape for the process.
 * Operators need atom-block projectors; the operators must outlive the
 * batch. precision 0 FP64, 1 FP32 arithmetic. */
int sdft_hip_batch_create(sdft_hip_context* ctx, size_t count, sdft_hip_operator* const* ops, size_t columns,
                          int precision, sdft_hip_batch** out);
int sdft_hip_batch_destroy(sdft_hip_batch* batch);
/* offsets[k] (count + 1 entries, complex elements): start of operator k's
 * n_k x columns column-major block in the concatenated input and outputs. */
int sdft_hip_batch_offsets(sdft_hip_batch* batch, long* offsets);
/* Host pointers, concatenated blocks: uploads x, replays the graph, returns
 * H x and S x. */
int sdft_hip_batch_apply(sdft_hip_batch* batch, const double* x, double* hy, double* sy);
/* Mean device time of one graph replay over `replays` replays (HIP events,
 * inputs resident). */
int sdft_hip_batch_benchmark(sdft_hip_batch* batch, int replays, double* milliseconds);
/* The tuned launch shape and its measured time; *from_cache = 1 when it came
 * from an earlier batch of the same shape. Any output may be NULL. */
int sdft_hip_batch_plan(sdft_hip_batch* batch, int* threads, int* planes, int* groups, double* tuned_ms,
                        int* from_cache);

/* One strict certificate-path step on a given block (the reference's
 * strict_residual_polish_step and strict_root_manifold_step primitives,
 * src/strict_residual_polish.hpp:68-197, as the device Davidson runs them;
 * exposed for tests): the block x (host, n x roots, in/out) gets a fresh FP64
 * certificate (S normalisation, H/S applications, energies, residuals), then
 * kind 0 tries the tangent polish along `direction` (host, n x roots), kind 1
 * the re-Ritz inside the block. A step is kept only when the fresh residual
 * improves and the S constraints hold; *accepted says whether it was. On
 * return x, energies and residual_norms describe the kept block. A nonfinite
 * direction is an invalid argument. */
int sdft_hip_strict_step(sdft_hip_operator* op, int kind, size_t roots, double tolerance, double* x,
                         const double* direction, double* energies, double* residual_norms, int* accepted);

/* Compensated per-atom onsite products W = [D_a P_a | Q_a P_a] on the device
 * (src/paw_compensated_product.hpp:110-135 order), host pointers; for tests. */
int sdft_hip_onsite_blocks(sdft_hip_context* ctx, size_t atoms, const long* rows, const double* d, const double* q,
                           const double* p, size_t columns, double* w, int precision);

#ifdef __cplusplus
}
#endif
#endif
