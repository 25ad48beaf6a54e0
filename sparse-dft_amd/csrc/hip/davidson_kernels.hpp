// Launchers of the Davidson column reductions and elementwise steps.
#pragma once
#include <hip/hip_runtime_api.h>

namespace sdft_hip {

// out[j] = sum_i Re(a_ij) Re(b_ij) + Im(a_ij) Im(b_ij), compensated
hipError_t launch_real_dots(const double* a, long lda, const double* b, long ldb, long n, long cols, double* out,
                            hipStream_t s);
// out(i, j) = sum_k conj(a_ki) b_kj, compensated, for i < ca, j < cb
hipError_t launch_complex_dots(const double* a, long lda, long ca, const double* b, long ldb, long cb, long n,
                               double* out, long ldo, hipStream_t s);
hipError_t launch_col_norms(const double* x, long ldx, long n, long cols, double* out, hipStream_t s);
// Davidson seed block, n x (warm_cols + guards) complex (ld ldo): warm columns
// from init (n x warm_cols, ld ldi) perturbed by pert |init_j| times the
// normalised splitmix64 vector of the column, then guard columns 0.05 b / |b| + e_order[g]
// (src/pw_core.cpp:1027-1112; the stream of uniform() calls of the host recipe).
hipError_t launch_davidson_seeds(long n, long warm_cols, const double* init, long ldi, double pert, long guards,
                                 const int* order, double* out, long ldo, hipStream_t s);
hipError_t launch_kinetic_weights(const double* x, long ldx, long n, long cols, const double* kinetic, double* out,
                                  hipStream_t s);
hipError_t launch_precondition(const double* r, long ldr, const int* src, long cols, const double* wk,
                               const double* kinetic, long n, double* out, long ldo, hipStream_t s);
hipError_t launch_ritz_residual(const double* hx, const double* sx, long ld, long n, long cols, const double* values,
                                double* r, hipStream_t s);
hipError_t launch_fma_residual(const double* hx, const double* sx, long ld, long n, long cols, const double* values,
                               double* r, hipStream_t s);
hipError_t launch_scale_columns(double* x, long ld, long n, long cols, const double* scales, hipStream_t s);
// values = energy / metric, roots = sqrt(metric), per column
hipError_t launch_rayleigh_quotients(const double* metric, const double* energy, long cols, double* values,
                                     double* roots, hipStream_t s);
// d = d where positive and finite, else 1
hipError_t launch_guard_divisors(double* d, long cols, hipStream_t s);
hipError_t launch_divide_columns(double* x, long ld, long n, long cols, const double* divisors, hipStream_t s);
hipError_t launch_polish_update(const double* x, const double* d, long ld, long n, long cols, const double* alpha,
                                double* trial, hipStream_t s);
hipError_t launch_hermitise(const double* m, long ldm, long size, double* out, long ldo, hipStream_t s);
// C(0:k, 0:k) = diag(theta) (identity for null theta), C(k:m, 0:k) = C(0:k, k:m)^H
hipError_t launch_ritz_border(double* c, long ld, long k, long m, const double* theta, hipStream_t s);
hipError_t launch_inv_sqrt_columns(double* u, long ld, long rows, long cols, const double* lambda, hipStream_t s);
// Block preconditioned CG of the Jacobi-Davidson correction equations (one
// CG per column, scalars in device arrays):
// out(:, c) = in(:, src[c])
hipError_t launch_gather_columns(const double* in, long ldi, const int* src, long n, long cols, double* out, long ldo,
                                 hipStream_t s);
// z(:, c) = K(kinetic, wk[src[c]]) r(:, c), K the kinetic preconditioner (positive)
hipError_t launch_jd_precondition(const double* r, long ldr, const int* src, long cols, const double* wk,
                                  const double* kinetic, long n, double* z, long ldz, hipStream_t s);
// a = rho / pq per column (0 unless pq > 0 and finite); t += a p; r -= a q
hipError_t launch_cg_update(double* t, double* r, const double* p, const double* q, long ld, long n, long cols,
                            const double* rho, const double* pq, hipStream_t s);
// b = rho_new / rho_old per column (0 unless rho_old > 0); p = z + b p
// fused CG step pass: alpha update of t and r, z = M r, partial sums of Re (r, z)
// (cg_step_parts() per column, part: cols x cg_step_parts())
hipError_t launch_cg_step(double* t, double* r, const double* p, const double* q, double* z, long ld, long n,
                          long cols, const double* rho, const double* pq, const int* src, const double* wk,
                          const double* kinetic, double* part, hipStream_t s);
// rho_new = sum of the partials in order, p = z + (rho_new / rho_old) p
hipError_t launch_cg_direction_parts(double* p, const double* z, long ld, long n, long cols, const double* part,
                                     const double* rho_old, double* rho_new, hipStream_t s);
int cg_step_parts();
hipError_t launch_cg_direction(double* p, const double* z, long ld, long n, long cols, const double* rho_new,
                               const double* rho_old, hipStream_t s);
hipError_t launch_finite_check(const double* x, long ld, long n, long cols, int* bad, hipStream_t s);
// max |B - I| (as the bits of a non-negative double, max-folded into *max_bits,
// which the caller zeroes); E = B - I written when e is not null
hipError_t launch_identity_defect(const double* b, long ldb, long m, double* e, long lde,
                                  unsigned long long* max_bits, hipStream_t s);
// out = I - E/2 + 3 E2/8
hipError_t launch_series_metric(const double* e, const double* e2, long ld, long m, double* out, long ldo,
                                hipStream_t s);
// out[0:m] = Re diag(a), out[m:2m] = Re diag(b)
hipError_t launch_diagonals(const double* a, long lda, const double* b, long ldb, long m, double* out,
                            hipStream_t s);
hipError_t launch_zero_lower(double* a, long lda, long m, hipStream_t s);

}  // namespace sdft_hip
