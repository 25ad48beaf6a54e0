// Column-wise reductions and elementwise steps of the block Davidson
// (src/pw_core.cpp:532-1437) on device-resident blocks. Blocks are column-major
// complex FP64 with leading dimension ld (complex elements).
#include <hip/hip_runtime.h>

#include "davidson_kernels.hpp"

namespace sdft_hip {

namespace {
constexpr int TPB = 256;
// column reductions: one workgroup of RPB threads per column (the columns of a
// block number a few hundred at most, so more threads per column is what puts
// enough loads in flight), two entries per thread and trip
constexpr int RPB = 1024;

// error-free transformation a + b = s + e
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
// double-double accumulate (hi, lo) += (h2, l2)
__device__ __forceinline__ void dd_add(double& hi, double& lo, double h2, double l2) {
#pragma clang fp contract(off)
  double s, e;
  two_sum(hi, h2, s, e);
  e += lo + l2;
  two_sum(s, e, hi, lo);
}

template <int NT = TPB>
__device__ __forceinline__ void block_reduce_dd(double& hi, double& lo, double* sh, double* sl) {
  const int t = threadIdx.x;
  sh[t] = hi;
  sl[t] = lo;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) {
      double h = sh[t], l = sl[t];
      dd_add(h, l, sh[t + s], sl[t + s]);
      sh[t] = h;
      sl[t] = l;
    }
    __syncthreads();
  }
  hi = sh[0];
  lo = sl[0];
}

// one block per column: sum_i re(a_i) re(b_i) + im(a_i) im(b_i), compensated
// (FMA-exact products accumulated in double-double; fixed reduction order)
__global__ __launch_bounds__(RPB) void k_real_dots(const double* __restrict__ a, long lda,
                                                   const double* __restrict__ b, long ldb, long n,
                                                   double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double sh[RPB], sl[RPB];
  const long col = blockIdx.x;
  const double2* ac = reinterpret_cast<const double2*>(a + 2 * col * lda);
  const double2* bc = reinterpret_cast<const double2*>(b + 2 * col * ldb);
  double hi = 0, lo = 0;
  long i = threadIdx.x;
  for (; i + RPB < n; i += 2 * RPB) {
    const double2 x0 = ac[i], y0 = bc[i], x1 = ac[i + RPB], y1 = bc[i + RPB];
    const double p1 = x0.x * y0.x, p2 = x0.y * y0.y, p3 = x1.x * y1.x, p4 = x1.y * y1.y;
    dd_add(hi, lo, p1, __builtin_fma(x0.x, y0.x, -p1));
    dd_add(hi, lo, p2, __builtin_fma(x0.y, y0.y, -p2));
    dd_add(hi, lo, p3, __builtin_fma(x1.x, y1.x, -p3));
    dd_add(hi, lo, p4, __builtin_fma(x1.y, y1.y, -p4));
  }
  if (i < n) {
    const double2 x0 = ac[i], y0 = bc[i];
    const double p1 = x0.x * y0.x, p2 = x0.y * y0.y;
    dd_add(hi, lo, p1, __builtin_fma(x0.x, y0.x, -p1));
    dd_add(hi, lo, p2, __builtin_fma(x0.y, y0.y, -p2));
  }
  block_reduce_dd<RPB>(hi, lo, sh, sl);
  if (threadIdx.x == 0) out[col] = hi + lo;
}

// one block per (i, j): <a_i, b_j> = sum conj(a) b, compensated
__global__ __launch_bounds__(TPB) void k_complex_dots(const double* __restrict__ a, long lda,
                                                      const double* __restrict__ b, long ldb, long n, int ca,
                                                      double* __restrict__ out, long ldo) {
#pragma clang fp contract(off)
  __shared__ double sh[TPB], sl[TPB];
  const long i = blockIdx.x, j = blockIdx.y;
  const double* ac = a + 2 * i * lda;
  const double* bc = b + 2 * j * ldb;
  double rh = 0, rl = 0, ih = 0, il = 0;
  for (long k = threadIdx.x; k < n; k += TPB) {
    const double ar = ac[2 * k], ai = ac[2 * k + 1], br = bc[2 * k], bi = bc[2 * k + 1];
    const double p1 = ar * br, p2 = ai * bi, p3 = ar * bi, p4 = ai * br;
    dd_add(rh, rl, p1, __builtin_fma(ar, br, -p1));
    dd_add(rh, rl, p2, __builtin_fma(ai, bi, -p2));
    dd_add(ih, il, p3, __builtin_fma(ar, bi, -p3));
    dd_add(ih, il, -p4, -__builtin_fma(ai, br, -p4));
  }
  block_reduce_dd(rh, rl, sh, sl);
  __syncthreads();
  block_reduce_dd(ih, il, sh, sl);
  if (threadIdx.x == 0) {
    double* o = out + 2 * (i + j * ldo);
    o[0] = rh + rl;
    o[1] = ih + il;
  }
  (void)ca;
}

// per column: sqrt(sum |x|^2)
__global__ __launch_bounds__(RPB) void k_col_norms(const double* __restrict__ x, long ldx, long n,
                                                   double* __restrict__ out) {
  __shared__ double sh[RPB];
  const long col = blockIdx.x;
  const double2* xc = reinterpret_cast<const double2*>(x + 2 * col * ldx);
  double s = 0, s1 = 0;
  long i = threadIdx.x;
  for (; i + RPB < n; i += 2 * RPB) {
    const double2 v0 = xc[i], v1 = xc[i + RPB];
    s += v0.x * v0.x + v0.y * v0.y;
    s1 += v1.x * v1.x + v1.y * v1.y;
  }
  if (i < n) s += xc[i].x * xc[i].x + xc[i].y * xc[i].y;
  s += s1;
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = RPB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[col] = sqrt(sh[0]);
}

// per column: sum_i |x_i|^2 T_i
__global__ __launch_bounds__(RPB) void k_kinetic_weights(const double* __restrict__ x, long ldx, long n,
                                                         const double* __restrict__ kinetic,
                                                         double* __restrict__ out) {
  __shared__ double sh[RPB];
  const long col = blockIdx.x;
  const double2* xc = reinterpret_cast<const double2*>(x + 2 * col * ldx);
  double s = 0, s1 = 0;
  long i = threadIdx.x;
  for (; i + RPB < n; i += 2 * RPB) {
    const double2 v0 = xc[i], v1 = xc[i + RPB];
    s += (v0.x * v0.x + v0.y * v0.y) * kinetic[i];
    s1 += (v1.x * v1.x + v1.y * v1.y) * kinetic[i + RPB];
  }
  if (i < n) s += (xc[i].x * xc[i].x + xc[i].y * xc[i].y) * kinetic[i];
  s += s1;
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = RPB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[col] = sh[0];
}

// src/pw_core.cpp:33-47
__device__ __forceinline__ double kinetic_preconditioner(double kinetic, double wk) {
  const double scale = 1.5 * fmax(wk, 2. / 27.211386245988);
  const double x = kinetic / scale;
  if (x > 1.) {
    const double inv = 1. / x;
    const double p = inv * (8. + inv * (12. + inv * (18. + 27. * inv)));
    return (2. / scale) * p / (p + 16.);
  }
  const double p = 27. + x * (18. + x * (12. + 8. * x));
  return (2. / scale) * p / (p + 16. * x * x * x * x);
}

// out(:, j) = -K(T, wk_j) * r(:, src[j])
__global__ void k_precondition(const double* __restrict__ r, long ldr, const int* __restrict__ src,
                               const double* __restrict__ wk, const double* __restrict__ kinetic, long n,
                               double* __restrict__ out, long ldo) {
  const long col = blockIdx.y;
  const int sc = src[col];
  const double w = wk[sc];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double k = -kinetic_preconditioner(kinetic[i], w);
    out[2 * (i + col * ldo)] = r[2 * (i + sc * ldr)] * k;
    out[2 * (i + col * ldo) + 1] = r[2 * (i + sc * ldr) + 1] * k;
  }
}

// residual = hx - sx * diag(values)   (Ritz step)
__global__ void k_ritz_residual(const double* __restrict__ hx, const double* __restrict__ sx, long ld, long n,
                                const double* __restrict__ values, double* __restrict__ r) {
  const long col = blockIdx.y;
  const double v = values[col];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long k = 2 * (i + col * ld);
    r[k] = hx[k] - sx[k] * v;
    r[k + 1] = hx[k + 1] - sx[k + 1] * v;
  }
}

// residual = fma(-value, sx, hx) per element (src/pw_core.cpp:587-593)
__global__ void k_fma_residual(const double* __restrict__ hx, const double* __restrict__ sx, long ld, long n,
                               const double* __restrict__ values, double* __restrict__ r) {
  const long col = blockIdx.y;
  const double v = values[col];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long k = 2 * (i + col * ld);
    r[k] = __builtin_fma(-v, sx[k], hx[k]);
    r[k + 1] = __builtin_fma(-v, sx[k + 1], hx[k + 1]);
  }
}

// x(:, j) *= s_j
__global__ void k_scale_columns(double* __restrict__ x, long ld, long n, const double* __restrict__ s) {
  const long col = blockIdx.y;
  const double f = s[col];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[2 * (i + col * ld)] *= f;
    x[2 * (i + col * ld) + 1] *= f;
  }
}

// x(:, j) /= s_j  (division, as the reference divides by the norm)
__global__ void k_divide_columns(double* __restrict__ x, long ld, long n, const double* __restrict__ s) {
  const long col = blockIdx.y;
  const double f = s[col];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[2 * (i + col * ld)] /= f;
    x[2 * (i + col * ld) + 1] /= f;
  }
}

// values = energy / metric and roots = sqrt(metric) per column (the Rayleigh
// quotients and S norms of a block, left on the device)
__global__ void k_rayleigh_quotients(const double* __restrict__ metric, const double* __restrict__ energy, long cols,
                                     double* __restrict__ values, double* __restrict__ roots) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < cols; c += (long)gridDim.x * blockDim.x) {
    values[c] = energy[c] / metric[c];
    roots[c] = sqrt(metric[c]);
  }
}

// divisors that are not positive and finite become 1 (zero columns stay zero)
__global__ void k_guard_divisors(double* __restrict__ d, long cols) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < cols; c += (long)gridDim.x * blockDim.x)
    if (!(d[c] > 0.) || !isfinite(d[c])) d[c] = 1.;
}

// trial(:, j) = fma(alpha_j, d(:, j), x(:, j)) where alpha_j != 0, else x
__global__ void k_polish_update(const double* __restrict__ x, const double* __restrict__ d, long ld, long n,
                                const double* __restrict__ alpha, double* __restrict__ t) {
  const long col = blockIdx.y;
  const double a = alpha[col];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long k = 2 * (i + col * ld);
    if (a != 0.0) {
      t[k] = __builtin_fma(a, d[k], x[k]);
      t[k + 1] = __builtin_fma(a, d[k + 1], x[k + 1]);
    } else {
      t[k] = x[k];
      t[k + 1] = x[k + 1];
    }
  }
}

// M(0:s, 0:s) <- (M + M^H) / 2, written into out (ld_out)
__global__ void k_hermitise(const double* __restrict__ m, long ldm, long s, double* __restrict__ out, long ldo) {
  const long total = s * s;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long i = e % s, j = e / s;
    const double ar = m[2 * (i + j * ldm)], ai = m[2 * (i + j * ldm) + 1];
    const double br = m[2 * (j + i * ldm)], bi = -m[2 * (j + i * ldm) + 1];
    out[2 * (i + j * ldo)] = 0.5 * (ar + br);
    out[2 * (i + j * ldo) + 1] = 0.5 * (ai + bi);
  }
}

// Projected matrix C (leading dimension ld, order m) whose first k basis
// columns are Ritz vectors of the previous Rayleigh-Ritz solve: their block is
// diag(theta) (the identity when theta is null) and the rows below it are the
// adjoint of the computed columns k..m, C(k:m, 0:k) = C(0:k, k:m)^H.
__global__ void k_ritz_border(double* __restrict__ c, long ld, long k, long m, const double* __restrict__ theta) {
  const long total = m * k;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long i = e % m, j = e / m;
    double re = 0., im = 0.;
    if (i >= k) {
      re = c[2 * (j + i * ld)];
      im = -c[2 * (j + i * ld) + 1];
    } else if (i == j) {
      re = theta ? theta[i] : 1.;
    }
    c[2 * (i + j * ld)] = re;
    c[2 * (i + j * ld) + 1] = im;
  }
}

// U(:, j) *= 1 / sqrt(lambda_j)  (block CGS2 eigen-normalisation)
__global__ void k_inv_sqrt_columns(double* __restrict__ u, long ld, long rows, const double* __restrict__ lambda) {
  const long col = blockIdx.y;
  const double f = 1. / sqrt(lambda[col]);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < rows; i += (long)gridDim.x * blockDim.x) {
    u[2 * (i + col * ld)] *= f;
    u[2 * (i + col * ld) + 1] *= f;
  }
}

__global__ void k_finite_check(const double* __restrict__ x, long ld, long n, long cols, int* __restrict__ bad) {
  const long total = n * cols;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long i = e % n, j = e / n;
    const double a = x[2 * (i + j * ld)], b = x[2 * (i + j * ld) + 1];
    if (!isfinite(a) || !isfinite(b)) *bad = 1;
  }
}

// max_ij |B_ij - delta_ij| over the m x m block (bits of a non-negative double,
// folded with an unsigned max so that any NaN dominates); E = B - I when e != 0.
__global__ __launch_bounds__(TPB) void k_identity_defect(const double* __restrict__ b, long ldb, long m,
                                                         double* __restrict__ e, long lde,
                                                         unsigned long long* __restrict__ max_bits) {
  __shared__ double red[TPB];
  double worst = 0.;
  const long total = m * m;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t % m, j = t / m;
    const double re = b[2 * (i + j * ldb)] - (i == j ? 1. : 0.), im = b[2 * (i + j * ldb) + 1];
    if (e) {
      e[2 * (i + j * lde)] = re;
      e[2 * (i + j * lde) + 1] = im;
    }
    const double a = sqrt(re * re + im * im);
    worst = (a > worst || a != a) ? a : worst;
  }
  red[threadIdx.x] = worst;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double o = red[threadIdx.x + w];
      if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(max_bits, (unsigned long long)__double_as_longlong(red[0]));
}

// M = I - E/2 + 3 E2/8 (E2 = E E): the series of (I + E)^(-1/2) to second order
__global__ void k_series_metric(const double* __restrict__ e, const double* __restrict__ e2, long ld, long m,
                                double* __restrict__ out, long ldo) {
  const long total = m * m;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t % m, j = t / m;
    out[2 * (i + j * ldo)] = (i == j ? 1. : 0.) - 0.5 * e[2 * (i + j * ld)] + 0.375 * e2[2 * (i + j * ld)];
    out[2 * (i + j * ldo) + 1] = -0.5 * e[2 * (i + j * ld) + 1] + 0.375 * e2[2 * (i + j * ld) + 1];
  }
}

// out[i] = Re a_ii, out[m + i] = Re b_ii
__global__ void k_diagonals(const double* __restrict__ a, long lda, const double* __restrict__ b, long ldb, long m,
                            double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m; i += (long)gridDim.x * blockDim.x) {
    out[i] = a[2 * (i + i * lda)];
    out[m + i] = b[2 * (i + i * ldb)];
  }
}

// zero the strictly lower triangle of an m x m block
__global__ void k_zero_lower(double* __restrict__ a, long lda, long m) {
  const long total = m * m;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t % m, j = t / m;
    if (i > j) {
      a[2 * (i + j * lda)] = 0.;
      a[2 * (i + j * lda) + 1] = 0.;
    }
  }
}

dim3 grid_cols(long n, long cols) {
  long bx = (n + 255) / 256;
  if (bx > 64) bx = 64;
  return dim3(unsigned(bx), unsigned(cols));
}

// ---- block preconditioned CG of the Jacobi-Davidson correction equations
// (one independent CG per column; its scalars live in device arrays, so an
// iteration needs no host round trip)

// out(:, c) = in(:, src[c])
__global__ void k_gather_columns(const double* __restrict__ in, long ldi, const int* __restrict__ src, long n,
                                 double* __restrict__ out, long ldo) {
  const long col = blockIdx.y;
  const long sc = src[col];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    out[2 * (i + col * ldo)] = in[2 * (i + sc * ldi)];
    out[2 * (i + col * ldo) + 1] = in[2 * (i + sc * ldi) + 1];
  }
}

// z(:, c) = K(kinetic_i, wk[src[c]]) r(:, c): the kinetic preconditioner as a
// positive diagonal (the Davidson correction is its negative)
__global__ void k_jd_precondition(const double* __restrict__ r, long ldr, const int* __restrict__ src,
                                  const double* __restrict__ wk, const double* __restrict__ kinetic, long n,
                                  double* __restrict__ z, long ldz) {
  const long col = blockIdx.y;
  const double w = wk[src[col]];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double k = kinetic_preconditioner(kinetic[i], w);
    z[2 * (i + col * ldz)] = r[2 * (i + col * ldr)] * k;
    z[2 * (i + col * ldz) + 1] = r[2 * (i + col * ldr) + 1] * k;
  }
}

// alpha = rho / (p, q) (0 when the curvature is not positive and finite:
// that column stops moving); t += alpha p; r -= alpha q
__global__ void k_cg_update(double* __restrict__ t, double* __restrict__ r, const double* __restrict__ p,
                            const double* __restrict__ q, long ld, long n, const double* __restrict__ rho,
                            const double* __restrict__ pq) {
  const long col = blockIdx.y;
  const double c = pq[col], num = rho[col];
  const double a = (c > 0. && isfinite(c) && isfinite(num)) ? num / c : 0.;
  if (a == 0.) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long k = 2 * (i + col * ld);
    t[k] += a * p[k];
    t[k + 1] += a * p[k + 1];
    r[k] -= a * q[k];
    r[k + 1] -= a * q[k + 1];
  }
}

// One pass of a CG step after its product q: alpha = rho / (p, q) (0 as in
// k_cg_update), t += alpha p, r -= alpha q, z = M r with the kinetic
// preconditioner of the column's root, and the partial sums of Re (r, z) over
// the rows of this workgroup (kCgParts workgroups per column, rows split into
// contiguous ranges, each summed in a fixed order): part[col * kCgParts + g].
constexpr int kCgParts = 8;
__global__ __launch_bounds__(256) void k_cg_step(double* __restrict__ t, double* __restrict__ r,
                                                 const double* __restrict__ p, const double* __restrict__ q,
                                                 double* __restrict__ z, long ld, long n,
                                                 const double* __restrict__ rho, const double* __restrict__ pq,
                                                 const int* __restrict__ src, const double* __restrict__ wk,
                                                 const double* __restrict__ kinetic, double* __restrict__ part) {
  __shared__ double sh[256];
  const long col = blockIdx.y;
  const int g = blockIdx.x;
  const double c = pq[col], num = rho[col];
  const double a = (c > 0. && isfinite(c) && isfinite(num)) ? num / c : 0.;
  const double w = wk[src[col]];
  const long chunk = (n + kCgParts - 1) / kCgParts;
  const long i0 = g * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
  double acc = 0.;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    const long k = 2 * (i + col * ld);
    double2 rv = *reinterpret_cast<const double2*>(r + k);
    if (a != 0.) {
      const double2 pv = *reinterpret_cast<const double2*>(p + k);
      const double2 qv = *reinterpret_cast<const double2*>(q + k);
      double2 tv = *reinterpret_cast<const double2*>(t + k);
      tv.x += a * pv.x;
      tv.y += a * pv.y;
      rv.x -= a * qv.x;
      rv.y -= a * qv.y;
      *reinterpret_cast<double2*>(t + k) = tv;
      *reinterpret_cast<double2*>(r + k) = rv;
    }
    const double kp = kinetic_preconditioner(kinetic[i], w);
    const double2 zv{rv.x * kp, rv.y * kp};
    *reinterpret_cast<double2*>(z + k) = zv;
    acc += rv.x * zv.x + rv.y * zv.y;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[col * kCgParts + g] = sh[0];
}

// rho_new = the column's kCgParts partial sums in order; beta = rho_new /
// rho_old (0 when rho_old is not positive); p = z + beta p; rho_new is stored
// for the next step's alpha
__global__ void k_cg_direction_parts(double* __restrict__ p, const double* __restrict__ z, long ld, long n,
                                     const double* __restrict__ part, const double* __restrict__ rho_old,
                                     double* __restrict__ rho_new) {
  const long col = blockIdx.y;
  double v = 0.;
#pragma unroll
  for (int g = 0; g < kCgParts; ++g) v += part[col * kCgParts + g];
  const double o = rho_old[col];
  const double b = (o > 0. && isfinite(o) && isfinite(v)) ? v / o : 0.;
  if (blockIdx.x == 0 && threadIdx.x == 0) rho_new[col] = v;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long k = 2 * (i + col * ld);
    p[k] = z[k] + b * p[k];
    p[k + 1] = z[k + 1] + b * p[k + 1];
  }
}

// beta = rho_new / rho_old (0 when rho_old is not positive); p = z + beta p
__global__ void k_cg_direction(double* __restrict__ p, const double* __restrict__ z, long ld, long n,
                               const double* __restrict__ rho_new, const double* __restrict__ rho_old) {
  const long col = blockIdx.y;
  const double o = rho_old[col], v = rho_new[col];
  const double b = (o > 0. && isfinite(o) && isfinite(v)) ? v / o : 0.;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long k = 2 * (i + col * ld);
    p[k] = z[k] + b * p[k];
    p[k + 1] = z[k + 1] + b * p[k + 1];
  }
}
}  // namespace

// Davidson seed block (src/pw_core.cpp:1027-1112) on the device, bit-equal to
// the sequential host recipe: uniform() call c returns splitmix64 of
// state0 + (c + 1) * golden, so every call is computed where it is needed;
// column c of the block consumes calls [2 n c, 2 n (c + 1)) (its broad()
// vector b). The two norms are summed in index order (one wave each), as
// the host loops do (this file is built without FMA contraction).
//   warm column (c < roots when warm): src + (pert |src|) b / |b|
//   guard column g: 0.05 b / |b|, plus 1 at row order[g]
__device__ inline double seed_uniform(unsigned long long c) {
  unsigned long long z = 0x6a09e667f3bcc909ULL + (c + 1) * 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return static_cast<double>(z >> 11) * 0x1.0p-53 - .5;
}
__global__ __launch_bounds__(TPB) void k_davidson_seeds(long n, long warm_cols, const double* __restrict__ init,
                                                        long ldi, double pert, const int* __restrict__ order,
                                                        double* __restrict__ out, long ldo) {
  __shared__ double norms[2];
  const long c = blockIdx.x;
  const bool warm = c < warm_cols;
  double* col = out + 2 * c * ldo;
  const double* src = warm ? init + 2 * c * ldi : nullptr;
  const unsigned long long base = static_cast<unsigned long long>(c) * 2ull * static_cast<unsigned long long>(n);
  for (long i = threadIdx.x; i < n; i += TPB) {
    col[2 * i] = seed_uniform(base + 2 * i);
    col[2 * i + 1] = seed_uniform(base + 2 * i + 1);
  }
  __syncthreads();
  // Each norm is summed in index order by one wave: its 64 lanes form
  // re^2 + im^2 of 64 consecutive entries in parallel, then every lane adds the
  // 64 terms in lane order from lane reads (the same terms in the same order as
  // one thread walking the column; entries past n contribute +0, which leaves
  // the non-negative sum unchanged bit for bit). Loads use clamped indices
  // without branches, and the next 4 x 64 entries are loaded while the current
  // ones are summed.
  auto ordered_norm = [n](const double* v, int lane) {
    constexpr int D = 4;
    auto fetch = [&](long i0, double* sq) {
#pragma unroll
      for (int u = 0; u < D; ++u) {
        const long i = i0 + 64 * u + lane;
        const long ic = i < n ? i : n - 1;
        const double2 z = *reinterpret_cast<const double2*>(v + 2 * ic);
        const double m = z.x * z.x + z.y * z.y;
        sq[u] = i < n ? m : 0.;
      }
    };
    double nn = 0., cur[D], nxt[D] = {0., 0., 0., 0.};
    fetch(0, cur);
    for (long i0 = 0; i0 < n; i0 += 64 * D) {
      if (i0 + 64 * D < n) fetch(i0 + 64 * D, nxt);
#pragma unroll
      for (int u = 0; u < D; ++u) {
        const unsigned long long bits = __builtin_bit_cast(unsigned long long, cur[u]);
        const int lo = int(bits & 0xffffffffull), hi = int(bits >> 32);
#pragma unroll
        for (int t = 0; t < 64; ++t) {
          const unsigned long long tb =
              (unsigned long long)(unsigned)__builtin_amdgcn_readlane(lo, t) |
              ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(hi, t) << 32);
          nn += __builtin_bit_cast(double, tb);
        }
      }
#pragma unroll
      for (int u = 0; u < D; ++u) cur[u] = nxt[u];
    }
    return sqrt(nn);
  };
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wave == 0) {
    const double r = ordered_norm(col, lane);
    if (lane == 0) norms[0] = r;
  } else if (wave == 1 && warm) {
    const double r = ordered_norm(src, lane);
    if (lane == 0) norms[1] = r;
  }
  __syncthreads();
  const double nb = norms[0];
  const double s = warm ? pert * norms[1] : 0.;
  for (long i = threadIdx.x; i < n; i += TPB) {
    const double br = col[2 * i] / nb, bi = col[2 * i + 1] / nb;
    if (warm) {
      col[2 * i] = src[2 * i] + s * br;
      col[2 * i + 1] = src[2 * i + 1] + s * bi;
    } else {
      col[2 * i] = .05 * br;
      col[2 * i + 1] = .05 * bi;
    }
  }
  if (!warm) {
    __syncthreads();
    if (threadIdx.x == 0) col[2 * static_cast<long>(order[c - warm_cols])] += 1.;
  }
}

#define SDFT_CHECKED(call) \
  do {                     \
    call;                  \
    return hipGetLastError(); \
  } while (0)

hipError_t launch_real_dots(const double* a, long lda, const double* b, long ldb, long n, long cols, double* out,
                            hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_real_dots, dim3(unsigned(cols)), dim3(RPB), 0, s, a, lda, b, ldb, n, out));
}
hipError_t launch_complex_dots(const double* a, long lda, long ca, const double* b, long ldb, long cb, long n,
                               double* out, long ldo, hipStream_t s) {
  if (ca <= 0 || cb <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_complex_dots, dim3(unsigned(ca), unsigned(cb)), dim3(TPB), 0, s, a, lda, b, ldb,
                                  n, 0, out, ldo));
}
hipError_t launch_col_norms(const double* x, long ldx, long n, long cols, double* out, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_col_norms, dim3(unsigned(cols)), dim3(RPB), 0, s, x, ldx, n, out));
}
hipError_t launch_davidson_seeds(long n, long warm_cols, const double* init, long ldi, double pert, long guards,
                                 const int* order, double* out, long ldo, hipStream_t s) {
  if (n <= 0 || warm_cols + guards <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_davidson_seeds, dim3(unsigned(warm_cols + guards)), dim3(TPB), 0, s, n, warm_cols,
                                  init, ldi, pert, order, out, ldo));
}
hipError_t launch_kinetic_weights(const double* x, long ldx, long n, long cols, const double* kinetic, double* out,
                                  hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_kinetic_weights, dim3(unsigned(cols)), dim3(RPB), 0, s, x, ldx, n, kinetic, out));
}
hipError_t launch_precondition(const double* r, long ldr, const int* src, long cols, const double* wk,
                               const double* kinetic, long n, double* out, long ldo, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_precondition, grid_cols(n, cols), dim3(256), 0, s, r, ldr, src, wk, kinetic, n,
                                  out, ldo));
}
hipError_t launch_gather_columns(const double* in, long ldi, const int* src, long n, long cols, double* out, long ldo,
                                 hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_gather_columns, grid_cols(n, cols), dim3(256), 0, s, in, ldi, src, n, out, ldo));
}
hipError_t launch_jd_precondition(const double* r, long ldr, const int* src, long cols, const double* wk,
                                  const double* kinetic, long n, double* z, long ldz, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_jd_precondition, grid_cols(n, cols), dim3(256), 0, s, r, ldr, src, wk, kinetic, n,
                                  z, ldz));
}
hipError_t launch_cg_update(double* t, double* r, const double* p, const double* q, long ld, long n, long cols,
                            const double* rho, const double* pq, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_cg_update, grid_cols(n, cols), dim3(256), 0, s, t, r, p, q, ld, n, rho, pq));
}
hipError_t launch_cg_direction(double* p, const double* z, long ld, long n, long cols, const double* rho_new,
                               const double* rho_old, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_cg_direction, grid_cols(n, cols), dim3(256), 0, s, p, z, ld, n, rho_new, rho_old));
}
hipError_t launch_cg_step(double* t, double* r, const double* p, const double* q, double* z, long ld, long n,
                          long cols, const double* rho, const double* pq, const int* src, const double* wk,
                          const double* kinetic, double* part, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  if (cols > 65535) return hipErrorInvalidValue;
  SDFT_CHECKED(hipLaunchKernelGGL(k_cg_step, dim3(kCgParts, unsigned(cols)), dim3(256), 0, s, t, r, p, q, z, ld, n,
                                  rho, pq, src, wk, kinetic, part));
}
hipError_t launch_cg_direction_parts(double* p, const double* z, long ld, long n, long cols, const double* part,
                                     const double* rho_old, double* rho_new, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_cg_direction_parts, grid_cols(n, cols), dim3(256), 0, s, p, z, ld, n, part,
                                  rho_old, rho_new));
}
int cg_step_parts() { return kCgParts; }
hipError_t launch_ritz_residual(const double* hx, const double* sx, long ld, long n, long cols, const double* values,
                                double* r, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_ritz_residual, grid_cols(n, cols), dim3(256), 0, s, hx, sx, ld, n, values, r));
}
hipError_t launch_fma_residual(const double* hx, const double* sx, long ld, long n, long cols, const double* values,
                               double* r, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_fma_residual, grid_cols(n, cols), dim3(256), 0, s, hx, sx, ld, n, values, r));
}
hipError_t launch_scale_columns(double* x, long ld, long n, long cols, const double* scales, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_scale_columns, grid_cols(n, cols), dim3(256), 0, s, x, ld, n, scales));
}
hipError_t launch_rayleigh_quotients(const double* metric, const double* energy, long cols, double* values,
                                     double* roots, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_rayleigh_quotients, dim3(unsigned((cols + 255) / 256)), dim3(256), 0, s, metric,
                                  energy, cols, values, roots));
}
hipError_t launch_guard_divisors(double* d, long cols, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_guard_divisors, dim3(unsigned((cols + 255) / 256)), dim3(256), 0, s, d, cols));
}
hipError_t launch_divide_columns(double* x, long ld, long n, long cols, const double* divisors, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_divide_columns, grid_cols(n, cols), dim3(256), 0, s, x, ld, n, divisors));
}
hipError_t launch_polish_update(const double* x, const double* d, long ld, long n, long cols, const double* alpha,
                                double* trial, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_polish_update, grid_cols(n, cols), dim3(256), 0, s, x, d, ld, n, alpha, trial));
}
hipError_t launch_ritz_border(double* c, long ld, long k, long m, const double* theta, hipStream_t s) {
  if (k <= 0 || m < k) return hipSuccess;
  const long total = m * k;
  const int blocks = int(total / 256 + 1 > 2048 ? 2048 : total / 256 + 1);
  SDFT_CHECKED(hipLaunchKernelGGL(k_ritz_border, dim3(blocks), dim3(256), 0, s, c, ld, k, m, theta));
}
hipError_t launch_hermitise(const double* m, long ldm, long size, double* out, long ldo, hipStream_t s) {
  if (size <= 0) return hipSuccess;
  const long total = size * size;
  const int blocks = int(total / 256 + 1 > 2048 ? 2048 : total / 256 + 1);
  SDFT_CHECKED(hipLaunchKernelGGL(k_hermitise, dim3(blocks), dim3(256), 0, s, m, ldm, size, out, ldo));
}
hipError_t launch_inv_sqrt_columns(double* u, long ld, long rows, long cols, const double* lambda, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_inv_sqrt_columns, grid_cols(rows, cols), dim3(256), 0, s, u, ld, rows, lambda));
}
hipError_t launch_finite_check(const double* x, long ld, long n, long cols, int* bad, hipStream_t s) {
  if (cols <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_finite_check, dim3(512), dim3(256), 0, s, x, ld, n, cols, bad));
}

namespace {
int square_blocks(long m) {
  const long total = m * m;
  return int(total / 256 + 1 > 2048 ? 2048 : total / 256 + 1);
}
}  // namespace
hipError_t launch_identity_defect(const double* b, long ldb, long m, double* e, long lde,
                                  unsigned long long* max_bits, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_identity_defect, dim3(square_blocks(m)), dim3(TPB), 0, s, b, ldb, m, e, lde,
                                  max_bits));
}
hipError_t launch_series_metric(const double* e, const double* e2, long ld, long m, double* out, long ldo,
                                hipStream_t s) {
  if (m <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_series_metric, dim3(square_blocks(m)), dim3(256), 0, s, e, e2, ld, m, out, ldo));
}
hipError_t launch_diagonals(const double* a, long lda, const double* b, long ldb, long m, double* out,
                            hipStream_t s) {
  if (m <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_diagonals, dim3(int((m + 255) / 256)), dim3(256), 0, s, a, lda, b, ldb, m, out));
}
hipError_t launch_zero_lower(double* a, long lda, long m, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  SDFT_CHECKED(hipLaunchKernelGGL(k_zero_lower, dim3(square_blocks(m)), dim3(256), 0, s, a, lda, m));
}


// Loads this file's code object on the calling thread (see module_prime.hpp).
hipError_t prime_module_davidson_kernels() {
  hipFuncAttributes a{};
  return hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_real_dots));
}

}  // namespace sdft_hip
