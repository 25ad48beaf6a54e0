// Symmetric tridiagonal eigensolver of the Rayleigh-Ritz step (see
// tridiag_eig.hpp). It replaces divide and conquer, which computes every
// eigenvector, by work on the wanted ones only:
//
//  * eigenvalues: one wavefront per eigenvalue index i narrows [lo, hi] with
//    64 Sturm counts per round (lane l evaluates lo + (l + 1)(hi - lo)/65), so
//    one round gains six bits; ten rounds reach the tolerance
//    max(2 eps max(|lo|, |hi|), eps |T|). The Gershgorin interval starts it.
//  * eigenvectors: one wavefront per cluster of the lowest k eigenvalues
//    (neighbours closer than kTridiagClusterGap |T|). For each member, lane 0
//    factorises T - x I (row pivoting, tiny pivots perturbed as LAPACK dlagts
//    does) and runs three inverse-iteration solves from a deterministic start
//    vector; after each solve the wavefront orthogonalises the iterate against
//    the cluster's earlier members (modified Gram-Schmidt) and normalises it.
//    Members of a cluster get shifts at least 10 eps |x| apart (LAPACK dstein).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "tridiag_core.hpp"
#include "tridiag_eig.hpp"
#include "wave_sum.hpp"

namespace sdft_hip {
namespace {

constexpr int kWave = 64;
constexpr int kValueWaves = 4;
constexpr int kRounds = 24;
constexpr int kSolves = 3;

__device__ inline double wave_sum(double v) { return wave::sum(v); }
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

// Gershgorin bounds of T, |T| and the pivot guard of the Sturm count, computed
// by one wavefront from d and e (both in LDS).
struct Bounds {
  double gl, gu, tnorm, pivmin;
};
__device__ inline Bounds gershgorin(const double* d, const double* e, int n) {
  const int lane = threadIdx.x % kWave;
  double lo = DBL_MAX, hi = -DBL_MAX, emax = 0.0;
  for (int i = lane; i < n; i += kWave) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i + 1 < n) emax = fmax(emax, e[i] * e[i]);
  }
  lo = -wave_max(-lo);
  hi = wave_max(hi);
  emax = wave_max(emax);
  Bounds b;
  b.tnorm = fmax(fabs(lo), fabs(hi));
  b.pivmin = DBL_MIN * fmax(1.0, emax);
  const double widen = 2.1 * (DBL_EPSILON * b.tnorm * n + 2.0 * b.pivmin);
  b.gl = lo - widen;
  b.gu = hi + widen;
  return b;
}

__global__ __launch_bounds__(kValueWaves * kWave) void k_tridiag_values(int n, const double* __restrict__ d,
                                                                        const double* __restrict__ e,
                                                                        int count, double* __restrict__ w,
                                                                        int* status) {
  extern __shared__ double lds[];
  double* sd = lds;
  double* se = sd + n;
  double* se2 = se + n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sd[i] = d[i];
    const double v = i + 1 < n ? e[i] : 0.0;
    se[i] = v;
    se2[i] = v * v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *status = 0;
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int index = blockIdx.x * kValueWaves + threadIdx.x / kWave;
  if (index >= count) return;
  const Bounds b = gershgorin(sd, se, n);
  double lo = b.gl, hi = b.gu;
  const double atol = DBL_EPSILON * b.tnorm;
  for (int round = 0; round < kRounds; ++round) {
    const double width = hi - lo;
    if (!(width > fmax(2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)), fmax(atol, b.pivmin)))) break;
    const double x = lo + width * (lane + 1) * (1.0 / 65.0);
    const int c = sturm_count(sd, se2, n, x, b.pivmin);
    const unsigned long long above = __ballot(c >= index + 1);
    const int first = above ? __ffsll(static_cast<long long>(above)) - 1 : kWave;
    const double xlo = __shfl(x, first > 0 ? first - 1 : 0, kWave);
    const double xhi = __shfl(x, first < kWave ? first : 0, kWave);
    if (first < kWave) {
      hi = xhi;
      if (first > 0) lo = xlo;
    } else {
      lo = __shfl(x, kWave - 1, kWave);
    }
  }
  if (lane == 0) w[index] = 0.5 * (lo + hi);
}

__global__ __launch_bounds__(kWave) void k_tridiag_vectors(int n, const double* __restrict__ d,
                                                           const double* __restrict__ e,
                                                           const double* __restrict__ w, int k,
                                                           double* __restrict__ z, long ldz, int* status) {
  const int start = blockIdx.x;
  if (start >= k) return;
  extern __shared__ double lds[];
  double* sd = lds;
  double* se = sd + n;
  double* inv0 = se + n;
  double* u1 = inv0 + n;
  double* u2 = u1 + n;
  double* lm = u2 + n;
  double* y = lm + n;
  unsigned char* piv = reinterpret_cast<unsigned char*>(y + n);
  const int lane = threadIdx.x;
  for (int i = lane; i < n; i += kWave) {
    sd[i] = d[i];
    se[i] = i + 1 < n ? e[i] : 0.0;
  }
  __syncthreads();
  const Bounds b = gershgorin(sd, se, n);
  const double gap = kTridiagClusterGap * b.tnorm;
  // only the first member of a cluster does work
  if (start > 0 && w[start] - w[start - 1] <= gap) return;
  int end = start;
  while (end + 1 < k && w[end + 1] - w[end] <= gap) ++end;
  // (a zero matrix keeps its pivots finite; 1/pert stays far from overflow)
  const double pert = fmax(DBL_EPSILON * b.tnorm, DBL_MIN / DBL_EPSILON);
  double xprev = 0.0;
  bool bad = false;
  for (int j = start; j <= end; ++j) {
    double x = w[j];
    if (j > start) {
      const double sep = 10.0 * DBL_EPSILON * fabs(x);
      if (x - xprev < sep) x = xprev + sep;
    }
    xprev = x;
    if (lane == 0) shifted_lu(sd, se, n, x, pert, inv0, u1, u2, lm, piv);
    for (int i = lane; i < n; i += kWave) y[i] = start_entry(j, i);
    __syncthreads();
    for (int it = 0; it < kSolves; ++it) {
      if (lane == 0) shifted_solve(inv0, u1, u2, lm, piv, n, y);
      __syncthreads();
      // scale by the largest entry first: the solve amplifies by up to 1/pert
      double big = 0.0;
      for (int i = lane; i < n; i += kWave) big = fmax(big, fabs(y[i]));
      big = wave_max(big);
      if (!(big > 0.0) || !(big < DBL_MAX)) {
        bad = true;
        break;
      }
      const double inv_big = 1.0 / big;
      for (int i = lane; i < n; i += kWave) y[i] *= inv_big;
      for (int p = start; p < j; ++p) {
        const double* zp = z + 2 * static_cast<long>(p) * ldz;
        double dot = 0.0;
        for (int i = lane; i < n; i += kWave) dot += y[i] * zp[2 * i];
        dot = wave_sum(dot);
        for (int i = lane; i < n; i += kWave) y[i] -= dot * zp[2 * i];
      }
      double nrm = 0.0;
      for (int i = lane; i < n; i += kWave) nrm += y[i] * y[i];
      nrm = wave_sum(nrm);
      if (!(nrm > 0.0)) {
        bad = true;
        break;
      }
      const double scale = 1.0 / sqrt(nrm);
      for (int i = lane; i < n; i += kWave) y[i] *= scale;
      __syncthreads();
    }
    if (bad) break;
    double* zj = z + 2 * static_cast<long>(j) * ldz;
    for (int i = lane; i < n; i += kWave) {
      zj[2 * i] = y[i];
      zj[2 * i + 1] = 0.0;
    }
    // the next member reads this column back (same workgroup)
    __syncthreads();
  }
  if (bad && lane == 0) *status = 1;
}

// Z <- Q Z with Q = H_0 H_1 ... H_{n-2} of the reduction (LAPACK zhetrd, uplo
// 'L'): H_j = I - tau_j v_j v_j^H, v_j zero above row j + 1, v_j(j + 1) = 1,
// rows j + 2 .. n - 1 in column j of refl. One wavefront per column of Z keeps
// the column in registers (row s * 64 + lane in slot s, NS slots) and applies
// H_{n-2}, ..., H_0 in turn. The workgroup stages `chunk` reflectors at a time
// in LDS as full NS * 64-row vectors (zeros above row j + 1 and below row
// n - 1); the loads of the next round are issued before the steps of the
// current one, so their latency is hidden behind `chunk` steps. A step reads
// all NS slots of the reflector from LDS at once, branch- and mask-free;
// columns are independent.
constexpr int kBackWaves = 4;
constexpr int kBackThreads = kBackWaves * kWave;
// reflectors per round: the register buffer of the next round (rows per
// thread x chunk) stays live during the current round's steps
__host__ __device__ constexpr int back_chunk(int ns) { return ns <= 8 ? 8 : (ns <= 12 ? 6 : 4); }

template <int NS>
__global__ __launch_bounds__(kBackThreads) void k_back_transform(int n, int k, int chunk,
                                                                 const double* __restrict__ refl, long ldr,
                                                                 const double* __restrict__ tau,
                                                                 double* __restrict__ z, long ldz) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  double2* stage = reinterpret_cast<double2*>(lds);
  constexpr int rows = NS * kWave;  // staged length of one reflector
  constexpr int per_thread = (rows + kBackThreads - 1) / kBackThreads;
  constexpr int max_chunk = back_chunk(NS);
  const int tid = threadIdx.x;
  const int lane = tid % kWave;
  const int col = blockIdx.x * kBackWaves + tid / kWave;
  const bool active = col < k;  // inactive wavefronts still stage reflectors
  double* zc = z + 2 * static_cast<long>(active ? col : 0) * ldz;
  double zr[NS], zi[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int i = s * kWave + lane;
    zr[s] = active && i < n ? zc[2 * i] : 0.0;
    zi[s] = active && i < n ? zc[2 * i + 1] : 0.0;
  }
  // the round starting at reflector j0 holds j0, j0 - 1, ..., j0 - count + 1
  double2 buf[max_chunk][per_thread];
  auto fetch = [&](int j0) {
    const int count = j0 + 1 < chunk ? j0 + 1 : chunk;
#pragma unroll
    for (int c = 0; c < max_chunk; ++c) {
      const int j = j0 - c;
      const double2* vc = reinterpret_cast<const double2*>(refl + 2 * static_cast<long>(j) * ldr);
#pragma unroll
      for (int r = 0; r < per_thread; ++r) {
        const int i = tid + r * kBackThreads;
        double2 v = make_double2(i == j + 1 ? 1.0 : 0.0, 0.0);
        if (c < count && i > j + 1 && i < n) v = vc[i];
        buf[c][r] = v;
      }
    }
  };
  fetch(n - 2);
  for (int j0 = n - 2; j0 >= 0; j0 -= chunk) {
    const int count = j0 + 1 < chunk ? j0 + 1 : chunk;
    __syncthreads();  // the previous round's reflectors are consumed
#pragma unroll
    for (int c = 0; c < max_chunk; ++c) {
      if (c < count) {
#pragma unroll
        for (int r = 0; r < per_thread; ++r) {
          const int i = tid + r * kBackThreads;
          if (i < rows) stage[static_cast<long>(c) * rows + i] = buf[c][r];
        }
      }
    }
    __syncthreads();
    if (j0 >= chunk) fetch(j0 - chunk);  // in flight during this round's steps
    if (!active) continue;
    for (int c = 0; c < count; ++c) {
      const int j = j0 - c;
      const double2* v = stage + static_cast<long>(c) * rows;
      const double tr = tau[2 * j], ti = tau[2 * j + 1];
      double2 vv[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) vv[s] = v[s * kWave + lane];
      double dr[4] = {0.0, 0.0, 0.0, 0.0}, di[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        dr[s % 4] += vv[s].x * zr[s] + vv[s].y * zi[s];
        di[s % 4] += vv[s].x * zi[s] - vv[s].y * zr[s];
      }
      const double hr0 = wave::sum((dr[0] + dr[1]) + (dr[2] + dr[3]));
      const double hi0 = wave::sum((di[0] + di[1]) + (di[2] + di[3]));
      const double hr = tr * hr0 - ti * hi0, hi = tr * hi0 + ti * hr0;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        zr[s] -= vv[s].x * hr - vv[s].y * hi;
        zi[s] -= vv[s].x * hi + vv[s].y * hr;
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int i = s * kWave + lane;
    if (i < n) {
      zc[2 * i] = zr[s];
      zc[2 * i + 1] = zi[s];
    }
  }
}

}  // namespace

hipError_t launch_back_transform(int n, int k, const double* refl, long ldr, const double* tau, double* z, long ldz,
                                 hipStream_t s) {
  if (n < 1 || n > kBackTransformMaxOrder || k < 0 || ldr < n || ldz < n) return hipErrorInvalidValue;
  if (k == 0 || n < 2) return hipSuccess;
  const int slots = (n + kWave - 1) / kWave + ((n + kWave - 1) / kWave) % 2;  // the instantiated NS
  const int chunk = back_chunk(slots);
  const size_t lds = 16 * size_t(chunk) * size_t(slots) * kWave;
  const dim3 grid((k + kBackWaves - 1) / kBackWaves), block(kBackThreads);
#define SDFT_BACK(NS)                                                                                      \
  if (slots <= NS) {                                                                                       \
    hipLaunchKernelGGL(k_back_transform<NS>, grid, block, lds, s, n, k, chunk, refl, ldr, tau, z, ldz);   \
    return hipGetLastError();                                                                              \
  }
  SDFT_BACK(2)
  SDFT_BACK(4)
  SDFT_BACK(6)
  SDFT_BACK(8)
  SDFT_BACK(10)
  SDFT_BACK(12)
  SDFT_BACK(14)
  SDFT_BACK(16)
#undef SDFT_BACK
  return hipGetLastError();
}

hipError_t launch_tridiag_values(int n, const double* d, const double* e, int count, double* w, int* status,
                                 hipStream_t s) {
  if (n < 1 || n > kTridiagEigMaxOrder || count < 1 || count > n) return hipErrorInvalidValue;
  const size_t lds = 3 * sizeof(double) * size_t(n);
  const int blocks = (count + kValueWaves - 1) / kValueWaves;
  hipLaunchKernelGGL(k_tridiag_values, dim3(blocks), dim3(kValueWaves * kWave), lds, s, n, d, e, count, w, status);
  return hipGetLastError();
}

hipError_t launch_tridiag_vectors(int n, const double* d, const double* e, const double* w, int k, double* z,
                                  long ldz, int* status, hipStream_t s) {
  if (n < 1 || n > kTridiagEigMaxOrder || k < 0 || k > n || ldz < n) return hipErrorInvalidValue;
  if (k == 0) return hipSuccess;
  const size_t lds = 7 * sizeof(double) * size_t(n) + size_t(n);
  hipLaunchKernelGGL(k_tridiag_vectors, dim3(k), dim3(kWave), lds, s, n, d, e, w, k, z, ldz, status);
  return hipGetLastError();
}


// Loads this file's code object on the calling thread (see module_prime.hpp).
hipError_t prime_module_tridiag_eig() {
  hipFuncAttributes a{};
  return hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_tridiag_values));
}

}  // namespace sdft_hip
