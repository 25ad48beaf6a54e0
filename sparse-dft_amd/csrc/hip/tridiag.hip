// Householder reduction of a Hermitian matrix to real symmetric tridiagonal
// form in one cooperative launch (the Rayleigh-Ritz eigensolve of the block
// Davidson). Output follows LAPACK zhetrd with uplo = 'L' (d, e, tau and the
// reflector tails v(k+2:n) in column k), so that a tridiagonal eigensolver and
// zunmtr complete the decomposition.
//
// Layout: every workgroup keeps the rows g, g + G, g + 2G, ... of the matrix in
// LDS for the whole reduction (the matrix never returns to HBM). Step k:
//   1. every workgroup forms the reflector of the current column redundantly
//      (identical code and data, so identical bits everywhere);
//   2. p_i = sum_j A_ij v_j on its own rows, published together with the
//      pre-update A_{i,k+1};
//   3. one grid barrier;
//   4. w = tau p - (tau/2)(tau p)^H v v, the rank-2 update of its rows, and the
//      next column for all rows from the published A_{i,k+1}.
// The published vectors are double-buffered by step parity, so one barrier per
// column suffices. The barrier spins on an agent-scope counter with a bounded
// wait: on expiry it raises an abort flag that every workgroup observes, and
// all workgroups leave the kernel.
#include <hip/hip_runtime.h>

#include "tridiag.hpp"

namespace sdft_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr long kSpinLimit = 1L << 20;

struct Z {
  double x, y;
};
__device__ inline Z zmul(Z a, Z b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline Z zconj(Z a) { return {a.x, -a.y}; }
__device__ inline Z zadd(Z a, Z b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline Z zsub(Z a, Z b) { return {a.x - b.x, a.y - b.y}; }
__device__ inline Z zscale(Z a, double s) { return {a.x * s, a.y * s}; }

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum of one double over the workgroup in a fixed order; every thread gets it.
__device__ double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double t = 0.;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) t += scratch[w];
  return t;
}

__device__ Z block_zsum(Z v, double* scratch) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    scratch[2 * wave] = v.x;
    scratch[2 * wave + 1] = v.y;
  }
  __syncthreads();
  Z t{0., 0.};
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    t.x += scratch[2 * w];
    t.y += scratch[2 * w + 1];
  }
  return t;
}

// Grid barrier number `k` (0-based): returns false when the wait expired or
// another workgroup gave up.
__device__ bool grid_barrier(unsigned* counter, int* abort_flag, unsigned target, int* ok_shared) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    for (long spin = 0;; ++spin) {
      if (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
      if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if (spin > kSpinLimit) {
        __hip_atomic_fetch_or(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *ok_shared = ok;
  }
  __syncthreads();
  __threadfence();
  return *ok_shared != 0;
}

__global__ __launch_bounds__(kThreads) void k_hetrd_lds(const Z* __restrict__ a, long lda, int n, int rpw,
                                                        Z* __restrict__ refl, long ldr, double* __restrict__ d,
                                                        double* __restrict__ e, Z* __restrict__ tau,
                                                        Z* __restrict__ pub_p, Z* __restrict__ pub_c,
                                                        unsigned* __restrict__ counter, int* __restrict__ abort_flag) {
  extern __shared__ double lds_raw[];
  Z* rows = reinterpret_cast<Z*>(lds_raw);  // rpw x n
  Z* v = rows + size_t(rpw) * n;            // reflector (trailing part), n
  Z* w = v + n;                             // p, then w
  Z* col = w + n;                           // current column, n
  double* scratch = reinterpret_cast<double*>(col + n);  // 2 * kWaves
  __shared__ int ok_shared;
  __shared__ Z sh_a2;

  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  // row i of the Hermitian input is the conjugate of column i (coalesced read)
  for (int r = 0; r < rpw; ++r) {
    const int i = g + r * G;
    if (i >= n) break;
    for (int j = tid; j < n; j += kThreads) rows[size_t(r) * n + j] = zconj(a[j + size_t(i) * lda]);
  }
  for (int j = tid; j < n; j += kThreads) col[j] = a[j];
  __syncthreads();

  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;  // order of the trailing block
    // ---- 1. reflector of column k (LAPACK zlarfg conventions)
    double part = 0.;
    for (int j = k + 2 + tid; j < n; j += kThreads) part += col[j].x * col[j].x + col[j].y * col[j].y;
    const double xn2 = block_sum(part, scratch);
    const Z alpha = col[k + 1];
    Z t{0., 0.}, scale{0., 0.};
    double beta;
    if (xn2 == 0. && alpha.y == 0.) {
      beta = alpha.x;
    } else {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      beta = alpha.x >= 0. ? -nrm : nrm;
      t = {(beta - alpha.x) / beta, -alpha.y / beta};
      // 1 / (alpha - beta)
      const Z den{alpha.x - beta, alpha.y};
      const double dd = den.x * den.x + den.y * den.y;
      scale = {den.x / dd, -den.y / dd};
    }
    for (int j = tid; j < m; j += kThreads) v[j] = j == 0 ? Z{1., 0.} : zmul(col[k + 1 + j], scale);
    if (g == 0 && tid == 0) {
      d[k] = col[k].x;
      e[k] = beta;
      tau[k] = t;
    }
    __syncthreads();
    if (g == 0)
      for (int j = 2 + tid; j <= m; j += kThreads) refl[(k + j) + size_t(k) * ldr] = v[j - 1];

    // ---- 2. p_i = sum_j A_ij v_j on own trailing rows; publish p_i, A_{i,k+1}
    Z* pp = pub_p + size_t(k & 1) * n;
    Z* pc = pub_c + size_t(k & 1) * n;
    {
      const int lane = tid & 63;
      for (int r = tid >> 6; r < rpw; r += kWaves) {  // one wave per row, no block barrier
        const int i = g + r * G;
        if (i >= n) break;
        if (i <= k) continue;
        const Z* row = rows + size_t(r) * n + (k + 1);
        Z acc{0., 0.};
        for (int j = lane; j < m; j += 64) acc = zadd(acc, zmul(row[j], v[j]));
        acc.x = wave_sum(acc.x);
        acc.y = wave_sum(acc.y);
        if (lane == 0) {
          pp[i] = acc;
          pc[i] = row[0];
        }
      }
    }
    // ---- 3. grid barrier
    if (!grid_barrier(counter, abort_flag, unsigned(G) * unsigned(k + 1), &ok_shared)) return;
    // ---- 4. w, rank-2 update of own rows, next column for all rows
    for (int j = tid; j < m; j += kThreads) {
      w[j] = pp[k + 1 + j];
      col[k + 1 + j] = pc[k + 1 + j];
    }
    __syncthreads();
    Z dot{0., 0.};  // (tau p)^H v
    for (int j = tid; j < m; j += kThreads) dot = zadd(dot, zmul(zconj(zmul(t, w[j])), v[j]));
    dot = block_zsum(dot, scratch);
    if (tid == 0) sh_a2 = zscale(zmul(t, dot), -0.5);
    __syncthreads();
    const Z a2 = sh_a2;
    for (int j = tid; j < m; j += kThreads) w[j] = zadd(zmul(t, w[j]), zmul(a2, v[j]));
    __syncthreads();
    for (int r = 0; r < rpw; ++r) {
      const int i = g + r * G;
      if (i >= n) break;
      if (i <= k) continue;
      const Z vi = v[i - k - 1], wi = w[i - k - 1];
      Z* row = rows + size_t(r) * n + (k + 1);
      for (int j = tid; j < m; j += kThreads)
        row[j] = zsub(row[j], zadd(zmul(vi, zconj(w[j])), zmul(wi, zconj(v[j]))));
    }
    const Z w0c = zconj(w[0]), v0c = zconj(v[0]);
    for (int j = tid; j < m; j += kThreads)
      col[k + 1 + j] = zsub(col[k + 1 + j], zadd(zmul(v[j], w0c), zmul(w[j], v0c)));
    __syncthreads();
  }
  if (g == 0 && tid == 0) d[n - 1] = col[n - 1].x;
}

int max_lds() {
  static int bytes = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
    return v;
  }();
  return bytes;
}

int device_cus() {
  static int cus = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return v;
  }();
  return cus;
}

size_t lds_bytes(int n, int rpw) { return (size_t(rpw) * n + 3 * size_t(n)) * sizeof(Z) + 2 * kWaves * sizeof(double); }
}  // namespace

bool hetrd_plan(int n, int* rows_per_group, int* groups) {
  if (n < 1 || n > kHetrdMaxOrder) return false;
  const int cus = device_cus();
  const size_t budget = size_t(max_lds());
  // smallest rows-per-group whose grid fits one workgroup per CU, then as many
  // rows as the LDS holds (fewer workgroups make a cheaper barrier)
  int rpw = 1;
  while (lds_bytes(n, rpw + 1) <= budget && rpw < n) ++rpw;
  if (lds_bytes(n, rpw) > budget) return false;
  const int g = (n + rpw - 1) / rpw;
  if (cus > 0 && g > cus) return false;
  *rows_per_group = rpw;
  *groups = g;
  return true;
}

size_t hetrd_workspace_bytes(int n) { return 4 * size_t(n) * sizeof(Z) + 64; }

hipError_t launch_hetrd(const double* a, long lda, int n, double* refl, long ldr, double* d, double* e, double* tau,
                        void* workspace, hipStream_t s) {
  int rpw = 0, groups = 0;
  if (!hetrd_plan(n, &rpw, &groups)) return hipErrorInvalidValue;
  if (n == 1) {
    // d = Re a; nothing to reduce
    return hipMemcpyAsync(d, a, sizeof(double), hipMemcpyDeviceToDevice, s);
  }
  Z* pub_p = reinterpret_cast<Z*>(workspace);
  Z* pub_c = pub_p + 2 * size_t(n);
  unsigned* counter = reinterpret_cast<unsigned*>(pub_c + 2 * size_t(n));
  int* abort_flag = reinterpret_cast<int*>(counter + 1);
  hipError_t err = hipMemsetAsync(counter, 0, 2 * sizeof(unsigned), s);
  if (err != hipSuccess) return err;
  const Z* az = reinterpret_cast<const Z*>(a);
  Z* rz = reinterpret_cast<Z*>(refl);
  Z* tz = reinterpret_cast<Z*>(tau);
  void* args[] = {(void*)&az, (void*)&lda, (void*)&n, (void*)&rpw, (void*)&rz, (void*)&ldr, (void*)&d,
                  (void*)&e, (void*)&tz, (void*)&pub_p, (void*)&pub_c, (void*)&counter, (void*)&abort_flag};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_hetrd_lds), dim3(groups), dim3(kThreads), args,
                                    static_cast<unsigned>(lds_bytes(n, rpw)), s);
}

const int* hetrd_abort_flag(const void* workspace, int n) {
  const Z* pub = reinterpret_cast<const Z*>(workspace);
  return reinterpret_cast<const int*>(reinterpret_cast<const unsigned*>(pub + 4 * size_t(n)) + 1);
}

}  // namespace sdft_hip
