// Householder reduction of a Hermitian matrix to real symmetric tridiagonal
// form in one cooperative launch (the Rayleigh-Ritz eigensolve of the block
// Davidson). Output follows LAPACK zhetrd with uplo = 'L' (d, e, tau and the
// reflector tails v(k+2:n) in column k), so that a tridiagonal eigensolver and
// zunmtr complete the decomposition.
//
// Layout: every workgroup keeps the rows g, g + G, g + 2G, ... of the matrix in
// LDS for the whole reduction (the matrix never returns to HBM). Step k starts
// with reflector k formed and p_i = sum_j A_ij v_j staged for the own rows,
// together with the pre-update A_{i,k+1}:
//   1. publication of the staged values and one grid barrier;
//   2. every workgroup loads all published values at once and reduces
//      (tau p)^H v;
//   3. w = tau p - (tau/2)(tau p)^H v v (in registers), the next column for all
//      rows from the published A_{i,k+1}, the rank-2 update of the own rows;
//   4. reflector k + 1 from the new column, formed redundantly by every
//      workgroup (identical code and data, so identical bits everywhere), and
//      the products of step k + 1 on the updated rows.
// The published vectors are double-buffered by step parity, so one barrier per
// column suffices. Each workgroup raises its own arrival flag (one cache line
// per workgroup, no read-modify-write contention) and polls all flags with one
// wave, with a bounded wait: on expiry it raises an abort flag that every
// workgroup observes, and all workgroups leave the kernel.
// Every published value is stored and loaded with agent-scope relaxed atomics
// (sc1 accesses, coherent across the XCDs' L2 caches); a workgroup raises its
// flag only after its stores have completed (vmcnt(0)), and readers load the
// values only after they have seen every flag. This needs no L2 write-back or
// invalidate, which an agent-scope fence would issue on every column.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "tridiag.hpp"
#include "wave_sum.hpp"

namespace sdft_hip {
namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr long kSpinLimit = 1L << 20;
constexpr int kPerThread = (kHetrdMaxOrder + kThreads - 1) / kThreads;  // trailing entries per thread

// Aligned to 16 bytes: one element is one 128-bit LDS access (the dynamic LDS
// base is declared 16-byte aligned below, past the static variables).
struct alignas(16) Z {
  double x, y;
};
__device__ inline Z zmul(Z a, Z b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline Z zconj(Z a) { return {a.x, -a.y}; }
__device__ inline Z zadd(Z a, Z b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline Z zsub(Z a, Z b) { return {a.x - b.x, a.y - b.y}; }
__device__ inline Z zscale(Z a, double s) { return {a.x * s, a.y * s}; }

__device__ inline double wave_sum(double v) { return wave::sum(v); }

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

__device__ inline void store_published(Z* p, Z v) {
  double* q = reinterpret_cast<double*>(p);
  __hip_atomic_store(q, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline Z load_published(const Z* p) {
  const double* q = reinterpret_cast<const double*>(p);
  return {__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
          __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
}
// s_waitcnt vmcnt(0) with the other counters unconstrained (gfx9 encoding)
constexpr int kWaitStores = 0x0F70;

// Grid barrier with publication: the workgroup's staged values (count entries
// of `src`, LDS, to dst[index[j]]) are stored by as many threads, each waits
// for its stores to complete, and thread 0 raises the workgroup's arrival flag
// to `step` (one flag per workgroup, each on its own 64-byte line, so arrivals
// never contend). Thread q then polls the flag of workgroup q (every flag is
// watched at once, so the wait ends one memory round trip after the last
// arrival). Returns false when the wait expired or another workgroup gave up.
constexpr int kFlagStride = 16;  // unsigned ints per flag line
__device__ bool publish_and_wait(unsigned* flags, int* abort_flag, unsigned step, int count, const int* index,
                                 const Z* src_p, const Z* src_c, Z* dst_p, Z* dst_c) {
  __syncthreads();
  const int tid = threadIdx.x;
  for (int j = tid; j < count; j += kThreads) {
    store_published(dst_p + index[j], src_p[j]);
    store_published(dst_c + index[j], src_c[j]);
  }
  __builtin_amdgcn_s_waitcnt(kWaitStores);
  __syncthreads();
  if (tid == 0)
    __hip_atomic_store(flags + blockIdx.x * kFlagStride, step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int ok = 1;
  for (int q = tid; q < int(gridDim.x) && ok; q += kThreads) {
    for (long spin = 0;; ++spin) {
      if (__hip_atomic_load(flags + q * kFlagStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= step) break;
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
  }
  return __syncthreads_and(ok) != 0;
}

// Reflector of the column held in `col` at rows c0 + 1 .. n - 1 (LAPACK zlarfg
// conventions, m = n - c0 - 1 entries): writes v[0..m) (v[0] = 1) and, for
// workgroup 0, d[c0], e[c0], tau[c0] and the reflector tail. `xn2` is the sum
// of |col[j]|^2 over j >= c0 + 2. Returns tau. Ends with a block barrier.
__device__ Z form_reflector(const Z* col, int c0, int n, double xn2, Z* v, Z* __restrict__ refl, long ldr,
                            double* __restrict__ d, double* __restrict__ e, Z* __restrict__ tau) {
  const int tid = threadIdx.x, m = n - c0 - 1;
  const Z alpha = col[c0 + 1];
  Z t{0., 0.}, scale{0., 0.};
  double beta;
  if (xn2 == 0. && alpha.y == 0.) {
    // column already reduced: H = I - 2 e1 e1^T (beta = -alpha) instead of
    // LAPACK's tau = 0; rocSOLVER zunmtr returned NaN columns for reflector
    // sets with tau = 0 at orders of a few hundred (real tridiagonal inputs)
    beta = -alpha.x;
    t = {2., 0.};
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= 0. ? -nrm : nrm;
    t = {(beta - alpha.x) / beta, -alpha.y / beta};
    const Z den{alpha.x - beta, alpha.y};  // 1 / (alpha - beta)
    const double dd = den.x * den.x + den.y * den.y;
    scale = {den.x / dd, -den.y / dd};
  }
  for (int j = tid; j < m; j += kThreads) v[j] = j == 0 ? Z{1., 0.} : zmul(col[c0 + 1 + j], scale);
  if (blockIdx.x == 0) {
    if (tid == 0) {
      d[c0] = col[c0].x;
      e[c0] = beta;
      tau[c0] = t;
    }
    for (int j = 2 + tid; j <= m; j += kThreads) refl[(c0 + j) + size_t(c0) * ldr] = zmul(col[c0 + j], scale);
  }
  __syncthreads();
  return t;
}

// p_i = sum_j A_ij v_j over the trailing columns c0 + 1 .. n - 1 for the own
// rows i > c0 (one wave per row), staged in LDS with A_{i,c0+1}; returns the
// number of staged rows through `staged`. Ends with a block barrier.
__device__ void stage_products(const Z* rows, int n, int rpw, int c0, const Z* v, Z* stage_p, Z* stage_c,
                               int* stage_i, int* staged) {
  const int tid = threadIdx.x, lane = tid & 63, g = blockIdx.x, G = gridDim.x, m = n - c0 - 1;
  const int first = c0 + 1 - g <= 0 ? 0 : (c0 + 1 - g + G - 1) / G;
  for (int r = first + (tid >> 6); r < rpw; r += kWaves) {
    const int i = g + r * G;
    if (i >= n) break;
    const Z* row = rows + size_t(r) * n + (c0 + 1);
    Z acc{0., 0.};
    for (int j = lane; j < m; j += 64) acc = zadd(acc, zmul(row[j], v[j]));
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (lane == 0) {
      stage_p[r - first] = acc;
      stage_c[r - first] = row[0];
      stage_i[r - first] = i;
    }
  }
  if (tid == 0) {
    int cnt = 0;
    for (int r = first; r < rpw && g + r * G < n; ++r) ++cnt;
    *staged = cnt;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_hetrd_lds(const Z* __restrict__ a, long lda, int n, int rpw,
                                                        Z* __restrict__ refl, long ldr, double* __restrict__ d,
                                                        double* __restrict__ e, Z* __restrict__ tau,
                                                        Z* __restrict__ pub_p, Z* __restrict__ pub_c,
                                                        unsigned* __restrict__ flags, int* __restrict__ abort_flag) {
  extern __shared__ __attribute__((aligned(16))) double lds_raw[];
  Z* rows = reinterpret_cast<Z*>(lds_raw);  // rpw x n
  Z* vbuf = rows + size_t(rpw) * n;         // two reflectors (steps of both parities), n each
  Z* col = vbuf + 2 * size_t(n);            // current column, n
  double* scratch = reinterpret_cast<double*>(col + n);  // 2 * kWaves
  Z* stage_p = reinterpret_cast<Z*>(scratch + 2 * kWaves);  // rpw staged p_i
  Z* stage_c = stage_p + rpw;                               // rpw staged A_{i,k+1}
  int* stage_i = reinterpret_cast<int*>(stage_c + rpw);     // rpw row indices
  __shared__ int staged;

  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  // row i of the Hermitian input is the conjugate of column i (coalesced read)
  for (int r = 0; r < rpw; ++r) {
    const int i = g + r * G;
    if (i >= n) break;
    for (int j = tid; j < n; j += kThreads) rows[size_t(r) * n + j] = zconj(a[j + size_t(i) * lda]);
  }
  for (int j = tid; j < n; j += kThreads) col[j] = a[j];
  __syncthreads();
  // prologue: reflector 0 and the products of step 0
  Z t;
  {
    double part = 0.;
    for (int j = 2 + tid; j < n; j += kThreads) part += col[j].x * col[j].x + col[j].y * col[j].y;
    const double xn2 = block_sum(part, scratch);
    t = form_reflector(col, 0, n, xn2, vbuf, refl, ldr, d, e, tau);
    stage_products(rows, n, rpw, 0, vbuf, stage_p, stage_c, stage_i, &staged);
  }
  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;  // order of the trailing block
    Z* v = vbuf + size_t(k & 1) * n;
    Z* vn = vbuf + size_t((k + 1) & 1) * n;
    Z* pp = pub_p + size_t(k & 1) * n;
    Z* pc = pub_c + size_t(k & 1) * n;
    const int first = k + 1 - g <= 0 ? 0 : (k + 1 - g + G - 1) / G;
    // ---- 1. publication and grid barrier
    if (!publish_and_wait(flags, abort_flag, unsigned(k + 1), staged, stage_i, stage_p, stage_c, pp, pc)) return;
    // ---- 2. all loads issued at once (one round trip); partial (tau p)^H v
    Z pw[kPerThread], pcv[kPerThread];
    const Z p0 = load_published(pp + k + 1);
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const int j = tid + q * kThreads;
      if (j < m) {
        pw[q] = load_published(pp + k + 1 + j);
        pcv[q] = load_published(pc + k + 1 + j);
      }
    }
    Z dot{0., 0.};
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const int j = tid + q * kThreads;
      if (j < m) dot = zadd(dot, zmul(zconj(zmul(t, pw[q])), v[j]));
    }
    dot = block_zsum(dot, scratch);
    const Z a2 = zscale(zmul(t, dot), -0.5);
    // ---- 3. w = tau p + a2 v (registers); next column for all rows; rank-2
    // update of the own rows; |col|^2 of the next reflector's tail
    const Z w0c = zconj(zadd(zmul(t, p0), a2));  // v[0] = 1
    double part = 0.;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const int j = tid + q * kThreads;
      if (j < m) {
        const Z vj = v[j];
        const Z wj = zadd(zmul(t, pw[q]), zmul(a2, vj));
        const Z c = zsub(pcv[q], zadd(zmul(vj, w0c), wj));
        col[k + 1 + j] = c;
        if (j >= 2) part += c.x * c.x + c.y * c.y;
        pw[q] = zconj(wj);   // conj(w_j)
        pcv[q] = zconj(vj);  // conj(v_j)
      }
    }
    for (int r = first; r < rpw; ++r) {
      const int i = g + r * G;
      if (i >= n) break;
      const Z vi = v[i - k - 1];
      const Z wi = zadd(zmul(t, stage_p[r - first]), zmul(a2, vi));
      Z* row = rows + size_t(r) * n + (k + 1);
#pragma unroll
      for (int q = 0; q < kPerThread; ++q) {
        const int j = tid + q * kThreads;
        if (j < m) row[j] = zsub(row[j], zadd(zmul(vi, pw[q]), zmul(wi, pcv[q])));
      }
    }
    const double xn2 = block_sum(part, scratch);  // also publishes col and the rows to the block
    if (k + 2 < n) {
      // ---- 4. next reflector and the products of step k + 1
      t = form_reflector(col, k + 1, n, xn2, vn, refl, ldr, d, e, tau);
      stage_products(rows, n, rpw, k + 1, vn, stage_p, stage_c, stage_i, &staged);
    } else if (g == 0 && tid == 0) {
      d[n - 1] = col[n - 1].x;
    }
  }
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

size_t lds_bytes(int n, int rpw) {
  return (size_t(rpw) * n + 3 * size_t(n) + 2 * size_t(rpw)) * sizeof(Z) + 2 * kWaves * sizeof(double) +
         size_t(rpw) * sizeof(int);
}


// Target rows per workgroup: 4 (measured at n = 320 / 640 / 960 against 6, 10,
// 16 and as many as the LDS holds: 5.2 / 5.8 / 6.4 us per column, the largest
// slabs 8.0 / 6.7 / 6.4); SDFT_HIP_HETRD_ROWS overrides. More rows are used
// when the workgroups would outnumber the CUs.
int target_rows_per_group() {
  static const int rows = [] {
    const char* e = std::getenv("SDFT_HIP_HETRD_ROWS");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : 4;
  }();
  return rows;
}

// LDS the kernel declares statically (the workgroup reductions of the
// runtime included), which the dynamic slab must leave room for.
size_t static_lds() {
  static const size_t bytes = [] {
    hipFuncAttributes attr{};
    if (hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(k_hetrd_lds)) != hipSuccess) return size_t(4096);
    return size_t(attr.sharedSizeBytes);
  }();
  return bytes;
}
}  // namespace

bool hetrd_plan(int n, int* rows_per_group, int* groups) {
  if (n < 1 || n > kHetrdMaxOrder) return false;
  const int cus = device_cus();
  const size_t total = size_t(max_lds()), fixed = static_lds();
  if (fixed >= total) return false;
  const size_t budget = total - fixed;
  // the target rows per workgroup, or more so that there is at most one
  // workgroup per CU
  int rpw = std::min(n, target_rows_per_group());
  if (cus > 0) rpw = std::max(rpw, (n + cus - 1) / cus);
  if (lds_bytes(n, rpw) > budget) return false;
  const int g = (n + rpw - 1) / rpw;
  if (cus > 0 && g > cus) return false;
  // every workgroup must be resident at once (the grid barrier waits for all)
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_hetrd_lds), kThreads,
                                                   lds_bytes(n, rpw)) != hipSuccess ||
      per_cu < 1 || (cus > 0 && long(per_cu) * cus < g))
    return false;
  *rows_per_group = rpw;
  *groups = g;
  return true;
}

// published vectors (4 n complex), then the abort flag line, then one flag line
// per workgroup (at most kHetrdMaxOrder workgroups)
size_t hetrd_workspace_bytes(int n) {
  return 4 * size_t(n) * sizeof(Z) + size_t(kFlagStride) * sizeof(unsigned) * (1 + size_t(kHetrdMaxOrder));
}

hipError_t launch_hetrd(const double* a, long lda, int n, double* refl, long ldr, double* d, double* e, double* tau,
                        void* workspace, hipStream_t s, bool force_expiry) {
  int rpw = 0, groups = 0;
  if (!hetrd_plan(n, &rpw, &groups)) return hipErrorInvalidValue;
  if (n == 1) {
    // d = Re a; nothing to reduce
    return hipMemcpyAsync(d, a, sizeof(double), hipMemcpyDeviceToDevice, s);
  }
  Z* pub_p = reinterpret_cast<Z*>(workspace);
  Z* pub_c = pub_p + 2 * size_t(n);
  int* abort_flag = reinterpret_cast<int*>(pub_c + 2 * size_t(n));
  unsigned* flags = reinterpret_cast<unsigned*>(abort_flag) + kFlagStride;
  hipError_t err = hipMemsetAsync(abort_flag, 0, size_t(kFlagStride) * sizeof(unsigned) * (1 + size_t(groups)), s);
  if (err != hipSuccess) return err;
  // host-forced expiry (fault injection): the abort flag is raised before the
  // launch, so every workgroup leaves at its first wait and the kernel ends
  if (force_expiry && (err = hipMemsetAsync(abort_flag, 1, 1, s)) != hipSuccess) return err;
  const Z* az = reinterpret_cast<const Z*>(a);
  Z* rz = reinterpret_cast<Z*>(refl);
  Z* tz = reinterpret_cast<Z*>(tau);
  // A plain launch: the plan keeps the grid within one workgroup per CU, so
  // on an idle device every workgroup is resident; when other work (another
  // stream or process) holds CUs, the bounded wait expires and the caller
  // redoes the solve by divide and conquer. A cooperative launch would
  // instead serialise the device for the reduction, and it is refused when
  // several processes share the device.
  hipLaunchKernelGGL(k_hetrd_lds, dim3(groups), dim3(kThreads), lds_bytes(n, rpw), s, az, lda, n, rpw, rz, ldr, d, e,
                     tz, pub_p, pub_c, flags, abort_flag);
  return hipGetLastError();
}

const int* hetrd_abort_flag(const void* workspace, int n) {
  const Z* pub = reinterpret_cast<const Z*>(workspace);
  return reinterpret_cast<const int*>(pub + 4 * size_t(n));
}


// Loads this file's code object on the calling thread (see module_prime.hpp).
hipError_t prime_module_tridiag() {
  hipFuncAttributes a{};
  return hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_hetrd_lds));
}

}  // namespace sdft_hip
