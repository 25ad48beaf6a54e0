// Fused persistent H psi / S psi of a batch of plane-wave operators sharing one
// FFT grid shape (the k-points of one crystal: uneven spheres, k-dependent
// projector phases, one grid). A launch holds `groups` workgroups; workgroup w
// takes the work items w, w + groups, ... of the (operator, column) list, and
// applies the whole operator to its column in LDS (the restatement of
// src/plane_wave_operator.cpp:229-305 per column):
//
//   1. psi into LDS, scattered into the (x, y) sticks of the operator's
//      sphere; z-FFT of the sticks (natural -> slot order);
//   2. per z slot (`planes` at a time): xy-plane, y- and x-FFT to slot order,
//      x V(r) (slot order), x- and y-FFT back, sticks gathered back;
//   3. z-FFT back to natural order;
//   4. projections P = B psi (lanes over projector rows, coalesced along the
//      rows of the column-major B), onsite D P and Q P per atom;
//   5. H psi = T psi + local / N + B^H (D P) and S psi = (1 + reg) psi + B^H (Q P)
//      (lanes over plane waves, coalesced along the transposed copy of B).
//
// No intermediate leaves the CU: per column the HBM traffic is psi in, H psi
// and S psi out, and the projector rows (L2-resident across the columns of
// one operator); V(r) and the twiddles are shared by every work item. The
// static schedule has no counters, so the launch can be captured once in a
// HIP graph and replayed.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "fft_device.hpp"
#include "fft_dispatch.hpp"

namespace sdft_hip {
namespace {

template <typename T>
__device__ __forceinline__ cx<T> load_cx(const double* p) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  return {T(v.x), T(v.y)};
}

// a * conj(b) accumulated: acc += conj(b) * a
template <typename T>
__device__ __forceinline__ void acc_conj(cx<T>& acc, cx<T> b, cx<T> a) {
  acc.x += b.x * a.x + b.y * a.y;
  acc.y += b.x * a.y - b.y * a.x;
}

// LDS layout of one workgroup (all cx<T>): sticks, planes, psi, twiddles
// (z, x, y), projections P, D P, Q P.
struct BatchLds {
  int sticks, planes, psi, tw, proj;
  __host__ __device__ BatchLds(const BatchLaunch& l) {
    sticks = 0;
    planes = sticks + l.nsticks_max * line_stride(l.n2);
    psi = planes + l.planes * l.n0 * line_stride(l.n1);
    tw = psi + l.n_max;
    proj = tw + l.n0 + l.n1 + l.n2;
  }
  __host__ __device__ int total(const BatchLaunch& l) const { return proj + 3 * l.np_max; }
};

template <typename T, int NT, int PB, int LEN>
__global__ __launch_bounds__(NT) void k_batch_apply(const BatchLaunch l, const BatchOpDev* __restrict__ ops,
                                                    const double* __restrict__ x, double* __restrict__ hy,
                                                    double* __restrict__ sy) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const BatchLds lay(l);
  cx<T>* sticks = lds + lay.sticks;
  cx<T>* planes = lds + lay.planes;
  cx<T>* psi = lds + lay.psi;
  cx<T>* twz = lds + lay.tw;
  cx<T>* twx = twz + l.n2;
  cx<T>* twy = twx + l.n0;
  cx<T>* P = lds + lay.proj;
  cx<T>* DP = P + l.np_max;
  cx<T>* QP = DP + l.np_max;
  const int tid = threadIdx.x;
  constexpr int kWaves = NT / 64;
  const int n0 = l.n0, n1 = l.n1, n2 = l.n2;
  const int ls1 = line_stride(n1), ls2 = line_stride(n2), plane_size = n0 * ls1;
  for (int e = tid; e < n2; e += NT) twz[e] = reinterpret_cast<const cx<T>*>(l.tw[2])[e];
  for (int e = tid; e < n0; e += NT) twx[e] = reinterpret_cast<const cx<T>*>(l.tw[0])[e];
  for (int e = tid; e < n1; e += NT) twy[e] = reinterpret_cast<const cx<T>*>(l.tw[1])[e];
  const double count = double(n0) * n1 * n2;
  const int items = l.ops * l.columns;
  for (int item = blockIdx.x; item < items; item += gridDim.x) {
    const int k = item / l.columns, c = item - k * l.columns;
    const BatchOpDev op = ops[k];
    const int n = op.n, ns = op.nsticks, np = op.np;
    const double* xc = x + 2 * (op.x_offset + (long)c * n);
    // 1. psi and the sticks; z-FFT to slots
    for (int g = tid; g < n; g += NT) psi[g] = load_cx<T>(xc + 2L * g);
    __syncthreads();
    for (int e = tid; e < ns * n2; e += NT) {
      const int s = e / n2, z = e - s * n2;
      const int g = op.stick_g[e];
      sticks[s * ls2 + z] = g >= 0 ? psi[g] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    LineGeom gz{ns, ls2, 1, false};
    gz.fast_group = true;
    line_fft<T, +1, true, LEN>(n2, sticks, gz, twz, tid, NT);
    // 2. xy-planes, PB z slots at a time
    const T* v = reinterpret_cast<const T*>(op.v);
    for (int z0 = 0; z0 < n2; z0 += PB) {
      const int pb = min(PB, n2 - z0);
      for (int e = tid; e < pb * plane_size; e += NT) planes[e] = {T(0), T(0)};
      __syncthreads();
      for (int e = tid; e < pb * ns; e += NT) {
        const int p = e / ns, s = e - p * ns;
        const int sx = op.stick_xy[s];
        const int xx = sx / n1, yy = sx - xx * n1;
        planes[p * plane_size + xx * ls1 + yy] = sticks[s * ls2 + z0 + p];
      }
      __syncthreads();
      LineGeom gy{pb * n0, ls1, 1, false};
      gy.fast_group = true;
      LineGeom gx{pb * n1, 1, ls1, true};
      gx.split = n1;
      gx.jump = plane_size - n1;
      line_fft<T, +1, true, LEN>(n1, planes, gy, twy, tid, NT);
      line_fft<T, +1, true, LEN>(n0, planes, gx, twx, tid, NT);
      for (int e = tid; e < pb * n0 * n1; e += NT) {
        const int p = e / (n0 * n1), r = e - p * (n0 * n1);
        const int xs = r / n1, ys = r - xs * n1;
        const T w = v[(long)(z0 + p) * n0 * n1 + r];
        cx<T>& q = planes[p * plane_size + xs * ls1 + ys];
        q.x *= w;
        q.y *= w;
      }
      __syncthreads();
      line_fft<T, -1, false, LEN>(n0, planes, gx, twx, tid, NT);
      line_fft<T, -1, false, LEN>(n1, planes, gy, twy, tid, NT);
      for (int e = tid; e < pb * ns; e += NT) {
        const int p = e / ns, s = e - p * ns;
        const int sx = op.stick_xy[s];
        const int xx = sx / n1, yy = sx - xx * n1;
        sticks[s * ls2 + z0 + p] = planes[p * plane_size + xx * ls1 + yy];
      }
      __syncthreads();
    }
    // 3. z-FFT back to natural order
    line_fft<T, -1, false, LEN>(n2, sticks, gz, twz, tid, NT);
    // 4. projections: wave w sums plane waves g = w, w + waves, ...; lane = projector row
    const cx<T>* b = reinterpret_cast<const cx<T>*>(op.b);
    cx<T>* red = planes;  // the planes are free now: kWaves x 64 partial sums
    const int wave = tid >> 6, lane = tid & 63;
    for (int i0 = 0; i0 < np; i0 += 64) {
      const int i = i0 + lane;
      cx<T> acc{T(0), T(0)};
      if (i < np) {
        // four plane waves per step: their projector rows are loaded together
        int g = wave;
        for (; g + 3 * kWaves < n; g += 4 * kWaves) {
          cx<T> bi[4], pg[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) bi[u] = b[i + (long)(g + u * kWaves) * np];
#pragma unroll
          for (int u = 0; u < 4; ++u) pg[u] = psi[g + u * kWaves];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            acc.x += bi[u].x * pg[u].x - bi[u].y * pg[u].y;
            acc.y += bi[u].x * pg[u].y + bi[u].y * pg[u].x;
          }
        }
        for (; g < n; g += kWaves) {
          const cx<T> bi = b[i + (long)g * np];
          const cx<T> pg = psi[g];
          acc.x += bi.x * pg.x - bi.y * pg.y;
          acc.y += bi.x * pg.y + bi.y * pg.x;
        }
      }
      red[wave * 64 + lane] = acc;
      __syncthreads();
      if (tid < 64 && i0 + tid < np) {
        cx<T> sum{T(0), T(0)};
        for (int w = 0; w < kWaves; ++w) sum = cadd(sum, red[w * 64 + tid]);
        P[i0 + tid] = sum;
      }
      __syncthreads();
    }
    // onsite blocks: row i of atom a, D_a and Q_a column-major rows x rows
    for (int i = tid; i < np; i += NT) {
      const int a = op.proj_atom[i], first = op.atom_first[a], rows = op.atom_rows[a], li = i - first;
      const cx<double>* d = reinterpret_cast<const cx<double>*>(op.dblk) + op.block_offset[a];
      const cx<double>* q = reinterpret_cast<const cx<double>*>(op.qblk) + op.block_offset[a];
      cx<T> dp{T(0), T(0)}, qp{T(0), T(0)};
      for (int j = 0; j < rows; ++j) {
        const cx<T> pj = P[first + j];
        const cx<double> dj = d[li + j * rows], qj = q[li + j * rows];
        dp.x += T(dj.x) * pj.x - T(dj.y) * pj.y;
        dp.y += T(dj.x) * pj.y + T(dj.y) * pj.x;
        qp.x += T(qj.x) * pj.x - T(qj.y) * pj.y;
        qp.y += T(qj.x) * pj.y + T(qj.y) * pj.x;
      }
      DP[i] = dp;
      QP[i] = qp;
    }
    __syncthreads();
    // 5. outputs: lane = plane wave (coalesced over the transposed projectors)
    const cx<T>* bt = reinterpret_cast<const cx<T>*>(op.bt);
    double* hc = hy + 2 * (op.x_offset + (long)c * n);
    double* sc = sy + 2 * (op.x_offset + (long)c * n);
    for (int g = tid; g < n; g += NT) {
      cx<T> ah{T(0), T(0)}, as{T(0), T(0)};
      int i = 0;
      for (; i + 4 <= np; i += 4) {
        cx<T> bg[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) bg[u] = bt[g + (long)(i + u) * n];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          acc_conj(ah, bg[u], DP[i + u]);
          acc_conj(as, bg[u], QP[i + u]);
        }
      }
      for (; i < np; ++i) {
        const cx<T> bg = bt[g + (long)i * n];
        acc_conj(ah, bg, DP[i]);
        acc_conj(as, bg, QP[i]);
      }
      const double2 xg = *reinterpret_cast<const double2*>(xc + 2L * g);
      const cx<T> loc = sticks[op.inverse[g]];
      const double t = op.kinetic[g];
      double2 h, s;
      h.x = t * xg.x + double(loc.x) / count + double(ah.x);
      h.y = t * xg.y + double(loc.y) / count + double(ah.y);
      s.x = op.s_diag * xg.x + double(as.x);
      s.y = op.s_diag * xg.y + double(as.y);
      *reinterpret_cast<double2*>(hc + 2L * g) = h;
      *reinterpret_cast<double2*>(sc + 2L * g) = s;
    }
    __syncthreads();
  }
}

__global__ void k_transpose_projectors(const double* __restrict__ in, int np, int n, double* __restrict__ out) {
  const long total = (long)np * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long g = e % n, i = e / n;
    out[2 * e] = in[2 * (i + g * np)];
    out[2 * e + 1] = in[2 * (i + g * np) + 1];
  }
}

__global__ void k_transpose_projectors32(const float* __restrict__ in, int np, int n, float* __restrict__ out) {
  const long total = (long)np * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long g = e % n, i = e / n;
    out[2 * e] = in[2 * (i + g * np)];
    out[2 * e + 1] = in[2 * (i + g * np) + 1];
  }
}

template <typename T, int NT, int PB, int LEN>
hipError_t launch_typed(const BatchLaunch& l, const BatchOpDev* ops, const double* x, double* hy, double* sy,
                        hipStream_t s) {
  const size_t lds = size_t(BatchLds(l).total(l)) * sizeof(cx<T>);
  hipLaunchKernelGGL((k_batch_apply<T, NT, PB, LEN>), dim3(l.groups), dim3(NT), lds, s, l, ops, x, hy, sy);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_threads(const BatchLaunch& l, const BatchOpDev* ops, const double* x, double* hy, double* sy,
                          hipStream_t s) {
  if (!batch_fixed_length(l)) {
    if (l.threads != 256) return hipErrorInvalidValue;
    return l.planes == 2 ? launch_typed<T, 256, 2, 0>(l, ops, x, hy, sy, s)
                         : launch_typed<T, 256, 1, 0>(l, ops, x, hy, sy, s);
  }
  switch (l.threads * 4 + l.planes) {
    case 256 * 4 + 1: return launch_typed<T, 256, 1, 32>(l, ops, x, hy, sy, s);
    case 256 * 4 + 2: return launch_typed<T, 256, 2, 32>(l, ops, x, hy, sy, s);
    case 512 * 4 + 1: return launch_typed<T, 512, 1, 32>(l, ops, x, hy, sy, s);
    case 512 * 4 + 2: return launch_typed<T, 512, 2, 32>(l, ops, x, hy, sy, s);
    case 1024 * 4 + 1: return launch_typed<T, 1024, 1, 32>(l, ops, x, hy, sy, s);
    default: return launch_typed<T, 1024, 2, 32>(l, ops, x, hy, sy, s);
  }
}

}  // namespace

bool batch_fixed_length(const BatchLaunch& l) { return l.n0 == 32 && l.n1 == 32 && l.n2 == 32; }

size_t batch_lds_bytes(const BatchLaunch& l, bool fp64) {
  if (l.n0 <= 0 || l.n1 <= 0 || l.n2 <= 0 || (l.planes != 1 && l.planes != 2)) return 0;
  // the projection reduction reuses the planes: waves x 64 partial sums
  if (l.planes * l.n0 * line_stride(l.n1) < l.threads) return 0;
  const size_t bytes = size_t(BatchLds(l).total(l)) * (fp64 ? 16 : 8);
  int dev = 0;
  hipDeviceProp_t p{};
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 0;
  return bytes <= size_t(p.sharedMemPerBlock) ? bytes : 0;
}

hipError_t launch_batch_apply(const BatchLaunch& l, const BatchOpDev* ops, bool fp64, const double* x, double* hy,
                              double* sy, hipStream_t s) {
  if (l.groups <= 0 || (l.threads != 256 && l.threads != 512 && l.threads != 1024) || (l.planes != 1 && l.planes != 2))
    return hipErrorInvalidValue;
  return fp64 ? launch_threads<double>(l, ops, x, hy, sy, s) : launch_threads<float>(l, ops, x, hy, sy, s);
}

hipError_t launch_transpose_projectors(const void* in, int np, int n, bool fp64, void* out, hipStream_t s) {
  const long total = (long)np * n;
  if (total <= 0) return hipSuccess;
  const int blocks = int(total / 256 + 1 > 4096 ? 4096 : total / 256 + 1);
  if (fp64)
    hipLaunchKernelGGL(k_transpose_projectors, dim3(blocks), dim3(256), 0, s, static_cast<const double*>(in), np, n,
                       static_cast<double*>(out));
  else
    hipLaunchKernelGGL(k_transpose_projectors32, dim3(blocks), dim3(256), 0, s, static_cast<const float*>(in), np, n,
                       static_cast<float*>(out));
  return hipGetLastError();
}


// Loads this file's code object on the calling thread (see module_prime.hpp).
hipError_t prime_module_batch() {
  hipFuncAttributes a{};
  return hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_transpose_projectors));
}

}  // namespace sdft_hip
