// Local Hamiltonian H_loc psi = T psi + FFT_fwd(V(r) . FFT_bwd(psi)) / N on
// gfx950, restating src/plane_wave_operator.cpp:229-258 (and the CUDA chain of
// src/compute_cuda.cu:141-177) as three launches that touch only the sphere:
//
//   A  k_sticks_backward : gather psi into the (x,y) sticks of the G-sphere,
//                          z-FFT (natural -> slot order), store stick buffer
//                          T1[column][z slot][stick]  (coalesced over sticks)
//   B  k_planes          : per (z slot, column) xy-plane in LDS: scatter the
//                          sticks, y- and x-FFT to slots, x V(r) in slot
//                          order, x- and y-FFT back to natural order, keep
//                          only the sticks -> T2
//   C  k_sticks_forward  : z-FFT from slots to natural order, gather onto the
//                          sphere, add the kinetic term and divide by N
//
// The full grid never exists in HBM; per column the traffic is the two stick
// buffers (nsticks * N2 values each) plus psi in/out.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "fft_device.hpp"
#include "fft_dispatch.hpp"
#include "fft_lengths.hpp"
#include "local_h.hpp"

namespace sdft_hip {

namespace {

template <typename T>
__device__ __forceinline__ cx<T> load_cx(const double* p) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  return {T(v.x), T(v.y)};
}

}  // namespace

// Elements of a stick workgroup per thread whose loads are issued together
// (sticks_per_block keeps a workgroup at <= 2048 elements, 8 per thread of
// 256): the sphere indices first, then every gather they address, so a thread
// waits for two memory round trips instead of two per element.
constexpr int kGatherPer = 8;

// A: grid (ceil(nsticks / spb), columns)
template <typename T>
__global__ __launch_bounds__(256) void k_sticks_backward(const double* __restrict__ x, long ldx,
                                                         const int* __restrict__ stick_g, int nsticks, int n2, int spb,
                                                         const cx<T>* __restrict__ tw_z, cx<T>* __restrict__ t1) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lines = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int first = blockIdx.x * spb;
  const int nl = min(spb, nsticks - first);
  const long col = blockIdx.y;
  const double* xc = x + 2 * col * ldx;
  const int ls = line_stride(n2);
  const int total = nl * n2;
  const int* sg = stick_g + (long)first * n2;
  for (int e0 = 0; e0 < total; e0 += kGatherPer * nt) {
    int g[kGatherPer];
#pragma unroll
    for (int u = 0; u < kGatherPer; ++u) {
      const int e = e0 + tid + u * nt;
      g[u] = e < total ? sg[e] : -1;
    }
    cx<T> v[kGatherPer];
#pragma unroll
    for (int u = 0; u < kGatherPer; ++u) v[u] = g[u] >= 0 ? load_cx<T>(xc + 2L * g[u]) : cx<T>{T(0), T(0)};
#pragma unroll
    for (int u = 0; u < kGatherPer; ++u) {
      const int e = e0 + tid + u * nt;
      if (e < total) {
        const int s = e / n2, z = e - s * n2;
        lines[s * ls + z] = v[u];
      }
    }
  }
  __syncthreads();
  LineGeom geom{nl, ls, 1, false};
  geom.fast_group = true;
  line_fft_dispatch<T, +1, true>(n2, lines, geom, tw_z, tid, nt);
  cx<T>* out = t1 + col * (long)n2 * nsticks;
  for (int e = tid; e < nl * n2; e += nt) {
    const int zs = e / nl, s = e - zs * nl;
    out[(long)zs * nsticks + first + s] = lines[s * ls + zs];
  }
}

// B: grid (n2 slots, columns). mode 0: local H (bwd, x V, fwd, keep sticks);
// mode 1: backward only, whole plane stored in slot order (potential setup).
// In mode 0 the y-transforms run only on the x-lines that hold sticks (the
// backward input and the forward output are zero / unused elsewhere).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_planes(const cx<T>* __restrict__ t1, cx<T>* __restrict__ t2,
                                                          const int* __restrict__ stick_xy, int nsticks, int n0,
                                                          int n1, int n2, const cx<T>* __restrict__ tw_x,
                                                          const cx<T>* __restrict__ tw_y,
                                                          const T* __restrict__ v_slots, int mode, int xs0, int xl0,
                                                          int xs1, int xl1) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* plane = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int zs = blockIdx.x;
  const long col = blockIdx.y;
  const int np = n0 * n1;
  // plane rows (fixed x) of odd stride n1p in LDS: element (x, y) at x n1p + y
  const int n1p = line_stride(n1), pad = n1p - n1;
  // twiddles of both line lengths staged in LDS after the plane
  cx<T>* twx = plane + n0 * n1p;
  cx<T>* twy = twx + n0;
  for (int e = tid; e < n0; e += nt) twx[e] = tw_x[e];
  for (int e = tid; e < n1; e += nt) twy[e] = tw_y[e];
  // y lines: one per x, contiguous elements; the two ranges of active x run
  // as one set of lines (x = xs0 .. xs0 + xl0 - 1, then xs1 .. xs1 + xl1 - 1)
  LineGeom gy{xl0 + xl1, n1p, 1, false};
  gy.fast_group = true;
  gy.split = xl0;
  gy.jump = (xs1 - xs0 - xl0) * n1p;
  cx<T>* ybase = plane + (xl0 ? xs0 : xs1) * n1p;
  if (!xl0) gy.split = 1 << 30;
  // x lines: one per y, elements strided by n1p
  const LineGeom gx{n1, 1, n1p, true};
  if (mode == 2) {
    // start from a full slot-order plane (real-space search images)
    const cx<T>* inp = t1 + (col * n2 + zs) * (long)np;
    for (int e = tid; e < np; e += nt) plane[e + (e / n1) * pad] = inp[e];
    __syncthreads();
  } else {
    for (int e = tid; e < n0 * n1p; e += nt) plane[e] = {T(0), T(0)};
    __syncthreads();
    const cx<T>* in = t1 + (col * n2 + zs) * (long)nsticks;
    for (int s = tid; s < nsticks; s += nt) {
      const int sx = stick_xy[s];
      plane[sx + (sx / n1) * pad] = in[s];
    }
    __syncthreads();
    if (xl0 + xl1) line_fft_dispatch<T, +1, true>(n1, ybase, gy, twy, tid, nt);
    line_fft_dispatch<T, +1, true>(n0, plane, gx, twx, tid, nt);
    if (mode == 1) {
      cx<T>* out = t2 + (col * n2 + zs) * (long)np;
      for (int e = tid; e < np; e += nt) out[e] = plane[e + (e / n1) * pad];
      return;
    }
    const T* v = v_slots + (long)zs * np;
    for (int e = tid; e < np; e += nt) {
      const T w = v[e];
      cx<T>& q = plane[e + (e / n1) * pad];
      q.x *= w;
      q.y *= w;
    }
    __syncthreads();
  }
  line_fft_dispatch<T, -1, false>(n0, plane, gx, twx, tid, nt);
  if (xl0 + xl1) line_fft_dispatch<T, -1, false>(n1, ybase, gy, twy, tid, nt);
  cx<T>* out = t2 + (col * n2 + zs) * (long)nsticks;
  for (int s = tid; s < nsticks; s += nt) {
    const int sx = stick_xy[s];
    out[s] = plane[sx + (sx / n1) * pad];
  }
}

// LDS of one plane workgroup: the plane and the x and y twiddles.
template <typename T>
size_t plane_lds_bytes(int n0, int n1) {
  return (size_t(n0) * line_stride(n1) + n0 + n1) * sizeof(cx<T>);
}

// Threads per plane workgroup (one xy-plane per workgroup): 1024 for FP64
// (one workgroup per CU, so more waves hide more latency) and 512 for FP32
// (several workgroups per CU), measured; SDFT_HIP_PLANE_THREADS (256, 512 or
// 1024) overrides both.
int plane_threads(bool fp64) {
  static const int forced = [] {
    const char* e = std::getenv("SDFT_HIP_PLANE_THREADS");
    const int t = e ? std::atoi(e) : 0;
    return t == 256 || t == 512 || t == 1024 ? t : 0;
  }();
  return forced ? forced : (fp64 ? 1024 : 512);
}

template <typename T, typename... Args>
void launch_k_planes(dim3 grid, size_t lds, hipStream_t stream, Args... args) {
  switch (plane_threads(sizeof(T) == 8)) {
    case 256:
      hipLaunchKernelGGL((k_planes<T, 256>), grid, dim3(256), lds, stream, args...);
      break;
    case 1024:
      hipLaunchKernelGGL((k_planes<T, 1024>), grid, dim3(1024), lds, stream, args...);
      break;
    default:
      hipLaunchKernelGGL((k_planes<T, 512>), grid, dim3(512), lds, stream, args...);
  }
}

// C: grid (ceil(nsticks / spb), columns). y = T x + local / N
template <typename T>
__global__ __launch_bounds__(256) void k_sticks_forward(const cx<T>* __restrict__ t2, const int* __restrict__ stick_g,
                                                        int nsticks, int n2, int spb, const cx<T>* __restrict__ tw_z,
                                                        const double* __restrict__ x, long ldx,
                                                        const double* __restrict__ kinetic, double* __restrict__ y,
                                                        long ldy, double inv_count_unused, double count) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lines = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int first = blockIdx.x * spb;
  const int nl = min(spb, nsticks - first);
  const long col = blockIdx.y;
  const cx<T>* in = t2 + col * (long)n2 * nsticks;
  const int ls = line_stride(n2);
  for (int e = tid; e < nl * n2; e += nt) {
    const int zs = e / nl, s = e - zs * nl;
    lines[s * ls + zs] = in[(long)zs * nsticks + first + s];
  }
  __syncthreads();
  LineGeom geom{nl, ls, 1, false};
  geom.fast_group = true;
  line_fft_dispatch<T, -1, false>(n2, lines, geom, tw_z, tid, nt);
  const double* xc = x + 2 * col * ldx;
  double* yc = y + 2 * col * ldy;
  const int total = nl * n2;
  const int* sg = stick_g + (long)first * n2;
  for (int e0 = 0; e0 < total; e0 += kGatherPer * nt) {
    int g[kGatherPer];
#pragma unroll
    for (int u = 0; u < kGatherPer; ++u) {
      const int e = e0 + tid + u * nt;
      g[u] = e < total ? sg[e] : -1;
    }
    double t[kGatherPer];
    double2 xi[kGatherPer];
#pragma unroll
    for (int u = 0; u < kGatherPer; ++u) {
      if (g[u] >= 0) {
        t[u] = kinetic[g[u]];
        xi[u] = *reinterpret_cast<const double2*>(xc + 2L * g[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kGatherPer; ++u) {
      if (g[u] < 0) continue;
      const int e = e0 + tid + u * nt;
      const int s = e / n2, z = e - s * n2;
      const cx<T> w = lines[s * ls + z];
      double2 r;
      r.x = t[u] * xi[u].x + double(w.x) / count;
      r.y = t[u] * xi[u].y + double(w.y) / count;
      *reinterpret_cast<double2*>(yc + 2L * g[u]) = r;
    }
  }
}

// S local part: y = scale * x (used when the operator has no projectors)
__global__ void k_scale_copy(const double* __restrict__ x, long ldx, double* __restrict__ y, long ldy, long rows,
                             double scale) {
  const long col = blockIdx.y;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < 2 * rows; i += (long)gridDim.x * blockDim.x)
    y[2 * col * ldy + i] = scale * x[2 * col * ldx + i];
}

// y(:, j) += scale * theta[j] * x(:, j), complex columns of `rows` entries.
__global__ void k_column_scaled_add(const double* __restrict__ x, long ldx, double* __restrict__ y, long ldy,
                                    long rows, const double* __restrict__ theta, double scale) {
  const long col = blockIdx.y;
  const double f = scale * theta[col];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < 2 * rows; i += (long)gridDim.x * blockDim.x)
    y[2 * col * ldy + i] += f * x[2 * col * ldx + i];
}

// Potential setup: natural-order complex grid -> real slot-order V (FP64 and
// FP32 copies) plus the largest |Im| / max(1, |Re|) ratio (block maxima).
__global__ void k_potential_finish(const cx<double>* __restrict__ planes, long count, double* __restrict__ v64,
                                   float* __restrict__ v32, double* __restrict__ block_ratio) {
  __shared__ double red[256];
  double worst = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    const cx<double> v = planes[i];
    v64[i] = v.x;
    v32[i] = float(v.x);
    worst = fmax(worst, fabs(v.y) / fmax(1.0, fabs(v.x)));
  }
  red[threadIdx.x] = worst;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) block_ratio[blockIdx.x] = red[0];
}

// natural real V(r) -> slot order (drop-in path that receives V(r) directly)
__global__ void k_potential_to_slots(const double* __restrict__ v_nat, int n0, int n1, int n2, int a0, int a1, int a2,
                                     double* __restrict__ v64, float* __restrict__ v32) {
  const long count = (long)n0 * n1 * n2;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    // slot layout [zs][xs][ys]
    const int ys = int(i % n1);
    const int xs = int((i / n1) % n0);
    const int zs = int(i / ((long)n0 * n1));
    const int x = slot_position(xs, a0, n0 / a0);
    const int y = slot_position(ys, a1, n1 / a1);
    const int z = slot_position(zs, a2, n2 / a2);
    const double v = v_nat[((long)x * n1 + y) * n2 + z];
    v64[i] = v;
    v32[i] = float(v);
  }
}

// slot order -> natural real V(r)
__global__ void k_potential_to_natural(const double* __restrict__ v64, int n0, int n1, int n2, int a0, int a1, int a2,
                                       double* __restrict__ v_nat) {
  const long count = (long)n0 * n1 * n2;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    const int ys = int(i % n1);
    const int xs = int((i / n1) % n0);
    const int zs = int(i / ((long)n0 * n1));
    const int x = slot_position(xs, a0, n0 / a0);
    const int y = slot_position(ys, a1, n1 / a1);
    const int z = slot_position(zs, a2, n2 / a2);
    v_nat[((long)x * n1 + y) * n2 + z] = v64[i];
  }
}

// ------------------------------------------------------------ host launchers
namespace {
int sticks_per_block(int n2) {
  int spb = 2048 / n2;
  if (spb < 1) spb = 1;
  if (spb > 64) spb = 64;
  return spb;
}
}  // namespace

int max_lds_bytes() {
  static int bytes = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
    return v;
  }();
  return bytes;
}

template <typename T>
hipError_t launch_local_h(const GridPlanDev& p, const double* x, long ldx, long columns, const double* kinetic,
                          double* y, long ldy, void* t1, void* t2, hipStream_t stream) {
  if (columns <= 0) return hipSuccess;
  // one xy-plane per workgroup must fit the LDS of one workgroup
  if (plane_lds_bytes<T>(p.n0, p.n1) > size_t(max_lds_bytes())) return hipErrorInvalidConfiguration;
  const int spb = sticks_per_block(p.n2);
  const dim3 ga((p.nsticks + spb - 1) / spb, (unsigned)columns);
  const size_t lds_lines = size_t(spb) * line_stride(p.n2) * sizeof(cx<T>);
  const cx<T>* twx = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[0] : p.tw32[0]);
  const cx<T>* twy = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[1] : p.tw32[1]);
  const cx<T>* twz = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[2] : p.tw32[2]);
  const T* v = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)p.v64 : (const void*)p.v32);
  hipLaunchKernelGGL(k_sticks_backward<T>, ga, dim3(256), lds_lines, stream, x, ldx, p.stick_g, p.nsticks, p.n2, spb,
                     twz, reinterpret_cast<cx<T>*>(t1));
  const size_t lds_plane = plane_lds_bytes<T>(p.n0, p.n1);
  launch_k_planes<T>(dim3(p.n2, (unsigned)columns), lds_plane, stream,
                     reinterpret_cast<const cx<T>*>(t1), reinterpret_cast<cx<T>*>(t2), p.stick_xy, p.nsticks, p.n0,
                     p.n1, p.n2, twx, twy, v, 0, p.xstart[0], p.xlines[0], p.xstart[1], p.xlines[1]);
  hipLaunchKernelGGL(k_sticks_forward<T>, ga, dim3(256), lds_lines, stream, reinterpret_cast<const cx<T>*>(t2),
                     p.stick_g, p.nsticks, p.n2, spb, twz, x, ldx, kinetic, y, ldy, 0.0, double(p.count));
  return hipGetLastError();
}

template hipError_t launch_local_h<double>(const GridPlanDev&, const double*, long, long, const double*, double*, long,
                                           void*, void*, hipStream_t);
template hipError_t launch_local_h<float>(const GridPlanDev&, const double*, long, long, const double*, double*, long,
                                          void*, void*, hipStream_t);

hipError_t launch_column_scaled_add(const double* x, long ldx, double* y, long ldy, long rows, long columns,
                                    const double* theta, double scale, hipStream_t stream) {
  if (rows <= 0 || columns <= 0) return hipSuccess;
  if (columns > 65535) return hipErrorInvalidValue;
  const long blocks = (2 * rows + 255) / 256;
  const unsigned bx = unsigned(blocks < 256 ? blocks : 256);
  hipLaunchKernelGGL(k_column_scaled_add, dim3(bx, unsigned(columns)), dim3(256), 0, stream, x, ldx, y, ldy, rows,
                     theta, scale);
  return hipGetLastError();
}

hipError_t launch_scale_copy(const double* x, long ldx, double* y, long ldy, long rows, long columns, double scale,
                             hipStream_t stream) {
  if (columns <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_copy, dim3(64, (unsigned)columns), dim3(256), 0, stream, x, ldx, y, ldy, rows, scale);
  return hipGetLastError();
}

// Full backward FFT of a natural complex grid (`grid`, count values) into the
// slot-ordered real potential. `full` is a plan whose sticks are every (x,y).
hipError_t launch_potential_setup(const GridPlanDev& full, const double* grid, void* t1, void* t2, double* v64,
                                  float* v32, double* block_ratio, int ratio_blocks, hipStream_t stream) {
  if (plane_lds_bytes<double>(full.n0, full.n1) > size_t(max_lds_bytes())) return hipErrorInvalidConfiguration;
  const int spb = sticks_per_block(full.n2);
  const dim3 ga((full.nsticks + spb - 1) / spb, 1);
  const size_t lds_lines = size_t(spb) * line_stride(full.n2) * sizeof(cx<double>);
  hipLaunchKernelGGL(k_sticks_backward<double>, ga, dim3(256), lds_lines, stream, grid, 0L, full.stick_g, full.nsticks,
                     full.n2, spb, reinterpret_cast<const cx<double>*>(full.tw64[2]), reinterpret_cast<cx<double>*>(t1));
  const size_t lds_plane = plane_lds_bytes<double>(full.n0, full.n1);
  launch_k_planes<double>(dim3(full.n2, 1), lds_plane, stream,
                     reinterpret_cast<const cx<double>*>(t1), reinterpret_cast<cx<double>*>(t2), full.stick_xy,
                     full.nsticks, full.n0, full.n1, full.n2, reinterpret_cast<const cx<double>*>(full.tw64[0]),
                     reinterpret_cast<const cx<double>*>(full.tw64[1]), (const double*)nullptr, 1, 0, full.n0, 0, 0);
  hipLaunchKernelGGL(k_potential_finish, dim3(ratio_blocks), dim3(256), 0, stream,
                     reinterpret_cast<const cx<double>*>(t2), (long)full.count, v64, v32, block_ratio);
  return hipGetLastError();
}

// ---- smooth density (src/density_functionals.cpp:406-460)
// density[i] += sum_c w_c |psi_c(i)|^2 over the columns of one chunk, in
// column order (deterministic); planes: columns x count complex, slot order.
__global__ void k_density_accumulate(const cx<double>* __restrict__ planes, long count, int columns,
                                     const double* __restrict__ weights, double* __restrict__ density) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    double acc = density[i];
    for (int c = 0; c < columns; ++c) {
      const cx<double> v = planes[(long)c * count + i];
      acc += weights[c] * (v.x * v.x + v.y * v.y);
    }
    density[i] = acc;
  }
}

// natural-order complex grid (re = slot-order real field, im = 0)
__global__ void k_slots_to_natural_complex(const double* __restrict__ v_slots, int n0, int n1, int n2, int a0, int a1,
                                           int a2, double* __restrict__ grid) {
  const long count = (long)n0 * n1 * n2;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    const int ys = int(i % n1);
    const int xs = int((i / n1) % n0);
    const int zs = int(i / ((long)n0 * n1));
    const int x = slot_position(xs, a0, n0 / a0);
    const int y = slot_position(ys, a1, n1 / a1);
    const int z = slot_position(zs, a2, n2 / a2);
    const long nat = ((long)x * n1 + y) * n2 + z;
    grid[2 * nat] = v_slots[i];
    grid[2 * nat + 1] = 0.0;
  }
}

// out[j] = conj(B[G_j]) / count, B the backward transform of a real field in
// slot order: the forward transform at G_j (real input).
__global__ void k_gather_support(const cx<double>* __restrict__ planes, int n0, int n1, int n2, int a0, int a1,
                                 int a2, const int* __restrict__ hkl, long nsupport, double inv_count,
                                 double* __restrict__ out) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < nsupport; j += (long)gridDim.x * blockDim.x) {
    int x = hkl[3 * j] % n0, y = hkl[3 * j + 1] % n1, z = hkl[3 * j + 2] % n2;
    x += x < 0 ? n0 : 0;
    y += y < 0 ? n1 : 0;
    z += z < 0 ? n2 : 0;
    // slot of natural position p on an axis L = A * (L / A): p = b + (L / A) a -> s = a + A b
    const int b0 = n0 / a0, b1 = n1 / a1, b2 = n2 / a2;
    const int xs = x / b0 + a0 * (x % b0), ys = y / b1 + a1 * (y % b1), zs = z / b2 + a2 * (z % b2);
    const cx<double> v = planes[((long)zs * n0 + xs) * n1 + ys];
    out[2 * j] = v.x * inv_count;
    out[2 * j + 1] = -v.y * inv_count;
  }
}

hipError_t launch_backward_planes(const GridPlanDev& p, const double* x, long ldx, long columns, void* t1, void* t2,
                                  hipStream_t stream) {
  if (columns <= 0) return hipSuccess;
  if (plane_lds_bytes<double>(p.n0, p.n1) > size_t(max_lds_bytes())) return hipErrorInvalidConfiguration;
  const int spb = sticks_per_block(p.n2);
  const dim3 ga((p.nsticks + spb - 1) / spb, (unsigned)columns);
  const size_t lds_lines = size_t(spb) * line_stride(p.n2) * sizeof(cx<double>);
  hipLaunchKernelGGL(k_sticks_backward<double>, ga, dim3(256), lds_lines, stream, x, ldx, p.stick_g, p.nsticks, p.n2,
                     spb, reinterpret_cast<const cx<double>*>(p.tw64[2]), reinterpret_cast<cx<double>*>(t1));
  const size_t lds_plane = plane_lds_bytes<double>(p.n0, p.n1);
  launch_k_planes<double>(dim3(p.n2, (unsigned)columns), lds_plane, stream,
                     reinterpret_cast<const cx<double>*>(t1), reinterpret_cast<cx<double>*>(t2), p.stick_xy,
                     p.nsticks, p.n0, p.n1, p.n2, reinterpret_cast<const cx<double>*>(p.tw64[0]),
                     reinterpret_cast<const cx<double>*>(p.tw64[1]), (const double*)nullptr, 1, p.xstart[0],
                     p.xlines[0], p.xstart[1], p.xlines[1]);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_backward_full(const GridPlanDev& p, const double* x, long ldx, long columns, void* t1, void* t2,
                                hipStream_t stream) {
  if (columns <= 0) return hipSuccess;
  if (plane_lds_bytes<T>(p.n0, p.n1) > size_t(max_lds_bytes())) return hipErrorInvalidConfiguration;
  const int spb = sticks_per_block(p.n2);
  const dim3 ga((p.nsticks + spb - 1) / spb, (unsigned)columns);
  const size_t lds_lines = size_t(spb) * line_stride(p.n2) * sizeof(cx<T>);
  const cx<T>* twx = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[0] : p.tw32[0]);
  const cx<T>* twy = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[1] : p.tw32[1]);
  const cx<T>* twz = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[2] : p.tw32[2]);
  hipLaunchKernelGGL(k_sticks_backward<T>, ga, dim3(256), lds_lines, stream, x, ldx, p.stick_g, p.nsticks, p.n2, spb,
                     twz, reinterpret_cast<cx<T>*>(t1));
  const size_t lds_plane = plane_lds_bytes<T>(p.n0, p.n1);
  launch_k_planes<T>(dim3(p.n2, (unsigned)columns), lds_plane, stream,
                     reinterpret_cast<const cx<T>*>(t1), reinterpret_cast<cx<T>*>(t2), p.stick_xy, p.nsticks, p.n0,
                     p.n1, p.n2, twx, twy, (const T*)nullptr, 1, p.xstart[0], p.xlines[0], p.xstart[1], p.xlines[1]);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_forward_planes(const GridPlanDev& p, const void* planes, long columns, const double* x, long ldx,
                                 const double* diag, double* y, long ldy, void* t_sticks, hipStream_t stream) {
  if (columns <= 0) return hipSuccess;
  if (plane_lds_bytes<T>(p.n0, p.n1) > size_t(max_lds_bytes())) return hipErrorInvalidConfiguration;
  const cx<T>* twx = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[0] : p.tw32[0]);
  const cx<T>* twy = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[1] : p.tw32[1]);
  const cx<T>* twz = reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? p.tw64[2] : p.tw32[2]);
  const size_t lds_plane = plane_lds_bytes<T>(p.n0, p.n1);
  launch_k_planes<T>(dim3(p.n2, (unsigned)columns), lds_plane, stream,
                     reinterpret_cast<const cx<T>*>(planes), reinterpret_cast<cx<T>*>(t_sticks), p.stick_xy,
                     p.nsticks, p.n0, p.n1, p.n2, twx, twy, (const T*)nullptr, 2, p.xstart[0], p.xlines[0],
                     p.xstart[1], p.xlines[1]);
  const int spb = sticks_per_block(p.n2);
  const dim3 ga((p.nsticks + spb - 1) / spb, (unsigned)columns);
  const size_t lds_lines = size_t(spb) * line_stride(p.n2) * sizeof(cx<T>);
  hipLaunchKernelGGL(k_sticks_forward<T>, ga, dim3(256), lds_lines, stream, reinterpret_cast<const cx<T>*>(t_sticks),
                     p.stick_g, p.nsticks, p.n2, spb, twz, x, ldx, diag, y, ldy, 0.0, double(p.count));
  return hipGetLastError();
}

template hipError_t launch_backward_full<double>(const GridPlanDev&, const double*, long, long, void*, void*,
                                                 hipStream_t);
template hipError_t launch_backward_full<float>(const GridPlanDev&, const double*, long, long, void*, void*,
                                                hipStream_t);
template hipError_t launch_forward_planes<double>(const GridPlanDev&, const void*, long, const double*, long,
                                                  const double*, double*, long, void*, hipStream_t);
template hipError_t launch_forward_planes<float>(const GridPlanDev&, const void*, long, const double*, long,
                                                 const double*, double*, long, void*, hipStream_t);

hipError_t launch_density_accumulate(const void* planes, long count, int columns, const double* weights,
                                     double* density, hipStream_t stream) {
  hipLaunchKernelGGL(k_density_accumulate, dim3(1024), dim3(256), 0, stream,
                     reinterpret_cast<const cx<double>*>(planes), count, columns, weights, density);
  return hipGetLastError();
}

hipError_t launch_slots_to_natural_complex(const GridPlanDev& p, const double* v_slots, double* grid,
                                           hipStream_t stream) {
  hipLaunchKernelGGL(k_slots_to_natural_complex, dim3(1024), dim3(256), 0, stream, v_slots, p.n0, p.n1, p.n2,
                     p.split[0], p.split[1], p.split[2], grid);
  return hipGetLastError();
}

hipError_t launch_gather_support(const GridPlanDev& p, const void* planes, const int* hkl, long nsupport,
                                 double* out, hipStream_t stream) {
  if (nsupport <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_support, dim3(256), dim3(256), 0, stream, reinterpret_cast<const cx<double>*>(planes),
                     p.n0, p.n1, p.n2, p.split[0], p.split[1], p.split[2], hkl, nsupport, 1.0 / double(p.count), out);
  return hipGetLastError();
}

hipError_t launch_potential_to_slots(const GridPlanDev& p, const double* v_nat, double* v64, float* v32,
                                     hipStream_t stream) {
  hipLaunchKernelGGL(k_potential_to_slots, dim3(1024), dim3(256), 0, stream, v_nat, p.n0, p.n1, p.n2, p.split[0],
                     p.split[1], p.split[2], v64, v32);
  return hipGetLastError();
}

hipError_t launch_potential_to_natural(const GridPlanDev& p, const double* v64, double* v_nat, hipStream_t stream) {
  hipLaunchKernelGGL(k_potential_to_natural, dim3(1024), dim3(256), 0, stream, v64, p.n0, p.n1, p.n2, p.split[0],
                     p.split[1], p.split[2], v_nat);
  return hipGetLastError();
}


// Loads this file's code object on the calling thread (see module_prime.hpp).
hipError_t prime_module_local_h() {
  hipFuncAttributes a{};
  return hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_sticks_backward<double>));
}

}  // namespace sdft_hip
