// Real-space projector search operator (src/plane_wave_operator.cpp:104-212):
// an approximate H or S for the search phase that evaluates the PAW
// projections on each atom's retained real-space points instead of B X.
//
//   psi(r)  = backward transform of x (full planes, slot order)
//   p_a     = W_a psi(points_a)              W_a: rows x points, w_i(r) = FFT_fwd(B_i) / N
//   c_a     = D_a p_a (H) or Q_a p_a (S)     (onsite kernel of gemm.hip)
//   f(r)    = V(r) psi(r) (H) or 0 (S);  f(points_a) += N W_a^H c_a
//   y       = T x + fwd(f)[sphere] / N (H) or (1 + reg) x + fwd(f)[sphere] / N (S)
//
// Points are stored as slot-order plane indices; the image accumulation is
// point-major (each point sums its atoms in atom order), so it is
// deterministic where atom spheres overlap.
#include <hip/hip_runtime.h>

#include "fft_device.hpp"
#include "realspace.hpp"

namespace sdft_hip {
namespace {

// Psi[j][c] = planes[c][point_j]   (j over all atoms' points, ld = total points)
template <typename T>
__global__ void k_rs_gather(const cx<T>* __restrict__ planes, long count, const int* __restrict__ points, long total,
                            int columns, cx<T>* __restrict__ psi) {
  const long n = total * columns;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long j = e % total;
    const int c = int(e / total);
    psi[e] = planes[(long)c * count + points[j]];
  }
}

// P[o_a + i][c] = sum_j W_a[i][j] psi_a[j][c]; one workgroup per (atom, column),
// the sum over points in a fixed order (per-thread strided partials, then a
// tree in LDS). W_a: rows x points, column-major per atom, concatenated.
template <typename T>
__global__ __launch_bounds__(256) void k_rs_project(const cx<T>* __restrict__ w, const long* __restrict__ w_offset,
                                                    const long* __restrict__ point_offset,
                                                    const int* __restrict__ atom_offset,
                                                    const int* __restrict__ atom_rows, const cx<T>* __restrict__ psi,
                                                    long total, double* __restrict__ p, long ldp) {
  __shared__ T red_re[256], red_im[256];
  const int a = blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
  const long p0 = point_offset[a], np_a = point_offset[a + 1] - p0;
  const int rows = atom_rows[a];
  const cx<T>* wa = w + w_offset[a];
  const cx<T>* ps = psi + (long)c * total + p0;
  for (int i = 0; i < rows; ++i) {
    T re = 0, im = 0;
    for (long j = tid; j < np_a; j += 256) {
      const cx<T> x = wa[i + j * rows], y = ps[j];
      re += x.x * y.x - x.y * y.y;
      im += x.x * y.y + x.y * y.x;
    }
    red_re[tid] = re;
    red_im[tid] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) {
        red_re[tid] += red_re[tid + s];
        red_im[tid] += red_im[tid + s];
      }
      __syncthreads();
    }
    if (tid == 0) {
      double* out = p + 2 * ((long)atom_offset[a] + i + (long)c * ldp);
      out[0] = double(red_re[0]);
      out[1] = double(red_im[0]);
    }
    __syncthreads();
  }
}

// f = V psi (H) or 0 (S) on every slot of every column's plane
template <typename T>
__global__ void k_rs_scale(cx<T>* __restrict__ planes, long count, int columns, const T* __restrict__ v_slots,
                           int overlap) {
  const long n = count * columns;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    cx<T>& f = planes[e];
    if (overlap) {
      f = {T(0), T(0)};
    } else {
      const T v = v_slots[e % count];
      f.x *= v;
      f.y *= v;
    }
  }
}

// per-point sum of |w_i|^2 over an atom's rows (planes: rows x count)
__global__ void k_rs_peak(const cx<double>* __restrict__ planes, long count, int rows, double* __restrict__ peak) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < count; r += (long)gridDim.x * blockDim.x) {
    double s = 0;
    for (int i = 0; i < rows; ++i) {
      const cx<double> v = planes[(long)i * count + r];
      s += v.x * v.x + v.y * v.y;
    }
    peak[r] = s;
  }
}

// W_a[i + j rows] = conj(planes[i][points[j]]) / N: w_i(r) = FFT_fwd(B_i)(r) / N is
// the conjugate of the backward transform of conj(B_i), divided by N
__global__ void k_rs_values(const cx<double>* __restrict__ planes, long count, int rows, const int* __restrict__ points,
                            long npoints, double inv_count, cx<double>* __restrict__ w64, cx<float>* __restrict__ w32) {
  const long n = npoints * rows;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = int(e % rows);
    const long j = e / rows;
    const cx<double> v = planes[(long)i * count + points[j]];
    const cx<double> w{v.x * inv_count, -v.y * inv_count};
    w64[e] = w;
    w32[e] = {float(w.x), float(w.y)};
  }
}

template <typename T>
__global__ void k_rs_add_images(cx<T>* __restrict__ planes, long count, int columns, const int* __restrict__ upoints,
                                long nunique, const int* __restrict__ entry_start, const int* __restrict__ entry_atom,
                                const int* __restrict__ entry_local, const cx<T>* __restrict__ w,
                                const long* __restrict__ w_offset, const int* __restrict__ atom_offset,
                                const int* __restrict__ atom_rows, const double* __restrict__ amp, long ldamp,
                                double scale) {
  const long n = nunique * columns;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long k = e % nunique;
    const int c = int(e / nunique);
    double re = 0, im = 0;
    for (int t = entry_start[k]; t < entry_start[k + 1]; ++t) {
      const int a = entry_atom[t], j = entry_local[t], rows = atom_rows[a];
      const cx<T>* wa = w + w_offset[a] + (long)j * rows;
      const double* ca = amp + 2 * ((long)atom_offset[a] + (long)c * ldamp);
      for (int i = 0; i < rows; ++i) {
        const double wr = double(wa[i].x), wi = -double(wa[i].y);  // conj(W_ij)
        const double cr = ca[2 * i], ci = ca[2 * i + 1];
        re += wr * cr - wi * ci;
        im += wr * ci + wi * cr;
      }
    }
    cx<T>& f = planes[(long)c * count + upoints[k]];
    f.x += T(scale * re);
    f.y += T(scale * im);
  }
}

}  // namespace

template <typename T>
hipError_t launch_rs_gather(const void* planes, long count, const int* points, long total, int columns, void* psi,
                            hipStream_t s) {
  if (total <= 0 || columns <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rs_gather<T>, dim3(1024), dim3(256), 0, s, reinterpret_cast<const cx<T>*>(planes), count,
                     points, total, columns, reinterpret_cast<cx<T>*>(psi));
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rs_project(const RealSpaceDev& rs, const void* psi, int columns, double* p, long ldp,
                             hipStream_t s) {
  if (columns <= 0 || rs.atoms <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rs_project<T>, dim3(rs.atoms, columns), dim3(256), 0, s,
                     reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? rs.w64 : rs.w32), rs.w_offset,
                     rs.point_offset, rs.atom_offset, rs.atom_rows, reinterpret_cast<const cx<T>*>(psi),
                     rs.total_points, p, ldp);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rs_images(const RealSpaceDev& rs, void* planes, long count, int columns, const T* v_slots,
                            int overlap, const double* amp, long ldamp, hipStream_t s) {
  if (columns <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rs_scale<T>, dim3(2048), dim3(256), 0, s, reinterpret_cast<cx<T>*>(planes), count, columns,
                     v_slots, overlap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || rs.unique <= 0) return e;
  hipLaunchKernelGGL(k_rs_add_images<T>, dim3(1024), dim3(256), 0, s, reinterpret_cast<cx<T>*>(planes), count,
                     columns, rs.unique_points, rs.unique, rs.entry_start, rs.entry_atom, rs.entry_local,
                     reinterpret_cast<const cx<T>*>(sizeof(T) == 8 ? rs.w64 : rs.w32), rs.w_offset, rs.atom_offset,
                     rs.atom_rows, amp, ldamp, double(count));
  return hipGetLastError();
}

hipError_t launch_rs_peak(const void* planes, long count, int rows, double* peak, hipStream_t s) {
  hipLaunchKernelGGL(k_rs_peak, dim3(1024), dim3(256), 0, s, reinterpret_cast<const cx<double>*>(planes), count, rows,
                     peak);
  return hipGetLastError();
}

hipError_t launch_rs_values(const void* planes, long count, int rows, const int* points, long npoints, void* w64,
                            void* w32, hipStream_t s) {
  if (npoints <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rs_values, dim3(1024), dim3(256), 0, s, reinterpret_cast<const cx<double>*>(planes), count,
                     rows, points, npoints, 1.0 / double(count), reinterpret_cast<cx<double>*>(w64),
                     reinterpret_cast<cx<float>*>(w32));
  return hipGetLastError();
}

template hipError_t launch_rs_gather<double>(const void*, long, const int*, long, int, void*, hipStream_t);
template hipError_t launch_rs_gather<float>(const void*, long, const int*, long, int, void*, hipStream_t);
template hipError_t launch_rs_project<double>(const RealSpaceDev&, const void*, int, double*, long, hipStream_t);
template hipError_t launch_rs_project<float>(const RealSpaceDev&, const void*, int, double*, long, hipStream_t);
template hipError_t launch_rs_images<double>(const RealSpaceDev&, void*, long, int, const double*, int,
                                             const double*, long, hipStream_t);
template hipError_t launch_rs_images<float>(const RealSpaceDev&, void*, long, int, const float*, int, const double*,
                                            long, hipStream_t);


// Loads this file's code object on the calling thread (see module_prime.hpp).
hipError_t prime_module_realspace() {
  hipFuncAttributes a{};
  return hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_rs_gather<double>));
}

}  // namespace sdft_hip
