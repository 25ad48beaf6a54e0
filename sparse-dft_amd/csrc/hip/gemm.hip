// Complex GEMM on CDNA4 matrix cores, column-major with explicit leading
// dimensions (the semantics of src/compute_backend.hpp:21-27):
//   C = op(A) op(B) + beta * D,  op in {Normal, Adjoint},  D may alias C.
// Complex products use four real MFMAs per 16x16x4 block:
//   re += Ar Br - Ai Bi,  im += Ar Bi + Ai Br.
// FP64 uses v_mfma_f64_16x16x4_f64, FP32 uses v_mfma_f32_16x16x4_f32 (inputs
// stored in FP64 or FP32 are converted while staging into LDS).
// Long contractions (K >> M, N: projections B X, Galerkin slabs) are split
// along K into a workspace and reduced in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include "gemm.hpp"

#include <algorithm>

namespace sdft_hip {

namespace {
constexpr int BM = 64, BN = 64, BK = 16, PAD = 4;
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Acc;
template <>
struct Acc<double> {
  using type = d4;
  static __device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Acc<float> {
  using type = f4;
  static __device__ __forceinline__ f4 mfma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename S>
__device__ __forceinline__ void load2(const S* p, double& re, double& im) {
  re = double(p[0]);
  im = double(p[1]);
}

// One 64x64 output tile per workgroup (4 waves, 2x2 of 32x32), K range of
// this split [k0, k1).
template <typename T, typename TA, typename TB, bool ADJ_A, bool ADJ_B>
__global__ __launch_bounds__(256) void k_zgemm(int M, int N, int K, const TA* __restrict__ A, long lda,
                                               const TB* __restrict__ B, long ldb, double* __restrict__ C, long ldc,
                                               const double* __restrict__ D, long ldd, double alpha, double beta,
                                               int kps, double* __restrict__ work) {
  __shared__ T as_re[BK][BM + PAD], as_im[BK][BM + PAD];
  __shared__ T bs_re[BK][BN + PAD], bs_im[BK][BN + PAD];
  using V = typename Acc<T>::type;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int k0 = blockIdx.z * kps, k1 = min(K, k0 + kps);
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  V acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_re[i][j] = V{0, 0, 0, 0};
      acc_im[i][j] = V{0, 0, 0, 0};
    }
  for (int kb = k0; kb < k1; kb += BK) {
    // stage A tile: A_eff[m][k], m in [m0, m0+BM), k in [kb, kb+BK)
#pragma unroll
    for (int it = 0; it < (BM * BK) / 256; ++it) {
      const int e = tid + it * 256;
      int mm, kk;
      if (!ADJ_A) { mm = e % BM; kk = e / BM; } else { kk = e % BK; mm = e / BK; }
      const int m = m0 + mm, k = kb + kk;
      double re = 0, im = 0;
      if (m < M && k < k1) {
        if (!ADJ_A) load2(A + 2 * (m + (long)k * lda), re, im);
        else { load2(A + 2 * (k + (long)m * lda), re, im); im = -im; }
      }
      as_re[kk][mm] = T(re);
      as_im[kk][mm] = T(im);
    }
#pragma unroll
    for (int it = 0; it < (BN * BK) / 256; ++it) {
      const int e = tid + it * 256;
      int nn, kk;
      if (!ADJ_B) { kk = e % BK; nn = e / BK; } else { nn = e % BN; kk = e / BN; }
      const int n = n0 + nn, k = kb + kk;
      double re = 0, im = 0;
      if (n < N && k < k1) {
        if (!ADJ_B) load2(B + 2 * (k + (long)n * ldb), re, im);
        else { load2(B + 2 * (n + (long)k * ldb), re, im); im = -im; }
      }
      bs_re[kk][nn] = T(re);
      bs_im[kk][nn] = T(im);
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4), c = lane & 15;
      T ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = as_re[kr][wm + 16 * i + c];
        ai[i] = as_im[kr][wm + 16 * i + c];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        br[j] = bs_re[kr][wn + 16 * j + c];
        bi[j] = bs_im[kr][wn + 16 * j + c];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc_re[i][j] = Acc<T>::mfma(ar[i], br[j], acc_re[i][j]);
          acc_re[i][j] = Acc<T>::mfma(-ai[i], bi[j], acc_re[i][j]);
          acc_im[i][j] = Acc<T>::mfma(ar[i], bi[j], acc_im[i][j]);
          acc_im[i][j] = Acc<T>::mfma(ai[i], br[j], acc_im[i][j]);
        }
    }
    __syncthreads();
  }
  // Accumulator layout of each 16x16 block, column lane % 16 in both cases:
  //   f32 16x16x4: register r holds row 4*(lane/16) + r
  //   f64 16x16x4: register r holds row 4*r + lane/16
  const bool split = work != nullptr;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = sizeof(T) == 8 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r;
        const int m = m0 + wm + 16 * i + row;
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m >= M || n >= N) continue;
        double re = alpha * double(acc_re[i][j][r]), im = alpha * double(acc_im[i][j][r]);
        if (split) {
          double* w = work + 2 * ((long)blockIdx.z * M * N + m + (long)n * M);
          w[0] = re;
          w[1] = im;
        } else {
          if (beta != 0.0) {
            const double* d = D + 2 * (m + (long)n * ldd);
            re += beta * d[0];
            im += beta * d[1];
          }
          double* c = C + 2 * (m + (long)n * ldc);
          c[0] = re;
          c[1] = im;
        }
      }
}

__global__ void k_splitk_reduce(int M, int N, int splits, const double* __restrict__ work, double* __restrict__ C,
                                long ldc, const double* __restrict__ D, long ldd, double beta) {
  // partials already carry alpha
  const long total = (long)M * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int m = int(e % M);
    const long n = e / M;
    double re = 0, im = 0;
    for (int s = 0; s < splits; ++s) {
      const double* w = work + 2 * ((long)s * total + e);
      re += w[0];
      im += w[1];
    }
    if (beta != 0.0) {
      const double* d = D + 2 * (m + n * ldd);
      re += beta * d[0];
      im += beta * d[1];
    }
    double* c = C + 2 * (m + n * ldc);
    c[0] = re;
    c[1] = im;
  }
}

// Per-atom onsite products W[:, 0:m] = D_a P_a, W[:, m:2m] = Q_a P_a.
// FP64: compensated (TwoSum of FMA-exact products) in the scalar order of
// src/paw_compensated_product.hpp:41-56; FP32: plain products.
__global__ void k_atom_blocks_comp(int np, int columns, int atoms, const int* __restrict__ atom_offset,
                                   const int* __restrict__ atom_rows, const long* __restrict__ block_offset,
                                   const double* __restrict__ dblk, const double* __restrict__ qblk,
                                   const double* __restrict__ P, long ldp, double* __restrict__ W, long ldw,
                                   int which_mask) {
#pragma clang fp contract(off)
  const long total = (long)np * columns * 2;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int row = int(e % np);
    const long rest = e / np;
    const int col = int(rest % columns);
    const int which = int(rest / columns);  // 0: D, 1: Q
    if (!((which_mask >> which) & 1)) continue;
    // find the atom of this row (atoms are few; linear scan over offsets)
    int a = 0;
    while (a + 1 < atoms && atom_offset[a + 1] <= row) ++a;
    const int r = atom_rows[a], o = atom_offset[a], i = row - o;
    const double* blk = (which ? qblk : dblk) + 2 * block_offset[a];  // column-major r x r
    double rs = 0, rc = 0, is = 0, ic = 0;
    auto add = [](double v, double& s, double& c) {
      const double next = s + v;
      const double back = next - s;
      c += (s - (next - back)) + (v - back);
      s = next;
    };
    auto product = [&](double x, double y, double& s, double& c) {
      const double rounded = x * y;
      add(rounded, s, c);
      add(__builtin_fma(x, y, -rounded), s, c);
    };
    const double* pc = P + 2 * (o + (long)col * ldp);
    for (int k = 0; k < r; ++k) {
      const double lr = blk[2 * (i + k * r)], li = blk[2 * (i + k * r) + 1];
      const double pr = pc[2 * k], pi = pc[2 * k + 1];
      product(lr, pr, rs, rc);
      product(-li, pi, rs, rc);
      product(lr, pi, is, ic);
      product(li, pr, is, ic);
    }
    double* w = W + 2 * (row + ((long)which * columns + col) * ldw);
    w[0] = rs + rc;
    w[1] = is + ic;
  }
}

__global__ void k_atom_blocks_f32(int np, int columns, int atoms, const int* __restrict__ atom_offset,
                                  const int* __restrict__ atom_rows, const long* __restrict__ block_offset,
                                  const double* __restrict__ dblk, const double* __restrict__ qblk,
                                  const double* __restrict__ P, long ldp, double* __restrict__ W, long ldw,
                                  int which_mask) {
  const long total = (long)np * columns * 2;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int row = int(e % np);
    const long rest = e / np;
    const int col = int(rest % columns);
    const int which = int(rest / columns);
    if (!((which_mask >> which) & 1)) continue;
    int a = 0;
    while (a + 1 < atoms && atom_offset[a + 1] <= row) ++a;
    const int r = atom_rows[a], o = atom_offset[a], i = row - o;
    const double* blk = (which ? qblk : dblk) + 2 * block_offset[a];
    const double* pc = P + 2 * (o + (long)col * ldp);
    float re = 0.f, im = 0.f;
    for (int k = 0; k < r; ++k) {
      const float lr = float(blk[2 * (i + k * r)]), li = float(blk[2 * (i + k * r) + 1]);
      const float pr = float(pc[2 * k]), pi = float(pc[2 * k + 1]);
      re += lr * pr - li * pi;
      im += lr * pi + li * pr;
    }
    double* w = W + 2 * (row + ((long)which * columns + col) * ldw);
    w[0] = re;
    w[1] = im;
  }
}

template <typename T, typename TA, typename TB>
hipError_t dispatch_ops(int aop, int bop, dim3 grid, int M, int N, int K, const TA* A, long lda, const TB* B, long ldb,
                        double* C, long ldc, const double* D, long ldd, double alpha, double beta, int kps,
                        double* work, hipStream_t s) {
#define SDFT_LAUNCH(AA, BB) \
  hipLaunchKernelGGL((k_zgemm<T, TA, TB, AA, BB>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, \
                     alpha, beta, kps, work)
  if (!aop && !bop) SDFT_LAUNCH(false, false);
  else if (aop && !bop) SDFT_LAUNCH(true, false);
  else if (!aop && bop) SDFT_LAUNCH(false, true);
  else SDFT_LAUNCH(true, true);
#undef SDFT_LAUNCH
  return hipGetLastError();
}
}  // namespace

std::size_t zgemm_workspace_bytes(int M, int N, int K) {
  const int splits = zgemm_splits(M, N, K);
  return splits > 1 ? std::size_t(splits) * M * N * 16 : 0;
}

int zgemm_splits(int M, int N, int K) {
  const long tiles = long((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  // aim for >= 512 workgroups on 256 CUs while keeping >= 256 K per split
  int splits = 1;
  while (tiles * splits < 512 && long(K) / (2 * splits) >= 256 && splits < 64) splits *= 2;
  return splits;
}

hipError_t zgemm(Precision prec, int aop, int bop, int M, int N, int K, double alpha, const double* A, long lda,
                 const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta, double* work,
                 std::size_t work_bytes, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (K <= 0) {
    // C = beta * D
    return launch_zaxpby(M, N, 0.0, nullptr, 0, beta, D, ldd, C, ldc, s);
  }
  int splits = zgemm_splits(M, N, K);
  if (splits > 1 && (work == nullptr || work_bytes < std::size_t(splits) * M * N * 16)) splits = 1;
  int kps = (K + splits - 1) / splits;
  kps = ((kps + BK - 1) / BK) * BK;
  splits = (K + kps - 1) / kps;
  const dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, splits);
  double* wk = splits > 1 ? work : nullptr;
  hipError_t e;
  if (prec == Precision::FP64)
    e = dispatch_ops<double, double, double>(aop, bop, grid, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha, beta, kps,
                                             wk, s);
  else
    e = dispatch_ops<float, double, double>(aop, bop, grid, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha, beta, kps,
                                            wk, s);
  if (e != hipSuccess || splits == 1) return e;
  hipLaunchKernelGGL(k_splitk_reduce, dim3(1024), dim3(256), 0, s, M, N, splits, work, C, ldc, D, ldd, beta);
  return hipGetLastError();
}

hipError_t zgemm_f32_storage(int aop, int bop, int M, int N, int K, double alpha, const float* A, long lda,
                             const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta,
                             double* work, std::size_t work_bytes, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  int splits = zgemm_splits(M, N, K);
  if (splits > 1 && (work == nullptr || work_bytes < std::size_t(splits) * M * N * 16)) splits = 1;
  int kps = (K + splits - 1) / splits;
  kps = ((kps + BK - 1) / BK) * BK;
  splits = (K + kps - 1) / kps;
  const dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, splits);
  double* wk = splits > 1 ? work : nullptr;
  hipError_t e = dispatch_ops<float, float, double>(aop, bop, grid, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha,
                                                    beta, kps, wk, s);
  if (e != hipSuccess || splits == 1) return e;
  hipLaunchKernelGGL(k_splitk_reduce, dim3(1024), dim3(256), 0, s, M, N, splits, work, C, ldc, D, ldd, beta);
  return hipGetLastError();
}

hipError_t launch_atom_blocks(Precision prec, int np, int columns, int atoms, const int* atom_offset,
                              const int* atom_rows, const long* block_offset, const double* dblk, const double* qblk,
                              const double* P, long ldp, double* W, long ldw, int which_mask, hipStream_t s) {
  const long total = long(np) * columns * 2;
  const int blocks = int(std::min<long>((total + 255) / 256, 4096));
  if (prec == Precision::FP64)
    hipLaunchKernelGGL(k_atom_blocks_comp, dim3(blocks), dim3(256), 0, s, np, columns, atoms, atom_offset, atom_rows,
                       block_offset, dblk, qblk, P, ldp, W, ldw, which_mask);
  else
    hipLaunchKernelGGL(k_atom_blocks_f32, dim3(blocks), dim3(256), 0, s, np, columns, atoms, atom_offset, atom_rows,
                       block_offset, dblk, qblk, P, ldp, W, ldw, which_mask);
  return hipGetLastError();
}

// C = alpha * X + beta * D (complex, column-major, real scalars)
__global__ void k_zaxpby(int M, int N, double alpha, const double* __restrict__ X, long ldx, double beta,
                         const double* __restrict__ D, long ldd, double* __restrict__ C, long ldc) {
  const long total = (long)M * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int m = int(e % M);
    const long n = e / M;
    double re = 0, im = 0;
    if (alpha != 0.0) { re = alpha * X[2 * (m + n * ldx)]; im = alpha * X[2 * (m + n * ldx) + 1]; }
    if (beta != 0.0) { re += beta * D[2 * (m + n * ldd)]; im += beta * D[2 * (m + n * ldd) + 1]; }
    C[2 * (m + n * ldc)] = re;
    C[2 * (m + n * ldc) + 1] = im;
  }
}

hipError_t launch_zaxpby(int M, int N, double alpha, const double* X, long ldx, double beta, const double* D, long ldd,
                         double* C, long ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const long total = long(M) * N;
  const int blocks = int(std::min<long>((total + 255) / 256, 8192));
  hipLaunchKernelGGL(k_zaxpby, dim3(blocks), dim3(256), 0, s, M, N, alpha, X, ldx, beta, D, ldd, C, ldc);
  return hipGetLastError();
}

}  // namespace sdft_hip
