// Complex GEMM on CDNA4 matrix cores, column-major with explicit leading
// dimensions (the semantics of src/compute_backend.hpp:21-27):
//   C = op(A) op(B) + beta * D,  op in {Normal, Adjoint},  D may alias C.
// Complex products use three real MFMAs per 16x16x4 block (Gauss form, below):
//   P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); re = P1 - P2, im = P3 - P1 - P2.
// FP64 uses v_mfma_f64_16x16x4_f64, FP32 uses v_mfma_f32_16x16x4_f32 (inputs
// stored in FP64 or FP32 are converted while staging into LDS).
// Long contractions (K >> M, N: projections B X, Galerkin slabs) are split
// along K into a workspace and reduced in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include "gemm.hpp"

#include <algorithm>

namespace sdft_hip {

namespace {
constexpr int BK = 16;  // K slab per LDS round
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

// One BM x BN output tile per 256-thread workgroup (4 waves, WGM along M),
// K range [k0, k1) of this split. Each wave computes (BM/WGM) x (BN/(4/WGM))
// as 16 x 16 blocks; per 4-deep K step it first issues the independent
// Ar*Br / Ar*Bi products of every block, then the dependent -Ai*Bi / Ai*Br
// ones, so that no MFMA waits on the accumulator of the one just issued.
// PF: the next K slab is loaded into registers while the current one is
// multiplied from LDS.
// XCD-aware tile order. Workgroups are dispatched round-robin over the eight
// XCDs (linear id L runs on XCD L % 8), each with its own L2. The bijection
// below hands every XCD a contiguous range of tiles in (split, m, n) order
// with n fastest, so the tiles that share an A panel (and, for a few M tiles,
// a B panel) run on the same XCD and reuse it from that XCD's L2.
constexpr int kXcds = 8;
struct TileId {
  int m, n, split;
};
__device__ __forceinline__ TileId xcd_tile(int mt, int nt, int splits) {
  const int total = mt * nt * splits;
  const int lin = blockIdx.x + mt * (blockIdx.y + nt * blockIdx.z);
  const int xcd = lin % kXcds, slot = lin / kXcds;
  const int per = (total + kXcds - 1) / kXcds;  // tiles of the fuller XCDs
  const int full = total - (per - 1) * kXcds;   // number of XCDs holding `per` tiles
  const int t = xcd < full ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
  return {(t / nt) % mt, t % nt, t / (nt * mt)};
}

// Complex products as three real MFMA products (Gauss / 3M: Ar Br, Ai Bi,
// (Ar + Ai)(Br + Bi)) instead of four, 25 % fewer MFMA cycles, in both
// precisions; SDFT_GEMM_GAUSS_FP64=0 / SDFT_GEMM_GAUSS_FP32=0 build the
// four-product form.
#ifndef SDFT_GEMM_GAUSS_FP64
#define SDFT_GEMM_GAUSS_FP64 1
#endif
#ifndef SDFT_GEMM_GAUSS_FP32
#define SDFT_GEMM_GAUSS_FP32 1
#endif

// WAVES: at least that many waves per SIMD. Left alone, the FP64 tiles keep
// their accumulators in AGPRs and the alpha/beta epilogue copies them all to
// VGPRs at once, which caps occupancy at two waves; the 3M tiles hold a third
// accumulator set, which fits three waves only in the split-K plan.
template <typename T, typename TA, typename TB, bool ADJ_A, bool ADJ_B, int TBM, int TBN, int WGM, bool PF, int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_zgemm(int M, int N, int K, const TA* __restrict__ A, long lda,
                                               const TB* __restrict__ B, long ldb, double* __restrict__ C, long ldc,
                                               const double* __restrict__ D, long ldd, double alpha, double beta,
                                               int kps, double* __restrict__ work) {
  constexpr int WGN = 4 / WGM, WM = TBM / WGM, WN = TBN / WGN, IM = WM / 16, JN = WN / 16;
  constexpr int EA = TBM * BK / 256, EB = TBN * BK / 256;
  static_assert(EA * 256 == TBM * BK && EB * 256 == TBN * BK, "staging must tile the workgroup");
  // LDS staging, conflict-free in both directions: an operand whose global
  // layout runs along K (adjoint A, normal B) is loaded by lanes along K and
  // kept [m][k] with row stride BK + 1; the other one is loaded along m and
  // kept [k][m] with row stride TBM + 16 (two K rows of one half-wave read
  // land on disjoint banks).
  constexpr int SA = ADJ_A ? BK + 1 : TBM + 16, SB = ADJ_B ? TBN + 16 : BK + 1;
  constexpr int LA = ADJ_A ? TBM * SA : BK * SA, LB = ADJ_B ? BK * SB : TBN * SB;
  __shared__ T as_re[LA], as_im[LA];
  __shared__ T bs_re[LB], bs_im[LB];
  auto aix = [](int kk, int mm) { return ADJ_A ? mm * SA + kk : kk * SA + mm; };
  auto bix = [](int kk, int nn) { return ADJ_B ? kk * SB + nn : nn * SB + kk; };
  using V = typename Acc<T>::type;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const TileId tile = xcd_tile(gridDim.x, gridDim.y, gridDim.z);
  const int m0 = tile.m * TBM, n0 = tile.n * TBN;
  const int k0 = tile.split * kps, k1 = min(K, k0 + kps);
  const int wm = (wave % WGM) * WM, wn = (wave / WGM) * WN;
  // kGauss: three real products per block (Gauss / 3M),
  // acc_re = sum Ar Br, acc_x = sum Ai Bi, acc_im = sum (Ar + Ai)(Br + Bi)
  constexpr bool kGauss = sizeof(T) == 8 ? SDFT_GEMM_GAUSS_FP64 : SDFT_GEMM_GAUSS_FP32;
  constexpr int GI = kGauss ? IM : 1, GJ = kGauss ? JN : 1;
  V acc_re[IM][JN], acc_im[IM][JN], acc_x[GI][GJ];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      acc_re[i][j] = V{0, 0, 0, 0};
      acc_im[i][j] = V{0, 0, 0, 0};
    }
#pragma unroll
  for (int i = 0; i < GI; ++i)
#pragma unroll
    for (int j = 0; j < GJ; ++j) acc_x[i][j] = V{0, 0, 0, 0};
  double pa[EA][2], pb[EB][2];
  // A_eff[m][k] and B_eff[k][n] of the K slab at kb into registers
  auto load = [&](int kb) {
#pragma unroll
    for (int it = 0; it < EA; ++it) {
      const int e = tid + it * 256;
      int mm, kk;
      if (!ADJ_A) { mm = e % TBM; kk = e / TBM; } else { kk = e % BK; mm = e / BK; }
      const int m = m0 + mm, k = kb + kk;
      double re = 0, im = 0;
      if (m < M && k < k1) {
        if (!ADJ_A) load2(A + 2 * (m + (long)k * lda), re, im);
        else { load2(A + 2 * (k + (long)m * lda), re, im); im = -im; }
      }
      pa[it][0] = re;
      pa[it][1] = im;
    }
#pragma unroll
    for (int it = 0; it < EB; ++it) {
      const int e = tid + it * 256;
      int nn, kk;
      if (!ADJ_B) { kk = e % BK; nn = e / BK; } else { nn = e % TBN; kk = e / TBN; }
      const int n = n0 + nn, k = kb + kk;
      double re = 0, im = 0;
      if (n < N && k < k1) {
        if (!ADJ_B) load2(B + 2 * (k + (long)n * ldb), re, im);
        else { load2(B + 2 * (n + (long)k * ldb), re, im); im = -im; }
      }
      pb[it][0] = re;
      pb[it][1] = im;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int it = 0; it < EA; ++it) {
      const int e = tid + it * 256;
      int mm, kk;
      if (!ADJ_A) { mm = e % TBM; kk = e / TBM; } else { kk = e % BK; mm = e / BK; }
      as_re[aix(kk, mm)] = T(pa[it][0]);
      as_im[aix(kk, mm)] = T(pa[it][1]);
    }
#pragma unroll
    for (int it = 0; it < EB; ++it) {
      const int e = tid + it * 256;
      int nn, kk;
      if (!ADJ_B) { kk = e % BK; nn = e / BK; } else { nn = e % TBN; kk = e / TBN; }
      bs_re[bix(kk, nn)] = T(pb[it][0]);
      bs_im[bix(kk, nn)] = T(pb[it][1]);
    }
  };
  if (PF && k0 < k1) load(k0);
  for (int kb = k0; kb < k1; kb += BK) {
    if (!PF) load(kb);
    store();
    __syncthreads();
    if (PF && kb + BK < k1) load(kb + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4), c = lane & 15;
      T ar[IM], ai[IM], br[JN], bi[JN];
#pragma unroll
      for (int i = 0; i < IM; ++i) {
        ar[i] = as_re[aix(kr, wm + 16 * i + c)];
        ai[i] = as_im[aix(kr, wm + 16 * i + c)];
      }
#pragma unroll
      for (int j = 0; j < JN; ++j) {
        br[j] = bs_re[bix(kr, wn + 16 * j + c)];
        bi[j] = bs_im[bix(kr, wn + 16 * j + c)];
      }
      if constexpr (kGauss) {
        T as[IM], bsum[JN];
#pragma unroll
        for (int i = 0; i < IM; ++i) as[i] = ar[i] + ai[i];
#pragma unroll
        for (int j = 0; j < JN; ++j) bsum[j] = br[j] + bi[j];
#pragma unroll
        for (int i = 0; i < IM; ++i)
#pragma unroll
          for (int j = 0; j < JN; ++j) {
            acc_re[i][j] = Acc<T>::mfma(ar[i], br[j], acc_re[i][j]);
            acc_im[i][j] = Acc<T>::mfma(as[i], bsum[j], acc_im[i][j]);
          }
#pragma unroll
        for (int i = 0; i < GI; ++i)
#pragma unroll
          for (int j = 0; j < GJ; ++j) acc_x[i][j] = Acc<T>::mfma(ai[i], bi[j], acc_x[i][j]);
      } else {
#pragma unroll
        for (int i = 0; i < IM; ++i)
#pragma unroll
          for (int j = 0; j < JN; ++j) {
            acc_re[i][j] = Acc<T>::mfma(ar[i], br[j], acc_re[i][j]);
            acc_im[i][j] = Acc<T>::mfma(ar[i], bi[j], acc_im[i][j]);
          }
#pragma unroll
        for (int i = 0; i < IM; ++i)
#pragma unroll
          for (int j = 0; j < JN; ++j) {
            acc_re[i][j] = Acc<T>::mfma(-ai[i], bi[j], acc_re[i][j]);
            acc_im[i][j] = Acc<T>::mfma(ai[i], br[j], acc_im[i][j]);
          }
      }
    }
    __syncthreads();
  }
  // Accumulator layout of each 16x16 block, column lane % 16 in both cases:
  //   f32 16x16x4: register r holds row 4*(lane/16) + r
  //   f64 16x16x4: register r holds row 4*r + lane/16
  const bool split = work != nullptr;
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = sizeof(T) == 8 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r;
        const int m = m0 + wm + 16 * i + row;
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m >= M || n >= N) continue;
        double re = double(acc_re[i][j][r]), im = double(acc_im[i][j][r]);
        if constexpr (kGauss) {
          const double s2 = double(acc_x[kGauss ? i : 0][kGauss ? j : 0][r]);
          im = im - re - s2;
          re = re - s2;
        }
        re *= alpha;
        im *= alpha;
        if (split) {
          double* w = work + 2 * ((long)tile.split * M * N + m + (long)n * M);
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
    // partials summed in split order (deterministic); eight loads issued
    // ahead of their additions
    double re = 0, im = 0;
    int s = 0;
    for (; s + 8 <= splits; s += 8) {
      double2 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const double2*>(work + 2 * ((long)(s + u) * total + e));
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        re += v[u].x;
        im += v[u].y;
      }
    }
    for (; s < splits; ++s) {
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

// Tile configurations, both precisions, with register prefetch: 0 = 128 x 32
// (tall, narrow outputs with normal A: rotations), 1 = 32 x 32 (every other
// shape: one 16 x 16 block per wave, the most workgroups). Measured on the Si64
// shapes, 32 x 32 / 64 x 32 / 64 x 64 (or 128 x 32), FP64 with the split counts
// of fp64_splits: slabs V^H W 800 x 160 x n 57.6 / 46.3 / 39.8 TFLOP/s,
// back-projections n x 160 x 512 63.4 / 57.9 / 42.2, Gram 160 x 160 x n
// 42.4 / 39.2 / 30.0, Rayleigh-Ritz 800^3 55.6 / 47.9 / 37.0, projections
// 512 x 160 x n 65.8 / 66.8 / 55.0, rotations n x 160 x 800 67.7 / - / 67.6;
// FP32 (four real products, best split count of each): slabs 75.3 / 69.9 /
// 68.3, back-projections 82.0 / 77.8 / 72.2, Gram 53.2 / 47.4 / 40.1,
// Rayleigh-Ritz 68.1 / 59.4 / 59.5, projections 82.6 / 88.1 / 79.7, rotations
// 86.0 / 89.2 / 77.8 (128 x 32: 90.5).
struct Tile {
  int bm, bn;
};
constexpr Tile kTiles[2] = {{128, 32}, {32, 32}};

bool tall_shape(int M, int N) { return M >= 2048 && N <= 192 && long(M) >= 8L * N; }

template <typename T, typename TA, typename TB, int CFG>
hipError_t launch_cfg(int aop, int bop, dim3 grid, int M, int N, int K, const TA* A, long lda, const TB* B, long ldb,
                      double* C, long ldc, const double* D, long ldd, double alpha, double beta, int kps,
                      double* work, hipStream_t s) {
  constexpr int TBM = kTiles[CFG].bm, TBN = kTiles[CFG].bn, WGM = CFG == 0 ? 4 : 2;
  constexpr bool PF = true;
  // measured on the Si64 shapes: the 3M FP64 tiles at two waves per SIMD
  // (three spill), the FP32 tiles at three
  constexpr int WAVES = (sizeof(T) == 8 && SDFT_GEMM_GAUSS_FP64) ? 2 : 3;
#define SDFT_LAUNCH(AA, BB)                                                                                      \
  hipLaunchKernelGGL((k_zgemm<T, TA, TB, AA, BB, TBM, TBN, WGM, PF, WAVES>), grid, dim3(256), 0, s, M, N, K, A, \
                     lda, B, ldb, C, ldc, D, ldd, alpha, beta, kps, work)
  if (!aop && !bop) SDFT_LAUNCH(false, false);
  else if (aop && !bop) SDFT_LAUNCH(true, false);
  else if (!aop && bop) SDFT_LAUNCH(false, true);
  else SDFT_LAUNCH(true, true);
#undef SDFT_LAUNCH
  return hipGetLastError();
}

template <typename T, typename TA, typename TB>
hipError_t dispatch_ops(int cfg, int aop, int bop, dim3 grid, int M, int N, int K, const TA* A, long lda, const TB* B,
                        long ldb, double* C, long ldc, const double* D, long ldd, double alpha, double beta, int kps,
                        double* work, hipStream_t s, const LaunchEvents* ev) {
  if (ev && ev->begin) {
    const hipError_t e = hipEventRecord(ev->begin, s);
    if (e != hipSuccess) return e;
  }
  hipError_t e;
  if (cfg == 0) e = launch_cfg<T, TA, TB, 0>(aop, bop, grid, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha, beta, kps, work, s);
  else e = launch_cfg<T, TA, TB, 1>(aop, bop, grid, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha, beta, kps, work, s);
  if (e != hipSuccess || !ev || !ev->end) return e;
  return hipEventRecord(ev->end, s);
}

// Split-K plan of a shape: tile configuration, splits and K per split.
struct Plan {
  int cfg, splits, kps;
  dim3 grid;
};
// splits along K: double while the workgroups number fewer than `target` and
// every split keeps >= 256 of K
int splits_for(long tiles, int K, long target) {
  int splits = 1;
  while (tiles * splits < target && long(K) / (2 * splits) >= 256 && splits < 64) splits *= 2;
  return splits;
}
long tile_count(int M, int N, int cfg) {
  return long((M + kTiles[cfg].bm - 1) / kTiles[cfg].bm) * ((N + kTiles[cfg].bn - 1) / kTiles[cfg].bn);
}
// 32 x 32 tiles split along K to >= 2048 (FP64) or >= 1024 (FP32)
// workgroups, tall shapes to >= 512 (they need no split); rotations (tall,
// normal A) 128 x 32
int plan_cfg(int M, int N, int aop) { return tall_shape(M, N) && !aop ? 0 : 1; }
int plan_splits(bool fp64, int M, int N, int K, int aop) {
  const int cfg = plan_cfg(M, N, aop);
  return splits_for(tile_count(M, N, cfg), K, tall_shape(M, N) ? 512 : fp64 ? 2048 : 1024);
}
Plan plan_for(bool fp64, int aop, int M, int N, int K, bool have_work) {
  Plan p{};
  p.splits = plan_splits(fp64, M, N, K, aop);
  if (p.splits > 1 && !have_work) p.splits = 1;
  p.kps = (K + p.splits - 1) / p.splits;
  p.kps = ((p.kps + BK - 1) / BK) * BK;
  p.splits = (K + p.kps - 1) / p.kps;
  p.cfg = plan_cfg(M, N, aop);
  p.grid = dim3((M + kTiles[p.cfg].bm - 1) / kTiles[p.cfg].bm, (N + kTiles[p.cfg].bn - 1) / kTiles[p.cfg].bn,
                p.splits);
  return p;
}
}  // namespace

int zgemm_real_products(Precision prec) {
  return (prec == Precision::FP64 ? SDFT_GEMM_GAUSS_FP64 : SDFT_GEMM_GAUSS_FP32) ? 3 : 4;
}

std::size_t zgemm_workspace_bytes(int M, int N, int K) {
  const int splits = zgemm_splits(M, N, K);
  return splits > 1 ? std::size_t(splits) * M * N * 16 : 0;
}

// the largest split count any plan of the shape uses (workspace sizing)
int zgemm_splits(int M, int N, int K) {
  int splits = 1;
  for (int aop = 0; aop < 2; ++aop)
    for (int fp64 = 0; fp64 < 2; ++fp64) splits = std::max(splits, plan_splits(fp64, M, N, K, aop));
  return splits;
}

hipError_t zgemm(Precision prec, int aop, int bop, int M, int N, int K, double alpha, const double* A, long lda,
                 const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta, double* work,
                 std::size_t work_bytes, hipStream_t s, const LaunchEvents* ev) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (K <= 0) {
    // C = beta * D
    return launch_zaxpby(M, N, 0.0, nullptr, 0, beta, D, ldd, C, ldc, s);
  }
  const bool room = work != nullptr && work_bytes >= zgemm_workspace_bytes(M, N, K);
  const Plan pl = plan_for(prec == Precision::FP64, aop, M, N, K, room);
  double* wk = pl.splits > 1 ? work : nullptr;
  hipError_t e;
  if (prec == Precision::FP64)
    e = dispatch_ops<double, double, double>(pl.cfg, aop, bop, pl.grid, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha,
                                             beta, pl.kps, wk, s, ev);
  else
    e = dispatch_ops<float, double, double>(pl.cfg, aop, bop, pl.grid, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha,
                                            beta, pl.kps, wk, s, ev);
  if (e != hipSuccess || pl.splits == 1) return e;
  hipLaunchKernelGGL(k_splitk_reduce, dim3(1024), dim3(256), 0, s, M, N, pl.splits, work, C, ldc, D, ldd, beta);
  return hipGetLastError();
}

hipError_t zgemm_f32_storage(int aop, int bop, int M, int N, int K, double alpha, const float* A, long lda,
                             const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta,
                             double* work, std::size_t work_bytes, hipStream_t s, const LaunchEvents* ev) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  const bool room = work != nullptr && work_bytes >= zgemm_workspace_bytes(M, N, K);
  const Plan pl = plan_for(false, aop, M, N, K, room);
  double* wk = pl.splits > 1 ? work : nullptr;
  hipError_t e = dispatch_ops<float, float, double>(pl.cfg, aop, bop, pl.grid, M, N, K, A, lda, B, ldb, C, ldc, D, ldd,
                                                    alpha, beta, pl.kps, wk, s, ev);
  if (e != hipSuccess || pl.splits == 1) return e;
  hipLaunchKernelGGL(k_splitk_reduce, dim3(1024), dim3(256), 0, s, M, N, pl.splits, work, C, ldc, D, ldd, beta);
  return hipGetLastError();
}

// ---- BF16 projector products (search arithmetic) ----
// C = alpha op(A) B + beta D with A the projector block stored as BF16 complex
// pairs (real part in the low 16 bits), B and C / D stored in FP64, products on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation. Lane l of that instruction
// holds 8 consecutive k of row (A) / column (B) l % 16, starting at
// k = 8 (l / 16); its accumulators are laid out as the f32 16x16 ones
// (register r: row 4 (l / 16) + r, column l % 16). Four real products per
// complex block: at BF16 precision the sums of the three-product form would
// lose the imaginary part to cancellation. One 128 x 32 tile per 256-thread
// workgroup, 4 waves of 32 x 32 along M, K slabs of 32 staged in LDS as
// [row][k] with a row stride of 40 elements (80 bytes, so the 8-element
// operand reads are aligned 16-byte LDS reads), the next slab prefetched into
// registers.
namespace {
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short s8v __attribute__((ext_vector_type(8)));
constexpr int BKH = 32, SHB = BKH + 8, HTBM = 128, HTBN = 32;

__device__ __forceinline__ __bf16 bf16_from_bits(unsigned short b) { return __builtin_bit_cast(__bf16, b); }
__device__ __forceinline__ bf8 bf8_negate(bf8 v) {
  s8v bits = __builtin_bit_cast(s8v, v);
  bits ^= s8v{short(0x8000), short(0x8000), short(0x8000), short(0x8000), short(0x8000), short(0x8000),
              short(0x8000), short(0x8000)};
  return __builtin_bit_cast(bf8, bits);
}

template <bool ADJ_A>
__global__ __launch_bounds__(256) void k_zgemm_bf16(int M, int N, int K, const unsigned* __restrict__ A, long lda,
                                                    const double* __restrict__ B, long ldb, double* __restrict__ C,
                                                    long ldc, const double* __restrict__ D, long ldd, double alpha,
                                                    double beta, int kps, double* __restrict__ work) {
  constexpr int WGM = 4, WM = HTBM / WGM, WN = HTBN, IM = WM / 16, JN = WN / 16;
  constexpr int EA = HTBM * BKH / 256, EB = HTBN * BKH / 256;
  __shared__ __attribute__((aligned(16))) __bf16 as_re[HTBM * SHB];
  __shared__ __attribute__((aligned(16))) __bf16 as_im[HTBM * SHB];
  __shared__ __attribute__((aligned(16))) __bf16 bs_re[HTBN * SHB];
  __shared__ __attribute__((aligned(16))) __bf16 bs_im[HTBN * SHB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const TileId tile = xcd_tile(gridDim.x, gridDim.y, gridDim.z);
  const int m0 = tile.m * HTBM, n0 = tile.n * HTBN;
  const int k0 = tile.split * kps, k1 = min(K, k0 + kps);
  const int wm = wave * WM;
  f4 acc_re[IM][JN], acc_im[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      acc_re[i][j] = f4{0, 0, 0, 0};
      acc_im[i][j] = f4{0, 0, 0, 0};
    }
  unsigned pa[EA];
  double2 pb[EB];
  auto load = [&](int kb) {
#pragma unroll
    for (int it = 0; it < EA; ++it) {
      const int e = tid + it * 256;
      int mm, kk;
      if (!ADJ_A) { mm = e % HTBM; kk = e / HTBM; } else { kk = e % BKH; mm = e / BKH; }
      const int m = m0 + mm, k = kb + kk;
      pa[it] = (m < M && k < k1) ? (ADJ_A ? A[k + (long)m * lda] : A[m + (long)k * lda]) : 0u;
    }
#pragma unroll
    for (int it = 0; it < EB; ++it) {
      const int e = tid + it * 256;
      const int kk = e % BKH, nn = e / BKH;
      const int n = n0 + nn, k = kb + kk;
      pb[it] = (n < N && k < k1) ? *reinterpret_cast<const double2*>(B + 2 * (k + (long)n * ldb)) : double2{0, 0};
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int it = 0; it < EA; ++it) {
      const int e = tid + it * 256;
      int mm, kk;
      if (!ADJ_A) { mm = e % HTBM; kk = e / HTBM; } else { kk = e % BKH; mm = e / BKH; }
      // op(A) = A^H conjugates: flip the sign bit of the imaginary part
      const unsigned short re = (unsigned short)(pa[it] & 0xffffu);
      const unsigned short im = (unsigned short)((pa[it] >> 16) ^ (ADJ_A ? 0x8000u : 0u));
      as_re[mm * SHB + kk] = bf16_from_bits(re);
      as_im[mm * SHB + kk] = bf16_from_bits(im);
    }
#pragma unroll
    for (int it = 0; it < EB; ++it) {
      const int e = tid + it * 256;
      const int kk = e % BKH, nn = e / BKH;
      bs_re[nn * SHB + kk] = __bf16(float(pb[it].x));
      bs_im[nn * SHB + kk] = __bf16(float(pb[it].y));
    }
  };
  if (k0 < k1) load(k0);
  for (int kb = k0; kb < k1; kb += BKH) {
    store();
    __syncthreads();
    if (kb + BKH < k1) load(kb + BKH);
    const int r = lane & 15, ko = 8 * (lane >> 4);
    bf8 ar[IM], ai[IM], br[JN], bi[JN], nbi[JN];
#pragma unroll
    for (int i = 0; i < IM; ++i) {
      ar[i] = *reinterpret_cast<const bf8*>(&as_re[(wm + 16 * i + r) * SHB + ko]);
      ai[i] = *reinterpret_cast<const bf8*>(&as_im[(wm + 16 * i + r) * SHB + ko]);
    }
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      br[j] = *reinterpret_cast<const bf8*>(&bs_re[(16 * j + r) * SHB + ko]);
      bi[j] = *reinterpret_cast<const bf8*>(&bs_im[(16 * j + r) * SHB + ko]);
      nbi[j] = bf8_negate(bi[j]);
    }
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j) {
        acc_re[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[i], br[j], acc_re[i][j], 0, 0, 0);
        acc_im[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[i], bi[j], acc_im[i][j], 0, 0, 0);
      }
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j) {
        acc_re[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai[i], nbi[j], acc_re[i][j], 0, 0, 0);
        acc_im[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai[i], br[j], acc_im[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
  const bool split = work != nullptr;
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int n = n0 + 16 * j + (lane & 15);
        if (m >= M || n >= N) continue;
        double re = alpha * double(acc_re[i][j][r]), im = alpha * double(acc_im[i][j][r]);
        if (split) {
          double* w = work + 2 * ((long)tile.split * M * N + m + (long)n * M);
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

// FP32 / FP64 complex pairs -> BF16 pairs (round to nearest even)
template <typename S>
__global__ void k_to_bf16_pairs(const S* __restrict__ x, long count, unsigned* __restrict__ out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x) {
    const unsigned short re = __builtin_bit_cast(unsigned short, __bf16(float(x[2 * e])));
    const unsigned short im = __builtin_bit_cast(unsigned short, __bf16(float(x[2 * e + 1])));
    out[e] = unsigned(re) | (unsigned(im) << 16);
  }
}
}  // namespace

hipError_t zgemm_bf16_storage(int aop, int M, int N, int K, double alpha, const unsigned* A, long lda,
                              const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta,
                              double* work, std::size_t work_bytes, hipStream_t s, const LaunchEvents* ev) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  const long tiles = long((M + HTBM - 1) / HTBM) * ((N + HTBN - 1) / HTBN);
  // long contractions split along K to >= 512 workgroups, within the workspace
  int splits = 1;
  while (tiles * splits < 512 && long(K) / (2 * splits) >= 256 && splits < 64) splits *= 2;
  while (splits > 1 && (!work || std::size_t(splits) * M * N * 16 > work_bytes)) splits /= 2;
  int kps = (K + splits - 1) / splits;
  kps = ((kps + BKH - 1) / BKH) * BKH;
  splits = (K + kps - 1) / kps;
  const dim3 grid((M + HTBM - 1) / HTBM, (N + HTBN - 1) / HTBN, splits);
  double* wk = splits > 1 ? work : nullptr;
  if (ev && ev->begin) {
    const hipError_t e = hipEventRecord(ev->begin, s);
    if (e != hipSuccess) return e;
  }
  if (aop)
    hipLaunchKernelGGL(k_zgemm_bf16<true>, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha,
                       beta, kps, wk);
  else
    hipLaunchKernelGGL(k_zgemm_bf16<false>, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, alpha,
                       beta, kps, wk);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && ev && ev->end) e = hipEventRecord(ev->end, s);
  if (e != hipSuccess || splits == 1) return e;
  hipLaunchKernelGGL(k_splitk_reduce, dim3(1024), dim3(256), 0, s, M, N, splits, work, C, ldc, D, ldd, beta);
  return hipGetLastError();
}

hipError_t launch_to_bf16_pairs(const float* x, long count, unsigned* out, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16_pairs<float>, dim3(2048), dim3(256), 0, s, x, count, out);
  return hipGetLastError();
}
hipError_t launch_to_bf16_pairs(const double* x, long count, unsigned* out, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16_pairs<double>, dim3(2048), dim3(256), 0, s, x, count, out);
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


// Loads this file's code object on the calling thread (see module_prime.hpp).
hipError_t prime_module_gemm() {
  hipFuncAttributes a{};
  return hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_splitk_reduce));
}

}  // namespace sdft_hip
