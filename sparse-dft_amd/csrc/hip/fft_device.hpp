// Device-side FFT building blocks for gfx950.
//
// * In-register DFTs of small smooth sizes (2..16) built from hand-written
//   radix-2/4 butterflies, symmetric-pair prime DFTs (3, 5, 7) and a
//   compile-time two-factor Cooley-Tukey split for composite sizes.
// * Line FFTs of length L = L1 * L2 held in LDS and computed by two
//   in-register sub-steps with an LDS exchange between them. The backward
//   transform takes natural order and leaves "slot order": slot a + L1*b holds
//   frequency/position b + L2*a. The forward transform takes slot order and
//   returns natural order. Pointwise work between them (x V(r)) therefore
//   never needs an explicit digit reversal.
//
// Sign convention: S = +1 is e^{+2 pi i nk/L} (FFTW_BACKWARD), S = -1 is
// e^{-2 pi i nk/L} (FFTW_FORWARD). Transforms are unnormalised.
#pragma once
#include <hip/hip_runtime.h>

#include "dft_constants.hpp"
#include "slot_order.hpp"

namespace sdft_hip {

// Aligned to its full size so that LDS and global accesses of one element are
// single 128-bit (FP64) or 64-bit (FP32) operations: split into two halves, a
// wave's consecutive elements would hit every LDS bank twice.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// a * (c + i s)
template <typename T>
__device__ __forceinline__ cx<T> cmul_cs(cx<T> a, T c, T s) { return {a.x * c - a.y * s, a.x * s + a.y * c}; }
// a * (S i)
template <int S, typename T>
__device__ __forceinline__ cx<T> mul_si(cx<T> a) {
  return S > 0 ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x};
}

// ------------------------------------------------------------ register DFTs
template <int N, int S, typename T>
struct Dft;

template <int S, typename T>
struct Dft<1, S, T> {
  static __device__ __forceinline__ void run(cx<T>*) {}
};

template <int S, typename T>
struct Dft<2, S, T> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    const cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int S, typename T>
struct Dft<4, S, T> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    const cx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const cx<T> c = cadd(v[1], v[3]), d = mul_si<S>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[1] = cadd(b, d);
    v[2] = csub(a, c);
    v[3] = csub(b, d);
  }
};

// Odd prime N: X_k = x0 + sum_{j=1}^{(N-1)/2} [(x_j + x_{N-j}) cos(2 pi jk/N)
//                                          + S i (x_j - x_{N-j}) sin(2 pi jk/N)]
template <int N, int S, typename T>
struct DftOddPrime {
  static __device__ __forceinline__ void run(cx<T>* v) {
    constexpr int H = (N - 1) / 2;
    cx<T> sum[H], dif[H];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      sum[j - 1] = cadd(v[j], v[N - j]);
      dif[j - 1] = csub(v[j], v[N - j]);
    }
    const cx<T> x0 = v[0];
    cx<T> total = x0;
#pragma unroll
    for (int j = 0; j < H; ++j) total = cadd(total, sum[j]);
    v[0] = total;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      cx<T> re = x0, im = {T(0), T(0)};
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const T c = T(DftTw<N>::c[(j * k) % N]);
        const T s = T(DftTw<N>::s[(j * k) % N]);
        re.x += c * sum[j - 1].x;
        re.y += c * sum[j - 1].y;
        im.x += s * dif[j - 1].x;
        im.y += s * dif[j - 1].y;
      }
      const cx<T> rot = mul_si<S>(im);  // S i * im
      v[k] = cadd(re, rot);
      v[N - k] = csub(re, rot);
    }
  }
};
template <int S, typename T>
struct Dft<3, S, T> : DftOddPrime<3, S, T> {};
template <int S, typename T>
struct Dft<5, S, T> : DftOddPrime<5, S, T> {};
template <int S, typename T>
struct Dft<7, S, T> : DftOddPrime<7, S, T> {};

// Composite N = A * B: n = n1 + A n2, k = k2 + B k1.
template <int N, int A, int B, int S, typename T>
struct DftSplit {
  static __device__ __forceinline__ void run(cx<T>* v) {
    cx<T> t[A][B];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n1][n2] = v[n1 + A * n2];
      Dft<B, S, T>::run(t[n1]);
#pragma unroll
      for (int k2 = 1; k2 < B; ++k2)
        if (n1 * k2) {
          const int j = (n1 * k2) % N;
          t[n1][k2] = cmul_cs(t[n1][k2], T(DftTw<N>::c[j]), T(S * DftTw<N>::s[j]));
        }
    }
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      cx<T> u[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) u[n1] = t[n1][k2];
      Dft<A, S, T>::run(u);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) v[k2 + B * k1] = u[k1];
    }
  }
};
template <int S, typename T>
struct Dft<6, S, T> : DftSplit<6, 2, 3, S, T> {};
template <int S, typename T>
struct Dft<8, S, T> : DftSplit<8, 2, 4, S, T> {};
template <int S, typename T>
struct Dft<9, S, T> : DftSplit<9, 3, 3, S, T> {};
template <int S, typename T>
struct Dft<10, S, T> : DftSplit<10, 2, 5, S, T> {};
template <int S, typename T>
struct Dft<12, S, T> : DftSplit<12, 3, 4, S, T> {};
template <int S, typename T>
struct Dft<14, S, T> : DftSplit<14, 2, 7, S, T> {};
template <int S, typename T>
struct Dft<15, S, T> : DftSplit<15, 3, 5, S, T> {};
template <int S, typename T>
struct Dft<16, S, T> : DftSplit<16, 4, 4, S, T> {};

// ------------------------------------------------------------ LDS line FFTs
// Lines live in LDS: element i of line l is at base[l * ls + i * es], plus
// `jump` for lines l >= split (two runs of lines in one pass; split = nlines
// and jump = 0 for a single run).
// tw[j] = (cos 2 pi j/L, sin 2 pi j/L) for j in [0, L).
// `fast_lines`: consecutive threads take consecutive lines in both passes
// (use when ls == 1). `fast_group`: the same in the digit-group pass only; with
// contiguous lines (es == 1) of odd stride ls this keeps that pass free of LDS
// bank conflicts (a thread of the group pass reads L1 consecutive elements, so
// threads of one line would sit L1 elements apart).
struct LineGeom {
  int nlines, ls, es;
  bool fast_lines;
  int split = 1 << 30, jump = 0;
  bool fast_group = false;
  __device__ __forceinline__ int line_offset(int line) const { return line * ls + (line >= split ? jump : 0); }
};

template <int NV, typename T>
__device__ __forceinline__ void item_split(int item, int per_line, const LineGeom& g, bool fast, int& line, int& idx) {
  if (fast) {
    line = item % g.nlines;
    idx = item / g.nlines;
  } else {
    line = item / per_line;
    idx = item % per_line;
  }
}

// Backward-style step A: per (line, n1): DFT_L2 over i = n1 + L1 n2, twiddle
// W^{S n1 k2}, store in place.  Used as step 1 of natural->slot and (with the
// roles of the digits exchanged) as step 2' of slot->natural without twiddles.
template <int L1, int L2, int S, bool TWIDDLE, typename T>
__device__ __forceinline__ void lds_strided_pass(cx<T>* base, const LineGeom& g, const cx<T>* tw, int tid,
                                                 int nthreads) {
  constexpr int L = L1 * L2;
  const int items = g.nlines * L1;
  for (int item = tid; item < items; item += nthreads) {
    int line, n1;
    item_split<L2, T>(item, L1, g, g.fast_lines, line, n1);
    cx<T>* p = base + g.line_offset(line) + n1 * g.es;
    cx<T> v[L2];
#pragma unroll
    for (int n2 = 0; n2 < L2; ++n2) v[n2] = p[n2 * L1 * g.es];
    Dft<L2, S, T>::run(v);
    if (TWIDDLE) {
#pragma unroll
      for (int k2 = 1; k2 < L2; ++k2) {
        const cx<T> w = tw[(n1 * k2) % L];
        v[k2] = cmul_cs(v[k2], w.x, T(S) * w.y);
      }
    }
#pragma unroll
    for (int k2 = 0; k2 < L2; ++k2) p[k2 * L1 * g.es] = v[k2];
  }
}

// Step over contiguous digit groups: per (line, b): DFT_L1 over i = a + L1 b
// (a in [0, L1)), optional twiddle W^{S b k} on output k, store at k + L1 b.
template <int L1, int L2, int S, bool TWIDDLE, typename T>
__device__ __forceinline__ void lds_group_pass(cx<T>* base, const LineGeom& g, const cx<T>* tw, int tid,
                                               int nthreads) {
  constexpr int L = L1 * L2;
  const int items = g.nlines * L2;
  for (int item = tid; item < items; item += nthreads) {
    int line, b;
    item_split<L1, T>(item, L2, g, g.fast_lines || g.fast_group, line, b);
    cx<T>* p = base + g.line_offset(line) + b * L1 * g.es;
    cx<T> v[L1];
#pragma unroll
    for (int a = 0; a < L1; ++a) v[a] = p[a * g.es];
    Dft<L1, S, T>::run(v);
    if (TWIDDLE) {
#pragma unroll
      for (int k = 1; k < L1; ++k) {
        const cx<T> w = tw[(b * k) % L];
        v[k] = cmul_cs(v[k], w.x, T(S) * w.y);
      }
    }
#pragma unroll
    for (int k = 0; k < L1; ++k) p[k * g.es] = v[k];
  }
}

// natural -> slot order, sign S (use S = +1 for the backward transform)
template <int L1, int L2, int S, typename T>
__device__ __forceinline__ void lds_fft_to_slots(cx<T>* base, const LineGeom& g, const cx<T>* tw, int tid, int nt) {
  lds_strided_pass<L1, L2, S, true, T>(base, g, tw, tid, nt);
  __syncthreads();
  lds_group_pass<L1, L2, S, false, T>(base, g, tw, tid, nt);
  __syncthreads();
}

// slot -> natural order, sign S (use S = -1 for the forward transform)
template <int L1, int L2, int S, typename T>
__device__ __forceinline__ void lds_fft_from_slots(cx<T>* base, const LineGeom& g, const cx<T>* tw, int tid,
                                                   int nt) {
  lds_group_pass<L1, L2, S, true, T>(base, g, tw, tid, nt);
  __syncthreads();
  lds_strided_pass<L1, L2, S, false, T>(base, g, tw, tid, nt);
  __syncthreads();
}

}  // namespace sdft_hip
