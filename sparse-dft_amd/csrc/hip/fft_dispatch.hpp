// Runtime-length dispatch of the LDS line FFTs (fft_device.hpp) over the
// supported axis lengths (fft_lengths.hpp), shared by the kernels that keep
// lines of several lengths in LDS.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_device.hpp"
#include "fft_lengths.hpp"

namespace sdft_hip {

// LDS stride of contiguous lines (z-sticks, plane rows): odd, so that threads
// one line apart hit different LDS banks.
__host__ __device__ inline int line_stride(int n) { return n | 1; }

// natural -> slot order (TO_SLOTS) or slot -> natural order, sign S, for a
// line length of the supported table; other lengths are left untouched.
// Ends with a block barrier.
template <typename T, int S, bool TO_SLOTS>
__device__ __forceinline__ void line_fft_dispatch(int length, cx<T>* base, const LineGeom& g, const cx<T>* tw, int tid,
                                                  int nt) {
  switch (length) {
#define SDFT_HIP_CASE(L, A, B)                                 \
  case L:                                                      \
    if (TO_SLOTS)                                              \
      lds_fft_to_slots<A, B, S, T>(base, g, tw, tid, nt);      \
    else                                                       \
      lds_fft_from_slots<A, B, S, T>(base, g, tw, tid, nt);    \
    break;
    SDFT_HIP_FOR_EACH_LENGTH(SDFT_HIP_CASE)
#undef SDFT_HIP_CASE
    default:
      break;
  }
}

// First factor L1 of the split of a supported length (0 otherwise), at
// compile time.
constexpr int length_l1(int length) {
#define SDFT_HIP_CASE(L, A, B) length == L ? A:
  return SDFT_HIP_FOR_EACH_LENGTH(SDFT_HIP_CASE) 0;
#undef SDFT_HIP_CASE
}

// The same transforms with the length fixed at compile time (LEN > 0), or the
// runtime dispatch (LEN = 0). A fixed length keeps only its own register DFTs
// in the kernel, so the kernel needs far fewer registers.
template <typename T, int S, bool TO_SLOTS, int LEN>
__device__ __forceinline__ void line_fft(int length, cx<T>* base, const LineGeom& g, const cx<T>* tw, int tid, int nt) {
  if constexpr (LEN > 0) {
    constexpr int A = length_l1(LEN), B = LEN / (A ? A : 1);
    static_assert(A > 0 && A * B == LEN, "unsupported fixed FFT length");
    if (TO_SLOTS)
      lds_fft_to_slots<A, B, S, T>(base, g, tw, tid, nt);
    else
      lds_fft_from_slots<A, B, S, T>(base, g, tw, tid, nt);
  } else {
    line_fft_dispatch<T, S, TO_SLOTS>(length, base, g, tw, tid, nt);
  }
}

}  // namespace sdft_hip
