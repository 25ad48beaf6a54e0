// Fixed-order 64-lane sums of doubles through DPP moves and lane reads, shared
// by the tridiagonalisation and the tridiagonal eigensolver kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace sdft_hip {
namespace wave {

// v from lane (lane ^ 1), (lane ^ 2) (quad permutations), the mirrored lane of
// the 8-lane half row, the mirrored lane of the 16-lane row (DPP controls)
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
__device__ inline double read_lane(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane(int(b), lane);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
  return __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
}
// Sum over the 64 lanes in a fixed order, every lane gets it: pairs, quads,
// half rows and rows through DPP moves, then the four row sums. Every call
// site runs it with all 64 lanes active (wave-uniform control flow).
__device__ inline double sum(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

}  // namespace wave
}  // namespace sdft_hip
