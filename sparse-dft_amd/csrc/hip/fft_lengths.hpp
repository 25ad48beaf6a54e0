// Supported FFT axis lengths and their two-factor split L = L1 * L2 (both
// factors are in-register DFT sizes). Shared by host planning and kernels so
// that the slot order (slot a + L1 b holds position b + L2 a) agrees.
#pragma once

namespace sdft_hip {

struct LengthSplit {
  int length, l1, l2;
};

// 2,3,5,7-smooth axis lengths up to 128 whose split fits in-register DFTs.
#define SDFT_HIP_FOR_EACH_LENGTH(X) \
  X(2, 1, 2)                        \
  X(3, 1, 3)                        \
  X(4, 2, 2)                        \
  X(5, 1, 5)                        \
  X(6, 2, 3)                        \
  X(7, 1, 7)                        \
  X(8, 2, 4)                        \
  X(9, 3, 3)                        \
  X(10, 2, 5)                       \
  X(12, 3, 4)                       \
  X(14, 2, 7)                       \
  X(15, 3, 5)                       \
  X(16, 4, 4)                       \
  X(18, 3, 6)                       \
  X(20, 4, 5)                       \
  X(21, 3, 7)                       \
  X(24, 4, 6)                       \
  X(25, 5, 5)                       \
  X(27, 3, 9)                       \
  X(28, 4, 7)                       \
  X(30, 5, 6)                       \
  X(32, 4, 8)                       \
  X(35, 5, 7)                       \
  X(36, 6, 6)                       \
  X(40, 5, 8)                       \
  X(42, 6, 7)                       \
  X(45, 5, 9)                       \
  X(48, 6, 8)                       \
  X(49, 7, 7)                       \
  X(50, 5, 10)                      \
  X(54, 6, 9)                       \
  X(56, 7, 8)                       \
  X(60, 6, 10)                      \
  X(63, 7, 9)                       \
  X(64, 8, 8)                       \
  X(70, 7, 10)                      \
  X(72, 8, 9)                       \
  X(75, 5, 15)                      \
  X(80, 8, 10)                      \
  X(81, 9, 9)                       \
  X(84, 7, 12)                      \
  X(90, 9, 10)                      \
  X(96, 8, 12)                      \
  X(98, 7, 14)                      \
  X(100, 10, 10)                    \
  X(105, 7, 15)                     \
  X(108, 9, 12)                     \
  X(112, 7, 16)                     \
  X(120, 10, 12)                    \
  X(126, 9, 14)                     \
  X(128, 8, 16)

inline bool length_split(int length, int& l1, int& l2) {
#define SDFT_HIP_CASE(L, A, B) \
  if (length == L) {           \
    l1 = A;                    \
    l2 = B;                    \
    return true;               \
  }
  SDFT_HIP_FOR_EACH_LENGTH(SDFT_HIP_CASE)
#undef SDFT_HIP_CASE
  return false;
}

}  // namespace sdft_hip
