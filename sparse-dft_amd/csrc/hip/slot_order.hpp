// Slot order of the two-step line FFT, shared by device and host code.
//
// A length L1 * L2 line transformed backward from natural order is left in
// slot order: slot a + L1*b holds natural index b + L2*a.
#pragma once

#if defined(__HIPCC__)
#define SDFT_HOST_DEVICE __host__ __device__
#else
#define SDFT_HOST_DEVICE
#endif

namespace sdft_hip {

// Position (natural index) held by slot s for a length L1 * L2 line.
SDFT_HOST_DEVICE inline int slot_position(int s, int L1, int L2) {
  const int a = s % L1, b = s / L1;
  return b + L2 * a;
}

}  // namespace sdft_hip
