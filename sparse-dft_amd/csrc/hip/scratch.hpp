// Fixed per-device staging memory for the entry point that allocates nothing
// (callers under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): one statically sized device array per GPU, alive until the
// process ends and shared by every such call on that GPU.
#pragma once
#include <cstddef>

namespace sdft_hip {

// 128 MiB: the stick buffers, projections and split-K partials of about 50
// FP64 columns of a 72^3-grid, 512-projector operator in one pass.
constexpr std::size_t kScratchBytes = std::size_t(128) << 20;

// Device address of `device`'s copy of the array. The first call for a device
// must be made with that device current (the context constructor does it);
// later calls return the address kept for it.
char* scratch_region(int device);

}  // namespace sdft_hip
