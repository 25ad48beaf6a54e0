// Statically allocated device staging (see scratch.hpp).
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <stdexcept>
#include <string>

#include "scratch.hpp"

namespace sdft_hip {
namespace {

// Each GPU that loads this code object holds its own copy. In host code the
// name is a host-side stand-in, never a device address: host code uses the
// address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
std::array<char*, kMaxDevices> g_scratch{};

}  // namespace

char* scratch_region(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("HIP: device index out of range for scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[std::size_t(device)]) {
    int current = -1;
    if (hipGetDevice(&current) != hipSuccess || current != device)
      throw std::logic_error("HIP: scratch address resolved with another device current");
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw std::runtime_error(std::string("HIP: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[std::size_t(device)] = static_cast<char*>(ptr);
  }
  return g_scratch[std::size_t(device)];
}

}  // namespace sdft_hip
