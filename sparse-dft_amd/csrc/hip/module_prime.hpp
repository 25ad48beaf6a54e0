// Code objects of the library's kernel files, loaded on the calling thread.
// The HIP runtime loads a file's code object at the first use of one of its
// kernels, and rocBLAS loads its GEMM kernel library at its first call; a
// context primes both on the thread that creates it (the first context of the
// process does), so that neither happens first on a short-lived worker thread.
#pragma once
#include <hip/hip_runtime_api.h>

namespace sdft_hip {

hipError_t prime_module_local_h();
hipError_t prime_module_gemm();
hipError_t prime_module_davidson_kernels();
hipError_t prime_module_tridiag();
hipError_t prime_module_tridiag_eig();
hipError_t prime_module_realspace();
hipError_t prime_module_batch();

}  // namespace sdft_hip
