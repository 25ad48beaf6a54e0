// Complex GEMM and PAW onsite-block launchers (MFMA, column-major, interleaved
// complex storage in FP64 unless stated otherwise).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace sdft_hip {

enum class Precision { FP64 = 0, FP32 = 1 };

// Optional events recorded on the launch stream immediately before and after
// the MFMA kernel of one GEMM (the split-K reduction is outside the pair).
struct LaunchEvents {
  hipEvent_t begin = nullptr, end = nullptr;
};

int zgemm_splits(int M, int N, int K);
std::size_t zgemm_workspace_bytes(int M, int N, int K);

// C = alpha op(A) op(B) + beta * D; aop/bop: 0 = Normal, 1 = Adjoint. D may
// alias C. `prec` selects the arithmetic (FP64 or FP32 MFMA); storage is FP64.
hipError_t zgemm(Precision prec, int aop, int bop, int M, int N, int K, double alpha, const double* A, long lda,
                 const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta, double* work,
                 std::size_t work_bytes, hipStream_t s, const LaunchEvents* ev = nullptr);
// Real MFMA products per complex multiply-add of the kernels: 3 (Gauss form)
// or 4. The conventional 8 m n k flop count of a complex GEMM corresponds to 4.
int zgemm_real_products(Precision prec);
// Same with A stored in FP32 (resident search copy of the projectors).
hipError_t zgemm_f32_storage(int aop, int bop, int M, int N, int K, double alpha, const float* A, long lda,
                             const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta,
                             double* work, std::size_t work_bytes, hipStream_t s, const LaunchEvents* ev = nullptr);

// C = alpha op(A) B + beta D (aop 0 = Normal, 1 = Adjoint) with A stored as BF16
// complex pairs (real part in the low 16 bits of each 32-bit word), B, C, D
// in FP64: BF16 MFMA products with FP32 accumulation (search arithmetic).
hipError_t zgemm_bf16_storage(int aop, int M, int N, int K, double alpha, const unsigned* A, long lda,
                              const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta,
                              double* work, std::size_t work_bytes, hipStream_t s, const LaunchEvents* ev = nullptr);
// count complex values (FP32 or FP64 pairs) -> BF16 pairs, round to nearest even
hipError_t launch_to_bf16_pairs(const float* x, long count, unsigned* out, hipStream_t s);
hipError_t launch_to_bf16_pairs(const double* x, long count, unsigned* out, hipStream_t s);

// W(:, 0:m) = D_a P_a and W(:, m:2m) = Q_a P_a per atom block (which_mask bit
// 0: D, bit 1: Q). FP64 uses the compensated scalar order of the reference.
hipError_t launch_atom_blocks(Precision prec, int np, int columns, int atoms, const int* atom_offset,
                              const int* atom_rows, const long* block_offset, const double* dblk, const double* qblk,
                              const double* P, long ldp, double* W, long ldw, int which_mask, hipStream_t s);

// C = alpha X + beta D
hipError_t launch_zaxpby(int M, int N, double alpha, const double* X, long ldx, double beta, const double* D, long ldd,
                         double* C, long ldc, hipStream_t s);

}  // namespace sdft_hip
