// Declarations of the reference's accelerator interface
// (src/compute_backend.hpp:13-91), restated so that the HIP backend compiles
// stand-alone in this repository. The class layout (virtual functions, their
// order and signatures) is kept identical so that hip_accelerator.cpp also
// compiles unchanged against the reference's own header, which is how a
// maintainer integrates it.
#pragma once
#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sdft {
using Complex64 = std::complex<double>;
enum class MatrixOp { Normal, Adjoint };
struct PbeXcPoint;  // src/pbe_xc.hpp; only passed by pointer here

class Accelerator {
 public:
  virtual ~Accelerator() = default;
  virtual std::string device() const = 0;
  virtual void gemm(MatrixOp, MatrixOp, std::size_t, std::size_t, std::size_t, const Complex64*, std::size_t,
                    const Complex64*, std::size_t, Complex64*, std::size_t, bool) = 0;
  virtual void gemm_immutable_left(std::uint64_t, MatrixOp ao, std::size_t m, std::size_t n, std::size_t k,
                                   const Complex64* a, std::size_t lda, const Complex64* b, std::size_t ldb,
                                   Complex64* c, std::size_t ldc, bool add) {
    gemm(ao, MatrixOp::Normal, m, n, k, a, lda, b, ldb, c, ldc, add);
  }
  virtual void gemm_triple_right(std::size_t m, std::size_t n, std::size_t k,
                                 const std::array<const Complex64*, 3>& a, std::size_t lda, const Complex64* b,
                                 std::size_t ldb, const std::array<Complex64*, 3>& c, std::size_t ldc) {
    for (std::size_t i = 0; i < 3; ++i)
      gemm(MatrixOp::Normal, MatrixOp::Normal, m, n, k, a[i], lda, b, ldb, c[i], ldc, false);
  }
  virtual void local_hamiltonian(const std::array<int, 3>&, const std::vector<std::size_t>&,
                                 const std::vector<double>&, const std::vector<double>&, const Complex64*,
                                 std::size_t, Complex64*, std::uint64_t) = 0;
  virtual void pbe_points(const double*, const double*, std::size_t, PbeXcPoint*) {
    throw std::runtime_error("native PBE is not implemented for this accelerator");
  }
  virtual bool paw_pair(std::uint64_t, std::size_t, std::size_t, std::size_t, const Complex64*, const Complex64*,
                        const Complex64*, const Complex64*, Complex64*, Complex64*, std::uint64_t) {
    return false;
  }
};

// The new factory a maintainer adds beside make_cuda_accelerator() /
// make_metal_accelerator() (src/compute_backend.hpp:90-91).
std::unique_ptr<Accelerator> make_hip_accelerator();
}  // namespace sdft
