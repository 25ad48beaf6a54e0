// The `hip` Accelerator: the reference's backend interface
// (src/compute_backend.hpp:64-89) implemented over the C ABI of
// include/sdft_hip.h. Semantics follow the CUDA backend
// (src/compute_cuda.cu:35-212): synchronous methods with results in host
// memory, identity-token caches for immutable operands, "HIP: " prefixed
// std::runtime_error / std::invalid_argument / std::overflow_error.
//
// Arithmetic precision of the search calls (the dispatcher only routes
// exact=false work here): SDFT_HIP_SEARCH_PRECISION=fp64 (default, same as the
// CUDA backend) or fp32 (FP32 MFMA / FP32 FFT, the FP32-search form).
#ifdef SDFT_HIP_WITH_REFERENCE_HEADER
#include "compute_backend.hpp"  // the reference's own header (integration build)
#else
#include "accelerator_iface.hpp"
#endif
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>

#include "../../../include/sdft_hip.h"

namespace sdft {
namespace {

void raise(int status) {
  if (status == SDFT_HIP_OK) return;
  const std::string msg = sdft_hip_last_error();
  if (status == SDFT_HIP_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  if (status == SDFT_HIP_OVERFLOW) throw std::overflow_error(msg);
  throw std::runtime_error(msg);
}

int search_precision() {
  const char* v = std::getenv("SDFT_HIP_SEARCH_PRECISION");
  if (!v || std::string(v) == "fp64") return 0;
  if (std::string(v) == "fp32") return 1;
  throw std::invalid_argument("SDFT_HIP_SEARCH_PRECISION must be fp64 or fp32");
}

const double* D(const Complex64* p) { return reinterpret_cast<const double*>(p); }
double* D(Complex64* p) { return reinterpret_cast<double*>(p); }

class HipAccelerator final : public Accelerator {
 public:
  HipAccelerator() {
    int device = 0;
    if (const char* v = std::getenv("SDFT_HIP_DEVICE")) device = std::atoi(v);
    raise(sdft_hip_context_create(device, &ctx_));
    char name[256];
    raise(sdft_hip_device_name(ctx_, name, sizeof name));
    device_ = name;
    precision_ = search_precision();
  }
  ~HipAccelerator() override { sdft_hip_context_destroy(ctx_); }
  HipAccelerator(const HipAccelerator&) = delete;
  HipAccelerator& operator=(const HipAccelerator&) = delete;

  std::string device() const override { return device_; }

  void gemm(MatrixOp ao, MatrixOp bo, std::size_t m, std::size_t n, std::size_t k, const Complex64* a,
            std::size_t lda, const Complex64* b, std::size_t ldb, Complex64* c, std::size_t ldc, bool add) override {
    raise(sdft_hip_zgemm(ctx_, ao == MatrixOp::Adjoint, bo == MatrixOp::Adjoint, m, n, k, D(a), lda, D(b), ldb, D(c),
                         ldc, add ? 1 : 0, 0, precision_));
  }

  void gemm_immutable_left(std::uint64_t identity, MatrixOp ao, std::size_t m, std::size_t n, std::size_t k,
                           const Complex64* a, std::size_t lda, const Complex64* b, std::size_t ldb, Complex64* c,
                           std::size_t ldc, bool add) override {
    raise(sdft_hip_zgemm(ctx_, ao == MatrixOp::Adjoint, 0, m, n, k, D(a), lda, D(b), ldb, D(c), ldc, add ? 1 : 0,
                         identity, precision_));
  }

  void gemm_triple_right(std::size_t m, std::size_t n, std::size_t k, const std::array<const Complex64*, 3>& a,
                         std::size_t lda, const Complex64* b, std::size_t ldb, const std::array<Complex64*, 3>& c,
                         std::size_t ldc) override {
    const double* aa[3] = {D(a[0]), D(a[1]), D(a[2])};
    double* cc[3] = {D(c[0]), D(c[1]), D(c[2])};
    raise(sdft_hip_zgemm_triple_right(ctx_, m, n, k, aa, lda, D(b), ldb, cc, ldc, precision_));
  }

  void local_hamiltonian(const std::array<int, 3>& shape, const std::vector<std::size_t>& indices,
                         const std::vector<double>& kinetic, const std::vector<double>& potential,
                         const Complex64* input, std::size_t columns, Complex64* output,
                         std::uint64_t operator_identity) override {
    if (kinetic.size() != indices.size())
      throw std::invalid_argument("HIP: accelerated FFT kinetic/index size mismatch");
    const std::size_t count = std::size_t(shape[0]) * shape[1] * shape[2];
    if (potential.size() != count) throw std::invalid_argument("HIP: accelerated FFT potential grid mismatch");
    static_assert(sizeof(std::size_t) == sizeof(std::uint64_t), "size_t must be 64-bit");
    raise(sdft_hip_local_hamiltonian(ctx_, shape.data(), reinterpret_cast<const std::uint64_t*>(indices.data()),
                                     indices.size(), kinetic.data(), potential.data(), D(input), columns, D(output),
                                     operator_identity, precision_));
  }

  bool paw_pair(std::uint64_t identity, std::size_t npw, std::size_t np, std::size_t columns, const Complex64* b,
                const Complex64* d, const Complex64* q, const Complex64* input, Complex64* h, Complex64* s,
                std::uint64_t projector_identity) override {
    int handled = 0;
    raise(sdft_hip_paw_pair(ctx_, identity, npw, np, columns, D(b), D(d), D(q), D(input), D(h), D(s),
                            projector_identity, precision_, &handled));
    return handled != 0;
  }

 private:
  sdft_hip_context* ctx_ = nullptr;
  std::string device_;
  int precision_ = 0;
};

}  // namespace

std::unique_ptr<Accelerator> make_hip_accelerator() { return std::make_unique<HipAccelerator>(); }

}  // namespace sdft
