// Dense Hermitian eigensolver of the Rayleigh-Ritz step (see dense_eigen.hpp).
#include "dense_eigen.hpp"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "../hip/tridiag.hpp"

namespace sdft_hip {

HermitianEigensolver::HermitianEigensolver(Context& ctx, long max_order) : ctx_(ctx), max_order_(max_order) {
  if (max_order < 1) throw std::invalid_argument("HIP: invalid eigensolver order");
  refl_.reserve(std::size_t(2 * max_order * max_order));
  z_.reserve(std::size_t(2 * max_order * max_order));
  small_.reserve(std::size_t(4 * max_order + 64));
  work_.reserve(hetrd_workspace_bytes(int(std::min<long>(max_order, kHetrdMaxOrder))));
}

std::vector<double> HermitianEigensolver::solve(const double* a, long m, long k, double* vecs) {
  if (m < 1 || m > max_order_ || k < 0 || k > m) throw std::invalid_argument("HIP: invalid Hermitian eigensolve shape");
  hipStream_t s = ctx_.stream();
  auto* Zm = reinterpret_cast<rocblas_double_complex*>(z_.p);
  int* info = reinterpret_cast<int*>(small_.p);
  double* d = small_.p + 8;
  double* e = d + m;
  double* tau = e + m;  // m complex
  int rpw = 0, groups = 0;
  const bool fused = hetrd_plan(int(m), &rpw, &groups);
  if (fused) {
    check(launch_hetrd(a, m, int(m), refl_.p, m, d, e, tau, work_.p, s), "tridiagonalisation");
    check(rocsolver_zstedc(ctx_.blas(), rocblas_evect_tridiagonal, rocblas_int(m), d, e, Zm, rocblas_int(m), info),
          "rocsolver_zstedc");
    if (k > 0 && m > 1)
      check(rocsolver_zunmtr(ctx_.blas(), rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                             rocblas_int(m), rocblas_int(k), reinterpret_cast<rocblas_double_complex*>(refl_.p),
                             rocblas_int(m), reinterpret_cast<rocblas_double_complex*>(tau), Zm, rocblas_int(m)),
            "rocsolver_zunmtr");
    ++fused_;
  } else {
    check(hipMemcpyAsync(z_.p, a, 16 * std::size_t(m) * m, hipMemcpyDeviceToDevice, s), "copy");
    check(rocsolver_zheevd(ctx_.blas(), rocblas_evect_original, rocblas_fill_lower, rocblas_int(m), Zm,
                           rocblas_int(m), d, e, info),
          "rocsolver_zheevd");
  }
  if (k > 0) check(hipMemcpyAsync(vecs, z_.p, 16 * std::size_t(m) * k, hipMemcpyDeviceToDevice, s), "vectors");
  std::vector<double> h(8 + m);
  check(hipMemcpyAsync(h.data(), small_.p, 8 * std::size_t(8 + m), hipMemcpyDeviceToHost, s), "eigenvalues");
  int aborted = 0;
  if (fused && m > 1)
    check(hipMemcpyAsync(&aborted, hetrd_abort_flag(work_.p, int(m)), sizeof(int), hipMemcpyDeviceToHost, s),
          "abort flag");
  ctx_.sync();
  if (aborted) throw std::runtime_error("HIP: tridiagonalisation grid barrier expired");
  int status = 0;
  std::memcpy(&status, h.data(), sizeof(int));
  if (status) throw std::runtime_error("HIP: Hermitian eigensolve failed");
  return std::vector<double>(h.begin() + 8, h.end());
}

}  // namespace sdft_hip
