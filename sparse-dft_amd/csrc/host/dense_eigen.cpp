// Dense Hermitian eigensolver of the Rayleigh-Ritz step (see dense_eigen.hpp).
#include "dense_eigen.hpp"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../hip/davidson_kernels.hpp"
#include "../hip/tridiag.hpp"
#include "../hip/tridiag_eig.hpp"

namespace sdft_hip {
namespace {

// SDFT_HIP_TRIDIAG_EIGEN=stedc selects rocSOLVER divide and conquer for the
// tridiagonal eigenproblem instead of multisection + inverse iteration.
bool use_stedc() {
  static const bool v = [] {
    const char* s = std::getenv("SDFT_HIP_TRIDIAG_EIGEN");
    return s && std::string(s) == "stedc";
  }();
  return v;
}

// Largest |G - I| accepted from the inverse-iteration vectors before the
// symmetric orthonormalisation; beyond it the divide-and-conquer path runs.
constexpr double kMaxDefect = 1e-6;

}  // namespace

HermitianEigensolver::HermitianEigensolver(Context& ctx, long max_order) : ctx_(ctx), max_order_(max_order) {
  if (max_order < 1) throw std::invalid_argument("HIP: invalid eigensolver order");
  refl_.reserve(std::size_t(2 * max_order * max_order));
  // entries outside the reflector tails are never written by the reduction
  check(hipMemsetAsync(refl_.p, 0, refl_.n * sizeof(double), ctx.stream()), "memset");
  z_.reserve(std::size_t(2 * max_order * max_order));
  small_.reserve(std::size_t(5 * max_order + 64));
  work_.reserve(hetrd_workspace_bytes(int(std::min<long>(max_order, kHetrdMaxOrder))));
}

std::vector<double> HermitianEigensolver::solve(const double* a, long m, long k, double* vecs, long values) {
  if (m < 1 || m > max_order_ || k < 0 || k > m) throw std::invalid_argument("HIP: invalid Hermitian eigensolve shape");
  const long count = values < 0 ? m : std::max(std::max(k, std::min(values, m)), 1L);
  hipStream_t s = ctx_.stream();
  auto* Zm = reinterpret_cast<rocblas_double_complex*>(z_.p);
  int* info = reinterpret_cast<int*>(small_.p);
  double* d = small_.p + 8;
  double* e = d + m;
  double* tau = e + m;  // m complex
  int rpw = 0, groups = 0;
  const bool fused = hetrd_plan(int(m), &rpw, &groups);
  if (!fused) return solve_zheevd(a, m, k, count, vecs);
  const bool force_expiry = ctx_.take_fault(Context::kFaultBarrierExpiry);
  check(launch_hetrd(a, m, int(m), refl_.p, m, d, e, tau, work_.p, s, force_expiry), "tridiagonalisation");
  ++fused_;
  if (!use_stedc() && m <= kTridiagEigMaxOrder) {
    std::vector<double> w;
    bool aborted = false;
    if (inverse_iteration(m, k, count, vecs, w, &aborted)) return w;
    if (aborted) {
      // `a` is unchanged: the whole solve is redone by divide and conquer
      ++barrier_fallbacks_;
      return solve_zheevd(a, m, k, count, vecs);
    }
    ++fallbacks_;
  }
  check(rocsolver_zstedc(ctx_.blas(), rocblas_evect_tridiagonal, rocblas_int(m), d, e, Zm, rocblas_int(m), info),
        "rocsolver_zstedc");
  ++stedc_;
  if (k > 0 && m > 1)
    check(rocsolver_zunmtr(ctx_.blas(), rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                           rocblas_int(m), rocblas_int(k), reinterpret_cast<rocblas_double_complex*>(refl_.p),
                           rocblas_int(m), reinterpret_cast<rocblas_double_complex*>(tau), Zm, rocblas_int(m)),
          "rocsolver_zunmtr");
  if (k > 0) check(hipMemcpyAsync(vecs, z_.p, 16 * std::size_t(m) * k, hipMemcpyDeviceToDevice, s), "vectors");
  std::vector<double> h(8 + m);
  check(hipMemcpyAsync(h.data(), small_.p, 8 * std::size_t(8 + m), hipMemcpyDeviceToHost, s), "eigenvalues");
  if (barrier_expired(m)) {
    ++barrier_fallbacks_;
    return solve_zheevd(a, m, k, count, vecs);
  }
  int status = 0;
  std::memcpy(&status, h.data(), sizeof(int));
  if (status) throw std::runtime_error("HIP: Hermitian eigensolve failed");
  return std::vector<double>(h.begin() + 8, h.begin() + 8 + count);
}

// Synchronises the stream; true when the reduction's bounded grid barrier
// expired (its outputs are then invalid).
bool HermitianEigensolver::barrier_expired(long m) {
  int aborted = 0;
  if (m > 1)
    check(hipMemcpyAsync(&aborted, hetrd_abort_flag(work_.p, int(m)), sizeof(int), hipMemcpyDeviceToHost,
                         ctx_.stream()),
          "abort flag");
  ctx_.sync();
  return aborted != 0;
}

// rocSOLVER zheevd on a copy of `a` (orders above the one-launch reduction's
// limit, and the recovery after an expired grid barrier).
std::vector<double> HermitianEigensolver::solve_zheevd(const double* a, long m, long k, long count, double* vecs) {
  hipStream_t s = ctx_.stream();
  auto* Zm = reinterpret_cast<rocblas_double_complex*>(z_.p);
  int* info = reinterpret_cast<int*>(small_.p);
  double* d = small_.p + 8;
  double* e = d + m;
  check(hipMemcpyAsync(z_.p, a, 16 * std::size_t(m) * m, hipMemcpyDeviceToDevice, s), "copy");
  check(rocsolver_zheevd(ctx_.blas(), rocblas_evect_original, rocblas_fill_lower, rocblas_int(m), Zm, rocblas_int(m),
                         d, e, info),
        "rocsolver_zheevd");
  if (k > 0) check(hipMemcpyAsync(vecs, z_.p, 16 * std::size_t(m) * k, hipMemcpyDeviceToDevice, s), "vectors");
  std::vector<double> h(8 + m);
  check(hipMemcpyAsync(h.data(), small_.p, 8 * std::size_t(8 + m), hipMemcpyDeviceToHost, s), "eigenvalues");
  ctx_.sync();
  int status = 0;
  std::memcpy(&status, h.data(), sizeof(int));
  if (status) throw std::runtime_error("HIP: Hermitian eigensolve failed");
  return std::vector<double>(h.begin() + 8, h.begin() + 8 + count);
}

// Tridiagonal eigenpairs by multisection and inverse iteration (the lowest
// `count` values, the lowest k vectors), the vectors back-transformed
// (register-resident kernel, zunmtr above its order) and symmetrically
// orthonormalised:
// vecs = Z (I + E)^(-1/2) to second order, E = Z^H Z - I. Mixing two
// eigenvectors by their overlap c, which inverse iteration bounds by
// O(eps |T| / gap), changes their residuals by c gap = O(eps |T|). Returns false
// (outputs invalid) when an iteration did not stay finite or |E| > kMaxDefect.
bool HermitianEigensolver::inverse_iteration(long m, long k, long count, double* vecs, std::vector<double>& values,
                                             bool* aborted_out) {
  hipStream_t s = ctx_.stream();
  int* status = reinterpret_cast<int*>(small_.p);
  auto* defect_bits = reinterpret_cast<unsigned long long*>(small_.p + 2);
  double* d = small_.p + 8;
  double* e = d + m;
  double* tau = e + m;
  double* w = tau + 2 * m;
  check(launch_tridiag_values(int(m), d, e, int(count), w, status, s), "tridiagonal eigenvalues");
  check(hipMemsetAsync(defect_bits, 0, sizeof(unsigned long long), s), "memset");
  if (k > 0) {
    check(launch_tridiag_vectors(int(m), d, e, w, int(k), z_.p, m, status, s), "tridiagonal eigenvectors");
    if (m <= kBackTransformMaxOrder)
      check(launch_back_transform(int(m), int(k), refl_.p, m, tau, z_.p, m, s), "back-transformation");
    else
      check(rocsolver_zunmtr(ctx_.blas(), rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                             rocblas_int(m), rocblas_int(k), reinterpret_cast<rocblas_double_complex*>(refl_.p),
                             rocblas_int(m), reinterpret_cast<rocblas_double_complex*>(tau),
                             reinterpret_cast<rocblas_double_complex*>(z_.p), rocblas_int(m)),
            "rocsolver_zunmtr");
    const std::size_t kk = std::size_t(2 * k * k);
    orth_.reserve(4 * kk);
    double* g = orth_.p;
    double* ge = g + kk;
    double* ge2 = ge + kk;
    double* gm = ge2 + kk;
    const Precision f64 = Precision::FP64;
    ctx_.zgemm(f64, 1, 0, k, k, m, 1.0, z_.p, m, z_.p, m, g, k, nullptr, 0, 0.0);
    check(launch_identity_defect(g, k, k, ge, k, defect_bits, s), "identity defect");
    ctx_.zgemm(f64, 0, 0, k, k, k, 1.0, ge, k, ge, k, ge2, k, nullptr, 0, 0.0);
    check(launch_series_metric(ge, ge2, k, k, gm, k, s), "series metric");
    ctx_.zgemm(f64, 0, 0, m, k, k, 1.0, z_.p, m, gm, k, vecs, m, nullptr, 0, 0.0);
  }
  values.assign(std::size_t(count), 0.0);
  check(hipMemcpyAsync(values.data(), w, 8 * std::size_t(count), hipMemcpyDeviceToHost, s), "eigenvalues");
  double head[3] = {0, 0, 0};
  check(hipMemcpyAsync(head, small_.p, sizeof(head), hipMemcpyDeviceToHost, s), "status");
  if (barrier_expired(m)) {
    *aborted_out = true;
    return false;
  }
  int st = 0;
  std::memcpy(&st, head, sizeof(int));
  unsigned long long bits = 0;
  std::memcpy(&bits, head + 2, sizeof(bits));
  double defect = 0.0;
  std::memcpy(&defect, &bits, sizeof(defect));
  bool finite = true;
  for (long i = 0; i < count; ++i) finite = finite && std::isfinite(values[std::size_t(i)]);
  if (st || !(defect <= kMaxDefect) || !finite) {
    if (std::getenv("SDFT_HIP_TRIDIAG_VERBOSE"))
      std::fprintf(stderr, "tridiagonal inverse iteration rejected: m=%ld k=%ld status=%d defect=%.3e finite=%d\n", m,
                   k, st, defect, int(finite));
    return false;
  }
  // a degenerate group may come back an ulp out of order
  for (long i = 1; i < count; ++i)
    values[std::size_t(i)] = std::max(values[std::size_t(i)], values[std::size_t(i - 1)]);
  return true;
}

}  // namespace sdft_hip
