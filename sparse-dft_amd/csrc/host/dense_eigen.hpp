// Dense Hermitian eigensolver of the Rayleigh-Ritz step: one-launch Householder
// tridiagonalisation (tridiag.hip), the tridiagonal eigenvalues by multisection
// and the wanted eigenvectors by inverse iteration (tridiag_eig.hip; rocSOLVER
// zstedc divide and conquer when that fails its checks or is selected), and the
// back-transformation of only the wanted columns (rocSOLVER zunmtr).
#pragma once
#include <vector>

#include "device.hpp"

namespace sdft_hip {

class HermitianEigensolver {
 public:
  HermitianEigensolver(Context& ctx, long max_order);
  // a: m x m device matrix (ld m), exactly Hermitian, left unchanged. Returns
  // the lowest max(k, values) eigenvalues ascending (all m when values < 0);
  // the eigenvectors of the lowest k go to vecs (device, ld m).
  std::vector<double> solve(const double* a, long m, long k, double* vecs, long values = -1);
  long launches_fused() const { return fused_; }
  // inverse-iteration solves that failed their checks and were redone by
  // divide and conquer; tridiagonal solves by divide and conquer in total
  long fallbacks() const { return fallbacks_; }
  long stedc_solves() const { return stedc_; }
  // solves whose one-launch reduction found its bounded grid barrier expired
  // (CUs held by other work past the wait bound) and that were redone by
  // rocSOLVER zheevd on the unchanged input
  long barrier_fallbacks() const { return barrier_fallbacks_; }
  std::size_t device_bytes() const { return refl_.bytes() + z_.bytes() + small_.bytes() + orth_.bytes() + work_.bytes(); }

 private:
  Context& ctx_;
  long max_order_;
  bool inverse_iteration(long m, long k, long count, double* vecs, std::vector<double>& values, bool* aborted);
  std::vector<double> solve_zheevd(const double* a, long m, long k, long count, double* vecs);
  bool barrier_expired(long m);
  DevBuf<double> refl_, z_, small_, orth_;
  DevBuf<char> work_;
  long fused_ = 0, fallbacks_ = 0, stedc_ = 0, barrier_fallbacks_ = 0;
};

}  // namespace sdft_hip
