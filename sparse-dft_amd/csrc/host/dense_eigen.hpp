// Dense Hermitian eigensolver of the Rayleigh-Ritz step: one-launch Householder
// tridiagonalisation (tridiag.hip), divide and conquer on the tridiagonal
// (rocSOLVER zstedc) and the back-transformation of only the wanted columns
// (rocSOLVER zunmtr).
#pragma once
#include <vector>

#include "device.hpp"

namespace sdft_hip {

class HermitianEigensolver {
 public:
  HermitianEigensolver(Context& ctx, long max_order);
  // a: m x m device matrix (ld m), exactly Hermitian, left unchanged. Returns
  // all m eigenvalues ascending; the eigenvectors of the lowest k go to vecs
  // (device, ld m).
  std::vector<double> solve(const double* a, long m, long k, double* vecs);
  long launches_fused() const { return fused_; }

 private:
  Context& ctx_;
  long max_order_;
  DevBuf<double> refl_, z_, small_;
  DevBuf<char> work_;
  long fused_ = 0;
};

}  // namespace sdft_hip
