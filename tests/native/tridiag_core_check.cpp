// Host build of the tridiagonal eigensolver's arithmetic (tridiag_core.hpp,
// shared with the device kernels of tridiag_eig.hip), driven the way the
// kernels drive it: multisection over 64 points per round for every
// eigenvalue, clusters at kTridiagClusterGap |T|, inverse iteration with three
// solves and Gram-Schmidt inside a cluster. Test helper only: loaded by
// tests/test_tridiag_core.py and checked against LAPACK (scipy).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "../../sparse-dft_amd/csrc/hip/tridiag_core.hpp"

namespace {

struct Bounds {
  double gl, gu, tnorm, pivmin;
};

Bounds gershgorin(const double* d, const double* e, int n) {
  double lo = DBL_MAX, hi = -DBL_MAX, emax = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::fmin(lo, d[i] - r);
    hi = std::fmax(hi, d[i] + r);
    if (i + 1 < n) emax = std::fmax(emax, e[i] * e[i]);
  }
  Bounds b;
  b.tnorm = std::fmax(std::fabs(lo), std::fabs(hi));
  b.pivmin = DBL_MIN * std::fmax(1.0, emax);
  const double widen = 2.1 * (DBL_EPSILON * b.tnorm * n + 2.0 * b.pivmin);
  b.gl = lo - widen;
  b.gu = hi + widen;
  return b;
}

}  // namespace

extern "C" {

// lowest `count` eigenvalues of T (d: n, e: n - 1) into w; returns the largest
// number of multisection rounds used
int tridiag_core_values(int n, const double* d, const double* e, int count, double* w) {
  std::vector<double> se(n), e2(n);
  for (int i = 0; i < n; ++i) {
    se[i] = i + 1 < n ? e[i] : 0.0;
    e2[i] = se[i] * se[i];
  }
  const Bounds b = gershgorin(d, se.data(), n);
  const double atol = DBL_EPSILON * b.tnorm;
  int most = 0;
  for (int index = 0; index < count; ++index) {
    double lo = b.gl, hi = b.gu;
    int round = 0;
    for (; round < 24; ++round) {
      const double width = hi - lo;
      if (!(width > std::fmax(2.0 * DBL_EPSILON * std::fmax(std::fabs(lo), std::fabs(hi)), std::fmax(atol, b.pivmin))))
        break;
      double x[64];
      int first = 64;
      for (int l = 0; l < 64; ++l) x[l] = lo + width * (l + 1) * (1.0 / 65.0);
      for (int l = 0; l < 64 && first == 64; ++l)
        if (sdft_hip::sturm_count(d, e2.data(), n, x[l], b.pivmin) >= index + 1) first = l;
      if (first < 64) {
        hi = x[first];
        if (first > 0) lo = x[first - 1];
      } else {
        lo = x[63];
      }
    }
    most = std::max(most, round);
    w[index] = 0.5 * (lo + hi);
  }
  return most;
}

// eigenvectors of the lowest k eigenvalues w into z (n x k, column-major);
// returns the largest cluster size, or -1 when an iterate left finite range
int tridiag_core_vectors(int n, const double* d, const double* e, const double* w, int k, double gap_fraction,
                         double* z) {
  std::vector<double> se(n), inv0(n), u1(n), u2(n), l(n), y(n);
  std::vector<unsigned char> piv(n);
  for (int i = 0; i < n; ++i) se[i] = i + 1 < n ? e[i] : 0.0;
  const Bounds b = gershgorin(d, se.data(), n);
  const double gap = gap_fraction * b.tnorm;
  const double pert = std::fmax(DBL_EPSILON * b.tnorm, DBL_MIN / DBL_EPSILON);
  int largest = 0;
  for (int start = 0; start < k; ++start) {
    if (start > 0 && w[start] - w[start - 1] <= gap) continue;
    int end = start;
    while (end + 1 < k && w[end + 1] - w[end] <= gap) ++end;
    largest = std::max(largest, end - start + 1);
    double xprev = 0.0;
    for (int j = start; j <= end; ++j) {
      double x = w[j];
      if (j > start) {
        const double sep = 10.0 * DBL_EPSILON * std::fabs(x);
        if (x - xprev < sep) x = xprev + sep;
      }
      xprev = x;
      sdft_hip::shifted_lu(d, se.data(), n, x, pert, inv0.data(), u1.data(), u2.data(), l.data(), piv.data());
      for (int i = 0; i < n; ++i) y[i] = sdft_hip::start_entry(j, i);
      for (int it = 0; it < 3; ++it) {
        sdft_hip::shifted_solve(inv0.data(), u1.data(), u2.data(), l.data(), piv.data(), n, y.data());
        double big = 0.0;
        for (int i = 0; i < n; ++i) big = std::fmax(big, std::fabs(y[i]));
        if (!(big > 0.0) || !(big < DBL_MAX)) return -1;
        for (int i = 0; i < n; ++i) y[i] *= 1.0 / big;
        for (int p = start; p < j; ++p) {
          const double* zp = z + static_cast<long>(p) * n;
          double dot = 0.0;
          for (int i = 0; i < n; ++i) dot += y[i] * zp[i];
          for (int i = 0; i < n; ++i) y[i] -= dot * zp[i];
        }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) nrm += y[i] * y[i];
        if (!(nrm > 0.0)) return -1;
        for (int i = 0; i < n; ++i) y[i] *= 1.0 / std::sqrt(nrm);
      }
      for (int i = 0; i < n; ++i) z[static_cast<long>(j) * n + i] = y[i];
    }
  }
  return largest;
}

}  // extern "C"
