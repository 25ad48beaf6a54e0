// CPU oracle: dense Hermitian eigensolvers used by the Davidson restatement.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Householder reduction to Hermitian tridiagonal form, a diagonal unitary
// making the off-diagonal real, implicit QL with Wilkinson-type shifts on the
// real tridiagonal matrix, and back-transformation. The generalized problem
// A v = lambda B v is reduced with B = L L^H to L^{-1} A L^{-H} (the itype=1
// form that LAPACK zhegv/zhegvd solve for src/pw_core.cpp:56-136).
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "oracle.hpp"

namespace orc {
namespace {
// Real symmetric tridiagonal QL with eigenvector accumulation into z
// (column-major n x n, z(i,j) = z[i + j*n]). d: diagonal, e: e[i] couples
// i and i+1 (e[n-1] unused).
void tridiagonal_ql(std::size_t n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& z) {
  if (n == 0) return;
  e[n - 1] = 0.;
  for (std::size_t l = 0; l < n; ++l) {
    int iter = 0;
    std::size_t m;
    do {
      for (m = l; m + 1 < n; ++m) {
        const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
        if (std::abs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
      }
      if (m != l) {
        if (++iter > 60) throw std::runtime_error("tridiagonal QL did not converge");
        double g = (d[l + 1] - d[l]) / (2. * e[l]);
        double r = std::hypot(g, 1.);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::abs(r) : -std::abs(r)));
        double s = 1., c = 1., p = 0.;
        bool early = false;
        for (std::size_t ii = m; ii-- > l;) {
          const std::size_t i = ii;
          double f = s * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.) {
            d[i + 1] -= p;
            e[m] = 0.;
            early = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2. * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          double* zi = z.data() + i * n;
          double* zj = z.data() + (i + 1) * n;
          for (std::size_t k = 0; k < n; ++k) {
            f = zj[k];
            zj[k] = s * zi[k] + c * f;
            zi[k] = c * zi[k] - s * f;
          }
        }
        if (early) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.;
      }
    } while (m != l);
  }
}
}  // namespace

void hermitian_eigen(std::size_t n, const C* a_in, double* values, C* vectors) {
  if (!n) return;
  {
    std::vector<C> a(a_in, a_in + n * n);
    if (lapack_heevd(n, a.data(), values)) {
      std::copy(a.begin(), a.end(), vectors);
      return;
    }
  }
  std::vector<C> a(a_in, a_in + n * n);  // column-major, lower triangle used after symmetrisation
  for (std::size_t j = 0; j < n; ++j)
    for (std::size_t i = j; i < n; ++i) {
      const C v = .5 * (a[i + j * n] + std::conj(a[j + i * n]));
      a[i + j * n] = v;
      a[j + i * n] = std::conj(v);
    }
  std::vector<std::vector<C>> reflectors;
  std::vector<double> taus;
  std::vector<C> p(n), w(n);
  for (std::size_t k = 0; k + 2 < n; ++k) {
    const std::size_t len = n - k - 1;
    C* x = a.data() + (k + 1) + k * n;
    double xnorm = 0;
    for (std::size_t i = 0; i < len; ++i) xnorm += std::norm(x[i]);
    xnorm = std::sqrt(xnorm);
    double tail = 0;
    for (std::size_t i = 1; i < len; ++i) tail += std::norm(x[i]);
    std::vector<C> v(len);
    if (xnorm == 0. || tail == 0.) {
      reflectors.push_back({});
      taus.push_back(0.);
      continue;
    }
    const double ax0 = std::abs(x[0]);
    const C phase = ax0 > 0 ? x[0] / ax0 : C(1);
    const C alpha = -phase * xnorm;
    for (std::size_t i = 0; i < len; ++i) v[i] = x[i];
    v[0] -= alpha;
    double vv = 0;
    for (std::size_t i = 0; i < len; ++i) vv += std::norm(v[i]);
    const double tau = 2. / vv;
    // trailing update A22 <- H A22 H, H = I - tau v v^H
    const std::size_t o = k + 1;
    for (std::size_t i = 0; i < len; ++i) {
      C s(0);
      for (std::size_t j = 0; j < len; ++j) s += a[(o + i) + (o + j) * n] * v[j];
      p[i] = tau * s;
    }
    C kk(0);
    for (std::size_t i = 0; i < len; ++i) kk += std::conj(v[i]) * p[i];
    kk *= .5 * tau;
    for (std::size_t i = 0; i < len; ++i) w[i] = p[i] - kk * v[i];
    for (std::size_t j = 0; j < len; ++j)
      for (std::size_t i = 0; i < len; ++i)
        a[(o + i) + (o + j) * n] -= v[i] * std::conj(w[j]) + w[i] * std::conj(v[j]);
    x[0] = alpha;
    for (std::size_t i = 1; i < len; ++i) x[i] = 0;
    for (std::size_t i = 0; i < len; ++i) a[k + (o + i) * n] = std::conj(x[i]);
    reflectors.push_back(std::move(v));
    taus.push_back(tau);
  }
  std::vector<double> d(n), e(n, 0.);
  std::vector<C> phase(n, C(1));
  for (std::size_t i = 0; i < n; ++i) d[i] = a[i + i * n].real();
  for (std::size_t i = 0; i + 1 < n; ++i) {
    const C sub = a[(i + 1) + i * n];
    const double m = std::abs(sub);
    e[i] = m;
    // T(i+1,i) = sub; with D = diag(phase), D^H T D has real off-diagonal m
    phase[i + 1] = m > 0 ? phase[i] * (sub / m) : phase[i];
  }
  std::vector<double> z(n * n, 0.);
  for (std::size_t i = 0; i < n; ++i) z[i + i * n] = 1.;
  tridiagonal_ql(n, d, e, z);
  std::vector<std::size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](std::size_t i, std::size_t j) { return d[i] < d[j]; });
  // vectors = Q D Z with Q = H_0 H_1 ... applied from the last reflector
  for (std::size_t c = 0; c < n; ++c) {
    const std::size_t src = order[c];
    values[c] = d[src];
    C* col = vectors + c * n;
    for (std::size_t i = 0; i < n; ++i) col[i] = phase[i] * z[i + src * n];
    for (std::size_t kk = reflectors.size(); kk-- > 0;) {
      const auto& v = reflectors[kk];
      if (v.empty()) continue;
      const std::size_t o = kk + 1;
      C s(0);
      for (std::size_t i = 0; i < v.size(); ++i) s += std::conj(v[i]) * col[o + i];
      s *= taus[kk];
      for (std::size_t i = 0; i < v.size(); ++i) col[o + i] -= s * v[i];
    }
  }
}

void hermitian_generalized_eigen(std::size_t n, const C* a, const C* b, double* values, C* vectors) {
  if (!n) return;
  {
    std::vector<C> aa(a, a + n * n), bb(b, b + n * n);
    if (lapack_hegvd(n, aa.data(), bb.data(), values)) {
      std::copy(aa.begin(), aa.end(), vectors);
      return;
    }
  }
  // Cholesky B = L L^H (lower, column-major)
  std::vector<C> l(n * n, C(0));
  for (std::size_t j = 0; j < n; ++j) {
    double diag = b[j + j * n].real();
    for (std::size_t k = 0; k < j; ++k) diag -= std::norm(l[j + k * n]);
    if (!(diag > 0)) throw std::runtime_error("generalized eigenproblem: metric is not positive definite");
    const double ljj = std::sqrt(diag);
    l[j + j * n] = ljj;
    for (std::size_t i = j + 1; i < n; ++i) {
      C s = .5 * (b[i + j * n] + std::conj(b[j + i * n]));
      for (std::size_t k = 0; k < j; ++k) s -= l[i + k * n] * std::conj(l[j + k * n]);
      l[i + j * n] = s / ljj;
    }
  }
  // W = L^{-1} A (forward substitution on columns), then Cm = L^{-1} W^H ... = L^{-1} A L^{-H}
  std::vector<C> w(a, a + n * n);
  auto lower_solve = [&](C* col) {
    for (std::size_t i = 0; i < n; ++i) {
      C s = col[i];
      for (std::size_t k = 0; k < i; ++k) s -= l[i + k * n] * col[k];
      col[i] = s / l[i + i * n];
    }
  };
  for (std::size_t j = 0; j < n; ++j) lower_solve(w.data() + j * n);
  std::vector<C> cm(n * n);
  for (std::size_t j = 0; j < n; ++j)
    for (std::size_t i = 0; i < n; ++i) cm[i + j * n] = std::conj(w[j + i * n]);  // W^H = A^H L^{-H} = A L^{-H}
  for (std::size_t j = 0; j < n; ++j) lower_solve(cm.data() + j * n);
  hermitian_eigen(n, cm.data(), values, vectors);
  // vectors <- L^{-H} Y (back substitution with L^H)
  for (std::size_t c = 0; c < n; ++c) {
    C* col = vectors + c * n;
    for (std::size_t ii = n; ii-- > 0;) {
      C s = col[ii];
      for (std::size_t k = ii + 1; k < n; ++k) s -= std::conj(l[k + ii * n]) * col[k];
      col[ii] = s / l[ii + ii * n];
    }
  }
}

}  // namespace orc
