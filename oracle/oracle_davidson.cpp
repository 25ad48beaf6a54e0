// CPU oracle: block Davidson in the S metric, restated from
// src/pw_core.cpp:532-1437 with the reference's default controls on the CPU
// backend, and src/strict_residual_polish.hpp:47-197.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <vector>

#include "oracle.hpp"

namespace orc {
namespace {

// Column-major dense complex matrix.
struct Mat {
  std::size_t rows{}, cols{};
  std::vector<C> v;
  Mat() = default;
  Mat(std::size_t r, std::size_t c) : rows(r), cols(c), v(r * c, C(0)) {}
  C* col(std::size_t j) { return v.data() + j * rows; }
  const C* col(std::size_t j) const { return v.data() + j * rows; }
  C& operator()(std::size_t i, std::size_t j) { return v[i + j * rows]; }
  const C& operator()(std::size_t i, std::size_t j) const { return v[i + j * rows]; }
};

// Compensated real inner product (src/pw_core.cpp:579-586 and
// src/strict_residual_polish.hpp:20-30: Neumaier branch form).
double real_dot(const C* a, const C* b, std::size_t n) {
  double sum = 0, corr = 0;
  for (std::size_t i = 0; i < n; ++i)
    for (const double term : {a[i].real() * b[i].real(), a[i].imag() * b[i].imag()}) {
      const double next = sum + term;
      corr += std::abs(sum) >= std::abs(term) ? (sum - next) + term : (term - next) + sum;
      sum = next;
    }
  return sum + corr;
}
// src/strict_residual_polish.hpp:32-43
C complex_dot(const C* a, const C* b, std::size_t n) {
  double re = 0, im = 0, re_e = 0, im_e = 0;
  auto add = [](double t, double& s, double& e) {
    const double next = s + t;
    e += std::abs(s) >= std::abs(t) ? (s - next) + t : (t - next) + s;
    s = next;
  };
  for (std::size_t i = 0; i < n; ++i) {
    add(a[i].real() * b[i].real(), re, re_e);
    add(a[i].imag() * b[i].imag(), re, re_e);
    add(a[i].real() * b[i].imag(), im, im_e);
    add(-a[i].imag() * b[i].real(), im, im_e);
  }
  return {re + re_e, im + im_e};
}
double col_norm(const C* a, std::size_t n) {
  double s = 0;
  for (std::size_t i = 0; i < n; ++i) s += std::norm(a[i]);
  return std::sqrt(s);
}
// C = op(A) * B for whole-matrix views
Mat mul(const Mat& a, const Mat& b, bool adjoint = false) {
  Mat r(adjoint ? a.cols : a.rows, b.cols);
  zgemm(adjoint ? Op::Adjoint : Op::Normal, Op::Normal, r.rows, r.cols, b.rows, a.v.data(), a.rows, b.v.data(),
        b.rows, r.v.data(), r.rows);
  return r;
}
Mat left_cols(const Mat& a, std::size_t c0, std::size_t count) {
  Mat r(a.rows, count);
  std::copy(a.col(c0), a.col(c0) + a.rows * count, r.v.begin());
  return r;
}
void hermitise(Mat& m) {
  for (std::size_t j = 0; j < m.cols; ++j)
    for (std::size_t i = j; i < m.rows; ++i) {
      const C v = .5 * (m(i, j) + std::conj(m(j, i)));
      m(i, j) = v;
      m(j, i) = std::conj(v);
    }
}
struct Ritz { std::vector<double> values; Mat vectors; };
Ritz rayleigh_ritz(const Mat& h, const Mat& s) {
  Ritz r;
  r.values.resize(h.rows);
  r.vectors = Mat(h.rows, h.rows);
  hermitian_generalized_eigen(h.rows, h.v.data(), s.v.data(), r.values.data(), r.vectors.v.data());
  return r;
}
// Householder QR of a (rows x cols, rows >= cols), returns the full Q (rows x rows).
Mat householder_q(const Mat& a) {
  const std::size_t m = a.rows, n = a.cols;
  Mat r = a, q(m, m);
  for (std::size_t i = 0; i < m; ++i) q(i, i) = 1;
  std::vector<std::vector<C>> vs;
  std::vector<double> taus;
  for (std::size_t k = 0; k < std::min(n, m); ++k) {
    const std::size_t len = m - k;
    double xn = 0;
    for (std::size_t i = 0; i < len; ++i) xn += std::norm(r(k + i, k));
    xn = std::sqrt(xn);
    std::vector<C> v(len);
    if (xn == 0) { vs.push_back({}); taus.push_back(0); continue; }
    const double a0 = std::abs(r(k, k));
    const C ph = a0 > 0 ? r(k, k) / a0 : C(1);
    const C alpha = -ph * xn;
    for (std::size_t i = 0; i < len; ++i) v[i] = r(k + i, k);
    v[0] -= alpha;
    double vv = 0;
    for (auto& x : v) vv += std::norm(x);
    const double tau = vv > 0 ? 2. / vv : 0.;
    for (std::size_t j = k; j < n; ++j) {
      C s(0);
      for (std::size_t i = 0; i < len; ++i) s += std::conj(v[i]) * r(k + i, j);
      s *= tau;
      for (std::size_t i = 0; i < len; ++i) r(k + i, j) -= s * v[i];
    }
    vs.push_back(std::move(v));
    taus.push_back(tau);
  }
  for (std::size_t kk = vs.size(); kk-- > 0;) {
    const auto& v = vs[kk];
    if (v.empty()) continue;
    for (std::size_t j = 0; j < m; ++j) {
      C s(0);
      for (std::size_t i = 0; i < v.size(); ++i) s += std::conj(v[i]) * q(kk + i, j);
      s *= taus[kk];
      for (std::size_t i = 0; i < v.size(); ++i) q(kk + i, j) -= s * v[i];
    }
  }
  return q;
}

class Solver {
 public:
  Solver(const Operator& op, std::size_t roots, std::size_t max_it, double tol, std::size_t factor)
      : op_(op), n_(op.n), R_(roots), max_it_(max_it), tol_(tol) {
    cap_ = std::min<std::size_t>(n_, std::max<std::size_t>(factor * roots, 64));
    basis_ = Mat(n_, cap_);
    hb_ = Mat(n_, cap_);
    sb_ = Mat(n_, cap_);
    cached_h_ = Mat(cap_, cap_);
    cached_s_ = Mat(cap_, cap_);
  }
  DavidsonResult run(const std::vector<C>& initial);

 private:
  Mat apply_s(const Mat& x) { Mat y(x.rows, x.cols); apply_block(op_, x.v.data(), x.cols, y.v.data(), true); return y; }
  Mat apply_h(const Mat& x) {
    Mat y(x.rows, x.cols);
    apply_block(op_, x.v.data(), x.cols, y.v.data(), false);
    h_columns_ += x.cols;
    return y;
  }
  void apply_pair_m(const Mat& x, Mat& hx, Mat& sx) {
    hx = Mat(x.rows, x.cols);
    sx = Mat(x.rows, x.cols);
    apply_pair(op_, x.v.data(), x.cols, hx.v.data(), sx.v.data());
    h_columns_ += x.cols;
  }
  Mat basis_view(std::size_t count) const { return left_cols(basis_, 0, count); }
  bool append(std::vector<C> v);
  bool append_block(Mat vectors);
  void update_projected(const Mat& applied, Mat& cached);
  void flush_h();
  void certify();
  std::vector<C> rayleigh_residual(const C* v, const C* hv, const C* sv, double& value) const {
    const double metric = real_dot(v, sv, n_);
    if (!(metric > 0)) throw std::runtime_error("invalid Davidson Rayleigh metric");
    value = real_dot(v, hv, n_) / metric;
    std::vector<C> r(n_);
    for (std::size_t i = 0; i < n_; ++i)
      r[i] = {std::fma(-value, sv[i].real(), hv[i].real()), std::fma(-value, sv[i].imag(), hv[i].imag())};
    return r;
  }
  double max_residual() const { return *std::max_element(res_norms_.begin(), res_norms_.end()); }
  void refresh_norms() {
    converged_ = true;
    for (std::size_t r = 0; r < R_; ++r) {
      res_norms_[r] = col_norm(residual_.col(r), n_);
      converged_ = converged_ && res_norms_[r] <= tol_;
    }
  }
  bool polish_step(const Mat& direction_in);
  bool manifold_step();

  const Operator& op_;
  std::size_t n_, R_, max_it_, cap_;
  double tol_;
  Mat basis_, hb_, sb_, cached_h_, cached_s_;
  std::size_t size_ = 0, computed_h_ = 0, projected_size_ = 0;
  bool incremental_ = true;
  Mat x_, sx_, residual_;
  std::vector<double> values_, res_norms_;
  bool converged_ = false;
  std::size_t h_columns_ = 0;
};

// src/pw_core.cpp:796-853 (serial CGS2 of one direction)
bool Solver::append(std::vector<C> v) {
  if (size_ >= cap_) return false;
  const double en = col_norm(v.data(), n_);
  if (!std::isfinite(en)) throw std::runtime_error("nonfinite Davidson direction");
  if (en == 0) return false;
  for (auto& z : v) z /= en;
  auto fresh_s = [&]() {
    Mat m(n_, 1);
    std::copy(v.begin(), v.end(), m.v.begin());
    return apply_s(m).v;
  };
  std::vector<C> sv = fresh_s();
  for (int pass = 0; pass < 2; ++pass) {
    if (pass) sv = fresh_s();
    if (size_ < 8) {
      for (std::size_t j = 0; j < size_; ++j) {
        C proj(0);
        const C* b = basis_.col(j);
        for (std::size_t i = 0; i < n_; ++i) proj += std::conj(b[i]) * sv[i];
        const C* s = sb_.col(j);
        for (std::size_t i = 0; i < n_; ++i) { v[i] -= proj * b[i]; sv[i] -= proj * s[i]; }
      }
    } else {
      std::vector<C> proj(size_), removal(n_);
      zgemm(Op::Adjoint, Op::Normal, size_, 1, n_, basis_.v.data(), n_, sv.data(), n_, proj.data(), size_);
      zgemm(Op::Normal, Op::Normal, n_, 1, size_, basis_.v.data(), n_, proj.data(), size_, removal.data(), n_);
      for (std::size_t i = 0; i < n_; ++i) v[i] -= removal[i];
    }
  }
  sv = fresh_s();
  C nn(0);
  for (std::size_t i = 0; i < n_; ++i) nn += std::conj(v[i]) * sv[i];
  const double norm2 = nn.real();
  if (!std::isfinite(norm2) || norm2 < -1e-10) throw std::runtime_error("Davidson encountered non-positive overlap direction");
  if (norm2 < 1e-24) return false;
  const double s = std::sqrt(norm2);
  for (std::size_t i = 0; i < n_; ++i) { basis_(i, size_) = v[i] / s; sb_(i, size_) = sv[i] / s; }
  ++size_;
  return true;
}

// src/pw_core.cpp:871-1013 (block CGS2 in the S metric, eigen-normalisation)
bool Solver::append_block(Mat vectors) {
  const std::size_t before = size_;
  if (vectors.cols > cap_ - size_) { vectors.cols = cap_ - size_; vectors.v.resize(n_ * vectors.cols); }
  if (!vectors.cols) return false;
  auto serial = [&]() {
    for (std::size_t j = 0; j < vectors.cols; ++j) append(std::vector<C>(vectors.col(j), vectors.col(j) + n_));
    return size_ > before;
  };
  if (vectors.cols == 1) return serial();
  for (std::size_t j = 0; j < vectors.cols; ++j) {
    const double len = col_norm(vectors.col(j), n_);
    if (!(len > 0) || !std::isfinite(len)) return serial();
    for (std::size_t i = 0; i < n_; ++i) vectors(i, j) /= len;
  }
  const Mat original = vectors;
  Mat metric = apply_s(vectors);
  for (int pass = 0; size_ && pass < 2; ++pass) {
    Mat proj(size_, vectors.cols), removal(n_, vectors.cols);
    zgemm(Op::Adjoint, Op::Normal, size_, vectors.cols, n_, basis_.v.data(), n_, metric.v.data(), n_, proj.v.data(), size_);
    zgemm(Op::Normal, Op::Normal, n_, vectors.cols, size_, basis_.v.data(), n_, proj.v.data(), size_, removal.v.data(), n_);
    for (std::size_t i = 0; i < vectors.v.size(); ++i) vectors.v[i] -= removal.v[i];
    metric = apply_s(vectors);
  }
  Mat gram = mul(vectors, metric, true);
  hermitise(gram);
  std::vector<double> ev(gram.rows);
  Mat evec(gram.rows, gram.rows);
  hermitian_eigen(gram.rows, gram.v.data(), ev.data(), evec.v.data());
  const double largest = *std::max_element(ev.begin(), ev.end()), smallest = *std::min_element(ev.begin(), ev.end());
  if (smallest < -1e-10) throw std::runtime_error("block Davidson encountered non-positive overlap");
  if (!(largest > 0) || smallest < 1e-10 * largest) { vectors = original; return serial(); }
  Mat norm(gram.rows, gram.rows);
  for (std::size_t j = 0; j < gram.rows; ++j)
    for (std::size_t i = 0; i < gram.rows; ++i) norm(i, j) = evec(i, j) * (1. / std::sqrt(ev[j]));
  const Mat nv = mul(vectors, norm), ns = mul(metric, norm);
  std::copy(nv.v.begin(), nv.v.end(), basis_.col(size_));
  std::copy(ns.v.begin(), ns.v.end(), sb_.col(size_));
  size_ += vectors.cols;
  return true;
}

// src/pw_core.cpp:771-785
void Solver::update_projected(const Mat& applied, Mat& cached) {
  const std::size_t first = incremental_ ? projected_size_ : 0;
  if (first == size_) return;
  const std::size_t cols = size_ - first;
  zgemm(Op::Adjoint, Op::Normal, size_, cols, n_, basis_.v.data(), n_, applied.col(first), n_,
        cached.col(first), cap_);
  if (first)
    zgemm(Op::Adjoint, Op::Normal, cols, first, n_, basis_.col(first), n_, applied.v.data(), n_,
          cached.v.data() + first, cap_);
}

void Solver::flush_h() {
  if (computed_h_ >= size_) return;
  const Mat b = left_cols(basis_, computed_h_, size_ - computed_h_);
  const Mat h = apply_h(b);
  std::copy(h.v.begin(), h.v.end(), hb_.col(computed_h_));
  computed_h_ = size_;
}

// src/pw_core.cpp:1125-1149 (certify_once=false)
void Solver::certify() {
  sx_ = apply_s(x_);
  for (std::size_t r = 0; r < R_; ++r) {
    C nn(0);
    for (std::size_t i = 0; i < n_; ++i) nn += std::conj(x_(i, r)) * sx_(i, r);
    const double norm2 = nn.real();
    if (!std::isfinite(norm2) || !(norm2 > 0)) throw std::runtime_error("invalid Davidson Ritz S norm");
    for (std::size_t i = 0; i < n_; ++i) x_(i, r) /= std::sqrt(norm2);
  }
  Mat hx;
  apply_pair_m(x_, hx, sx_);
  residual_ = Mat(n_, R_);
  for (std::size_t r = 0; r < R_; ++r) {
    auto res = rayleigh_residual(x_.col(r), hx.col(r), sx_.col(r), values_[r]);
    std::copy(res.begin(), res.end(), residual_.col(r));
  }
  refresh_norms();
}

// src/strict_residual_polish.hpp:125-197 (no locked bands)
bool Solver::polish_step(const Mat& pre) {
  const double old_max = max_residual();
  if (old_max <= tol_) return false;
  Mat direction = pre;
  for (int pass = 0; pass < 2; ++pass) {
    const Mat c = mul(sx_, direction, true);  // sx^H d
    const Mat rem = mul(x_, c);
    for (std::size_t i = 0; i < direction.v.size(); ++i) direction.v[i] -= rem.v[i];
  }
  Mat dh, ds;
  apply_pair_m(direction, dh, ds);
  Mat trial = x_;
  bool moved = false;
  for (std::size_t r = 0; r < R_; ++r) {
    if (col_norm(residual_.col(r), n_) <= .5 * tol_) continue;
    const double gradient = real_dot(direction.col(r), residual_.col(r), n_);
    const double curvature = real_dot(direction.col(r), dh.col(r), n_) - values_[r] * real_dot(direction.col(r), ds.col(r), n_);
    if (!(curvature > 0.) || !std::isfinite(curvature) || !std::isfinite(gradient)) continue;
    const double alpha = -gradient / curvature;
    if (!(alpha > 0.) || !std::isfinite(alpha) || alpha * col_norm(direction.col(r), n_) > 1e-4) continue;
    for (std::size_t i = 0; i < n_; ++i)
      trial(i, r) = {std::fma(alpha, direction(i, r).real(), x_(i, r).real()),
                     std::fma(alpha, direction(i, r).imag(), x_(i, r).imag())};
    moved = true;
  }
  if (!moved) return false;
  Mat th, ts;
  apply_pair_m(trial, th, ts);
  for (std::size_t r = 0; r < R_; ++r) {
    const double metric = real_dot(trial.col(r), ts.col(r), n_);
    if (!(metric > 0.) || !std::isfinite(metric)) return false;
    for (std::size_t i = 0; i < n_; ++i) trial(i, r) /= std::sqrt(metric);
  }
  apply_pair_m(trial, th, ts);
  std::vector<double> energies(R_);
  Mat errors(n_, R_);
  for (std::size_t r = 0; r < R_; ++r) {
    const double metric = real_dot(trial.col(r), ts.col(r), n_);
    energies[r] = real_dot(trial.col(r), th.col(r), n_) / metric;
    for (std::size_t i = 0; i < n_; ++i)
      errors(i, r) = {std::fma(-energies[r], ts(i, r).real(), th(i, r).real()),
                      std::fma(-energies[r], ts(i, r).imag(), th(i, r).imag())};
  }
  double new_max = 0;
  for (std::size_t r = 0; r < R_; ++r) new_max = std::max(new_max, col_norm(errors.col(r), n_));
  const Mat gram = mul(trial, ts, true);
  double gerr = 0;
  for (std::size_t j = 0; j < R_; ++j)
    for (std::size_t i = 0; i < R_; ++i) gerr = std::max(gerr, std::abs(gram(i, j) - (i == j ? C(1) : C(0))));
  if (!std::isfinite(new_max) || !(gerr <= 1e-12) || !(new_max <= tol_ || new_max < .999 * old_max)) return false;
  x_ = std::move(trial);
  sx_ = std::move(ts);
  values_ = energies;
  residual_ = std::move(errors);
  return true;
}

// src/strict_residual_polish.hpp:68-123 (no locked bands)
bool Solver::manifold_step() {
  const double old_max = max_residual();
  if (old_max <= tol_) return false;
  Mat hx, os;
  apply_pair_m(x_, hx, os);
  Mat h(R_, R_), s(R_, R_);
  for (std::size_t i = 0; i < R_; ++i)
    for (std::size_t j = 0; j < R_; ++j) {
      h(i, j) = complex_dot(x_.col(i), hx.col(j), n_);
      s(i, j) = complex_dot(x_.col(i), os.col(j), n_);
    }
  // (h + h^H)/2 computed from copies as in the reference
  Mat h2 = h, s2 = s;
  for (std::size_t i = 0; i < R_; ++i)
    for (std::size_t j = 0; j < R_; ++j) {
      h2(i, j) = .5 * (h(i, j) + std::conj(h(j, i)));
      s2(i, j) = .5 * (s(i, j) + std::conj(s(j, i)));
    }
  const Ritz small = rayleigh_ritz(h2, s2);
  Mat trial = mul(x_, small.vectors), th, ts;
  apply_pair_m(trial, th, ts);
  for (std::size_t r = 0; r < R_; ++r) {
    const double norm = real_dot(trial.col(r), ts.col(r), n_);
    if (!std::isfinite(norm) || norm <= 0) return false;
    for (std::size_t i = 0; i < n_; ++i) trial(i, r) /= std::sqrt(norm);
  }
  apply_pair_m(trial, th, ts);
  std::vector<double> energies(R_);
  Mat errors(n_, R_);
  for (std::size_t r = 0; r < R_; ++r) {
    energies[r] = real_dot(trial.col(r), th.col(r), n_) / real_dot(trial.col(r), ts.col(r), n_);
    for (std::size_t i = 0; i < n_; ++i)
      errors(i, r) = {std::fma(-energies[r], ts(i, r).real(), th(i, r).real()),
                      std::fma(-energies[r], ts(i, r).imag(), th(i, r).imag())};
  }
  double new_max = 0;
  for (std::size_t r = 0; r < R_; ++r) new_max = std::max(new_max, col_norm(errors.col(r), n_));
  const Mat gram = mul(trial, ts, true);
  double gerr = 0;
  for (std::size_t j = 0; j < R_; ++j)
    for (std::size_t i = 0; i < R_; ++i) gerr = std::max(gerr, std::abs(gram(i, j) - (i == j ? C(1) : C(0))));
  if (!std::isfinite(new_max) || gerr > 1e-12 || !(new_max <= tol_ || new_max < .999 * old_max)) return false;
  x_ = std::move(trial);
  sx_ = std::move(ts);
  values_ = energies;
  residual_ = std::move(errors);
  return true;
}

DavidsonResult Solver::run(const std::vector<C>& initial) {
  const std::size_t n = n_, R = R_;
  if (!n || !R || R > n || !max_it_ || (!initial.empty() && initial.size() != n * R) || !(tol_ > 0))
    throw std::invalid_argument("invalid matrix-free Davidson problem");
  incremental_ = tol_ >= 1e-12;
  bool strict_projection = false;
  unsigned polish_attempts = 0, manifold_attempts = 0;
  double strict_entry = std::numeric_limits<double>::infinity();
  // canonical order by H/S diagonal (src/pw_core.cpp:1016-1021)
  std::vector<std::size_t> order(n);
  for (std::size_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](std::size_t i, std::size_t j) {
    return op_.h_diagonal[i] / op_.s_diagonal[i] < op_.h_diagonal[j] / op_.s_diagonal[j];
  });
  std::uint64_t state = 0x6a09e667f3bcc909ULL;  // src/pw_core.cpp:1027-1034 (splitmix64)
  auto uniform = [&]() {
    std::uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * 0x1.0p-53 - .5;
  };
  auto broad = [&]() {
    std::vector<C> s(n);
    for (std::size_t i = 0; i < n; ++i) {
      const double re = uniform();
      const double im = uniform();
      s[i] = C(re, im);
    }
    const double nn = col_norm(s.data(), n);
    for (auto& z : s) z /= nn;
    return s;
  };
  const double warm_pert = std::clamp(100. * tol_, 1e-10, 1e-6);
  const bool warm = !initial.empty() && 2 * R < n;
  std::vector<std::vector<C>> seeds;
  if (warm)
    for (std::size_t j = 0; j < R; ++j) {
      std::vector<C> s(initial.begin() + j * n, initial.begin() + (j + 1) * n);
      const double nn = col_norm(s.data(), n);
      const auto b = broad();
      for (std::size_t i = 0; i < n; ++i) s[i] += warm_pert * nn * b[i];
      seeds.push_back(std::move(s));
    }
  const std::size_t guards = std::min(n, 2 * R);
  for (std::size_t j = 0; j < guards && size_ < cap_; ++j) {
    auto s = broad();
    for (auto& z : s) z *= .05;
    s[order[j]] += 1.;
    seeds.push_back(std::move(s));
  }
  Mat block(n, seeds.size());
  for (std::size_t j = 0; j < seeds.size(); ++j) std::copy(seeds[j].begin(), seeds[j].end(), block.col(j));
  append_block(std::move(block));

  values_.assign(R, 0.);
  res_norms_.assign(R, 0.);
  bool certified = false;
  std::size_t iterations = 0;
  Ritz rr;
  for (std::size_t it = 0; it < max_it_; ++it) {
    flush_h();
    update_projected(hb_, cached_h_);
    update_projected(sb_, cached_s_);
    projected_size_ = size_;
    Mat ph(size_, size_), ps(size_, size_);
    for (std::size_t j = 0; j < size_; ++j)
      for (std::size_t i = 0; i < size_; ++i) { ph(i, j) = cached_h_(i, j); ps(i, j) = cached_s_(i, j); }
    hermitise(ph);
    hermitise(ps);
    rr = rayleigh_ritz(ph, ps);
    for (std::size_t r = 0; r < R; ++r) values_[r] = rr.values[r];
    const Mat rot = left_cols(rr.vectors, 0, R);
    x_ = mul(basis_view(size_), rot);
    sx_ = mul(left_cols(sb_, 0, size_), rot);
    residual_ = mul(left_cols(hb_, 0, size_), rot);
    for (std::size_t r = 0; r < R; ++r)
      for (std::size_t i = 0; i < n; ++i) residual_(i, r) -= sx_(i, r) * values_[r];
    refresh_norms();
    iterations = it + 1;
    bool fresh = false;
    if (max_residual() <= 8 * tol_) { certify(); fresh = true; }
    if (converged_ && !fresh) { certify(); fresh = true; }
    if (!converged_ && tol_ <= 1e-12 && !strict_projection && max_residual() <= 1e-10) {
      if (!fresh) { certify(); fresh = true; }
      if (!converged_) {  // src/pw_core.cpp:1247-1267 strict projected-pair rebuild
        Mat fh, fs;
        apply_pair_m(basis_view(size_), fh, fs);
        std::copy(fh.v.begin(), fh.v.end(), hb_.v.begin());
        std::copy(fs.v.begin(), fs.v.end(), sb_.v.begin());
        computed_h_ = size_;
        projected_size_ = 0;
        incremental_ = false;
        strict_projection = true;
        continue;
      }
    }
    bool polished = false;
    if (!converged_ && tol_ <= 1e-12 && max_residual() <= 1e-10 && max_residual() < .5 * strict_entry &&
        (polish_attempts < 16 || manifold_attempts < 4)) {
      if (!fresh) certify();
      strict_entry = 0;
      for (std::size_t r = 0; r < R; ++r) strict_entry = std::max(strict_entry, col_norm(residual_.col(r), n));
      if (!converged_ && manifold_attempts < 4 && it + 1 < max_it_) {
        ++manifold_attempts;
        ++it;
        iterations = it + 1;
        polished |= manifold_step();
        refresh_norms();
      }
      for (; !converged_ && polish_attempts < 16 && it + 1 < max_it_; ++polish_attempts) {
        Mat dir = residual_;
        for (std::size_t r = 0; r < R; ++r) {
          double wk = 0.;
          for (std::size_t i = 0; i < n; ++i) wk += std::norm(x_(i, r)) * op_.kinetic[i];
          for (std::size_t i = 0; i < n; ++i) dir(i, r) *= -kinetic_preconditioner(op_.kinetic[i], wk);
        }
        ++it;
        iterations = it + 1;
        const bool improved = polish_step(dir);
        polished |= improved;
        if (!improved) { ++polish_attempts; break; }
        refresh_norms();
      }
    }
    if (converged_) { certified = true; break; }
    if (polished && it + 1 < max_it_) {
      const bool at_cap = size_ + R > cap_;
      const std::size_t keep = at_cap ? std::min(size_, 2 * R) : size_;
      Mat retained(n, keep);
      std::copy(x_.v.begin(), x_.v.end(), retained.v.begin());
      if (at_cap) {
        if (keep > R) {
          const Mat extra = mul(basis_view(size_), left_cols(rr.vectors, R, keep - R));
          std::copy(extra.v.begin(), extra.v.end(), retained.col(R));
        }
      } else if (size_ > R) {  // strict_preserved_search_subspace (src/strict_residual_polish.hpp:47-63)
        const Mat b = basis_view(size_);
        const Mat cross = mul(b, sx_, true);  // (sx^H B)^H = B^H sx
        const Mat q = householder_q(cross);
        const Mat extra = mul(b, left_cols(q, R, size_ - R));
        std::copy(extra.v.begin(), extra.v.end(), retained.col(R));
      }
      size_ = 0;
      computed_h_ = 0;
      projected_size_ = 0;
      append_block(std::move(retained));
    }
    if (size_ + R > cap_) {  // thick restart (src/pw_core.cpp:1357-1383)
      const std::size_t keep = std::min(size_, 2 * R);
      Mat retained = mul(basis_view(size_), left_cols(rr.vectors, 0, keep));
      size_ = 0;
      computed_h_ = 0;
      projected_size_ = 0;
      append_block(std::move(retained));
    }
    const std::size_t previous = size_;
    std::vector<std::vector<C>> dirs;
    for (std::size_t r = 0; r < R && size_ < cap_; ++r) {
      if (res_norms_[r] <= .5 * tol_) continue;
      std::vector<C> d(residual_.col(r), residual_.col(r) + n);
      double wk = 0;
      for (std::size_t i = 0; i < n; ++i) wk += std::norm(x_(i, r)) * op_.kinetic[i];
      for (std::size_t i = 0; i < n; ++i) d[i] *= -kinetic_preconditioner(op_.kinetic[i], wk);
      dirs.push_back(std::move(d));
    }
    Mat corr(n, dirs.size());
    for (std::size_t j = 0; j < dirs.size(); ++j) std::copy(dirs[j].begin(), dirs[j].end(), corr.col(j));
    append_block(std::move(corr));
    if (size_ == previous) {
      bool added = false;
      for (std::size_t j = 0; j < n && !added; ++j) {
        std::vector<C> s(n, C(0));
        s[order[j]] = 1;
        added = append(std::move(s));
      }
      if (!added) break;
    }
  }
  if (!certified) certify();
  DavidsonResult out;
  out.eigenvalues = values_;
  out.residual_norms = res_norms_;
  out.eigenvectors = x_.v;
  out.converged = converged_;
  out.iterations = iterations;
  out.h_columns = h_columns_;
  return out;
}

}  // namespace

DavidsonResult davidson(const Operator& op, std::size_t roots, std::size_t max_iterations, double tolerance,
                        const std::vector<C>& initial, std::size_t subspace_factor) {
  Solver s(op, roots, max_iterations, tolerance, subspace_factor);
  return s.run(initial);
}

}  // namespace orc
