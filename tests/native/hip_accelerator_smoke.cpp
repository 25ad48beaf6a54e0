// Smoke test of the `hip` Accelerator through the reference's interface, in
// the style of tests/native/compute_backend_smoke.cpp and
// tests/native/fft_backend_smoke.cpp of the reference: independent naive
// references, padded strides, beta = 0 / 1, identity-token reuse, and the
// closed-form local-operator stencil. Distinct non-zero exit code per gate.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <limits>
#include <vector>

#include "accelerator_iface.hpp"

using namespace sdft;
using C = Complex64;

static double gemm_check(Accelerator& acc, std::size_t m, std::size_t n, std::size_t k) {
  double worst = 0;
  for (auto ao : {MatrixOp::Normal, MatrixOp::Adjoint})
    for (auto bo : {MatrixOp::Normal, MatrixOp::Adjoint}) {
      const std::size_t ar = ao == MatrixOp::Normal ? m : k, ac = ao == MatrixOp::Normal ? k : m;
      const std::size_t br = bo == MatrixOp::Normal ? k : n, bc = bo == MatrixOp::Normal ? n : k;
      const std::size_t lda = ar + 3, ldb = br + 5, ldc = m + 7;
      const C sentinel(9876., -5432.);
      std::vector<C> a(lda * ac, sentinel), b(ldb * bc, sentinel), ref(m * n), c(ldc * n, sentinel);
      for (std::size_t j = 0; j < ac; ++j)
        for (std::size_t i = 0; i < ar; ++i) a[i + j * lda] = {std::sin(.13 * (i + 2) * (j + 1)), std::cos(.17 * (i + 4) * (j + 2))};
      for (std::size_t j = 0; j < bc; ++j)
        for (std::size_t i = 0; i < br; ++i) b[i + j * ldb] = {std::cos(.21 * (i + 1) * (j + 3)), std::sin(.31 * (i + 3) * (j + 2))};
      double scale = 1;
      for (std::size_t j = 0; j < n; ++j)
        for (std::size_t i = 0; i < m; ++i) {
          C s{};
          for (std::size_t p = 0; p < k; ++p)
            s += (ao == MatrixOp::Normal ? a[i + p * lda] : std::conj(a[p + i * lda])) *
                 (bo == MatrixOp::Normal ? b[p + j * ldb] : std::conj(b[j + p * ldb]));
          ref[i + j * m] = s;
          scale = std::max(scale, std::abs(s));
        }
      for (std::size_t j = 0; j < n; ++j)
        for (std::size_t i = 0; i < m; ++i) c[i + j * ldc] = {std::numeric_limits<double>::quiet_NaN(), 0.};
      acc.gemm(ao, bo, m, n, k, a.data(), lda, b.data(), ldb, c.data(), ldc, false);
      for (std::size_t j = 0; j < n; ++j) {
        for (std::size_t i = 0; i < m; ++i) worst = std::max(worst, std::abs(c[i + j * ldc] - ref[i + j * m]) / scale);
        for (std::size_t i = m; i < ldc; ++i)
          if (c[i + j * ldc] != sentinel) return 1e300;
      }
      acc.gemm(ao, bo, m, n, k, a.data(), lda, b.data(), ldb, c.data(), ldc, true);
      for (std::size_t j = 0; j < n; ++j)
        for (std::size_t i = 0; i < m; ++i) worst = std::max(worst, std::abs(c[i + j * ldc] - 2. * ref[i + j * m]) / scale);
    }
  return worst;
}

static double stencil_check(Accelerator& acc, std::array<int, 3> shape) {
  const std::size_t count = std::size_t(shape[0]) * shape[1] * shape[2], npw = std::min<std::size_t>(73, count);
  const std::size_t columns = 5;
  auto index = [&](int x, int y, int z) {
    auto wrap = [](int a, int b) { return (a % b + b) % b; };
    return (std::size_t(wrap(x, shape[0])) * shape[1] + wrap(y, shape[1])) * shape[2] + wrap(z, shape[2]);
  };
  std::vector<std::size_t> idx(npw);
  std::vector<int> lookup(count, -1);
  std::vector<double> pot(count), kin(npw);
  for (std::size_t g = 0; g < npw; ++g) { idx[g] = g; lookup[g] = int(g); kin[g] = .013 * (1 + g % 19); }
  for (int x = 0; x < shape[0]; ++x)
    for (int y = 0; y < shape[1]; ++y)
      for (int z = 0; z < shape[2]; ++z) {
        const double phase = 6.2831853071795864769 * (double(x) / shape[0] + 2. * y / shape[1]);
        pot[index(x, y, z)] = .4 + .13 * std::cos(phase) + .07 * std::sin(12.5663706143591729538 * z / shape[2]);
      }
  std::vector<C> in(npw * columns), out(npw * columns), expected(npw * columns);
  for (std::size_t b = 0; b < columns; ++b)
    for (std::size_t g = 0; g < npw; ++g) in[b * npw + g] = {std::sin(.17 * (g + 1) * (b + 2)), std::cos(.13 * (g + 2) * (b + 1))};
  double scale = 1;
  for (std::size_t b = 0; b < columns; ++b)
    for (std::size_t g = 0; g < npw; ++g) {
      const int z = int(idx[g] % shape[2]), y = int((idx[g] / shape[2]) % shape[1]), x = int(idx[g] / shape[2] / shape[1]);
      auto at = [&](int a, int c, int d) { const int j = lookup[index(a, c, d)]; return j < 0 ? C{} : in[b * npw + j]; };
      expected[b * npw + g] = (kin[g] + .4) * in[b * npw + g] + .065 * (at(x - 1, y - 2, z) + at(x + 1, y + 2, z)) +
                              C(0, -.035) * at(x, y, z - 2) + C(0, .035) * at(x, y, z + 2);
      scale = std::max(scale, std::abs(expected[b * npw + g]));
    }
  acc.local_hamiltonian(shape, idx, kin, pot, in.data(), columns, out.data(), 0);
  double worst = 0;
  for (std::size_t i = 0; i < in.size(); ++i) worst = std::max(worst, std::abs(out[i] - expected[i]) / scale);
  return worst;
}

static double paw_check(Accelerator& acc) {
  const std::size_t npw = 400, np = 12, m = 6;
  std::vector<C> b(np * npw), d(np * np), q(np * np), x(npw * m), h(npw * m), s(npw * m);
  for (std::size_t g = 0; g < npw; ++g)
    for (std::size_t p = 0; p < np; ++p) b[p + g * np] = {.04 * std::sin(.3 * (p + 1) * (g + 2)), .03 * std::cos(.7 * (p + 3) * (g + 1))};
  for (std::size_t i = 0; i < np; ++i)
    for (std::size_t j = 0; j <= i; ++j) {
      const C v(std::sin(i + 2. * j), i == j ? 0 : std::cos(i - 3. * j));
      d[i + j * np] = v; d[j + i * np] = std::conj(v);
      const C w(.1 * std::cos(i * j + 1.), 0);
      q[i + j * np] = w; q[j + i * np] = w;
    }
  for (std::size_t c = 0; c < m; ++c)
    for (std::size_t g = 0; g < npw; ++g) x[g + c * npw] = {std::cos(.11 * (g + 1) * (c + 1)), std::sin(.07 * (g + 3) * (c + 2))};
  if (!acc.paw_pair(7, npw, np, m, b.data(), d.data(), q.data(), x.data(), h.data(), s.data(), 8)) return 1e300;
  double worst = 0, scale = 1;
  for (std::size_t c = 0; c < m; ++c) {
    std::vector<C> p(np, C{}), dp(np, C{}), qp(np, C{});
    for (std::size_t i = 0; i < np; ++i)
      for (std::size_t g = 0; g < npw; ++g) p[i] += b[i + g * np] * x[g + c * npw];
    for (std::size_t i = 0; i < np; ++i)
      for (std::size_t j = 0; j < np; ++j) { dp[i] += d[i + j * np] * p[j]; qp[i] += q[i + j * np] * p[j]; }
    for (std::size_t g = 0; g < npw; ++g) {
      C eh{}, es{};
      for (std::size_t i = 0; i < np; ++i) { eh += std::conj(b[i + g * np]) * dp[i]; es += std::conj(b[i + g * np]) * qp[i]; }
      scale = std::max({scale, std::abs(eh), std::abs(es)});
      worst = std::max({worst, std::abs(h[g + c * npw] - eh), std::abs(s[g + c * npw] - es)});
    }
  }
  return worst / scale;
}

int main() {
  try {
    auto acc = make_hip_accelerator();
    const double g1 = std::max({gemm_check(*acc, 23, 19, 37), gemm_check(*acc, 127, 89, 367), gemm_check(*acc, 4097, 1, 1031)});
    if (!(g1 < 2e-11)) { std::printf("{\"gate\":\"gemm\",\"error\":%g}\n", g1); return 3; }
    double f = 0;
    for (auto shape : {std::array<int, 3>{8, 15, 14}, std::array<int, 3>{27, 25, 16}, std::array<int, 3>{4, 6, 7}})
      f = std::max(f, stencil_check(*acc, shape));
    if (!(f < 2e-11)) { std::printf("{\"gate\":\"fft\",\"error\":%g}\n", f); return 4; }
    const double p = paw_check(*acc);
    if (!(p < 2e-11)) { std::printf("{\"gate\":\"paw_pair\",\"error\":%g}\n", p); return 5; }
    bool rejected = false;
    try {
      std::vector<std::size_t> idx{0, 1};
      std::vector<double> kin{1, 1}, pot(10, 1.0);
      std::vector<C> in(2), out(2);
      acc->local_hamiltonian({2, 2, 2}, idx, kin, pot, in.data(), 1, out.data(), 0);
    } catch (const std::invalid_argument&) {
      rejected = true;
    }
    if (!rejected) return 6;
    std::printf("{\"device\":\"%s\",\"gemm_scaled_error\":%.3e,\"fft_scaled_error\":%.3e,\"paw_scaled_error\":%.3e,\"passed\":true}\n",
                acc->device().c_str(), g1, f, p);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 2;
  }
}
