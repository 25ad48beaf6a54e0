// CPU oracle: G-sphere, FFT, operator construction and application.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp). Built with -ffp-contract=off.
#include "oracle.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {
namespace {
constexpr double kPi = 3.1415926535897932384626433832795;  // src/scf_core.cpp:62
constexpr double kTwoPi = 2.0 * kPi;
Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 scale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
}  // namespace

double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double determinant(const std::array<Vec3, 3>& l) { return dot(l[0], cross(l[1], l[2])); }

std::array<Vec3, 3> reciprocal(const std::array<Vec3, 3>& a) {
  const double volume = determinant(a);
  if (!std::isfinite(volume) || std::abs(volume) < 1e-12) throw std::invalid_argument("singular lattice");
  return {scale(cross(a[1], a[2]), kTwoPi / volume), scale(cross(a[2], a[0]), kTwoPi / volume),
          scale(cross(a[0], a[1]), kTwoPi / volume)};
}

// Restates src/pw_core.cpp:140-171 expression by expression so that the FP64
// sort key rounds identically (this decides ties inside degenerate shells).
std::vector<GVector> enumerate_g_vectors(const std::array<Vec3, 3>& r, double cutoff_ev, Vec3 kf,
                                         double hartree_to_ev) {
  if (cutoff_ev <= 0.0 || !std::isfinite(cutoff_ev)) throw std::invalid_argument("cutoff");
  const double gmax = std::sqrt(2.0 * cutoff_ev / hartree_to_ev);
  const double det = determinant(r);
  if (std::abs(det) < 1e-12 || !std::isfinite(det)) throw std::invalid_argument("singular reciprocal");
  const Vec3 d0 = scale(cross(r[1], r[2]), 1.0 / det);
  const Vec3 d1 = scale(cross(r[2], r[0]), 1.0 / det);
  const Vec3 d2 = scale(cross(r[0], r[1]), 1.0 / det);
  const int hmax = static_cast<int>(std::ceil(gmax * std::sqrt(dot(d0, d0)) + std::abs(kf.x))) + 1;
  const int kmax = static_cast<int>(std::ceil(gmax * std::sqrt(dot(d1, d1)) + std::abs(kf.y))) + 1;
  const int lmax = static_cast<int>(std::ceil(gmax * std::sqrt(dot(d2, d2)) + std::abs(kf.z))) + 1;
  const Vec3 shift = add(add(scale(r[0], kf.x), scale(r[1], kf.y)), scale(r[2], kf.z));
  std::vector<GVector> out;
  for (int h = -hmax; h <= hmax; ++h)
    for (int k = -kmax; k <= kmax; ++k)
      for (int l = -lmax; l <= lmax; ++l) {
        const Vec3 g = add(add(scale(r[0], h), scale(r[1], k)), scale(r[2], l));
        const Vec3 s = add(g, shift);
        const double g2 = dot(s, s);
        if (0.5 * g2 <= cutoff_ev / hartree_to_ev + 1e-12) out.push_back({h, k, l, g, 0.5 * g2 * hartree_to_ev});
      }
  std::sort(out.begin(), out.end(), [](const GVector& a, const GVector& b) {
    if (a.kinetic_ev != b.kinetic_ev) return a.kinetic_ev < b.kinetic_ev;
    if (a.h != b.h) return a.h < b.h;
    if (a.k != b.k) return a.k < b.k;
    return a.l < b.l;
  });
  return out;
}

int fast_fft_length(int minimum) {
  if (minimum < 1) throw std::invalid_argument("invalid FFT length");
  for (int n = minimum;; ++n) {
    int rem = n;
    for (int p : {2, 3, 5, 7})
      while (rem % p == 0) rem /= p;
    if (rem == 1) return n;
  }
}

// ---------------------------------------------------------------- FFT
Fft1d::Fft1d(int n) : n_(n) {
  if (n < 1) throw std::invalid_argument("FFT length");
  int rem = n;
  while (rem % 4 == 0) { factors_.push_back(4); rem /= 4; }
  for (int p : {2, 3, 5, 7})
    while (rem % p == 0) { factors_.push_back(p); rem /= p; }
  for (int p = 11; rem > 1; p += 2)
    while (rem % p == 0) { factors_.push_back(p); rem /= p; }
  tw_fwd_.resize(n);
  tw_bwd_.resize(n);
  for (int j = 0; j < n; ++j) {
    const double a = kTwoPi * double(j) / double(n);
    tw_fwd_[j] = C(std::cos(a), -std::sin(a));
    tw_bwd_[j] = C(std::cos(a), std::sin(a));
  }
}

// Recursive decimation in time: X[k + m s] = sum_q W_p^{qs} (W_n^{qk} Y_q[k]).
void Fft1d::rec(const C* in, C* out, int n, int stride, int fi, int sign, int tw_stride) const {
  if (n == 1) { out[0] = in[0]; return; }
  const int p = factors_[fi], m = n / p;
  for (int q = 0; q < p; ++q) rec(in + q * stride, out + q * m, m, stride * p, fi + 1, sign, tw_stride * p);
  const std::vector<C>& tw = sign < 0 ? tw_fwd_ : tw_bwd_;
  C t[16], r[16];
  std::vector<C> big;
  C* tp = t;
  C* rp = r;
  if (p > 16) { big.resize(2 * p); tp = big.data(); rp = big.data() + p; }
  const int wp = n_ / p;  // W_p^{j} = tw[j * wp]
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < p; ++q) tp[q] = q == 0 ? out[k] : tw[(std::size_t(q) * k * tw_stride) % n_] * out[q * m + k];
    if (p == 2) {
      rp[0] = tp[0] + tp[1];
      rp[1] = tp[0] - tp[1];
    } else if (p == 4) {
      const C a = tp[0] + tp[2], b = tp[0] - tp[2], c = tp[1] + tp[3], d = tp[1] - tp[3];
      const C di = sign < 0 ? C(d.imag(), -d.real()) : C(-d.imag(), d.real());  // d * (sign i)
      rp[0] = a + c;
      rp[1] = b + di;
      rp[2] = a - c;
      rp[3] = b - di;
    } else {
      for (int s = 0; s < p; ++s) {
        C acc = tp[0];
        for (int q = 1; q < p; ++q) acc += tw[((q * s) % p) * wp] * tp[q];
        rp[s] = acc;
      }
    }
    for (int s = 0; s < p; ++s) out[k + m * s] = rp[s];
  }
}

void Fft1d::exec(C* data, C* work, int sign) const {
  if (n_ == 1) return;
  std::copy(data, data + n_, work);
  rec(work, data, n_, 1, 0, sign, 1);
}

Fft3d::Fft3d(std::array<int, 3> s)
    : shape_(s), count_(std::size_t(s[0]) * s[1] * s[2]), f0_(s[0]), f1_(s[1]), f2_(s[2]) {}

void Fft3d::exec(C* g, int sign) const {
  const int n0 = shape_[0], n1 = shape_[1], n2 = shape_[2];
  const int lmax = std::max({n0, n1, n2});
  std::vector<C> line(lmax), work(lmax);
  for (std::size_t xy = 0; xy < std::size_t(n0) * n1; ++xy) f2_.exec(g + xy * n2, work.data(), sign);
  for (int x = 0; x < n0; ++x)
    for (int z = 0; z < n2; ++z) {
      C* base = g + std::size_t(x) * n1 * n2 + z;
      for (int y = 0; y < n1; ++y) line[y] = base[std::size_t(y) * n2];
      f1_.exec(line.data(), work.data(), sign);
      for (int y = 0; y < n1; ++y) base[std::size_t(y) * n2] = line[y];
    }
  const std::size_t plane = std::size_t(n1) * n2;
  for (std::size_t yz = 0; yz < plane; ++yz) {
    for (int x = 0; x < n0; ++x) line[x] = g[x * plane + yz];
    f0_.exec(line.data(), work.data(), sign);
    for (int x = 0; x < n0; ++x) g[x * plane + yz] = line[x];
  }
}

// ---------------------------------------------------------------- BLAS
// Optional CBLAS/LAPACKE (ILP64, as bundled with numpy) bound at run time. The
// reference's CPU path calls cblas_zgemm and zhegvd when they are available
// (src/compute_backend.cpp:182-184, src/pw_core.cpp:104-122); the timed CPU
// baseline binds them, the parity tests keep the plain loops below.
namespace {
using zgemm64_t = void (*)(int, int, int, std::int64_t, std::int64_t, std::int64_t, const void*, const void*,
                           std::int64_t, const void*, std::int64_t, const void*, void*, std::int64_t);
using zhegvd64_t = std::int64_t (*)(int, std::int64_t, char, char, std::int64_t, void*, std::int64_t, void*,
                                    std::int64_t, double*);
using zheevd64_t = std::int64_t (*)(int, char, char, std::int64_t, void*, std::int64_t, double*);
zgemm64_t g_zgemm = nullptr;
zhegvd64_t g_zhegvd = nullptr;
zheevd64_t g_zheevd = nullptr;
}  // namespace

namespace {
using set_threads_t = void (*)(int);
set_threads_t g_blas_threads = nullptr;
}  // namespace

void set_threads(int n) {
  if (n < 1) throw std::invalid_argument("thread count must be positive");
  omp_set_num_threads(n);
  if (g_blas_threads) g_blas_threads(n);
}

bool bind_blas(const char* path) {
  if (!path || !*path) {
    g_blas_threads = nullptr;
    g_zgemm = nullptr;
    g_zhegvd = nullptr;
    g_zheevd = nullptr;
    return true;
  }
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return false;
  g_zgemm = reinterpret_cast<zgemm64_t>(dlsym(h, "scipy_cblas_zgemm64_"));
  g_zhegvd = reinterpret_cast<zhegvd64_t>(dlsym(h, "scipy_LAPACKE_zhegvd64_"));
  g_zheevd = reinterpret_cast<zheevd64_t>(dlsym(h, "scipy_LAPACKE_zheevd64_"));
  g_blas_threads = reinterpret_cast<set_threads_t>(dlsym(h, "scipy_openblas_set_num_threads64_"));
  return g_zgemm && g_zhegvd && g_zheevd;
}

bool lapack_hegvd(std::size_t n, C* a, C* b, double* w) {
  if (!g_zhegvd) return false;
  const auto info = g_zhegvd(102, 1, 'V', 'L', std::int64_t(n), a, std::int64_t(n), b, std::int64_t(n), w);
  if (info) throw std::runtime_error("LAPACK zhegvd failed");
  return true;
}

bool lapack_heevd(std::size_t n, C* a, double* w) {
  if (!g_zheevd) return false;
  const auto info = g_zheevd(102, 'V', 'L', std::int64_t(n), a, std::int64_t(n), w);
  if (info) throw std::runtime_error("LAPACK zheevd failed");
  return true;
}

// ---------------------------------------------------------------- GEMM
void zgemm(Op ao, Op bo, std::size_t m, std::size_t n, std::size_t k, const C* a, std::size_t lda,
           const C* b, std::size_t ldb, C* c, std::size_t ldc, bool add) {
  if (!m || !n || !k || lda < (ao == Op::Normal ? m : k) || ldb < (bo == Op::Normal ? k : n) || ldc < m)
    throw std::invalid_argument("invalid complex GEMM shape/leading dimension");
  if (g_zgemm) {
    const C one(1), beta(add ? 1. : 0.);
    g_zgemm(102, ao == Op::Normal ? 111 : 113, bo == Op::Normal ? 111 : 113, std::int64_t(m), std::int64_t(n),
            std::int64_t(k), &one, a, std::int64_t(lda), b, std::int64_t(ldb), &beta, c, std::int64_t(ldc));
    return;
  }
#pragma omp parallel for schedule(static) if (double(m) * n * k > 1e5)
  for (std::ptrdiff_t j = 0; j < std::ptrdiff_t(n); ++j) {
    std::vector<C> col(m, C(0));
    for (std::size_t p = 0; p < k; ++p) {
      const C bv = bo == Op::Normal ? b[p + j * ldb] : std::conj(b[j + p * ldb]);
      if (ao == Op::Normal) {
        const C* ac = a + p * lda;
        for (std::size_t i = 0; i < m; ++i) col[i] += ac[i] * bv;
      } else {
        for (std::size_t i = 0; i < m; ++i) col[i] += std::conj(a[p + i * lda]) * bv;
      }
    }
    for (std::size_t i = 0; i < m; ++i) c[i + j * ldc] = add ? c[i + j * ldc] + col[i] : col[i];
  }
}

// src/paw_compensated_product.hpp:22-56: TwoSum accumulation of FMA-exact
// products, scalar order c, then per k: re*re, -im*im (real) and re*im, im*re.
namespace {
struct CompSum {
  double sum = 0., corr = 0.;
  void add(double v) {
    const double next = sum + v;
    const double back = next - sum;
    corr += (sum - (next - back)) + (v - back);
    sum = next;
  }
  void product(double a, double b) {
    const double r = a * b;
    add(r);
    add(std::fma(a, b, -r));
  }
  double value() const { return sum + corr; }
};
}  // namespace

void compensated_product(Op ao, std::size_t rows, std::size_t cols, std::size_t inner, const C* a,
                         std::size_t lda, const C* b, std::size_t ldb, C* c, std::size_t ldc, bool add) {
  if (!rows || !cols || !inner || ldb < inner || ldc < rows || lda < (ao == Op::Normal ? rows : inner))
    throw std::invalid_argument("invalid compensated PAW product dimensions");
  for (std::size_t col = 0; col < cols; ++col)
    for (std::size_t row = 0; row < rows; ++row) {
      CompSum re, im;
      if (add) { re.add(c[row + col * ldc].real()); im.add(c[row + col * ldc].imag()); }
      for (std::size_t k = 0; k < inner; ++k) {
        const C l = ao == Op::Normal ? a[row + k * lda] : std::conj(a[k + row * lda]);
        const C r = b[k + col * ldb];
        re.product(l.real(), r.real());
        re.product(-l.imag(), r.imag());
        im.product(l.real(), r.imag());
        im.product(l.imag(), r.real());
      }
      c[row + col * ldc] = C(re.value(), im.value());
    }
}

double kinetic_preconditioner(double kinetic, double wk) {
  if (!std::isfinite(kinetic) || kinetic < 0 || !std::isfinite(wk) || wk < 0)
    throw std::invalid_argument("kinetic preconditioner requires nonnegative Hartree energies");
  const double scale = 1.5 * std::max(wk, 2. / 27.211386245988);
  const double x = kinetic / scale;
  if (x > 1.) {
    const double inv = 1. / x;
    const double p = inv * (8. + inv * (12. + inv * (18. + 27. * inv)));
    return (2. / scale) * p / (p + 16.);
  }
  const double p = 27. + x * (18. + x * (12. + 8. * x));
  return (2. / scale) * p / (p + 16. * x * x * x * x);
}

// ---------------------------------------------------------------- operator
Operator::~Operator() { delete fft; }

namespace {
std::size_t wrap_index(const std::array<int, 3>& s, int h, int k, int l) {
  auto wrap = [](int x, int size) { int r = x % size; return r < 0 ? r + size : r; };
  return (std::size_t(wrap(h, s[0])) * s[1] + wrap(k, s[1])) * s[2] + wrap(l, s[2]);
}
void finish_paw(Operator& op, const std::vector<AtomBlock>& paw) {
  std::size_t total = 0;
  for (const auto& a : paw) total += a.projectors;
  op.np = total;
  op.all_b.assign(total * op.n, C(0));
  std::size_t off = 0;
  for (const auto& raw : paw) {
    const std::size_t r = raw.projectors;
    if (!r || raw.amplitudes.size() != r * op.n || raw.d.size() != r * r || raw.q.size() != r * r)
      throw std::invalid_argument("invalid PAW low-rank operator dimensions");
    std::vector<C> d(r * r), q(r * r);  // column-major, Hermitised (src/plane_wave_operator.cpp:494-499)
    for (std::size_t i = 0; i < r; ++i)
      for (std::size_t j = 0; j < r; ++j) {
        d[i + j * r] = .5 * (raw.d[i * r + j] + std::conj(raw.d[j * r + i]));
        q[i + j * r] = .5 * (raw.q[i * r + j] + std::conj(raw.q[j * r + i]));
      }
    for (std::size_t i = 0; i < r; ++i)
      for (std::size_t g = 0; g < op.n; ++g) op.all_b[(off + i) + g * total] = raw.amplitudes[i * op.n + g];
    // diagonals: b_g^H (D b)_g (src/plane_wave_operator.cpp:501-502)
    for (std::size_t g = 0; g < op.n; ++g) {
      C hd(0), sd(0);
      for (std::size_t i = 0; i < r; ++i) {
        C db(0), qb(0);
        for (std::size_t j = 0; j < r; ++j) {
          db += d[i + j * r] * raw.amplitudes[j * op.n + g];
          qb += q[i + j * r] * raw.amplitudes[j * op.n + g];
        }
        hd += std::conj(raw.amplitudes[i * op.n + g]) * db;
        sd += std::conj(raw.amplitudes[i * op.n + g]) * qb;
      }
      op.h_diagonal[g] += hd.real();
      op.s_diagonal[g] += sd.real();
    }
    op.atom_offset.push_back(off);
    op.atom_rows.push_back(r);
    op.d.push_back(std::move(d));
    op.q.push_back(std::move(q));
    off += r;
  }
}
}  // namespace

Operator* make_operator(const std::vector<GVector>& basis, const FourierDensity& potential,
                        const std::vector<AtomBlock>& paw, double hartree_to_ev, double regularization,
                        std::array<int, 3> shape) {
  auto op = new Operator();
  try {
    op->n = basis.size();
    op->regularization = regularization;
    if (!op->n) throw std::invalid_argument("empty basis");
    int low[3]{basis[0].h, basis[0].k, basis[0].l}, high[3]{low[0], low[1], low[2]}, radius[3]{};
    for (const auto& g : basis) {
      const int key[]{g.h, g.k, g.l};
      for (int a = 0; a < 3; ++a) { low[a] = std::min(low[a], key[a]); high[a] = std::max(high[a], key[a]); }
    }
    double vmax = 0;
    for (const auto& [key, v] : potential) {
      const int k[]{std::get<0>(key), std::get<1>(key), std::get<2>(key)};
      bool retained = true;
      for (int a = 0; a < 3; ++a) retained &= std::abs(k[a]) <= high[a] - low[a];
      if (retained)
        for (int a = 0; a < 3; ++a) radius[a] = std::max(radius[a], std::abs(k[a]));
      vmax = std::max(vmax, std::abs(v));
    }
    for (const auto& [key, v] : potential) {  // src/plane_wave_operator.cpp:439-443
      const auto other = potential.find({-std::get<0>(key), -std::get<1>(key), -std::get<2>(key)});
      if (std::abs(v - std::conj(other == potential.end() ? C{} : other->second)) > 1e-10 * std::max(1., vmax))
        throw std::invalid_argument("local Fourier potential is not Hermitian");
    }
    op->count = 1;
    for (int a = 0; a < 3; ++a) {  // src/plane_wave_operator.cpp:446-457
      const int minimum = std::max(3, 2 * std::max(radius[a], high[a] - low[a]) + 1);
      op->shape[a] = shape[a] > 0 ? shape[a] : fast_fft_length(minimum);
      op->count *= op->shape[a];
    }
    op->fft = new Fft3d(op->shape);
    std::vector<C> buf(op->count, C(0));
    for (const auto& [key, v] : potential) {
      const int k[]{std::get<0>(key), std::get<1>(key), std::get<2>(key)};
      bool retained = true;
      for (int a = 0; a < 3; ++a) retained &= std::abs(k[a]) <= high[a] - low[a];
      if (!retained) continue;
      buf[wrap_index(op->shape, k[0], k[1], k[2])] = v;
    }
    op->fft->exec(buf.data(), +1);
    op->local_real.resize(op->count);
    for (std::size_t i = 0; i < op->count; ++i) {
      if (std::abs(buf[i].imag()) > 1e-9 * std::max(1., std::abs(buf[i].real())))
        throw std::runtime_error("complex real-space local potential");
      op->local_real[i] = buf[i].real();
    }
    const auto zero = potential.find({0, 0, 0});
    const double v0 = zero == potential.end() ? 0. : zero->second.real();
    for (const auto& g : basis) {
      op->indices.push_back(wrap_index(op->shape, g.h, g.k, g.l));
      op->kinetic.push_back(g.kinetic_ev / hartree_to_ev);
      op->h_diagonal.push_back(g.kinetic_ev / hartree_to_ev + v0);
      op->s_diagonal.push_back(1 + regularization);
    }
    auto sorted = op->indices;
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      throw std::invalid_argument("duplicate plane-wave index");
    finish_paw(*op, paw);
  } catch (...) {
    delete op;
    throw;
  }
  return op;
}

Operator* make_operator_arrays(std::size_t n, std::array<int, 3> shape, const std::size_t* indices,
                               const double* kinetic, const double* local_real, double regularization,
                               std::size_t atoms, const std::size_t* rows, const C* all_b_colmajor,
                               const C* d_blocks, const C* q_blocks, double v0) {
  auto op = new Operator();
  op->n = n;
  op->shape = shape;
  op->count = std::size_t(shape[0]) * shape[1] * shape[2];
  op->regularization = regularization;
  op->indices.assign(indices, indices + n);
  op->kinetic.assign(kinetic, kinetic + n);
  op->local_real.assign(local_real, local_real + op->count);
  op->fft = new Fft3d(shape);
  for (std::size_t g = 0; g < n; ++g) {
    op->h_diagonal.push_back(kinetic[g] + v0);
    op->s_diagonal.push_back(1 + regularization);
  }
  std::size_t np = 0;
  for (std::size_t a = 0; a < atoms; ++a) np += rows[a];
  std::vector<AtomBlock> paw(atoms);
  std::size_t off = 0, doff = 0;
  for (std::size_t a = 0; a < atoms; ++a) {
    const std::size_t r = rows[a];
    paw[a].projectors = r;
    paw[a].amplitudes.resize(r * n);
    for (std::size_t i = 0; i < r; ++i)
      for (std::size_t g = 0; g < n; ++g) paw[a].amplitudes[i * n + g] = all_b_colmajor[(off + i) + g * np];
    paw[a].d.resize(r * r);
    paw[a].q.resize(r * r);
    for (std::size_t i = 0; i < r; ++i)
      for (std::size_t j = 0; j < r; ++j) {  // input blocks column-major
        paw[a].d[i * r + j] = d_blocks[doff + i + j * r];
        paw[a].q[i * r + j] = q_blocks[doff + i + j * r];
      }
    off += r;
    doff += r * r;
  }
  finish_paw(*op, paw);
  return op;
}

// src/plane_wave_operator.cpp:218-306 (exact=true CPU path; per-atom D/Q
// products compensated, B X and B^H P through the plain GEMM).
void apply_block(const Operator& op, const C* x, std::size_t columns, C* y, bool overlap,
                 const C* shared) {
  if (!columns || !x || !y || x == y) throw std::invalid_argument("invalid direct plane-wave buffers");
  const std::size_t n = op.n, count = op.count;
  if (overlap) {
    for (std::size_t c = 0; c < columns; ++c)
      for (std::size_t i = 0; i < n; ++i) y[c * n + i] = (1 + op.regularization) * x[c * n + i];
  } else {
#pragma omp parallel if (count * columns >= 32768)
    {
      std::vector<C> work(count);
#pragma omp for schedule(static)
      for (std::ptrdiff_t c = 0; c < std::ptrdiff_t(columns); ++c) {
        const C* in = x + c * n;
        C* out = y + c * n;
        std::fill(work.begin(), work.end(), C(0));
        for (std::size_t i = 0; i < n; ++i) work[op.indices[i]] = in[i];
        op.fft->exec(work.data(), +1);
        for (std::size_t i = 0; i < count; ++i) work[i] = C(work[i].real() * op.local_real[i], work[i].imag() * op.local_real[i]);
        op.fft->exec(work.data(), -1);
        for (std::size_t i = 0; i < n; ++i) out[i] = op.kinetic[i] * in[i] + work[op.indices[i]] / double(count);
      }
    }
  }
  if (!op.np) return;
  std::vector<C> owned;
  if (!shared) {
    owned.resize(op.np * columns);
    zgemm(Op::Normal, Op::Normal, op.np, columns, n, op.all_b.data(), op.np, x, n, owned.data(), op.np);
    shared = owned.data();
  }
  std::vector<C> projected(op.np * columns);
  for (std::size_t a = 0; a < op.atom_rows.size(); ++a) {
    const auto& onsite = overlap ? op.q[a] : op.d[a];
    const std::size_t r = op.atom_rows[a], off = op.atom_offset[a];
    compensated_product(Op::Normal, r, columns, r, onsite.data(), r, shared + off, op.np, projected.data() + off,
                        op.np, false);
  }
  zgemm(Op::Adjoint, Op::Normal, n, columns, op.np, op.all_b.data(), op.np, projected.data(), op.np, y, n, true);
}

void apply_pair(const Operator& op, const C* x, std::size_t columns, C* hy, C* sy) {
  if (!columns || !x || !hy || !sy || x == hy || x == sy || hy == sy)
    throw std::invalid_argument("invalid joint plane-wave buffers");
  std::vector<C> amplitudes;
  const C* shared = nullptr;
  if (op.np) {
    amplitudes.resize(op.np * columns);
    zgemm(Op::Normal, Op::Normal, op.np, columns, op.n, op.all_b.data(), op.np, x, op.n, amplitudes.data(), op.np);
    shared = amplitudes.data();
  }
  apply_block(op, x, columns, hy, false, shared);
  apply_block(op, x, columns, sy, true, shared);
}

}  // namespace orc
