// Solver snapshot reader/writer (format and validation of
// src/plane_wave_operator.cpp:520-653; see snapshot.hpp).
#include "snapshot.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <stdexcept>

namespace sdft_hip {
namespace {
using C = std::complex<double>;
constexpr std::uint64_t kLimit = std::uint64_t(1) << 40;  // bound on every derived size

static_assert(sizeof(C) == 2 * sizeof(double), "interleaved complex layout");
static_assert(std::numeric_limits<double>::is_iec559, "IEEE754 doubles");

void require_little_endian() {
  if (__BYTE_ORDER__ != __ORDER_LITTLE_ENDIAN__)
    throw std::runtime_error("solver snapshot currently requires little-endian IEEE754 host");
}

std::uint64_t mul(std::uint64_t a, std::uint64_t b) {
  if (b && a > kLimit / b) throw std::runtime_error("solver snapshot dimension overflow");
  return a * b;
}

class Reader {
 public:
  explicit Reader(std::ifstream& f) : f_(f) {}
  template <class T>
  void read(T* out, std::uint64_t count) {
    if (count) f_.read(reinterpret_cast<char*>(out), std::streamsize(count * sizeof(T)));
    if (!f_) throw std::runtime_error("truncated solver snapshot");
  }
  template <class T>
  std::vector<T> vec(std::uint64_t count) {
    std::vector<T> v(count);
    read(v.data(), count);
    return v;
  }

 private:
  std::ifstream& f_;
};

bool all_finite(const std::vector<double>& v) {
  return std::all_of(v.begin(), v.end(), [](double x) { return std::isfinite(x); });
}
bool all_finite(const C* v, std::size_t count) {
  for (std::size_t i = 0; i < count; ++i)
    if (!std::isfinite(v[i].real()) || !std::isfinite(v[i].imag())) return false;
  return true;
}
// ||M - M^H||_F <= 1e-12 max(1, ||M||_F) for a column-major r x r block
bool hermitian(const C* m, std::size_t r) {
  double diff = 0, norm = 0;
  for (std::size_t j = 0; j < r; ++j)
    for (std::size_t i = 0; i < r; ++i) {
      norm += std::norm(m[i + j * r]);
      diff += std::norm(m[i + j * r] - std::conj(m[j + i * r]));
    }
  return std::sqrt(diff) <= 1e-12 * std::max(1.0, std::sqrt(norm));
}

template <class T>
void put(std::ofstream& f, const T* v, std::size_t count) {
  if (count) f.write(reinterpret_cast<const char*>(v), std::streamsize(count * sizeof(T)));
  if (!f) throw std::runtime_error("solver snapshot write failed");
}
}  // namespace

SnapshotArrays read_snapshot(const std::string& path) {
  require_little_endian();
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open solver snapshot: " + path);
  const std::streamoff length = f.tellg();
  f.seekg(0);
  Reader in(f);
  char magic[8];
  in.read(magic, 8);
  const std::string format(magic, 8);
  SnapshotArrays s;
  if (format == "SDFTPWO1") s.form = SnapshotArrays::Folded;
  else if (format == "SDFTPWO2") s.form = SnapshotArrays::Dense;
  else if (format == "SDFTPWO3") s.form = SnapshotArrays::Blocks;
  else throw std::runtime_error("unsupported solver snapshot format");
  std::uint64_t h[9];
  in.read(h, 9);
  s.n = h[0];
  s.count = h[1];
  for (int a = 0; a < 3; ++a) s.shape[a] = h[2 + a];
  s.projectors = h[5];
  s.roots = h[6];
  s.iterations = h[7];
  const std::uint64_t ni = h[8];
  bool ok = s.n && s.count && s.n <= s.count && s.roots && s.roots <= s.n && s.iterations &&
            s.iterations <= std::uint64_t(INT_MAX);
  for (int a = 0; a < 3; ++a) ok = ok && s.shape[a] >= 3 && s.shape[a] <= std::uint64_t(INT_MAX);
  if (ok) ok = mul(mul(s.shape[0], s.shape[1]), s.shape[2]) == s.count && (!ni || ni == mul(s.n, s.roots));
  if (!ok) throw std::runtime_error("invalid solver snapshot dimensions");
  double settings[2];
  in.read(settings, 2);
  std::uint64_t block_square = 0, extra = 0;
  if (s.form == SnapshotArrays::Blocks) {
    std::uint64_t atoms = 0;
    in.read(&atoms, 1);
    if (!atoms || atoms > s.projectors || length < 0 || atoms > std::uint64_t(length) / 8)
      throw std::runtime_error("invalid solver snapshot atom count");
    s.channels = in.vec<std::uint64_t>(atoms);
    std::uint64_t spanned = 0;
    for (std::uint64_t c : s.channels) {
      if (!c || c > s.projectors - spanned) throw std::runtime_error("invalid solver snapshot atom boundaries");
      spanned += c;
      const std::uint64_t sq = mul(c, c);
      if (sq > kLimit - block_square) throw std::runtime_error("solver snapshot atom storage overflow");
      block_square += sq;
    }
    if (spanned != s.projectors) throw std::runtime_error("solver snapshot atom boundaries do not span projectors");
    extra = 8 * (1 + atoms);
  }
  const std::uint64_t paw = s.form == SnapshotArrays::Folded
                                ? mul(6, mul(s.n, s.projectors))
                                : mul(2, mul(s.n, s.projectors)) +
                                      mul(4, s.form == SnapshotArrays::Blocks ? block_square
                                                                              : mul(s.projectors, s.projectors));
  const std::uint64_t doubles = mul(4, s.n) + s.count + paw + mul(2, ni);
  if (doubles > kLimit / 8 || length < 0 || std::uint64_t(length) != 96 + extra + mul(8, doubles))
    throw std::runtime_error("solver snapshot size mismatch");
  s.regularization = settings[0];
  s.tolerance = settings[1];
  if (!std::isfinite(s.regularization) || s.regularization < 0 || !std::isfinite(s.tolerance) || s.tolerance <= 0)
    throw std::runtime_error("invalid solver snapshot scalars");
  s.indices = in.vec<std::uint64_t>(s.n);
  {
    std::vector<std::uint64_t> sorted = s.indices;
    std::sort(sorted.begin(), sorted.end());
    if (sorted.back() >= s.count || std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      throw std::runtime_error("invalid solver snapshot FFT indices");
  }
  for (auto* v : {&s.kinetic, &s.h_diagonal, &s.s_diagonal, &s.local_real}) {
    *v = in.vec<double>(v == &s.local_real ? s.count : s.n);
    if (!all_finite(*v)) throw std::runtime_error("nonfinite solver snapshot array");
  }
  for (double x : s.s_diagonal)
    if (x <= 0) throw std::runtime_error("invalid solver snapshot overlap diagonal");
  s.all_b = in.vec<C>(mul(s.projectors, s.n));
  if (!all_finite(s.all_b.data(), s.all_b.size())) throw std::runtime_error("nonfinite solver snapshot projector array");
  auto onsite = [&](std::vector<C>& dst, std::uint64_t rows, std::uint64_t cols, bool square) {
    const std::size_t at = dst.size();
    dst.resize(at + rows * cols);
    in.read(dst.data() + at, rows * cols);
    if (!all_finite(dst.data() + at, rows * cols)) throw std::runtime_error("nonfinite solver snapshot PAW array");
    if (square && !hermitian(dst.data() + at, rows))
      throw std::runtime_error("non-Hermitian solver snapshot onsite matrix");
  };
  if (s.form == SnapshotArrays::Blocks) {
    for (std::uint64_t c : s.channels) {
      onsite(s.onsite_d, c, c, true);
      onsite(s.onsite_q, c, c, true);
    }
  } else if (s.form == SnapshotArrays::Dense) {
    onsite(s.onsite_d, s.projectors, s.projectors, true);
    onsite(s.onsite_q, s.projectors, s.projectors, true);
  } else {
    onsite(s.onsite_d, s.projectors, s.n, false);
    onsite(s.onsite_q, s.projectors, s.n, false);
  }
  s.initial = in.vec<C>(ni);
  if (!all_finite(s.initial.data(), s.initial.size()))
    throw std::runtime_error("nonfinite solver snapshot warm vector");
  return s;
}

void write_snapshot(const std::string& path, const SnapshotArrays& s) {
  require_little_endian();
  if (!s.roots || s.roots > s.n || !s.iterations || !std::isfinite(s.tolerance) || s.tolerance <= 0 ||
      (!s.initial.empty() && s.initial.size() != s.roots * s.n))
    throw std::invalid_argument("invalid solver snapshot settings");
  if (!all_finite(s.initial.data(), s.initial.size()))
    throw std::invalid_argument("nonfinite snapshot initial vector");
  if (std::filesystem::exists(path) || std::filesystem::exists(path + ".tmp"))
    throw std::runtime_error("refusing to overwrite solver snapshot: " + path);
  std::ofstream f(path + ".tmp", std::ios::binary);
  if (!f) throw std::runtime_error("cannot open solver snapshot: " + path);
  const char* magic = s.form == SnapshotArrays::Blocks ? "SDFTPWO3" : s.form == SnapshotArrays::Dense ? "SDFTPWO2"
                                                                                                     : "SDFTPWO1";
  put(f, magic, 8);
  const std::uint64_t h[9]{s.n, s.count, s.shape[0], s.shape[1], s.shape[2], s.projectors,
                           s.roots, s.iterations, std::uint64_t(s.initial.size())};
  put(f, h, 9);
  const double settings[2]{s.regularization, s.tolerance};
  put(f, settings, 2);
  if (s.form == SnapshotArrays::Blocks) {
    const std::uint64_t atoms = s.channels.size();
    put(f, &atoms, 1);
    put(f, s.channels.data(), s.channels.size());
  }
  put(f, s.indices.data(), s.indices.size());
  for (const auto* v : {&s.kinetic, &s.h_diagonal, &s.s_diagonal, &s.local_real}) put(f, v->data(), v->size());
  put(f, s.all_b.data(), s.all_b.size());
  if (s.form == SnapshotArrays::Blocks) {
    std::size_t at = 0;
    for (std::uint64_t c : s.channels) {
      put(f, s.onsite_d.data() + at, c * c);
      put(f, s.onsite_q.data() + at, c * c);
      at += c * c;
    }
  } else {
    put(f, s.onsite_d.data(), s.onsite_d.size());
    put(f, s.onsite_q.data(), s.onsite_q.size());
  }
  put(f, s.initial.data(), s.initial.size());
  f.close();
  if (!f) throw std::runtime_error("cannot finalize solver snapshot");
  std::filesystem::rename(path + ".tmp", path);
}

}  // namespace sdft_hip
