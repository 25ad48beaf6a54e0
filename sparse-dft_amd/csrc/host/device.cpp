// Device context, grid planning and the device-resident plane-wave operator.
#include "device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <sstream>
#include <tuple>

#include "../hip/fft_lengths.hpp"

namespace sdft_hip {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}
void check(rocblas_status s, const char* what) {
  if (s != rocblas_status_success)
    throw std::runtime_error(std::string("HIP: ") + what + ": rocBLAS/rocSOLVER status " + std::to_string(int(s)));
}

std::uint64_t grid_index(const std::array<int, 3>& s, int h, int k, int l) {
  auto wrap = [](int x, int size) { int r = x % size; return r < 0 ? r + size : r; };
  return (std::uint64_t(wrap(h, s[0])) * s[1] + wrap(k, s[1])) * s[2] + wrap(l, s[2]);
}

// src/fft_planning.hpp:16-23
int fast_fft_length(int minimum) {
  if (minimum < 1) throw std::invalid_argument("HIP: invalid FFT length");
  for (int n = minimum;; ++n) {
    int rem = n;
    for (int p : {2, 3, 5, 7})
      while (rem % p == 0) rem /= p;
    if (rem == 1) return n;
    if (n == std::numeric_limits<int>::max()) throw std::overflow_error("HIP: FFT length overflow");
  }
}

namespace {
std::array<int, 3> splits_for(const std::array<int, 3>& shape) {
  std::array<int, 3> out{};
  for (int a = 0; a < 3; ++a) {
    int l1 = 0, l2 = 0;
    if (!length_split(shape[a], l1, l2))
      throw std::invalid_argument("HIP: unsupported FFT axis length " + std::to_string(shape[a]) +
                                  " (supported: 2,3,5,7-smooth lengths up to 128 with in-register factors)");
    out[a] = l1;
  }
  return out;
}
}  // namespace

HostGridPlan build_grid_plan(std::array<int, 3> shape, const std::uint64_t* indices, std::size_t n) {
  HostGridPlan p;
  p.shape = shape;
  p.split = splits_for(shape);
  p.count = long(shape[0]) * shape[1] * shape[2];
  const int n2 = shape[2];
  std::vector<std::pair<int, int>> keys;  // (xy, g)
  keys.reserve(n);
  for (std::size_t g = 0; g < n; ++g) {
    if (indices[g] >= std::uint64_t(p.count)) throw std::invalid_argument("HIP: plane-wave index out of range");
    keys.push_back({int(indices[g] / n2), int(g)});
  }
  std::vector<int> xy;
  xy.reserve(n);
  for (auto& k : keys) xy.push_back(k.first);
  std::sort(xy.begin(), xy.end());
  xy.erase(std::unique(xy.begin(), xy.end()), xy.end());
  p.stick_xy = xy;
  p.stick_g.assign(xy.size() * n2, -1);
  for (std::size_t g = 0; g < n; ++g) {
    const int key = int(indices[g] / n2), z = int(indices[g] % n2);
    const int sid = int(std::lower_bound(xy.begin(), xy.end(), key) - xy.begin());
    if (p.stick_g[std::size_t(sid) * n2 + z] >= 0) throw std::invalid_argument("HIP: duplicate plane-wave index");
    p.stick_g[std::size_t(sid) * n2 + z] = int(g);
  }
  return p;
}

HostGridPlan build_full_plan(std::array<int, 3> shape) {
  HostGridPlan p;
  p.shape = shape;
  p.split = splits_for(shape);
  p.count = long(shape[0]) * shape[1] * shape[2];
  const int planes = shape[0] * shape[1];
  p.stick_xy.resize(planes);
  p.stick_g.resize(std::size_t(p.count));
  for (int s = 0; s < planes; ++s) p.stick_xy[s] = s;
  for (long i = 0; i < p.count; ++i) p.stick_g[i] = int(i);
  return p;
}

void DeviceGrid::upload(const HostGridPlan& h, hipStream_t s) {
  stick_g.upload(h.stick_g.data(), h.stick_g.size(), s);
  stick_xy.upload(h.stick_xy.data(), h.stick_xy.size(), s);
  dev.n0 = h.shape[0];
  dev.n1 = h.shape[1];
  dev.n2 = h.shape[2];
  dev.count = h.count;
  for (int a = 0; a < 3; ++a) dev.split[a] = h.split[a];
  dev.nsticks = h.nsticks();
  dev.stick_g = stick_g.p;
  dev.stick_xy = stick_xy.p;
  for (int a = 0; a < 3; ++a) {
    const int L = h.shape[a];
    std::vector<double> t64(2 * L);
    std::vector<float> t32(2 * L);
    for (int j = 0; j < L; ++j) {
      const long double ang = 6.283185307179586476925286766559L * (long double)j / (long double)L;
      t64[2 * j] = double(std::cos(ang));
      t64[2 * j + 1] = double(std::sin(ang));
      t32[2 * j] = float(t64[2 * j]);
      t32[2 * j + 1] = float(t64[2 * j + 1]);
    }
    tw64[a].upload(t64.data(), t64.size(), s);
    tw32[a].upload(t32.data(), t32.size(), s);
    dev.tw64[a] = tw64[a].p;
    dev.tw32[a] = tw32[a].p;
  }
  // host vectors are temporaries of the caller: finish the copies now
  check(hipStreamSynchronize(s), "grid upload");
}

// ------------------------------------------------------------ Context
Context::Context(int device) : device_(device) {
  check(hipSetDevice(device), "hipSetDevice");
  hipDeviceProp_t props{};
  check(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties");
  name_ = props.name;
  if (props.gcnArchName[0]) name_ += std::string(" (") + props.gcnArchName + ")";
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  if (rocblas_create_handle(&blas_) != rocblas_status_success) {
    (void)hipStreamDestroy(stream_);
    throw std::runtime_error("HIP: rocblas_create_handle failed");
  }
  check(rocblas_set_stream(blas_, stream_), "rocblas_set_stream");
}

Context::~Context() {
  (void)hipStreamSynchronize(stream_);
  if (blas_) rocblas_destroy_handle(blas_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Context::record(const std::string& name, double seconds, std::size_t calls) {
  auto& c = counters[name];
  c.seconds += seconds;
  c.calls += calls;
}

void Context::zgemm(Precision prec, int aop, int bop, long M, long N, long K, double alpha, const double* A, long lda,
                    const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta) {
  if (M <= 0 || N <= 0) return;
  if (std::max({M, N, K}) > long(std::numeric_limits<int>::max()))
    throw std::overflow_error("HIP: GEMM dimension overflow");
  const std::size_t bytes = zgemm_workspace_bytes(int(M), int(N), int(K));
  double* w = bytes ? gemm_work(bytes) : nullptr;
  check(sdft_hip::zgemm(prec, aop, bop, int(M), int(N), int(K), alpha, A, lda, B, ldb, C, ldc, D, ldd, beta, w,
                        bytes ? gemm_work_bytes() : 0, stream_),
        "zgemm launch");
}

// ------------------------------------------------------------ operator
DeviceOperator::DeviceOperator(Context& ctx, const OperatorDesc& d) : ctx_(ctx) {
  using C = std::complex<double>;
  n_ = d.n;
  reg_ = d.regularization;
  if (!n_ || !std::isfinite(d.ev_per_hartree) || d.ev_per_hartree <= 0 || !std::isfinite(reg_) || reg_ < 0)
    throw std::invalid_argument("HIP: invalid plane-wave operator dimensions or units");
  // retained Galerkin box and grid rule (src/plane_wave_operator.cpp:400-457)
  int low[3]{d.hkl[0], d.hkl[1], d.hkl[2]}, high[3]{low[0], low[1], low[2]}, radius[3]{};
  for (std::size_t g = 0; g < n_; ++g)
    for (int a = 0; a < 3; ++a) {
      low[a] = std::min(low[a], d.hkl[3 * g + a]);
      high[a] = std::max(high[a], d.hkl[3 * g + a]);
    }
  std::map<std::tuple<int, int, int>, C> vmap;
  double vmax = 0;
  for (std::size_t i = 0; i < d.nv; ++i) {
    const C v(d.v_values[2 * i], d.v_values[2 * i + 1]);
    if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) throw std::invalid_argument("HIP: nonfinite local potential");
    const int k[3]{d.v_hkl[3 * i], d.v_hkl[3 * i + 1], d.v_hkl[3 * i + 2]};
    vmap[{k[0], k[1], k[2]}] = v;
    bool retained = true;
    for (int a = 0; a < 3; ++a) retained &= std::abs(k[a]) <= high[a] - low[a];
    if (retained)
      for (int a = 0; a < 3; ++a) radius[a] = std::max(radius[a], std::abs(k[a]));
    vmax = std::max(vmax, std::abs(v));
  }
  for (const auto& [key, v] : vmap) {
    const auto other = vmap.find({-std::get<0>(key), -std::get<1>(key), -std::get<2>(key)});
    if (std::abs(v - std::conj(other == vmap.end() ? C{} : other->second)) > 1e-10 * std::max(1., vmax))
      throw std::invalid_argument("HIP: local Fourier potential is not Hermitian");
  }
  count_ = 1;
  for (int a = 0; a < 3; ++a) {
    const int minimum = std::max(3, 2 * std::max(radius[a], high[a] - low[a]) + 1);
    shape_[a] = d.shape[a] > 0 ? d.shape[a] : fast_fft_length(minimum);
    count_ *= shape_[a];
  }
  hipStream_t s = ctx_.stream();
  // V(r) = Re(backward FFT of V(G)) on the device, FP64, slot order
  {
    std::vector<double> grid(2 * std::size_t(count_), 0.0);
    for (const auto& [key, v] : vmap) {
      const int k[]{std::get<0>(key), std::get<1>(key), std::get<2>(key)};
      bool retained = true;
      for (int a = 0; a < 3; ++a) retained &= std::abs(k[a]) <= high[a] - low[a];
      if (!retained) continue;
      const auto i = grid_index(shape_, k[0], k[1], k[2]);
      grid[2 * i] = v.real();
      grid[2 * i + 1] = v.imag();
    }
    const HostGridPlan full = build_full_plan(shape_);
    DeviceGrid fg;
    fg.upload(full, s);
    DevBuf<double> g, t1, t2, ratio;
    g.upload(grid.data(), grid.size(), s);
    t1.reserve(2 * std::size_t(count_));
    t2.reserve(2 * std::size_t(count_));
    const int blocks = 256;
    ratio.reserve(blocks);
    v64_.reserve(std::size_t(count_));
    v32_.reserve(std::size_t(count_));
    check(launch_potential_setup(fg.dev, g.p, t1.p, t2.p, v64_.p, v32_.p, ratio.p, blocks, s), "potential setup");
    std::vector<double> r(blocks);
    check(hipMemcpyAsync(r.data(), ratio.p, blocks * sizeof(double), hipMemcpyDeviceToHost, s), "ratio copy");
    ctx_.sync();
    if (*std::max_element(r.begin(), r.end()) > 1e-9) throw std::runtime_error("HIP: complex real-space local potential");
  }
  const auto zero = vmap.find({0, 0, 0});
  const double v0 = zero == vmap.end() ? 0. : zero->second.real();
  indices_.resize(n_);
  kinetic_.resize(n_);
  h_diag_.resize(n_);
  s_diag_.assign(n_, 1 + reg_);
  for (std::size_t g = 0; g < n_; ++g) {
    if (!std::isfinite(d.kinetic_ev[g])) throw std::invalid_argument("HIP: nonfinite kinetic energy");
    indices_[g] = grid_index(shape_, d.hkl[3 * g], d.hkl[3 * g + 1], d.hkl[3 * g + 2]);
    kinetic_[g] = d.kinetic_ev[g] / d.ev_per_hartree;
    h_diag_[g] = d.kinetic_ev[g] / d.ev_per_hartree + v0;
  }
  const HostGridPlan plan = build_grid_plan(shape_, indices_.data(), n_);  // rejects duplicates
  grid_.upload(plan, s);
  grid_.dev.v64 = v64_.p;
  grid_.dev.v32 = v32_.p;
  kinetic_d_.upload(kinetic_.data(), n_, s);
  // PAW blocks (src/plane_wave_operator.cpp:484-516)
  atoms_ = int(d.atoms);
  np_ = 0;
  for (std::size_t a = 0; a < d.atoms; ++a) np_ += std::size_t(d.rows[a]);
  if (np_) {
    std::vector<C> ball(np_ * n_), dbl, qbl;
    std::vector<int> offs, rws;
    std::vector<long> boffs;
    const C* amp = reinterpret_cast<const C*>(d.amplitudes);
    const C* dd = reinterpret_cast<const C*>(d.d);
    const C* qq = reinterpret_cast<const C*>(d.q);
    std::size_t off = 0, ao = 0, mo = 0;
    for (std::size_t a = 0; a < d.atoms; ++a) {
      const std::size_t r = std::size_t(d.rows[a]);
      if (!r) throw std::invalid_argument("HIP: invalid PAW low-rank operator dimensions");
      std::vector<C> dh(r * r), qh(r * r);  // column-major Hermitised
      double dn = 0, qn = 0, de = 0, qe = 0;
      for (std::size_t i = 0; i < r; ++i)
        for (std::size_t j = 0; j < r; ++j) {
          const C dij = dd[mo + i * r + j], dji = dd[mo + j * r + i];
          const C qij = qq[mo + i * r + j], qji = qq[mo + j * r + i];
          if (!std::isfinite(dij.real()) || !std::isfinite(dij.imag()) || !std::isfinite(qij.real()) ||
              !std::isfinite(qij.imag()))
            throw std::invalid_argument("HIP: nonfinite PAW projector/D/Q data");
          dn += std::norm(dij);
          qn += std::norm(qij);
          de += std::norm(dij - std::conj(dji));
          qe += std::norm(qij - std::conj(qji));
          dh[i + j * r] = .5 * (dij + std::conj(dji));
          qh[i + j * r] = .5 * (qij + std::conj(qji));
        }
      if (std::sqrt(de) > 1e-10 * std::max(1., std::sqrt(dn)) || std::sqrt(qe) > 1e-10 * std::max(1., std::sqrt(qn))) {
        std::ostringstream m;
        m << "HIP: invalid Hermitian PAW blocks: atom=" << a;
        throw std::invalid_argument(m.str());
      }
      for (std::size_t i = 0; i < r; ++i)
        for (std::size_t g = 0; g < n_; ++g) {
          const C b = amp[ao + i * n_ + g];
          if (!std::isfinite(b.real()) || !std::isfinite(b.imag()))
            throw std::invalid_argument("HIP: nonfinite PAW projector/D/Q data");
          ball[(off + i) + g * np_] = b;
        }
      // diagonals b^H (D b) per G (src/plane_wave_operator.cpp:501-502)
      for (std::size_t g = 0; g < n_; ++g) {
        C hd(0), sd(0);
        for (std::size_t i = 0; i < r; ++i) {
          C db(0), qb(0);
          for (std::size_t j = 0; j < r; ++j) {
            db += dh[i + j * r] * amp[ao + j * n_ + g];
            qb += qh[i + j * r] * amp[ao + j * n_ + g];
          }
          hd += std::conj(amp[ao + i * n_ + g]) * db;
          sd += std::conj(amp[ao + i * n_ + g]) * qb;
        }
        h_diag_[g] += hd.real();
        s_diag_[g] += sd.real();
      }
      offs.push_back(int(off));
      rws.push_back(int(r));
      boffs.push_back(long(dbl.size()));
      dbl.insert(dbl.end(), dh.begin(), dh.end());
      qbl.insert(qbl.end(), qh.begin(), qh.end());
      off += r;
      ao += r * n_;
      mo += r * r;
    }
    b64_.upload(reinterpret_cast<const double*>(ball.data()), 2 * ball.size(), s);
    std::vector<float> b32(2 * ball.size());
    for (std::size_t i = 0; i < b32.size(); ++i) b32[i] = float(reinterpret_cast<const double*>(ball.data())[i]);
    b32_.upload(b32.data(), b32.size(), s);
    dblk_.upload(reinterpret_cast<const double*>(dbl.data()), 2 * dbl.size(), s);
    qblk_.upload(reinterpret_cast<const double*>(qbl.data()), 2 * qbl.size(), s);
    atom_offset_.upload(offs.data(), offs.size(), s);
    atom_rows_.upload(rws.data(), rws.size(), s);
    block_offset_.upload(boffs.data(), boffs.size(), s);
    ctx_.sync();
  }
  // chunk of columns per launch: bounds the stick buffers (~512 MiB each)
  const std::size_t per_col = std::size_t(grid_.dev.nsticks) * shape_[2] * 16;
  chunk_ = std::max<std::size_t>(1, std::min<std::size_t>(512, (std::size_t(512) << 20) / std::max<std::size_t>(1, per_col)));
  ctx_.sync();
}

std::vector<double> DeviceOperator::local_real_natural() {
  DevBuf<double> nat;
  nat.reserve(std::size_t(count_));
  check(launch_potential_to_natural(grid_.dev, v64_.p, nat.p, ctx_.stream()), "potential to natural");
  std::vector<double> out(static_cast<std::size_t>(count_));
  check(hipMemcpyAsync(out.data(), nat.p, out.size() * sizeof(double), hipMemcpyDeviceToHost, ctx_.stream()),
        "potential copy");
  ctx_.sync();
  return out;
}

void* DeviceOperator::stick_buffer(int which, std::size_t columns, Precision prec) {
  if (columns > chunk_) throw std::invalid_argument("HIP: column block exceeds the stick-buffer chunk");
  const std::size_t bytes = columns * std::size_t(grid_.dev.nsticks) * shape_[2] * (prec == Precision::FP64 ? 16 : 8);
  return which ? static_cast<void*>(t2_.reserve(bytes)) : static_cast<void*>(t1_.reserve(bytes));
}

void DeviceOperator::apply(int which, const double* x, long columns, double* hy, double* sy, Precision prec) {
  if (columns <= 0 || !x) throw std::invalid_argument("HIP: invalid plane-wave block operator buffers");
  if ((which != 1 && (!hy || hy == x)) || (which != 0 && (!sy || sy == x)) || (which == 2 && hy == sy))
    throw std::invalid_argument("HIP: invalid direct plane-wave input/output buffers");
  for (long c0 = 0; c0 < columns; c0 += long(chunk_)) {
    const long c = std::min<long>(long(chunk_), columns - c0);
    const long off = 2 * c0 * long(n_);
    apply_chunk(which, x + off, c, hy ? hy + off : nullptr, sy ? sy + off : nullptr, prec);
  }
}

// src/plane_wave_operator.cpp:218-374: local part through the stick pipeline,
// P = B X once for the pair, onsite D/Q blocks, then B^H (.) added with the
// kinetic/local H or the (1+reg) identity of S folded into the GEMM epilogue.
void DeviceOperator::apply_chunk(int which, const double* x, long c, double* hy, double* sy, Precision prec) {
  hipStream_t s = ctx_.stream();
  const long n = long(n_), np = long(np_);
  const bool want_h = which != 1, want_s = which != 0;
  if (np) {
    p_.reserve(std::size_t(2 * np * c));
    w_.reserve(std::size_t(2 * np * 2 * c));
    if (prec == Precision::FP64)
      ctx_.zgemm(prec, 0, 0, np, c, n, 1.0, b64_.p, np, x, n, p_.p, np, nullptr, 0, 0.0);
    else {
      const std::size_t bytes = zgemm_workspace_bytes(int(np), int(c), int(n));
      double* wk = bytes ? ctx_.gemm_work(bytes) : nullptr;
      check(zgemm_f32_storage(0, 0, int(np), int(c), int(n), 1.0, b32_.p, np, x, n, p_.p, np, nullptr, 0, 0.0, wk,
                              bytes ? ctx_.gemm_work_bytes() : 0, s),
            "projection");
    }
    const int mask = (want_h ? 1 : 0) | (want_s ? 2 : 0);
    check(launch_atom_blocks(prec, int(np), int(c), atoms_, atom_offset_.p, atom_rows_.p, block_offset_.p, dblk_.p,
                             qblk_.p, p_.p, np, w_.p, np, mask, s),
          "onsite blocks");
  }
  if (want_h) {
    const std::size_t bytes = std::size_t(c) * grid_.dev.nsticks * shape_[2] * (prec == Precision::FP64 ? 16 : 8);
    t1_.reserve(bytes);
    t2_.reserve(bytes);
    if (prec == Precision::FP64)
      check(launch_local_h<double>(grid_.dev, x, n, c, kinetic_d_.p, hy, n, t1_.p, t2_.p, s), "local H");
    else
      check(launch_local_h<float>(grid_.dev, x, n, c, kinetic_d_.p, hy, n, t1_.p, t2_.p, s), "local H");
  }
  auto backproject = [&](const double* w, double* out, const double* d, double beta) {
    if (prec == Precision::FP64)
      ctx_.zgemm(prec, 1, 0, n, c, np, 1.0, b64_.p, np, w, np, out, n, d, n, beta);
    else {
      const std::size_t bytes = zgemm_workspace_bytes(int(n), int(c), int(np));
      double* wk = bytes ? ctx_.gemm_work(bytes) : nullptr;
      check(zgemm_f32_storage(1, 0, int(n), int(c), int(np), 1.0, b32_.p, np, w, np, out, n, d, n, beta, wk,
                              bytes ? ctx_.gemm_work_bytes() : 0, s),
            "back projection");
    }
  };
  if (np) {
    if (want_h) backproject(w_.p, hy, hy, 1.0);
    if (want_s) backproject(w_.p + 2 * np * c, sy, x, 1.0 + reg_);
  } else if (want_s) {
    check(launch_scale_copy(x, n, sy, n, n, c, 1.0 + reg_, s), "overlap identity");
  }
}

}  // namespace sdft_hip
