// Device context, grid planning and the device-resident plane-wave operator.
#include "device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

#include "../hip/slot_order.hpp"
#include "../hip/fft_lengths.hpp"
#include "../hip/module_prime.hpp"
#include "../hip/scratch.hpp"
#include "snapshot.hpp"

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
  // x-lines holding sticks: the signed Miller range [lo, hi] of x, wrapped into
  // at most two ranges of natural x (the whole axis when they would overlap)
  const int n0 = shape[0], n1 = shape[1];
  int lo = n0, hi = -n0;
  for (int key : xy) {
    const int x = key / n1, h = x <= n0 / 2 ? x : x - n0;
    lo = std::min(lo, h);
    hi = std::max(hi, h);
  }
  if (xy.empty() || hi - lo + 1 >= n0) {
    p.xstart[0] = 0;
    p.xlines[0] = n0;
  } else {
    if (hi >= 0) {
      p.xstart[0] = std::max(lo, 0);
      p.xlines[0] = hi + 1 - p.xstart[0];
    }
    if (lo < 0) {
      p.xstart[1] = n0 + lo;
      p.xlines[1] = std::min(hi, -1) - lo + 1;
    }
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
  p.xlines[0] = shape[0];
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
  for (int r = 0; r < 2; ++r) {
    dev.xstart[r] = h.xstart[r];
    dev.xlines[r] = h.xlines[r];
  }
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
namespace {
// The kernel code objects of this library and rocBLAS's GEMM kernel library
// are loaded once per process, on the thread that creates the first context,
// rather than lazily on whichever thread first launches a kernel: a solve
// service's worker threads end before the process does, and a process that
// first used the libraries on such a thread ended in a segmentation fault at
// exit. Per device (each device loads its own code objects).
void prime_libraries(int device) {
  static std::mutex m;
  static std::vector<int> primed;
  std::lock_guard<std::mutex> lock(m);
  if (std::find(primed.begin(), primed.end(), device) != primed.end()) return;
  rocblas_initialize();
  for (hipError_t (*f)() : {prime_module_local_h, prime_module_gemm, prime_module_davidson_kernels,
                            prime_module_tridiag, prime_module_tridiag_eig, prime_module_realspace,
                            prime_module_batch})
    check(f(), "kernel code object");
  primed.push_back(device);
}
}  // namespace

Context::Context(int device) : device_(device) {
  check(hipSetDevice(device), "hipSetDevice");
  prime_libraries(device);
  scratch_ = scratch_region(device);
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
  check(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking), "hipStreamCreate");
  check(hipEventCreateWithFlags(&fork_event_, hipEventDisableTiming), "hipEventCreate");
  check(hipEventCreateWithFlags(&join_event_, hipEventDisableTiming), "hipEventCreate");
}

PinnedStaging::PinnedStaging(std::size_t bytes) : cap_(bytes) {
  check(hipHostMalloc(reinterpret_cast<void**>(&base_), bytes, hipHostMallocDefault), "hipHostMalloc");
}

PinnedStaging::~PinnedStaging() {
  if (base_) (void)hipHostFree(base_);
}

void* PinnedStaging::take(std::size_t bytes) {
  const std::size_t aligned = (bytes + 255) / 256 * 256;
  if (aligned > cap_ - used_) return nullptr;
  void* p = base_ + used_;
  used_ += aligned;
  return p;
}

PinnedStaging& Context::staging() {
  if (!staging_) staging_ = std::make_unique<PinnedStaging>(std::size_t(4) << 20);
  return *staging_;
}

namespace {
constexpr std::size_t kBulkChunk = std::size_t(8) << 20;
}

void Context::bulk_buffers() {
  if (bulk_[0]) return;
  for (int i = 0; i < 2; ++i) {
    check(hipHostMalloc(reinterpret_cast<void**>(&bulk_[i]), kBulkChunk, hipHostMallocDefault), "hipHostMalloc");
    check(hipEventCreateWithFlags(&bulk_done_[i], hipEventDisableTiming), "hipEventCreate");
  }
}

void Context::upload_bulk(void* device_dst, const void* host_src, std::size_t bytes) {
  bulk_buffers();
  const char* src = static_cast<const char*>(host_src);
  char* dst = static_cast<char*>(device_dst);
  bool pending[2] = {false, false};
  for (std::size_t off = 0, k = 0; off < bytes; off += kBulkChunk, ++k) {
    const int b = int(k & 1);
    const std::size_t len = std::min(kBulkChunk, bytes - off);
    if (pending[b]) check(hipEventSynchronize(bulk_done_[b]), "upload chunk");
    std::memcpy(bulk_[b], src + off, len);
    check(hipMemcpyAsync(dst + off, bulk_[b], len, hipMemcpyHostToDevice, stream_), "upload chunk");
    check(hipEventRecord(bulk_done_[b], stream_), "hipEventRecord");
    pending[b] = true;
  }
  sync();
}

void Context::download_bulk(void* host_dst, const void* device_src, std::size_t bytes) {
  bulk_buffers();
  const char* src = static_cast<const char*>(device_src);
  char* dst = static_cast<char*>(host_dst);
  const std::size_t chunks = (bytes + kBulkChunk - 1) / kBulkChunk;
  auto issue = [&](std::size_t k) {
    const std::size_t off = k * kBulkChunk, len = std::min(kBulkChunk, bytes - off);
    check(hipMemcpyAsync(bulk_[k & 1], src + off, len, hipMemcpyDeviceToHost, stream_), "download chunk");
    check(hipEventRecord(bulk_done_[k & 1], stream_), "hipEventRecord");
  };
  if (chunks) issue(0);
  for (std::size_t k = 0; k < chunks; ++k) {
    // chunk k + 1 lands in the other buffer while chunk k is copied out
    if (k + 1 < chunks) issue(k + 1);
    check(hipEventSynchronize(bulk_done_[k & 1]), "download chunk");
    const std::size_t off = k * kBulkChunk, len = std::min(kBulkChunk, bytes - off);
    std::memcpy(dst + off, bulk_[k & 1], len);
  }
}

void Context::fork() {
  check(hipEventRecord(fork_event_, stream_), "fork record");
  check(hipStreamWaitEvent(aux_, fork_event_, 0), "fork wait");
}

void Context::join() {
  check(hipEventRecord(join_event_, aux_), "join record");
  check(hipStreamWaitEvent(stream_, join_event_, 0), "join wait");
}

Context::~Context() {
  if (aux_) (void)hipStreamSynchronize(aux_);
  (void)hipStreamSynchronize(stream_);
  for (hipEvent_t e : events_) (void)hipEventDestroy(e);
  if (fork_event_) (void)hipEventDestroy(fork_event_);
  if (join_event_) (void)hipEventDestroy(join_event_);
  for (int i = 0; i < 2; ++i) {
    if (bulk_[i]) (void)hipHostFree(bulk_[i]);
    if (bulk_done_[i]) (void)hipEventDestroy(bulk_done_[i]);
  }
  if (blas_) rocblas_destroy_handle(blas_);
  if (aux_) (void)hipStreamDestroy(aux_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Context::inject_fault(int kind, long count) {
  if (kind < 0 || kind >= kFaultKinds || count < 0) throw std::invalid_argument("HIP: invalid fault injection");
  faults_[kind].store(count, std::memory_order_relaxed);
}

bool Context::take_fault(int kind) {
  long v = faults_[kind].load(std::memory_order_relaxed);
  while (v > 0)
    if (faults_[kind].compare_exchange_weak(v, v - 1, std::memory_order_relaxed)) return true;
  return false;
}

void Context::record(const std::string& name, double seconds, std::size_t calls) {
  auto& c = counters[name];
  c.seconds += seconds;
  c.calls += calls;
}

void Context::clock_enable(bool on) {
  sync();
  clock_on_ = on;
  pending_.clear();
  events_used_ = 0;
  for (auto& c : clock_) c = ClockStat{};
}

LaunchEvents Context::clock_launch(int kind, double work, double bytes) {
  if (!clock_on_) return {};
  if (kind < 0 || kind >= kClockKinds) throw std::invalid_argument("HIP: invalid kernel clock kind");
  while (events_.size() < events_used_ + 2) {
    hipEvent_t e{};
    check(hipEventCreate(&e), "hipEventCreate");
    events_.push_back(e);
  }
  pending_.push_back({kind, work, bytes, events_used_});
  LaunchEvents ev{events_[events_used_], events_[events_used_ + 1]};
  events_used_ += 2;
  return ev;
}

Context::ClockStat Context::clock_stat(int kind) {
  if (kind < 0 || kind >= kClockKinds) throw std::invalid_argument("HIP: invalid kernel clock kind");
  sync();
  for (const auto& p : pending_) {
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, events_[p.first_event], events_[p.first_event + 1]), "hipEventElapsedTime");
    auto& c = clock_[p.kind];
    c.launches += 1;
    c.milliseconds += ms;
    c.work += p.work;
    c.bytes += p.bytes;
  }
  pending_.clear();
  events_used_ = 0;
  return clock_[kind];
}

// Algorithmic HBM bytes of one complex GEMM: A (M x K, a_bytes per element as
// stored), B (K x N, FP64 complex) and C (M x N) once each, plus D when beta != 0.
double gemm_algorithmic_bytes(long M, long N, long K, int a_bytes, bool reads_d) {
  return double(M) * K * a_bytes + 16.0 * double(K) * N + 16.0 * double(M) * N * (reads_d ? 2 : 1);
}

void Context::zgemm(Precision prec, int aop, int bop, long M, long N, long K, double alpha, const double* A, long lda,
                    const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta) {
  if (M <= 0 || N <= 0) return;
  if (std::max({M, N, K}) > long(std::numeric_limits<int>::max()))
    throw std::overflow_error("HIP: GEMM dimension overflow");
  const std::size_t bytes = zgemm_workspace_bytes(int(M), int(N), int(K));
  double* w = bytes ? gemm_work(bytes) : nullptr;
  zgemm_into(prec, aop, bop, M, N, K, alpha, A, lda, B, ldb, C, ldc, D, ldd, beta, w, bytes ? gemm_work_bytes() : 0);
}

void Context::zgemm_into(Precision prec, int aop, int bop, long M, long N, long K, double alpha, const double* A,
                         long lda, const double* B, long ldb, double* C, long ldc, const double* D, long ldd,
                         double beta, double* work, std::size_t work_bytes) {
  if (M <= 0 || N <= 0) return;
  if (std::max({M, N, K}) > long(std::numeric_limits<int>::max()))
    throw std::overflow_error("HIP: GEMM dimension overflow");
  const LaunchEvents ev = K > 0 ? clock_launch(prec == Precision::FP64 ? kGemmFP64 : kGemmFP32, 8.0 * M * N * K,
                                               gemm_algorithmic_bytes(M, N, K, 16, beta != 0.0))
                                : LaunchEvents{};
  check(sdft_hip::zgemm(prec, aop, bop, int(M), int(N), int(K), alpha, A, lda, B, ldb, C, ldc, D, ldd, beta, work,
                        work_bytes, stream_, &ev),
        "zgemm launch");
}

// Algorithmic bytes of one local-Hamiltonian launch over m columns:
// m (2 npw c + 4 N c) + N r + npw (4 + r), c/r = complex/real element bytes of
// the arithmetic (each 3-D transform reads and writes the grid once; V(r), the
// sphere indices and T are read once per launch).
double local_algorithmic_bytes(long npw, long count, long m, bool fp64) {
  const double c = fp64 ? 16.0 : 8.0, r = fp64 ? 8.0 : 4.0;
  return double(m) * (2.0 * npw * c + 4.0 * count * c) + count * r + npw * (4.0 + r);
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
    atom_rows_host_ = rws;
    block_offset_.upload(boffs.data(), boffs.size(), s);
    ctx_.sync();
  }
  set_chunk();
  ctx_.sync();
}

// Columns per local-H launch: the stick buffers stay below 512 MiB each;
// SDFT_HIP_APPLY_CHUNK lowers the cap (a smaller chunk keeps the stick
// buffers resident in the last-level cache between the three kernels).
void DeviceOperator::set_chunk() {
  const std::size_t per_col = std::size_t(grid_.dev.nsticks) * shape_[2] * 16;
  std::size_t cap = 512;
  if (const char* e = std::getenv("SDFT_HIP_APPLY_CHUNK")) {
    const long v = std::atol(e);
    if (v > 0) cap = std::min<std::size_t>(cap, std::size_t(v));
  }
  chunk_ = std::max<std::size_t>(1, std::min<std::size_t>(cap, (std::size_t(512) << 20) / std::max<std::size_t>(1, per_col)));
}

DeviceOperator::DeviceOperator(Context& ctx, const SnapshotArrays& sa) : ctx_(ctx) {
  n_ = std::size_t(sa.n);
  np_ = std::size_t(sa.projectors);
  count_ = long(sa.count);
  reg_ = sa.regularization;
  form_ = int(sa.form);
  for (int a = 0; a < 3; ++a) {
    shape_[a] = int(sa.shape[a]);
    int l1 = 0, l2 = 0;
    if (!length_split(shape_[a], l1, l2)) throw std::invalid_argument("HIP: snapshot FFT length has no device transform");
  }
  if (sa.indices.size() != n_ || sa.kinetic.size() != n_ || sa.h_diagonal.size() != n_ ||
      sa.s_diagonal.size() != n_ || sa.local_real.size() != sa.count || sa.all_b.size() != np_ * n_)
    throw std::invalid_argument("HIP: inconsistent snapshot arrays");
  hipStream_t s = ctx_.stream();
  indices_ = sa.indices;
  kinetic_ = sa.kinetic;
  h_diag_ = sa.h_diagonal;
  s_diag_ = sa.s_diagonal;
  const HostGridPlan plan = build_grid_plan(shape_, indices_.data(), n_);
  grid_.upload(plan, s);
  {
    DevBuf<double> nat;
    nat.upload(sa.local_real.data(), sa.local_real.size(), s);
    v64_.reserve(std::size_t(count_));
    v32_.reserve(std::size_t(count_));
    check(launch_potential_to_slots(grid_.dev, nat.p, v64_.p, v32_.p, s), "potential to slots");
    ctx_.sync();
  }
  grid_.dev.v64 = v64_.p;
  grid_.dev.v32 = v32_.p;
  kinetic_d_.upload(kinetic_.data(), n_, s);
  if (np_) {
    const double* b = reinterpret_cast<const double*>(sa.all_b.data());
    b64_.upload(b, 2 * sa.all_b.size(), s);
    std::vector<float> b32(2 * sa.all_b.size());
    for (std::size_t i = 0; i < b32.size(); ++i) b32[i] = float(b[i]);
    b32_.upload(b32.data(), b32.size(), s);
    if (form_ == SnapshotArrays::Blocks) {
      std::vector<int> offs;
      std::vector<long> boffs;
      std::size_t off = 0, sq = 0;
      for (std::uint64_t c : sa.channels) {
        offs.push_back(int(off));
        atom_rows_host_.push_back(int(c));
        boffs.push_back(long(sq));
        off += c;
        sq += c * c;
      }
      if (off != np_ || sa.onsite_d.size() != sq || sa.onsite_q.size() != sq)
        throw std::invalid_argument("HIP: inconsistent snapshot atom blocks");
      atoms_ = int(sa.channels.size());
      dblk_.upload(reinterpret_cast<const double*>(sa.onsite_d.data()), 2 * sq, s);
      qblk_.upload(reinterpret_cast<const double*>(sa.onsite_q.data()), 2 * sq, s);
      atom_offset_.upload(offs.data(), offs.size(), s);
      atom_rows_.upload(atom_rows_host_.data(), atom_rows_host_.size(), s);
      block_offset_.upload(boffs.data(), boffs.size(), s);
    } else {
      const std::size_t cols = form_ == SnapshotArrays::Dense ? np_ : n_;
      if (sa.onsite_d.size() != np_ * cols || sa.onsite_q.size() != np_ * cols)
        throw std::invalid_argument("HIP: inconsistent snapshot onsite matrices");
      onsite_d_.upload(reinterpret_cast<const double*>(sa.onsite_d.data()), 2 * sa.onsite_d.size(), s);
      onsite_q_.upload(reinterpret_cast<const double*>(sa.onsite_q.data()), 2 * sa.onsite_q.size(), s);
    }
  }
  set_chunk();
  ctx_.sync();
}

SnapshotArrays DeviceOperator::snapshot(std::size_t roots, std::size_t iterations, double tolerance,
                                        const std::complex<double>* initial) {
  using C = std::complex<double>;
  SnapshotArrays sa;
  sa.form = SnapshotArrays::Form(form_);
  sa.n = n_;
  sa.count = std::uint64_t(count_);
  for (int a = 0; a < 3; ++a) sa.shape[a] = std::uint64_t(shape_[a]);
  sa.projectors = np_;
  sa.roots = roots;
  sa.iterations = iterations;
  sa.regularization = reg_;
  sa.tolerance = tolerance;
  sa.indices = indices_;
  sa.kinetic = kinetic_;
  sa.h_diagonal = h_diag_;
  sa.s_diagonal = s_diag_;
  sa.local_real = local_real_natural();
  hipStream_t s = ctx_.stream();
  auto down = [&](std::vector<C>& dst, const double* src, std::size_t count) {
    dst.resize(count);
    if (count) check(hipMemcpyAsync(dst.data(), src, 16 * count, hipMemcpyDeviceToHost, s), "snapshot copy");
  };
  if (np_) {
    down(sa.all_b, b64_.p, np_ * n_);
    if (form_ == SnapshotArrays::Blocks) {
      std::size_t sq = 0;
      for (int r : atom_rows_host_) {
        sa.channels.push_back(std::uint64_t(r));
        sq += std::size_t(r) * r;
      }
      down(sa.onsite_d, dblk_.p, sq);
      down(sa.onsite_q, qblk_.p, sq);
    } else {
      const std::size_t cols = form_ == SnapshotArrays::Dense ? np_ : n_;
      down(sa.onsite_d, onsite_d_.p, np_ * cols);
      down(sa.onsite_q, onsite_q_.p, np_ * cols);
    }
  } else {
    sa.form = SnapshotArrays::Dense;  // an operator without projectors is stored as PWO2
  }
  ctx_.sync();
  if (initial) sa.initial.assign(initial, initial + n_ * roots);
  return sa;
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

std::size_t DeviceOperator::device_bytes() const {
  std::size_t b = grid_.stick_g.bytes() + grid_.stick_xy.bytes();
  for (int a = 0; a < 3; ++a) b += grid_.tw64[a].bytes() + grid_.tw32[a].bytes();
  b += onsite_d_.bytes() + onsite_q_.bytes() + v64_.bytes() + kinetic_d_.bytes() + b64_.bytes() + dblk_.bytes() +
       qblk_.bytes() + v32_.bytes() + b32_.bytes() + atom_offset_.bytes() + atom_rows_.bytes() + block_offset_.bytes() +
       t1_.bytes() + t2_.bytes() + p_.bytes() + w_.bytes();
  b += rs_w64_.bytes() + rs_sdiag_.bytes() + rs_w32_.bytes() + rs_woff_.bytes() + rs_poff_.bytes() +
       rs_points_.bytes() + rs_upoints_.bytes() + rs_estart_.bytes() + rs_eatom_.bytes() + rs_elocal_.bytes() +
       rs_aoff_.bytes() + rs_planes_.bytes() + rs_psi_.bytes();
  return b;
}

void DeviceOperator::apply(int which, const double* x, long columns, double* hy, double* sy, Precision prec) {
  if (columns <= 0 || !x) throw std::invalid_argument("HIP: invalid plane-wave block operator buffers");
  if ((which != 1 && (!hy || hy == x)) || (which != 0 && (!sy || sy == x)) || (which == 2 && hy == sy))
    throw std::invalid_argument("HIP: invalid direct plane-wave input/output buffers");
  if (rs_enabled_ && prec == Precision::FP32) {
    // search applications through the real-space projectors (src/plane_wave_operator.cpp:221-223)
    if (which != 1) apply_real_space(0, x, columns, hy, prec);
    if (which != 0) apply_real_space(1, x, columns, sy, prec);
    return;
  }
  for (long c0 = 0; c0 < columns; c0 += long(chunk_)) {
    const long c = std::min<long>(long(chunk_), columns - c0);
    const long off = 2 * c0 * long(n_);
    apply_chunk(which, x + off, c, hy ? hy + off : nullptr, sy ? sy + off : nullptr, prec,
                owned_staging(which, c, prec));
  }
  if (which != 1 && ctx_.take_fault(Context::kFaultNonfiniteApply)) {
    // injected fault: all-ones bits are a NaN double
    check(hipMemsetAsync(hy, 0xFF, sizeof(double), ctx_.stream()), "fault injection");
  }
}

const unsigned* DeviceOperator::projectors_bf16() {
  if (!b16_.p) {
    const std::size_t count = np_ * n_;
    b16_.reserve(count);
    check(launch_to_bf16_pairs(b32_.p, long(count), b16_.p, ctx_.stream()), "BF16 projectors");
  }
  return b16_.p;
}

void DeviceOperator::apply_shifted(const double* x, long columns, const double* theta, double* y, Precision prec,
                                   bool bf16_projectors) {
  if (columns <= 0 || !x || !y || y == x || !theta)
    throw std::invalid_argument("HIP: invalid shifted plane-wave application buffers");
  if (rs_enabled_ && prec == Precision::FP32) {
    // the real-space projectors apply H and S separately
    double* sx = shift_work_.reserve(2 * n_ * std::size_t(columns));
    apply_real_space(0, x, columns, y, prec);
    apply_real_space(1, x, columns, sx, prec);
    check(launch_column_scaled_add(sx, long(n_), y, long(n_), long(n_), columns, theta, -1.0, ctx_.stream()),
          "shift");
  } else {
    for (long c0 = 0; c0 < columns; c0 += long(chunk_)) {
      const long c = std::min<long>(long(chunk_), columns - c0);
      const long off = 2 * c0 * long(n_);
      apply_chunk(2, x + off, c, y + off, nullptr, prec, owned_staging(2, c, prec), theta + c0,
                  bf16_projectors && prec != Precision::FP64);
    }
  }
  if (ctx_.take_fault(Context::kFaultNonfiniteApply)) {
    // injected fault: all-ones bits are a NaN double
    check(hipMemsetAsync(y, 0xFF, sizeof(double), ctx_.stream()), "fault injection");
  }
}

namespace {
std::size_t align256(std::size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

std::size_t DeviceOperator::StagingBytes::total() const {
  return 2 * align256(sticks) + align256(p) + align256(w) + align256(work);
}

// The buffers apply_chunk touches for `c` columns: stick buffers only when H
// is applied, P when it is formed (atom-block and dense forms), W for any
// nonlocal part, and the largest split-K workspace of the pass's GEMMs.
DeviceOperator::StagingBytes DeviceOperator::staging_bytes(int which, long c, Precision prec) const {
  StagingBytes z;
  const long n = long(n_), np = long(np_);
  const std::size_t csz = prec == Precision::FP64 ? 16 : 8;
  if (which != 1) z.sticks = std::size_t(c) * std::size_t(grid_.dev.nsticks) * std::size_t(shape_[2]) * csz;
  if (!np) return z;
  const int cc = int(c);
  z.w = 32 * std::size_t(np) * std::size_t(c);
  if (form_ == 3 || form_ == 2) {
    z.p = 16 * std::size_t(np) * std::size_t(c);
    z.work = zgemm_workspace_bytes(int(np), cc, int(n));
  }
  if (form_ != 3) z.work = std::max(z.work, zgemm_workspace_bytes(int(np), cc, int(form_ == 2 ? np : n)));
  z.work = std::max(z.work, zgemm_workspace_bytes(int(n), cc, int(np)));
  return z;
}

DeviceOperator::Staging DeviceOperator::owned_staging(int which, long c, Precision prec) {
  const StagingBytes z = staging_bytes(which, c, prec);
  Staging st;
  if (z.sticks) {
    st.t1 = t1_.reserve(z.sticks);
    st.t2 = t2_.reserve(z.sticks);
  }
  if (z.p) st.p = p_.reserve(z.p / sizeof(double));
  if (z.w) st.w = w_.reserve(z.w / sizeof(double));
  if (z.work) {
    st.work = ctx_.gemm_work(z.work);
    st.work_bytes = ctx_.gemm_work_bytes();
  }
  return st;
}

std::size_t DeviceOperator::noalloc_bytes(int which, long columns, Precision prec) const {
  return staging_bytes(which, columns, prec).total();
}

void DeviceOperator::apply_noalloc(int which, const double* x, long columns, double* hy, double* sy, Precision prec) {
  if (columns <= 0 || !x) throw std::invalid_argument("HIP: invalid plane-wave block operator buffers");
  if ((which != 1 && (!hy || hy == x)) || (which != 0 && (!sy || sy == x)) || (which == 2 && hy == sy))
    throw std::invalid_argument("HIP: invalid direct plane-wave input/output buffers");
  if (rs_enabled_ && prec == Precision::FP32)
    throw std::invalid_argument("HIP: the allocation-free apply does not take the real-space projector search");
  char* base = ctx_.scratch();
  if (!base) throw std::logic_error("HIP: context without a scratch array");
  const std::size_t one = noalloc_bytes(which, 1, prec);
  if (one > kScratchBytes)
    throw std::overflow_error("HIP: one column of this operator needs " + std::to_string(one) +
                            " bytes of staging, more than the " + std::to_string(kScratchBytes) +
                            "-byte scratch array of the allocation-free apply");
  for (long c0 = 0; c0 < columns;) {
    // the most columns (up to apply()'s chunk) whose staging fits
    long c = std::min<long>(long(chunk_), columns - c0);
    while (c > 1 && noalloc_bytes(which, c, prec) > kScratchBytes) {
      const double ratio = double(kScratchBytes) / double(noalloc_bytes(which, c, prec));
      c = std::max<long>(1, std::min<long>(c - 1, long(double(c) * ratio)));
    }
    const StagingBytes z = staging_bytes(which, c, prec);
    Staging st;
    char* q = base;
    auto carve = [&q](std::size_t bytes) {
      char* r = bytes ? q : nullptr;
      q += align256(bytes);
      return r;
    };
    st.t1 = carve(z.sticks);
    st.t2 = carve(z.sticks);
    st.p = reinterpret_cast<double*>(carve(z.p));
    st.w = reinterpret_cast<double*>(carve(z.w));
    st.work = reinterpret_cast<double*>(carve(z.work));
    st.work_bytes = z.work;
    const long off = 2 * c0 * long(n_);
    apply_chunk(which, x + off, c, hy ? hy + off : nullptr, sy ? sy + off : nullptr, prec, st);
    c0 += c;
  }
}

long DeviceOperator::enable_real_space_search(double factor, const double* lattice, const double* fractional,
                                              const double* radius, double* retained) {
  using C = std::complex<double>;
  if (factor == 0) {
    rs_enabled_ = false;
    return 0;
  }
  if (!std::isfinite(factor) || factor < 1. || factor > 2.)
    throw std::invalid_argument("SDFT_PROJECTOR_RADIUS_FACTOR must be in [1,2]");
  if (!np_ || form_ != 3) throw std::invalid_argument("HIP: real-space search needs atom-block PAW projectors");
  const double threshold = std::pow(10., -2. - 5. * (factor - 1.));
  hipStream_t s = ctx_.stream();
  const long n = long(n_), count = count_;
  std::vector<C> b(np_ * n_);
  check(hipMemcpyAsync(b.data(), b64_.p, 16 * b.size(), hipMemcpyDeviceToHost, s), "projector copy");
  ctx_.sync();
  int rmax = 0;
  for (int r : atom_rows_host_) rmax = std::max(rmax, r);
  DevBuf<double> xin, peak, w64a;
  DevBuf<float> w32a;
  DevBuf<char> t1, planes;
  DevBuf<int> pts;
  xin.reserve(std::size_t(2 * n * rmax));
  peak.reserve(std::size_t(count));
  t1.reserve(std::size_t(rmax) * grid_.dev.nsticks * shape_[2] * 16);
  planes.reserve(std::size_t(rmax) * count * 16);
  std::vector<long> woff, poff{0};
  std::vector<int> aoff, all_points;
  std::vector<C> w_all;
  std::vector<std::array<int, 3>> entries;  // (slot point, atom, local index)
  std::vector<C> xa;
  std::vector<double> pk(count);
  long row0 = 0;
  double kept_power = 0, total_power = 0;
  const int n0 = shape_[0], n1 = shape_[1], n2 = shape_[2];
  const int a0 = grid_.dev.split[0], a1 = grid_.dev.split[1], a2 = grid_.dev.split[2];
  // natural grid coordinates of slot-order plane index i ([zs][xs][ys])
  auto natural_xyz = [&](long i, int& x, int& y, int& z) {
    const int ys = int(i % n1), xs = int((i / n1) % n0), zs = int(i / (long(n0) * n1));
    x = slot_position(xs, a0, n0 / a0);
    y = slot_position(ys, a1, n1 / a1);
    z = slot_position(zs, a2, n2 / a2);
  };
  for (std::size_t a = 0; a < atom_rows_host_.size(); ++a) {
    const int rows = atom_rows_host_[a];
    // w_i(r) = FFT_fwd(B_i)(r) / N = conj(backward(conj B_i))(r) / N
    xa.assign(std::size_t(n) * rows, C{});
    for (int i = 0; i < rows; ++i)
      for (long g = 0; g < n; ++g) xa[std::size_t(i) * n + g] = std::conj(b[std::size_t(row0 + i) + std::size_t(g) * np_]);
    check(hipMemcpyAsync(xin.p, xa.data(), 16 * xa.size(), hipMemcpyHostToDevice, s), "projector rows");
    check(launch_backward_full<double>(grid_.dev, xin.p, n, rows, t1.p, planes.p, s), "projector transforms");
    check(launch_rs_peak(planes.p, count, rows, peak.p, s), "projector peak");
    check(hipMemcpyAsync(pk.data(), peak.p, 8 * std::size_t(count), hipMemcpyDeviceToHost, s), "peak copy");
    ctx_.sync();
    // peak holds N^2 sum_i |w_i|^2: keep |w(r)| > threshold max |w|, or the
    // points of the atom's sphere of radius factor * radius
    const double top = *std::max_element(pk.begin(), pk.end());
    const bool sphere = lattice && fractional && radius && radius[a] > 0;
    const double cut = sphere ? factor * radius[a] : 0.;
    std::vector<int> keep;
    double kept_a = 0, total_a = 0;
    for (long r = 0; r < count; ++r) {
      total_a += pk[r];
      bool k;
      if (sphere) {
        int ix, iy, iz;
        natural_xyz(r, ix, iy, iz);
        double f[3] = {double(ix) / n0 - fractional[3 * a], double(iy) / n1 - fractional[3 * a + 1],
                       double(iz) / n2 - fractional[3 * a + 2]};
        for (double& v : f) v -= std::round(v);
        double d2 = 0;
        for (int c = 0; c < 3; ++c) {
          const double u = f[0] * lattice[c] + f[1] * lattice[3 + c] + f[2] * lattice[6 + c];
          d2 += u * u;
        }
        k = d2 <= cut * cut;
      } else {
        k = pk[r] > threshold * threshold * top;
      }
      if (k) {
        keep.push_back(int(r));
        kept_a += pk[r];
      }
    }
    kept_power += kept_a;
    total_power += total_a;
    const long np_a = long(keep.size());
    pts.upload(keep.data(), keep.size(), s);
    w64a.reserve(std::size_t(2 * rows * std::max<long>(np_a, 1)));
    w32a.reserve(std::size_t(2 * rows * std::max<long>(np_a, 1)));
    check(launch_rs_values(planes.p, count, rows, pts.p, np_a, w64a.p, w32a.p, s), "projector values");
    std::vector<C> wa(std::size_t(rows) * np_a);
    check(hipMemcpyAsync(wa.data(), w64a.p, 16 * wa.size(), hipMemcpyDeviceToHost, s), "values copy");
    ctx_.sync();
    woff.push_back(long(w_all.size()));
    w_all.insert(w_all.end(), wa.begin(), wa.end());
    for (long j = 0; j < np_a; ++j) entries.push_back({keep[j], int(a), int(j)});
    all_points.insert(all_points.end(), keep.begin(), keep.end());
    poff.push_back(poff.back() + np_a);
    aoff.push_back(int(row0));
    row0 += rows;
  }
  // point-major CSR of (atom, local index), atoms ascending within a point
  std::sort(entries.begin(), entries.end());
  std::vector<int> upoints, estart, eatom, elocal;
  for (std::size_t t = 0; t < entries.size(); ++t) {
    if (t == 0 || entries[t][0] != entries[t - 1][0]) {
      upoints.push_back(entries[t][0]);
      estart.push_back(int(t));
    }
    eatom.push_back(entries[t][1]);
    elocal.push_back(entries[t][2]);
  }
  estart.push_back(int(entries.size()));
  rs_w64_.upload(reinterpret_cast<const double*>(w_all.data()), 2 * w_all.size(), s);
  std::vector<float> w32(2 * w_all.size());
  for (std::size_t i = 0; i < w32.size(); ++i) w32[i] = float(reinterpret_cast<const double*>(w_all.data())[i]);
  rs_w32_.upload(w32.data(), w32.size(), s);
  rs_woff_.upload(woff.data(), woff.size(), s);
  rs_poff_.upload(poff.data(), poff.size(), s);
  rs_points_.upload(all_points.data(), all_points.size(), s);
  rs_upoints_.upload(upoints.data(), upoints.size(), s);
  rs_estart_.upload(estart.data(), estart.size(), s);
  rs_eatom_.upload(eatom.data(), eatom.size(), s);
  rs_elocal_.upload(elocal.data(), elocal.size(), s);
  rs_aoff_.upload(aoff.data(), aoff.size(), s);
  std::vector<double> sd(n_, 1.0 + reg_);
  rs_sdiag_.upload(sd.data(), sd.size(), s);
  ctx_.sync();
  rs_ = RealSpaceDev{};
  rs_.atoms = int(atom_rows_host_.size());
  rs_.total_points = long(all_points.size());
  rs_.unique = long(upoints.size());
  rs_.w64 = rs_w64_.p;
  rs_.w32 = rs_w32_.p;
  rs_.w_offset = rs_woff_.p;
  rs_.point_offset = rs_poff_.p;
  rs_.atom_offset = rs_aoff_.p;
  rs_.atom_rows = atom_rows_.p;
  rs_.points = rs_points_.p;
  rs_.unique_points = rs_upoints_.p;
  rs_.entry_start = rs_estart_.p;
  rs_.entry_atom = rs_eatom_.p;
  rs_.entry_local = rs_elocal_.p;
  // natural indices of the retained points, for export
  rs_natural_.clear();
  rs_counts_.clear();
  for (std::size_t a = 0; a + 1 < poff.size(); ++a) {
    rs_counts_.push_back(poff[a + 1] - poff[a]);
    for (long j = poff[a]; j < poff[a + 1]; ++j) {
      int x, y, z;
      natural_xyz(all_points[j], x, y, z);
      rs_natural_.push_back((long(x) * n1 + y) * n2 + z);
    }
  }
  if (retained) *retained = total_power > 0 ? kept_power / total_power : 1.;
  rs_enabled_ = true;
  return rs_.total_points;
}

void DeviceOperator::real_space_points(std::vector<long>& natural, std::vector<long>& per_atom) const {
  natural = rs_natural_;
  per_atom = rs_counts_;
}

// src/plane_wave_operator.cpp:169-212 with the FFTs on the device
void DeviceOperator::apply_real_space(int which, const double* x, long columns, double* y, Precision prec) {
  if (!rs_enabled_) throw std::invalid_argument("HIP: real-space search is not enabled");
  if ((which != 0 && which != 1) || columns <= 0 || !x || !y || x == y)
    throw std::invalid_argument("HIP: invalid real-space search buffers");
  hipStream_t s = ctx_.stream();
  const long n = long(n_), np = long(np_), count = count_;
  const bool f64 = prec == Precision::FP64;
  const std::size_t csz = f64 ? 16 : 8;
  const long chunk = std::max<long>(1, std::min<long>(columns, (256L << 20) / (count * long(csz))));
  rs_planes_.reserve(std::size_t(chunk) * count * csz);
  rs_psi_.reserve(std::size_t(std::max<long>(rs_.total_points, 1)) * chunk * csz);
  t1_.reserve(std::size_t(chunk) * grid_.dev.nsticks * shape_[2] * csz);
  p_.reserve(std::size_t(2 * np * chunk));
  w_.reserve(std::size_t(2 * np * 2 * chunk));
  const double* diag = which ? rs_sdiag_.p : kinetic_d_.p;
  for (long c0 = 0; c0 < columns; c0 += chunk) {
    const long c = std::min(chunk, columns - c0);
    const double* xc = x + 2 * c0 * n;
    double* yc = y + 2 * c0 * n;
    if (f64) {
      check(launch_backward_full<double>(grid_.dev, xc, n, c, t1_.p, rs_planes_.p, s), "real-space backward");
      check(launch_rs_gather<double>(rs_planes_.p, count, rs_.points, rs_.total_points, int(c), rs_psi_.p, s), "gather");
      check(launch_rs_project<double>(rs_, rs_psi_.p, int(c), p_.p, np, s), "projection");
    } else {
      check(launch_backward_full<float>(grid_.dev, xc, n, c, t1_.p, rs_planes_.p, s), "real-space backward");
      check(launch_rs_gather<float>(rs_planes_.p, count, rs_.points, rs_.total_points, int(c), rs_psi_.p, s), "gather");
      check(launch_rs_project<float>(rs_, rs_psi_.p, int(c), p_.p, np, s), "projection");
    }
    check(launch_atom_blocks(prec, int(np), int(c), atoms_, atom_offset_.p, atom_rows_.p, block_offset_.p, dblk_.p,
                             qblk_.p, p_.p, np, w_.p, np, which ? 2 : 1, s),
          "onsite blocks");
    const double* amp = w_.p + (which ? 2 * np * c : 0);
    if (f64) {
      check(launch_rs_images<double>(rs_, rs_planes_.p, count, int(c), v64_.p, which, amp, np, s), "images");
      check(launch_forward_planes<double>(grid_.dev, rs_planes_.p, c, xc, n, diag, yc, n, t2_.reserve(std::size_t(c) * grid_.dev.nsticks * shape_[2] * csz), s),
            "real-space forward");
    } else {
      check(launch_rs_images<float>(rs_, rs_planes_.p, count, int(c), v32_.p, which, amp, np, s), "images");
      check(launch_forward_planes<float>(grid_.dev, rs_planes_.p, c, xc, n, diag, yc, n, t2_.reserve(std::size_t(c) * grid_.dev.nsticks * shape_[2] * csz), s),
            "real-space forward");
    }
  }
}

void DeviceOperator::project(const double* x, long columns, double* p, Precision prec) {
  if (columns <= 0 || !x || !p || x == p || !np_) throw std::invalid_argument("HIP: invalid PAW projection buffers");
  ctx_.zgemm(prec, 0, 0, long(np_), columns, long(n_), 1.0, b64_.p, long(np_), x, long(n_), p, long(np_), nullptr, 0,
             0.0);
}

// src/density_functionals.cpp:406-460. density(r) = sum_b occ_b / volume
// |psi_b(r)|^2 with psi_b the unnormalised backward transform of the band's
// coefficients; result(G) = forward transform / N at the support points. The
// grid is alias free for every G - G' and every support point (any such grid
// gives the same coefficients; the reference notes the same for its
// SDFT_FAST_DENSITY_GRID enlargement, src/density_functionals.cpp:356-366).
std::vector<std::complex<double>> DeviceOperator::density(const double* x, long bands, const double* occupations,
                                                          double volume, const int* support, long nsupport) {
  if (bands <= 0 || !x || !occupations || !support || nsupport <= 0 || !std::isfinite(volume) || volume <= 0)
    throw std::invalid_argument("invalid plane-wave FFT density dimensions");
  for (long b = 0; b < bands; ++b)
    if (!std::isfinite(occupations[b]) || occupations[b] < 0 || occupations[b] > 2)
      throw std::invalid_argument("invalid FFT spin-summed occupation");
  // Miller indices of the basis on the operator grid
  std::vector<int> hkl(3 * n_);
  int lo[3]{0, 0, 0}, hi[3]{0, 0, 0}, m[3]{0, 0, 0};
  for (std::size_t g = 0; g < n_; ++g) {
    const std::uint64_t i = indices_[g];
    const int c[3]{int(i / (std::uint64_t(shape_[1]) * shape_[2])), int((i / shape_[2]) % shape_[1]),
                   int(i % shape_[2])};
    for (int a = 0; a < 3; ++a) {
      const int h = c[a] <= shape_[a] / 2 ? c[a] : c[a] - shape_[a];
      hkl[3 * g + a] = h;
      lo[a] = std::min(lo[a], h);
      hi[a] = std::max(hi[a], h);
    }
  }
  for (long j = 0; j < nsupport; ++j)
    for (int a = 0; a < 3; ++a) m[a] = std::max(m[a], std::abs(support[3 * j + a]));
  std::array<int, 3> shape{};
  for (int a = 0; a < 3; ++a) {
    m[a] = std::max(m[a], hi[a] - lo[a]);
    int len = std::max(3, 2 * m[a] + 1), l1 = 0, l2 = 0;
    while (len <= 128 && !length_split(len, l1, l2)) ++len;
    if (len > 128) throw std::invalid_argument("HIP: density grid beyond the device FFT lengths");
    shape[a] = len;
  }
  if ((std::size_t(shape[0]) * (shape[1] | 1) + shape[0] + shape[1]) * 16 > std::size_t(max_lds_bytes()))
    throw std::invalid_argument("HIP: density grid plane exceeds the LDS of one workgroup");
  const long count = long(shape[0]) * shape[1] * shape[2];
  std::vector<std::uint64_t> idx(n_);
  for (std::size_t g = 0; g < n_; ++g) idx[g] = grid_index(shape, hkl[3 * g], hkl[3 * g + 1], hkl[3 * g + 2]);
  hipStream_t s = ctx_.stream();
  DeviceGrid gd, gf;
  gd.upload(build_grid_plan(shape, idx.data(), n_), s);
  gf.upload(build_full_plan(shape), s);
  const long chunk = std::max<long>(1, std::min<long>(bands, (256L << 20) / (count * 16)));
  DevBuf<char> t1, t2;
  t1.reserve(std::size_t(std::max<long>(chunk * gd.dev.nsticks * shape[2], count)) * 16);
  t2.reserve(std::size_t(chunk * count) * 16);
  DevBuf<double> rho, nat, weights, out;
  rho.reserve(std::size_t(count));
  nat.reserve(std::size_t(2 * count));
  weights.reserve(std::size_t(chunk));
  out.reserve(std::size_t(2 * nsupport));
  check(hipMemsetAsync(rho.p, 0, sizeof(double) * std::size_t(count), s), "zero density");
  std::vector<double> w(chunk);
  for (long b0 = 0; b0 < bands; b0 += chunk) {
    const long c = std::min(chunk, bands - b0);
    for (long j = 0; j < c; ++j) w[j] = occupations[b0 + j] / volume;
    check(hipMemcpyAsync(weights.p, w.data(), sizeof(double) * std::size_t(c), hipMemcpyHostToDevice, s), "weights");
    check(launch_backward_planes(gd.dev, x + 2 * b0 * long(n_), long(n_), c, t1.p, t2.p, s), "band transforms");
    check(launch_density_accumulate(t2.p, count, int(c), weights.p, rho.p, s), "density accumulation");
    ctx_.sync();  // the host weights are reused by the next chunk
  }
  check(launch_slots_to_natural_complex(gf.dev, rho.p, nat.p, s), "density to natural order");
  check(launch_backward_planes(gf.dev, nat.p, 0, 1, t1.p, t2.p, s), "density transform");
  DevBuf<int> keys;
  keys.upload(support, std::size_t(3 * nsupport), s);
  check(launch_gather_support(gf.dev, t2.p, keys.p, nsupport, out.p, s), "support gather");
  std::vector<std::complex<double>> result(nsupport);
  check(hipMemcpyAsync(result.data(), out.p, 16 * std::size_t(nsupport), hipMemcpyDeviceToHost, s), "density copy");
  ctx_.sync();
  return result;
}

// Per-atom Q_a P_a in the compensated FP64 order (atom blocks) or the dense
// Q P (PWO2); the folded form keeps no Q of its own.
void DeviceOperator::projector_metric(const double* p, long columns, double* qp) {
  if (columns <= 0 || !p || !qp || p == qp) throw std::invalid_argument("HIP: invalid projector metric buffers");
  if (!np_) return;
  const long np = long(np_);
  if (form_ == 3) {
    // the onsite kernel writes Q P into the second column block of its output
    w_.reserve(std::size_t(2 * np * 2 * columns));
    check(launch_atom_blocks(Precision::FP64, int(np), int(columns), atoms_, atom_offset_.p, atom_rows_.p,
                             block_offset_.p, dblk_.p, qblk_.p, p, np, w_.p, np, 2, ctx_.stream()),
          "projector metric");
    check(hipMemcpyAsync(qp, w_.p + 2 * np * columns, 16 * std::size_t(np) * columns, hipMemcpyDeviceToDevice,
                         ctx_.stream()),
          "projector metric copy");
  } else if (form_ == 2) {
    ctx_.zgemm(Precision::FP64, 0, 0, np, columns, np, 1.0, onsite_q_.p, np, p, np, qp, np, nullptr, 0, 0.0);
  } else {
    throw std::invalid_argument("HIP: a folded (PWO1) operator has no projector metric");
  }
}

// src/plane_wave_operator.cpp:218-374: local part through the stick pipeline,
// P = B X once for the pair, onsite D/Q blocks, then B^H (.) added with the
// kinetic/local H or the (1+reg) identity of S folded into the GEMM epilogue.
void DeviceOperator::apply_chunk(int which, const double* x, long c, double* hy, double* sy, Precision prec,
                                 const Staging& st, const double* shift, bool bf16) {
  hipStream_t s = ctx_.stream();
  const long n = long(n_), np = long(np_);
  const bool want_h = which != 1, want_s = which != 0;
  // The local Hamiltonian (LDS/latency bound FFT pipeline) runs on the
  // context's second stream beside the projector chain (MFMA GEMMs) on the
  // main stream; the H back-projection adds to its output after the join.
  const bool overlap = want_h && np;
  if (want_h) {
    hipStream_t ls = overlap ? ctx_.aux_stream() : s;
    if (overlap) ctx_.fork();
    const bool f64 = prec == Precision::FP64;
    const double local_bytes = local_algorithmic_bytes(n, count_, c, f64);
    const LaunchEvents ev = ctx_.clock_launch(f64 ? Context::kLocalFP64 : Context::kLocalFP32, local_bytes, local_bytes);
    if (ev.begin) check(hipEventRecord(ev.begin, ls), "hipEventRecord");
    if (f64)
      check(launch_local_h<double>(grid_.dev, x, n, c, kinetic_d_.p, hy, n, st.t1, st.t2, ls), "local H");
    else
      check(launch_local_h<float>(grid_.dev, x, n, c, kinetic_d_.p, hy, n, st.t1, st.t2, ls), "local H");
    if (ev.end) check(hipEventRecord(ev.end, ls), "hipEventRecord");
  }
  auto gemm = [&](int aop, long M, long K, double alpha, const double* A, long lda, const double* B, long ldb,
                  double* C, long ldc, const double* D, long ldd, double beta) {
    ctx_.zgemm_into(prec, aop, 0, M, c, K, alpha, A, lda, B, ldb, C, ldc, D, ldd, beta, st.work, st.work_bytes);
  };
  if (np && form_ != 3) {
    // PWO2 dense D/Q on P = B X, or PWO1 folded DB/QB on X
    // (src/plane_wave_operator.cpp:290-301)
    const bool dense = form_ == 2;
    if (dense) gemm(0, np, n, 1.0, b64_.p, np, x, n, st.p, np, nullptr, 0, 0.0);
    const double* rhs = dense ? st.p : x;
    const long inner = dense ? np : n, ldr = dense ? np : n;
    if (want_h) gemm(0, np, inner, 1.0, onsite_d_.p, np, rhs, ldr, st.w, np, nullptr, 0, 0.0);
    if (want_s) gemm(0, np, inner, 1.0, onsite_q_.p, np, rhs, ldr, st.w + 2 * np * c, np, nullptr, 0, 0.0);
  } else if (np) {
    if (prec == Precision::FP64)
      gemm(0, np, n, 1.0, b64_.p, np, x, n, st.p, np, nullptr, 0, 0.0);
    else if (bf16) {
      const LaunchEvents ev =
          ctx_.clock_launch(Context::kGemmBF16, 8.0 * np * c * n, gemm_algorithmic_bytes(np, c, n, 4, false));
      check(zgemm_bf16_storage(0, int(np), int(c), int(n), 1.0, projectors_bf16(), np, x, n, st.p, np, nullptr, 0, 0.0,
                               st.work, st.work_bytes, s, &ev),
            "projection");
    } else {
      const LaunchEvents ev =
          ctx_.clock_launch(Context::kGemmFP32, 8.0 * np * c * n, gemm_algorithmic_bytes(np, c, n, 8, false));
      check(zgemm_f32_storage(0, 0, int(np), int(c), int(n), 1.0, b32_.p, np, x, n, st.p, np, nullptr, 0, 0.0, st.work,
                              st.work_bytes, s, &ev),
            "projection");
    }
    const int mask = (want_h ? 1 : 0) | (want_s ? 2 : 0);
    check(launch_atom_blocks(prec, int(np), int(c), atoms_, atom_offset_.p, atom_rows_.p, block_offset_.p, dblk_.p,
                             qblk_.p, st.p, np, st.w, np, mask, s),
          "onsite blocks");
  }
  auto backproject = [&](const double* w, double* out, const double* d, double beta) {
    if (prec == Precision::FP64)
      gemm(1, n, np, 1.0, b64_.p, np, w, np, out, n, d, n, beta);
    else if (bf16) {
      const LaunchEvents ev =
          ctx_.clock_launch(Context::kGemmBF16, 8.0 * n * c * np, gemm_algorithmic_bytes(n, c, np, 4, beta != 0.0));
      check(zgemm_bf16_storage(1, int(n), int(c), int(np), 1.0, projectors_bf16(), np, w, np, out, n, d, n, beta,
                               st.work, st.work_bytes, s, &ev),
            "back projection");
    } else {
      const LaunchEvents ev =
          ctx_.clock_launch(Context::kGemmFP32, 8.0 * n * c * np, gemm_algorithmic_bytes(n, c, np, 8, beta != 0.0));
      check(zgemm_f32_storage(1, 0, int(n), int(c), int(np), 1.0, b32_.p, np, w, np, out, n, d, n, beta, st.work,
                              st.work_bytes, s, &ev),
            "back projection");
    }
  };
  if (shift) {
    // (H - theta S) x = local H x + B^H (D - theta Q) P - theta (1 + reg) x
    if (np) {
      check(launch_column_scaled_add(st.w + 2 * np * c, np, st.w, np, np, c, shift, -1.0, s), "onsite shift");
      if (overlap) ctx_.join();
      backproject(st.w, hy, hy, 1.0);
    }
    check(launch_column_scaled_add(x, n, hy, n, n, c, shift, -(1.0 + reg_), s), "identity shift");
    return;
  }
  if (np) {
    if (want_s) backproject(st.w + 2 * np * c, sy, x, 1.0 + reg_);
    if (want_h) {
      if (overlap) ctx_.join();
      backproject(st.w, hy, hy, 1.0);
    }
  } else if (want_s) {
    check(launch_scale_copy(x, n, sy, n, n, c, 1.0 + reg_, s), "overlap identity");
  }
}

}  // namespace sdft_hip
