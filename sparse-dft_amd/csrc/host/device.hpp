// Device context and the device-resident plane-wave operator (the HIP-side
// state behind the C ABI). One Context per host thread, as the reference keeps
// one thread_local Accelerator per thread (src/compute_backend.cpp:62-73).
#pragma once
#include <hip/hip_runtime_api.h>
#include <rocblas/rocblas.h>

#include <array>
#include <atomic>
#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hip/gemm.hpp"
#include "../hip/local_h.hpp"
#include "../hip/realspace.hpp"

namespace sdft_hip {

void check(hipError_t e, const char* what);
void check(rocblas_status s, const char* what);

double gemm_algorithmic_bytes(long M, long N, long K, int a_bytes, bool reads_d);

// Pinned host staging for the small control transfers of one stream: an
// upload is copied into a fresh slot and queued without waiting, a download
// lands in one; slots are handed out again only after reset(), which the
// owner calls right after a full synchronisation of the stream.
class PinnedStaging {
 public:
  explicit PinnedStaging(std::size_t bytes);
  ~PinnedStaging();
  PinnedStaging(const PinnedStaging&) = delete;
  PinnedStaging& operator=(const PinnedStaging&) = delete;
  // a slot of `bytes`, or nullptr when the rest of the buffer is too small
  void* take(std::size_t bytes);
  void reset() { used_ = 0; }
  std::size_t capacity() const { return cap_; }

 private:
  char* base_ = nullptr;
  std::size_t cap_ = 0, used_ = 0;
};

// Grow-only device buffer (src/compute_cuda.cu:72-79 reserve semantics).
template <class T>
struct DevBuf {
  T* p = nullptr;
  std::size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  std::size_t bytes() const { return n * sizeof(T); }
  T* reserve(std::size_t count) {
    if (count > n) {
      release();
      check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)), "hipMalloc");
      n = count;
    }
    return p;
  }
  void upload(const T* host, std::size_t count, hipStream_t s) {
    reserve(count);
    if (count) check(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
};

struct PerfCounter {
  double seconds = 0;
  std::size_t calls = 0;
};

// Host description of the sphere-restricted grid plan.
struct HostGridPlan {
  std::array<int, 3> shape{};
  std::array<int, 3> split{};
  std::vector<int> stick_xy;  // nsticks
  std::vector<int> stick_g;   // nsticks * n2
  long count = 0;
  int xstart[2]{}, xlines[2]{};  // x-lines holding sticks (two ranges of natural x)
  int nsticks() const { return int(stick_xy.size()); }
};
HostGridPlan build_grid_plan(std::array<int, 3> shape, const std::uint64_t* indices, std::size_t n);
HostGridPlan build_full_plan(std::array<int, 3> shape);

// Device copy of a grid plan with twiddles (potential buffers owned separately).
struct DeviceGrid {
  GridPlanDev dev;
  DevBuf<int> stick_g, stick_xy;
  DevBuf<double> tw64[3];
  DevBuf<float> tw32[3];
  void upload(const HostGridPlan& h, hipStream_t s);
};

class Context {
 public:
  explicit Context(int device);
  ~Context();
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  // Second stream of the context (default priority) for work that overlaps the
  // main stream inside one operation: fork() makes it wait for everything
  // queued on the main stream so far, join() makes the main stream wait for
  // everything queued on it.
  hipStream_t aux_stream() const { return aux_; }
  // pinned staging of the main stream's control transfers (allocated once)
  PinnedStaging& staging();
  // bulk transfers between pageable host memory and the device on the main
  // stream through two pinned chunks: the DMA of one chunk overlaps the host
  // copy of the other. Both return with the transfer complete.
  void upload_bulk(void* device_dst, const void* host_src, std::size_t bytes);
  void download_bulk(void* host_dst, const void* device_src, std::size_t bytes);
  void fork();
  void join();
  rocblas_handle blas() const { return blas_; }
  const std::string& name() const { return name_; }
  void sync() { check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }
  double* gemm_work(std::size_t bytes) { return work_.reserve((bytes + 15) / 16 * 2); }
  std::size_t gemm_work_bytes() const { return work_.n * sizeof(double); }
  // C = alpha op(A) op(B) + beta D on device pointers, workspace managed here.
  void zgemm(Precision prec, int aop, int bop, long M, long N, long K, double alpha, const double* A, long lda,
             const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta);
  // The same with the split-K workspace supplied by the caller (work_bytes
  // >= zgemm_workspace_bytes(M, N, K) keeps the plan zgemm above would take).
  void zgemm_into(Precision prec, int aop, int bop, long M, long N, long K, double alpha, const double* A, long lda,
                  const double* B, long ldb, double* C, long ldc, const double* D, long ldd, double beta, double* work,
                  std::size_t work_bytes);
  // this device's statically allocated scratch array (kScratchBytes)
  char* scratch() const { return scratch_; }
  void record(const std::string& name, double seconds, std::size_t calls = 1);
  std::map<std::string, PerfCounter> counters;

  // Fault injection (tests of failure handling; every kernel still exits
  // normally): the next `count` events of a kind misbehave on purpose.
  //   kFaultBarrierExpiry: the one-launch tridiagonalisation finds its grid
  //     barrier expired (the abort flag is raised from the host before launch);
  //   kFaultNonfiniteApply: an operator application that produces H x writes
  //     a NaN into the first entry of that output.
  enum FaultKind { kFaultBarrierExpiry = 0, kFaultNonfiniteApply = 1, kFaultKinds = 2 };
  void inject_fault(int kind, long count);
  // consumes one pending fault of `kind`; true when one was pending
  bool take_fault(int kind);

  // Kernel clock: event pairs on this context's stream around every MFMA GEMM
  // kernel and every local-Hamiltonian pipeline launch, summed per kind.
  enum ClockKind { kGemmFP64 = 0, kGemmFP32 = 1, kLocalFP64 = 2, kLocalFP32 = 3, kGemmBF16 = 4, kClockKinds = 5 };
  struct ClockStat {
    long launches = 0;
    double milliseconds = 0, work = 0;  // work: flops (GEMM) or algorithmic bytes (local H)
    double bytes = 0;                     // algorithmic HBM bytes (operands read once, result written once)
  };
  void clock_enable(bool on);  // clears the sums either way
  // Events to record for one launch (both null when the clock is off).
  LaunchEvents clock_launch(int kind, double work, double bytes);
  ClockStat clock_stat(int kind);  // synchronises the stream and folds pending pairs

  // caches for the drop-in host-pointer entry points (identity tokens)
  struct Resident {
    std::uint64_t identity = 0;
    std::size_t rows = 0, cols = 0, ld = 0;
    DevBuf<double> values;
  } resident_left[2];
  DevBuf<double> ga, gb, gc;
  struct LocalCache {
    std::uint64_t identity = 0;
    std::array<int, 3> shape{};
    std::size_t npw = 0;
    DeviceGrid grid;
    DevBuf<double> v64, kinetic, vnat;
    DevBuf<float> v32;
    DevBuf<char> t1, t2;
    DevBuf<double> in, out;
  } local;
  struct PawCache {
    std::uint64_t identity = 0, projector_identity = 0;
    std::size_t npw = 0, np = 0;
    DevBuf<double> b, d, q, in, p, dq, h, s;
  } paw;

 private:
  int device_;
  hipStream_t stream_{};
  hipStream_t aux_{};
  hipEvent_t fork_event_{}, join_event_{};
  std::unique_ptr<PinnedStaging> staging_;
  char* bulk_[2] = {nullptr, nullptr};
  hipEvent_t bulk_done_[2] = {};
  void bulk_buffers();
  rocblas_handle blas_{};
  std::string name_;
  char* scratch_ = nullptr;
  DevBuf<double> work_;
  bool clock_on_ = false;
  std::vector<hipEvent_t> events_;
  std::size_t events_used_ = 0;
  struct PendingLaunch {
    int kind;
    double work, bytes;
    std::size_t first_event;
  };
  std::vector<PendingLaunch> pending_;
  ClockStat clock_[kClockKinds];
  std::atomic<long> faults_[kFaultKinds] = {};
};

double local_algorithmic_bytes(long npw, long count, long m, bool fp64);
struct SnapshotArrays;

struct OperatorDesc {
  std::size_t n = 0;
  const int* hkl = nullptr;            // n x 3 Miller indices (reference G order)
  const double* kinetic_ev = nullptr;  // n
  double ev_per_hartree = 27.211386245988;
  std::size_t nv = 0;
  const int* v_hkl = nullptr;          // nv x 3
  const double* v_values = nullptr;    // nv complex (re, im)
  std::size_t atoms = 0;
  const long* rows = nullptr;          // projectors per atom
  const double* amplitudes = nullptr;  // per atom row-major rows x n complex, concatenated
  const double* d = nullptr;           // per atom row-major rows x rows complex, concatenated
  const double* q = nullptr;
  double regularization = 1e-12;
  std::array<int, 3> shape{};          // <= 0: reference grid rule
};

// Device-resident restatement of PlaneWaveOperator::Impl
// (src/plane_wave_operator.cpp:29-516): index map, V(r) on the grid, kinetic
// and H/S diagonals, stacked projectors B (np x n, column-major), Hermitised
// per-atom D and Q.
class DeviceOperator {
 public:
  DeviceOperator(Context& ctx, const OperatorDesc& desc);
  // From the arrays of a solver snapshot (src/plane_wave_operator.cpp:587-653):
  // V(r), indices, kinetic and diagonals as stored; the nonlocal part in the
  // snapshot's form (PWO3 atom blocks, PWO2 dense D/Q, PWO1 folded DB/QB).
  DeviceOperator(Context& ctx, const SnapshotArrays& snapshot);
  // The operator's arrays in its snapshot form, with a Davidson request.
  SnapshotArrays snapshot(std::size_t roots, std::size_t iterations, double tolerance,
                          const std::complex<double>* initial) ;
  // which: 0 = H, 1 = S, 2 = joint pair. x, hy, sy: device pointers, ld = n.
  void apply(int which, const double* x, long columns, double* hy, double* sy, Precision prec);
  // The same application staged entirely in the context's static scratch
  // array: no device allocation. Columns go in passes of the most that fit
  // (at most the chunk apply() uses, so a block apply() takes in one pass
  // gives identical results here when it fits). Throws std::overflow_error
  // when one column does not fit, std::invalid_argument for FP32 while the
  // real-space projector search is enabled (its staging is not bounded here).
  void apply_noalloc(int which, const double* x, long columns, double* hy, double* sy, Precision prec);
  // scratch bytes apply_noalloc needs for one pass of `columns` columns
  std::size_t noalloc_bytes(int which, long columns, Precision prec) const;
  // y(:, j) = (H - theta[j] S) x(:, j): one back-projection of the shifted
  // onsite products (D - theta_j Q) P instead of one each for H and S (the
  // correction equations' operator); theta: device array of `columns` shifts.
  // bf16_projectors (FP32 only): the projection and back-projection GEMMs on
  // BF16 MFMA with FP32 accumulation, the projectors as a resident BF16 copy.
  void apply_shifted(const double* x, long columns, const double* theta, double* y, Precision prec,
                     bool bf16_projectors = false);
  // P = B X (np x columns, ld np) and QP = Q P with the onsite metric
  // (src/plane_wave_operator.cpp:375-389); device pointers.
  void project(const double* x, long columns, double* p, Precision prec);
  // Smooth density Fourier coefficients on `support` (Miller triples) from the
  // band columns x (device, n x bands) with spin-summed occupations (host),
  // restating plane_wave_density_fft (src/density_functionals.cpp:406-460) on
  // the smallest supported alias-free grid.
  std::vector<std::complex<double>> density(const double* x, long bands, const double* occupations, double volume,
                                            const int* support, long nsupport);
  void projector_metric(const double* p, long columns, double* qp);
  // Real-space projector search (src/plane_wave_operator.cpp:104-212). Per
  // atom, an atom with a radius > 0 (geometry given) keeps the grid points
  // within factor * radius of its centre (minimum image); otherwise the points
  // where the projector amplitude exceeds 10^(-2 - 5 (factor - 1)) of its
  // maximum. factor 0 disables. While enabled, FP32 (search) applications use
  // it. Returns the retained points; *retained gets the kept fraction of
  // sum_r sum_i |w_i(r)|^2. lattice: 3 rows of cell vectors (bohr);
  // fractional: atoms x 3; radius: atoms (any of them may be null).
  void set_chunk();
  long enable_real_space_search(double factor, const double* lattice = nullptr, const double* fractional = nullptr,
                                const double* radius = nullptr, double* retained = nullptr);
  bool real_space_search() const { return rs_enabled_; }
  // which 0: H, 1: S through the real-space projectors, device pointers
  void apply_real_space(int which, const double* x, long columns, double* y, Precision prec);
  // retained points per atom (natural grid indices, atom by atom)
  void real_space_points(std::vector<long>& natural, std::vector<long>& per_atom) const;
  std::size_t n() const { return n_; }
  std::size_t np() const { return np_; }
  long count() const { return count_; }
  std::array<int, 3> shape() const { return shape_; }
  double regularization() const { return reg_; }
  const std::vector<std::uint64_t>& indices() const { return indices_; }
  const std::vector<double>& kinetic() const { return kinetic_; }
  const std::vector<double>& h_diagonal() const { return h_diag_; }
  const std::vector<double>& s_diagonal() const { return s_diag_; }
  const double* kinetic_device() const { return kinetic_d_.p; }
  std::vector<double> local_real_natural();
  int nsticks() const { return grid_.dev.nsticks; }
  Context& context() { return ctx_; }
  // device buffers of the Davidson solver, kept between solves of this
  // operator (type known to davidson.cpp only)
  std::shared_ptr<void>& davidson_cache() { return davidson_cache_; }
  // device bytes held by the operator's own arrays and work buffers (the
  // Davidson buffers kept for the next solve are counted by the solver)
  std::size_t device_bytes() const;
  std::size_t max_chunk() const { return chunk_; }
  const GridPlanDev& grid_plan() const { return grid_.dev; }
  // device arrays of the nonlocal part (atom-block form 3): projectors
  // np x n column-major (FP64 and FP32 copies), Hermitised per-atom D and Q
  // blocks (column-major, at block_offset per atom), first projector and
  // projector count per atom
  int nonlocal_form() const { return form_; }
  int atoms() const { return atoms_; }
  const double* projectors64() const { return b64_.p; }
  const float* projectors32() const { return b32_.p; }
  const double* onsite_d_blocks() const { return dblk_.p; }
  const double* onsite_q_blocks() const { return qblk_.p; }
  const long* block_offsets() const { return block_offset_.p; }
  const int* atom_offsets() const { return atom_offset_.p; }
  const int* atom_row_counts() const { return atom_rows_.p; }
  const std::vector<int>& atom_rows_host() const { return atom_rows_host_; }
  // stick buffer 0 or 1 sized for `columns` columns at precision `prec`
  void* stick_buffer(int which, std::size_t columns, Precision prec);

 private:
  // Device staging of one apply_chunk pass: the two stick buffers of the
  // local-Hamiltonian pipeline, P = B X, W = (D P | Q P) and the split-K
  // partials of the pass's GEMMs.
  struct Staging {
    void* t1 = nullptr;
    void* t2 = nullptr;
    double* p = nullptr;
    double* w = nullptr;
    double* work = nullptr;
    std::size_t work_bytes = 0;
  };
  struct StagingBytes {
    std::size_t sticks = 0, p = 0, w = 0, work = 0;
    std::size_t total() const;  // with 256-byte alignment of each region
  };
  StagingBytes staging_bytes(int which, long columns, Precision prec) const;
  Staging owned_staging(int which, long columns, Precision prec);
  // shift != nullptr (which 2, sy unused): hy = (H - shift_j S) x_j
  void apply_chunk(int which, const double* x, long columns, double* hy, double* sy, Precision prec,
                   const Staging& st, const double* shift = nullptr, bool bf16_projectors = false);
  // the projectors as BF16 complex pairs (np x n), made from the FP32 copy on first use
  const unsigned* projectors_bf16();
  std::shared_ptr<void> davidson_cache_;
  Context& ctx_;
  std::size_t n_ = 0, np_ = 0, chunk_ = 256;
  long count_ = 0;
  std::array<int, 3> shape_{};
  double reg_ = 0;
  std::vector<std::uint64_t> indices_;
  std::vector<double> kinetic_, h_diag_, s_diag_;
  int form_ = 3;  // nonlocal form: 3 atom blocks, 2 dense D/Q, 1 folded DB/QB
  std::vector<int> atom_rows_host_;
  DevBuf<double> onsite_d_, onsite_q_;  // dense np x np or folded np x n (forms 2, 1)
  DeviceGrid grid_;
  DevBuf<double> v64_, kinetic_d_, b64_, dblk_, qblk_;
  DevBuf<float> v32_, b32_;
  DevBuf<unsigned> b16_;
  DevBuf<int> atom_offset_, atom_rows_;
  DevBuf<long> block_offset_;
  int atoms_ = 0;
  DevBuf<char> t1_, t2_;
  DevBuf<double> p_, w_;
  DevBuf<double> shift_work_;  // S x of a shifted application through the real-space search
  bool rs_enabled_ = false;
  RealSpaceDev rs_{};
  DevBuf<double> rs_w64_, rs_sdiag_;
  DevBuf<float> rs_w32_;
  DevBuf<long> rs_woff_, rs_poff_;
  DevBuf<int> rs_points_, rs_upoints_, rs_estart_, rs_eatom_, rs_elocal_, rs_aoff_;
  DevBuf<char> rs_planes_, rs_psi_;
  std::vector<long> rs_natural_, rs_counts_;
};

// reference index map (src/plane_wave_operator.cpp:94-97)
std::uint64_t grid_index(const std::array<int, 3>& shape, int h, int k, int l);
int fast_fft_length(int minimum);

}  // namespace sdft_hip
