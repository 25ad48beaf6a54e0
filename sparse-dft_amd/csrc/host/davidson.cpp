// Device-resident block Davidson (see davidson.hpp).
#include "davidson.hpp"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstring>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <numeric>

#include "../hip/davidson_kernels.hpp"
#include "dense_eigen.hpp"

namespace sdft_hip {
namespace {
using C = std::complex<double>;
double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// Device buffers of a solve; the operator keeps one set between solves
// (DevBuf grows on demand), so repeated solves allocate nothing.
struct Workspace {
  DevBuf<double> basis, hb, sb, ch, cs, x, sx, hx, res;
  DevBuf<double> small, w1, w2, w3, w4, m1, m2, m3, rr, rrb, qbuf, corr, retained;
  std::unique_ptr<HermitianEigensolver> eig;
  long eig_order = 0;
  bool busy = false;
};

std::size_t workspace_bytes(const Workspace& w) {
  std::size_t b = 0;
  for (const DevBuf<double>* d : {&w.basis, &w.hb, &w.sb, &w.ch, &w.cs, &w.x, &w.sx, &w.hx, &w.res, &w.small, &w.w1,
                                  &w.w2, &w.w3, &w.w4, &w.m1, &w.m2, &w.m3, &w.rr, &w.rrb, &w.qbuf, &w.corr,
                                  &w.retained})
    b += d->bytes();
  if (w.eig) b += w.eig->device_bytes();
  return b;
}

std::shared_ptr<Workspace> acquire_workspace(DeviceOperator& op) {
  std::shared_ptr<void>& slot = op.davidson_cache();
  auto ws = std::static_pointer_cast<Workspace>(slot);
  if (!ws || ws->busy) {
    // a solve already holds the operator's set: this one gets its own
    ws = std::make_shared<Workspace>();
    if (!slot) slot = ws;
  }
  ws->busy = true;
  return ws;
}

class Solver {
 public:
  Solver(DeviceOperator& op, const DavidsonParams& p)
      : op_(op), ctx_(op.context()), s_(ctx_.stream()), p_(p), n_(long(op.n())), R_(p.roots), tol_(p.tolerance),
        ws_(acquire_workspace(op)), basis_(ws_->basis), hb_(ws_->hb), sb_(ws_->sb), ch_(ws_->ch), cs_(ws_->cs),
        x_(ws_->x), sx_(ws_->sx), hx_(ws_->hx), res_(ws_->res), small_(ws_->small), w1_(ws_->w1), w2_(ws_->w2),
        w3_(ws_->w3), w4_(ws_->w4), m1_(ws_->m1), m2_(ws_->m2), m3_(ws_->m3), rr_(ws_->rr), rrb_(ws_->rrb),
        qbuf_(ws_->qbuf), corr_(ws_->corr), retained_(ws_->retained) {
    if (!n_ || R_ <= 0 || R_ > n_ || p.max_iterations <= 0 || !std::isfinite(tol_) || tol_ <= 0)
      throw std::invalid_argument("HIP: invalid matrix-free Davidson problem");
    if (p.subspace_factor < 4 || p.subspace_factor > 16)
      throw std::invalid_argument("HIP: Davidson subspace factor must be in [4,16]");
    cap_ = std::min<long>(n_, std::max<long>(p.subspace_factor * R_, 64));
    if (p.capacity) {
      // src/pw_core.cpp:733-734
      if (p.capacity < std::min<long>(n_, 2 * R_))
        throw std::invalid_argument("HIP: Davidson local subspace must hold roots and guard states");
      cap_ = std::min<long>(n_, p.capacity);
    }
    exact_ = !p.search_fp32;
    basis_.reserve(2 * n_ * cap_);
    hb_.reserve(2 * n_ * cap_);
    sb_.reserve(2 * n_ * cap_);
    ch_.reserve(2 * cap_ * cap_);
    cs_.reserve(2 * cap_ * cap_);
    x_.reserve(2 * n_ * R_);
    sx_.reserve(2 * n_ * R_);
    hx_.reserve(2 * n_ * R_);
    res_.reserve(2 * n_ * R_);
    // All work buffers are sized once here so that no buffer is reallocated
    // while kernels that use it are queued.
    small_.reserve(std::size_t(std::max(16 * cap_, 64 * R_)) + 1024);
    w1_.reserve(2 * n_ * R_);
    w2_.reserve(2 * n_ * R_);
    w3_.reserve(2 * n_ * cap_);
    w4_.reserve(2 * n_ * cap_);
    m1_.reserve(8 * R_ * R_ + 2 * cap_ + 64);
    m2_.reserve(2 * cap_ * cap_);
    m3_.reserve(4 * cap_ * cap_);
    rr_.reserve(2 * cap_ * cap_ + 16);
    rrb_.reserve(2 * cap_ * cap_ + 16);
    qbuf_.reserve(2 * cap_ * cap_ + 16);
    corr_.reserve(2 * n_ * R_);
    retained_.reserve(2 * n_ * cap_);
    if (!ws_->eig || ws_->eig_order < cap_) {
      ws_->eig = std::make_unique<HermitianEigensolver>(ctx_, cap_);
      ws_->eig_order = cap_;
    }
    eig_ = ws_->eig.get();
    stedc_base_ = eig_->stedc_solves();
    barrier_base_ = eig_->barrier_fallbacks();
    if (p.excluded_count < 0 || (p.excluded_count && !p.excluded))
      throw std::invalid_argument("invalid completed-band dimensions");
    excl_ = p.excluded_count;
    if (excl_) setup_excluded(p.excluded);
  }
  ~Solver() {
    // queued work may still use the buffers; the next solve on this stream
    // is ordered after it
    ws_->busy = false;
  }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;
  DavidsonOutput run(const C* initial);
  // One strict step on a given block, from a fresh FP64 certificate of it (the
  // reference's strict_residual_polish_step / strict_root_manifold_step
  // primitives, src/strict_residual_polish.hpp:68-197, without locked bands).
  StrictStepOutput strict_primitive(int kind, const C* x, const C* direction);

 private:
  Precision search_prec() const { return exact_ ? Precision::FP64 : Precision::FP32; }
  // guard Ritz vectors of the refinement block (R/2 unless set)
  long refinement_guards() const {
    return p_.refinement_guards > 0 ? std::min(p_.refinement_guards, R_) : std::max(1L, R_ / 2);
  }
  Precision prec(bool exact) const { return exact ? Precision::FP64 : search_prec(); }
  double* col(DevBuf<double>& b, long j, long ld) { return b.p + 2 * j * ld; }
  void sync() {
    ctx_.sync();
    stage_.reset();
  }
  void gemm(bool exact, int aop, int bop, long M, long N, long K, double alpha, const double* A, long lda,
            const double* B, long ldb, double* Cm, long ldc, const double* D = nullptr, long ldd = 0,
            double beta = 0.0) {
    const double t0 = p_.timing ? (sync(), now()) : 0;
    ctx_.zgemm(prec(exact), aop, bop, M, N, K, alpha, A, lda, B, ldb, Cm, ldc, D, ldd, beta);
    if (p_.timing) { sync(); stats_.seconds_gemm += now() - t0; }
  }
  void apply(int which, const double* in, long cols, double* hy, double* sy, bool exact) {
    const double t0 = p_.timing ? (sync(), now()) : 0;
    const Precision pr = prec(exact);
    op_.apply(which, in, cols, hy, sy, pr);
    if (which != 1) (pr == Precision::FP64 ? stats_.h_columns_exact : stats_.h_columns_search) += cols;
    if (p_.timing) { sync(); stats_.seconds_apply += now() - t0; }
  }
  void copy(double* dst, const double* src, long complex_count) {
    check(hipMemcpyAsync(dst, src, 16 * std::size_t(complex_count), hipMemcpyDeviceToDevice, s_), "copy");
  }
  std::vector<double> to_host(const double* d, long count) {
    std::vector<double> h(count);
    const std::size_t bytes = 8 * std::size_t(count);
    void* slot = stage_.take(bytes);
    if (!slot && bytes <= stage_.capacity()) {
      sync();
      slot = stage_.take(bytes);
    }
    check(hipMemcpyAsync(slot ? slot : h.data(), d, bytes, hipMemcpyDeviceToHost, s_), "to host");
    sync();
    if (slot) std::memcpy(h.data(), slot, bytes);
    return h;
  }
  // small host -> device transfers of control scalars through pinned slots:
  // the host source may be reused immediately and nothing waits
  void upload_bytes(void* d, const void* h, std::size_t bytes) {
    void* slot = stage_.take(bytes);
    if (!slot && bytes <= stage_.capacity()) {
      sync();
      slot = stage_.take(bytes);
    }
    if (!slot) {
      check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s_), "upload");
      sync();
      return;
    }
    std::memcpy(slot, h, bytes);
    check(hipMemcpyAsync(d, slot, bytes, hipMemcpyHostToDevice, s_), "upload");
  }
  void upload(double* d, const double* h, long count) { upload_bytes(d, h, 8 * std::size_t(count)); }
  std::vector<double> col_norms(const double* x, long ld, long cols) {
    double* o = small_.p;
    check(launch_col_norms(x, ld, n_, cols, o, s_), "norms");
    return to_host(o, cols);
  }
  std::vector<double> real_dots(const double* a, const double* b, long cols) {
    double* o = small_.p;
    check(launch_real_dots(a, n_, b, n_, n_, cols, o, s_), "dots");
    return to_host(o, cols);
  }
  void check_finite(const double* x, long cols, const char* what) {
    int* flag = reinterpret_cast<int*>(small_.p);
    check(hipMemsetAsync(flag, 0, sizeof(int), s_), "memset");
    check(launch_finite_check(x, n_, n_, cols, flag, s_), "finite");
    int bad = 0;
    check(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s_), "finite copy");
    sync();
    if (bad) throw std::runtime_error(std::string("HIP: nonfinite ") + what);
  }
  // Hermitian eigen-decomposition of a device matrix (overwritten by vectors)
  std::vector<double> heevd(double* a, long m, long lda) {
    const double t0 = p_.timing ? (sync(), now()) : 0;
    std::vector<double> w = heevd_device(a, m, lda);
    if (p_.timing) stats_.seconds_ritz += now() - t0;
    return w;
  }
  std::vector<double> heevd_device(double* a, long m, long lda) {
    double* w = small_.p + 8;
    double* e = w + m;
    int* info = reinterpret_cast<int*>(small_.p);
    check(rocsolver_zheevd(ctx_.blas(), rocblas_evect_original, rocblas_fill_lower, rocblas_int(m),
                           reinterpret_cast<rocblas_double_complex*>(a), rocblas_int(lda), w, e, info),
          "rocsolver_zheevd");
    std::vector<double> h = to_host(small_.p, 8 + m);
    if (*reinterpret_cast<int*>(h.data())) throw std::runtime_error("HIP: Hermitian eigensolve failed");
    return std::vector<double>(h.begin() + 8, h.begin() + 8 + m);
  }
  // max_ij |B_ij - delta_ij| of an m x m device block; E = B - I into e if given
  double identity_defect(const double* b, long m, long ld, double* e) {
    auto* bits = reinterpret_cast<unsigned long long*>(small_.p);
    check(hipMemsetAsync(bits, 0, sizeof(unsigned long long), s_), "memset");
    check(launch_identity_defect(b, ld, m, e, m, bits, s_), "identity defect");
    return to_host(small_.p, 1)[0];  // the bits of a non-negative double (or NaN)
  }
  // Rayleigh-Ritz of the projected pencil (A, B) of the S-orthonormal basis,
  // B = I + E: congruence with M = I - E/2 + 3E^2/8, so that M B M = I + O(E^3),
  // then the Hermitian eigensolve of M A M and vectors M Y. This solves the
  // same pencil as zhegvd (src/pw_core.cpp:104-122) to O(|E|^3); a basis whose
  // metric is further from the identity than 1e-4 takes the Cholesky reduction.
  // A returns the first `vectors` B-orthonormal eigenvectors, ascending.
  std::vector<double> rr_solve(double* a, double* b, long m, long vectors) {
    const double t0 = p_.timing ? (sync(), now()) : 0;
    double* e = m3_.p;
    double* e2 = m3_.p + 2 * m * m;
    double* mm = qbuf_.p;
    double* t = m2_.p;
    // max |B - I| and, as a finiteness check of the projected Hamiltonian,
    // max |A - I| (a NaN dominates both maxima), read in one transfer
    auto* bits = reinterpret_cast<unsigned long long*>(small_.p);
    check(hipMemsetAsync(bits, 0, 2 * sizeof(unsigned long long), s_), "memset");
    check(launch_identity_defect(b, m, m, e, m, bits, s_), "identity defect");
    check(launch_identity_defect(a, m, m, nullptr, m, bits + 1, s_), "finite check");
    const std::vector<double> defects = to_host(small_.p, 2);
    const double defect = defects[0];
    if (!std::isfinite(defect) || !std::isfinite(defects[1]))
      throw std::runtime_error("HIP: nonfinite Davidson projected matrices (an operator application was not finite)");
    if (!(defect <= 1e-4)) {
      ++stats_.rr_cholesky;
      if (p_.timing) stats_.seconds_ritz += now() - t0;
      return hegvd(a, b, m, m);
    }
    const Precision f64 = Precision::FP64;
    ctx_.zgemm(f64, 0, 0, m, m, m, 1.0, e, m, e, m, e2, m, nullptr, 0, 0.0);
    check(launch_series_metric(e, e2, m, m, mm, m, s_), "series metric");
    ctx_.zgemm(f64, 0, 0, m, m, m, 1.0, mm, m, a, m, t, m, nullptr, 0, 0.0);
    ctx_.zgemm(f64, 0, 0, m, m, m, 1.0, t, m, mm, m, e2, m, nullptr, 0, 0.0);
    check(launch_hermitise(e2, m, m, e, m, s_), "hermitise");
    const long k = std::min(vectors, m);
    std::vector<double> w = eig_->solve(e, m, k, t, k);
    ctx_.zgemm(f64, 0, 0, m, k, m, 1.0, mm, m, t, m, a, m, nullptr, 0, 0.0);
    if (p_.timing) { sync(); stats_.seconds_ritz += now() - t0; }
    return w;
  }
  // Generalized A v = l B v (itype 1), A overwritten by B-orthonormal vectors:
  // B = L L^H (potrf), C = L^{-1} A L^{-H} (two BLAS3 triangular solves),
  // C = Y diag(l) Y^H (divide and conquer), v = L^{-H} Y. The same reduction as
  // LAPACK zhegv/zhegvd, with the two-sided solve done by ztrsm.
  std::vector<double> hegvd(double* a, double* b, long m, long ld) {
    double* w = small_.p + 8;
    double* e = w + m;
    int* info = reinterpret_cast<int*>(small_.p);
    const double t0 = p_.timing ? (sync(), now()) : 0;
    auto* A = reinterpret_cast<rocblas_double_complex*>(a);
    auto* B = reinterpret_cast<rocblas_double_complex*>(b);
    const rocblas_double_complex one{1.0, 0.0};
    check(rocsolver_zpotrf(ctx_.blas(), rocblas_fill_lower, rocblas_int(m), B, rocblas_int(ld), info),
          "rocsolver_zpotrf");
    int pinfo = 0;
    check(hipMemcpyAsync(&pinfo, info, sizeof(int), hipMemcpyDeviceToHost, s_), "info copy");
    sync();
    if (pinfo) throw std::runtime_error("HIP: Davidson Rayleigh-Ritz metric is not positive definite");
    check(rocblas_ztrsm(ctx_.blas(), rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                        rocblas_diagonal_non_unit, rocblas_int(m), rocblas_int(m), &one, B, rocblas_int(ld), A,
                        rocblas_int(ld)),
          "ztrsm");
    check(rocblas_ztrsm(ctx_.blas(), rocblas_side_right, rocblas_fill_lower, rocblas_operation_conjugate_transpose,
                        rocblas_diagonal_non_unit, rocblas_int(m), rocblas_int(m), &one, B, rocblas_int(ld), A,
                        rocblas_int(ld)),
          "ztrsm");
    check(rocsolver_zheevd(ctx_.blas(), rocblas_evect_original, rocblas_fill_lower, rocblas_int(m), A,
                           rocblas_int(ld), w, e, info),
          "rocsolver_zheevd");
    check(rocblas_ztrsm(ctx_.blas(), rocblas_side_left, rocblas_fill_lower, rocblas_operation_conjugate_transpose,
                        rocblas_diagonal_non_unit, rocblas_int(m), rocblas_int(m), &one, B, rocblas_int(ld), A,
                        rocblas_int(ld)),
          "ztrsm");
    std::vector<double> h = to_host(small_.p, 8 + m);
    if (p_.timing) stats_.seconds_ritz += now() - t0;
    if (*reinterpret_cast<int*>(h.data())) throw std::runtime_error("HIP: Davidson Rayleigh-Ritz failed");
    return std::vector<double>(h.begin() + 8, h.begin() + 8 + m);
  }

  bool append(const double* v_in);
  bool append_block(double* v, long cols);
  // completed bands E (S-orthonormal) and their S images; v -= E (SE)^H v
  void setup_excluded(const double* e);
  void remove_excluded(double* v, long cols);
  bool cholesky_normalise(const double* v, const double* metric, const double* gram, double* gh, long cols,
                          bool exact);
  void update_projected(const double* applied, double* cached);
  void flush_h();
  void certify();
  void refresh_norms() {
    res_norms_ = col_norms(res_.p, n_, R_);
    converged_ = true;
    for (long r = 0; r < R_; ++r) converged_ = converged_ && res_norms_[r] <= tol_;
  }
  double max_res() const { return *std::max_element(res_norms_.begin(), res_norms_.end()); }
  void rebuild_exact() {  // fresh FP64 pair images of the whole basis
    apply(2, basis_.p, size_, hb_.p, sb_.p, true);
    computed_h_ = size_;
    projected_size_ = 0;
  }
  bool polish_step(double* direction);
  bool refine(long& it, long& iterations);
  void correction_cg(const std::vector<int>& active, double* t);
  bool manifold_step();
  void preserved_subspace(long keep, double* retained);

  DeviceOperator& op_;
  Context& ctx_;
  hipStream_t s_;
  DavidsonParams p_;
  long n_, R_, cap_ = 0;
  double tol_;
  bool exact_ = true, incremental_ = true, converged_ = false;
  std::shared_ptr<Workspace> ws_;
  DevBuf<double>&basis_, &hb_, &sb_, &ch_, &cs_, &x_, &sx_, &hx_, &res_;
  DevBuf<double>&small_, &w1_, &w2_, &w3_, &w4_, &m1_, &m2_, &m3_, &rr_, &rrb_, &qbuf_, &corr_, &retained_;
  HermitianEigensolver* eig_ = nullptr;
  long stedc_base_ = 0;  // divide-and-conquer solves of the shared eigensolver before this solve
  long barrier_base_ = 0;  // its barrier-expiry fallbacks before this solve
  PinnedStaging& stage_ = ctx_.staging();
  long excl_ = 0;
  const double* excluded_ = nullptr;
  DevBuf<double> excluded_s_, excl_coef_;
  long size_ = 0, computed_h_ = 0, projected_size_ = 0;
  double search_residual_ = -1;  // max residual of the search arithmetic at the hand-off, until confirmed
  std::vector<double> values_, res_norms_;
  DavidsonStats stats_;
};

// src/pw_core.cpp:796-853: serial CGS2 of one direction in exact FP64
bool Solver::append(const double* v_in) {
  if (size_ >= cap_) return false;
  double* v = w1_.p;
  double* sv = w2_.p;
  copy(v, v_in, n_);
  const double en = col_norms(v, n_, 1)[0];
  if (!std::isfinite(en)) throw std::runtime_error("HIP: nonfinite Davidson direction");
  if (en == 0) return false;
  upload(small_.p, &en, 1);
  check(launch_divide_columns(v, n_, n_, 1, small_.p, s_), "scale");
  apply(1, v, 1, nullptr, sv, true);
  for (int pass = 0; pass < 2; ++pass) {
    if (excl_) remove_excluded(v, 1);  // src/pw_core.cpp:810
    if (pass || excl_) apply(1, v, 1, nullptr, sv, true);
    if (size_ < 8) {
      for (long j = 0; j < size_; ++j) {
        double* proj = m1_.p;
        gemm(true, 1, 0, 1, 1, n_, 1.0, col(basis_, j, n_), n_, sv, n_, proj, 1);
        gemm(true, 0, 0, n_, 1, 1, -1.0, col(basis_, j, n_), n_, proj, 1, v, n_, v, n_, 1.0);
        gemm(true, 0, 0, n_, 1, 1, -1.0, col(sb_, j, n_), n_, proj, 1, sv, n_, sv, n_, 1.0);
      }
    } else {
      double* proj = m1_.p;
      gemm(true, 1, 0, size_, 1, n_, 1.0, basis_.p, n_, sv, n_, proj, size_);
      gemm(true, 0, 0, n_, 1, size_, -1.0, basis_.p, n_, proj, size_, v, n_, v, n_, 1.0);
    }
  }
  apply(1, v, 1, nullptr, sv, true);
  double* nn = m1_.p;
  gemm(true, 1, 0, 1, 1, n_, 1.0, v, n_, sv, n_, nn, 1);
  const double norm2 = to_host(nn, 1)[0];
  if (!std::isfinite(norm2) || norm2 < -1e-10) throw std::runtime_error("HIP: Davidson encountered non-positive overlap direction");
  if (norm2 < 1e-24) return false;
  const double r = std::sqrt(norm2);
  upload(small_.p, &r, 1);
  check(launch_divide_columns(v, n_, n_, 1, small_.p, s_), "scale");
  check(launch_divide_columns(sv, n_, n_, 1, small_.p, s_), "scale");
  copy(col(basis_, size_, n_), v, n_);
  copy(col(sb_, size_, n_), sv, n_);
  ++size_;
  return true;
}

// src/pw_core.cpp:871-1013: block CGS2 in the S metric with eigen-normalisation.
// `v` (n x cols, ld n) is a scratch block owned by the caller.
bool Solver::append_block(double* v, long cols) {
  const long before = size_;
  cols = std::min(cols, cap_ - size_);
  if (cols <= 0) return false;
  auto serial = [&]() {
    for (long j = 0; j < cols; ++j) append(v + 2 * j * n_);
    return size_ > before;
  };
  if (cols == 1) return serial();
  const double t0 = p_.timing ? (sync(), now()) : 0;
  const bool exact = exact_;  // exact_orthogonalize: search arithmetic before the hand-off
  const std::vector<double> len = col_norms(v, n_, cols);
  for (double l : len)
    if (!(l > 0) || !std::isfinite(l)) return serial();
  upload(small_.p + 8, len.data(), cols);
  check(launch_divide_columns(v, n_, n_, cols, small_.p + 8, s_), "normalise");
  double* original = w3_.p;
  double* metric = w4_.p;
  copy(original, v, n_ * cols);
  apply(1, v, cols, nullptr, metric, exact);
  for (int pass = 0; (size_ || excl_) && pass < 2; ++pass) {
    if (size_) {
      double* proj = m2_.p;
      gemm(exact, 1, 0, size_, cols, n_, 1.0, basis_.p, n_, metric, n_, proj, size_);
      gemm(exact, 0, 0, n_, cols, size_, -1.0, basis_.p, n_, proj, size_, v, n_, v, n_, 1.0);
    }
    remove_excluded(v, cols);  // src/pw_core.cpp:909-922
    apply(1, v, cols, nullptr, metric, exact);
  }
  double* gram = m3_.p;
  double* gh = m3_.p + 2 * cols * cols;
  gemm(exact, 1, 0, cols, cols, n_, 1.0, v, n_, metric, n_, gram, cols);
  check(launch_hermitise(gram, cols, cols, gh, cols, s_), "hermitise");
  if (cholesky_normalise(v, metric, gram, gh, cols, exact)) {
    size_ += cols;
    if (p_.timing) { sync(); stats_.seconds_ortho += now() - t0; }
    return true;
  }
  check(launch_hermitise(gram, cols, cols, gh, cols, s_), "hermitise");
  const std::vector<double> lam = heevd(gh, cols, cols);
  const double largest = *std::max_element(lam.begin(), lam.end());
  const double smallest = *std::min_element(lam.begin(), lam.end());
  if (smallest < -1e-10) throw std::runtime_error("HIP: block Davidson encountered non-positive overlap");
  if (!(largest > 0) || smallest < 1e-10 * largest) {
    copy(v, original, n_ * cols);
    if (p_.timing) stats_.seconds_ortho += now() - t0;
    return serial();
  }
  upload(small_.p + 8, lam.data(), cols);
  check(launch_inv_sqrt_columns(gh, cols, cols, cols, small_.p + 8, s_), "normalisation");
  gemm(exact, 0, 0, n_, cols, cols, 1.0, v, n_, gh, cols, col(basis_, size_, n_), n_);
  gemm(exact, 0, 0, n_, cols, cols, 1.0, metric, n_, gh, cols, col(sb_, size_, n_), n_);
  size_ += cols;
  if (p_.timing) { sync(); stats_.seconds_ortho += now() - t0; }
  return true;
}

// src/pw_core.cpp:675-711: completed bands of a band-group local solve are
// checked S-orthonormal and projected out of every new search direction.
void Solver::setup_excluded(const double* e) {
  excluded_ = e;
  excluded_s_.reserve(std::size_t(2 * n_ * excl_));
  excl_coef_.reserve(std::size_t(2 * excl_ * std::max(cap_, R_)));
  apply(1, e, excl_, nullptr, excluded_s_.p, true);
  DevBuf<double> gram;
  gram.reserve(std::size_t(2 * excl_ * excl_));
  gemm(true, 1, 0, excl_, excl_, n_, 1.0, e, n_, excluded_s_.p, n_, gram.p, excl_);
  if (!(identity_defect(gram.p, excl_, excl_, nullptr) <= 1e-10))
    throw std::invalid_argument("completed bands must be S orthonormal");
}

void Solver::remove_excluded(double* v, long cols) {
  if (!excl_ || cols <= 0) return;
  gemm(true, 1, 0, excl_, cols, n_, 1.0, excluded_s_.p, n_, v, n_, excl_coef_.p, excl_);
  gemm(true, 0, 0, n_, cols, excl_, -1.0, excluded_, n_, excl_coef_.p, excl_, v, n_, v, n_, 1.0);
}

// Cholesky form of the block normalisation: G = U^H U, W = V U^{-1} spans the
// same space as the eigen-normalised block V U diag(l)^(-1/2) of the reference
// (src/pw_core.cpp:960-1000), so the Galerkin pencil is unchanged. It is taken
// only when the Cholesky pivots bound cond(G) by 1e8 and the result is
// S-orthonormal to 1e-10 (FP64) or 1e-4 (FP32 search arithmetic); otherwise the
// caller runs the reference's eigen-normalisation and its dependency test.
// On success W and S W are written after the current basis.
bool Solver::cholesky_normalise(const double* v, const double* metric, const double* gram, double* gh, long cols,
                                bool exact) {
  int* info = reinterpret_cast<int*>(small_.p);
  auto* U = reinterpret_cast<rocblas_double_complex*>(gh);
  // everything is queued first and checked after one synchronisation: a
  // failed factorisation only leaves garbage in columns that are not kept
  check(rocsolver_zpotrf(ctx_.blas(), rocblas_fill_upper, rocblas_int(cols), U, rocblas_int(cols), info),
        "rocsolver_zpotrf");
  double* diag = small_.p + 8;
  check(launch_diagonals(gram, cols, gh, cols, cols, diag, s_), "diagonals");
  int* trtri_info = reinterpret_cast<int*>(small_.p + 1);
  check(rocsolver_ztrtri(ctx_.blas(), rocblas_fill_upper, rocblas_diagonal_non_unit, rocblas_int(cols), U,
                         rocblas_int(cols), trtri_info),
        "rocsolver_ztrtri");
  check(launch_zero_lower(gh, cols, cols, s_), "zero lower");
  double* w = col(basis_, size_, n_);
  double* sw = col(sb_, size_, n_);
  gemm(exact, 0, 0, n_, cols, cols, 1.0, v, n_, gh, cols, w, n_);
  gemm(exact, 0, 0, n_, cols, cols, 1.0, metric, n_, gh, cols, sw, n_);
  double* g2 = m2_.p;
  gemm(exact, 1, 0, cols, cols, n_, 1.0, w, n_, sw, n_, g2, cols);
  auto* bits = reinterpret_cast<unsigned long long*>(small_.p + 2);
  check(hipMemsetAsync(bits, 0, sizeof(unsigned long long), s_), "memset");
  check(launch_identity_defect(g2, cols, cols, nullptr, cols, bits, s_), "identity defect");
  const std::vector<double> h = to_host(small_.p, 8 + 2 * cols);
  double gmax = 0, umin = std::numeric_limits<double>::infinity();
  for (long i = 0; i < cols; ++i) {
    gmax = std::max(gmax, h[8 + i]);
    umin = std::min(umin, h[8 + cols + i] * h[8 + cols + i]);
  }
  const double defect = h[2];  // the bits of a non-negative double (or NaN)
  if (p_.verbose > 1)
    std::fprintf(stderr, "  cholesky block of %ld: potrf info %d, trtri info %d, min pivot^2 / max diagonal %.3e, "
                 "defect %.3e\n", cols, *reinterpret_cast<const int*>(h.data()),
                 *reinterpret_cast<const int*>(h.data() + 1), umin / gmax, defect);
  if (*reinterpret_cast<const int*>(h.data())) return false;
  if (!(gmax > 0) || !(umin >= 1e-8 * gmax) || !std::isfinite(gmax)) return false;
  if (*reinterpret_cast<const int*>(h.data() + 1)) return false;
  if (!(defect <= (exact ? 1e-10 : 1e-4))) return false;
  ++stats_.cholesky_blocks;
  return true;
}

// src/pw_core.cpp:771-785
void Solver::update_projected(const double* applied, double* cached) {
  const long first = incremental_ ? projected_size_ : 0;
  if (first == size_) return;
  const long cols = size_ - first;
  gemm(exact_, 1, 0, size_, cols, n_, 1.0, basis_.p, n_, applied + 2 * first * n_, n_, cached + 2 * first * cap_,
       cap_);
  if (first)
    gemm(exact_, 1, 0, cols, first, n_, 1.0, basis_.p + 2 * first * n_, n_, applied, n_, cached + 2 * first, cap_);
}

void Solver::flush_h() {
  if (computed_h_ >= size_) return;
  apply(0, col(basis_, computed_h_, n_), size_ - computed_h_, col(hb_, computed_h_, n_), nullptr, exact_);
  computed_h_ = size_;
}

// src/pw_core.cpp:1125-1149 (certify_once = false)
void Solver::certify() {
  apply(1, x_.p, R_, nullptr, sx_.p, true);
  const std::vector<double> norm2 = real_dots(x_.p, sx_.p, R_);
  std::vector<double> root(R_);
  for (long r = 0; r < R_; ++r) {
    if (!std::isfinite(norm2[r]) || !(norm2[r] > 0)) throw std::runtime_error("HIP: invalid Davidson Ritz S norm");
    root[r] = std::sqrt(norm2[r]);
  }
  upload(small_.p + 8, root.data(), R_);
  check(launch_divide_columns(x_.p, n_, n_, R_, small_.p + 8, s_), "normalise");
  apply(2, x_.p, R_, hx_.p, sx_.p, true);
  const std::vector<double> metric = real_dots(x_.p, sx_.p, R_);
  const std::vector<double> energy = real_dots(x_.p, hx_.p, R_);
  for (long r = 0; r < R_; ++r) {
    if (!(metric[r] > 0)) throw std::runtime_error("HIP: invalid Davidson Rayleigh metric");
    values_[r] = energy[r] / metric[r];
  }
  upload(small_.p + 8, values_.data(), R_);
  check(launch_fma_residual(hx_.p, sx_.p, n_, n_, R_, small_.p + 8, res_.p, s_), "residual");
  refresh_norms();
}

// src/strict_residual_polish.hpp:125-197 (no locked bands)
bool Solver::polish_step(double* direction) {
  const double old_max = max_res();
  if (old_max <= tol_) return false;
  for (int pass = 0; pass < 2; ++pass) {
    gemm(true, 1, 0, R_, R_, n_, 1.0, sx_.p, n_, direction, n_, m1_.p, R_);
    gemm(true, 0, 0, n_, R_, R_, -1.0, x_.p, n_, m1_.p, R_, direction, n_, direction, n_, 1.0);
  }
  apply(2, direction, R_, w1_.p, w2_.p, true);  // dh, ds
  const std::vector<double> grad = real_dots(direction, res_.p, R_);
  const std::vector<double> dhd = real_dots(direction, w1_.p, R_);
  const std::vector<double> dsd = real_dots(direction, w2_.p, R_);
  const std::vector<double> dn = col_norms(direction, n_, R_);
  std::vector<double> alpha(R_, 0.0);
  bool moved = false;
  for (long r = 0; r < R_; ++r) {
    if (res_norms_[r] <= .5 * tol_) continue;
    const double curvature = dhd[r] - values_[r] * dsd[r];
    if (!(curvature > 0.) || !std::isfinite(curvature) || !std::isfinite(grad[r])) continue;
    const double a = -grad[r] / curvature;
    if (!(a > 0.) || !std::isfinite(a) || a * dn[r] > 1e-4) continue;
    alpha[r] = a;
    moved = true;
  }
  if (!moved) return false;
  double* trial = w3_.p;
  upload(small_.p + 8, alpha.data(), R_);
  check(launch_polish_update(x_.p, direction, n_, n_, R_, small_.p + 8, trial, s_), "polish update");
  apply(2, trial, R_, w1_.p, w2_.p, true);
  const std::vector<double> metric = real_dots(trial, w2_.p, R_);
  std::vector<double> root(R_);
  for (long r = 0; r < R_; ++r) {
    if (!(metric[r] > 0.) || !std::isfinite(metric[r])) return false;
    root[r] = std::sqrt(metric[r]);
  }
  upload(small_.p + 8, root.data(), R_);
  check(launch_divide_columns(trial, n_, n_, R_, small_.p + 8, s_), "normalise");
  apply(2, trial, R_, w1_.p, w2_.p, true);
  const std::vector<double> m2 = real_dots(trial, w2_.p, R_);
  const std::vector<double> e2 = real_dots(trial, w1_.p, R_);
  std::vector<double> energies(R_);
  for (long r = 0; r < R_; ++r) energies[r] = e2[r] / m2[r];
  upload(small_.p + 8, energies.data(), R_);
  double* errors = w4_.p;
  check(launch_fma_residual(w1_.p, w2_.p, n_, n_, R_, small_.p + 8, errors, s_), "errors");
  const std::vector<double> en = col_norms(errors, n_, R_);
  const double new_max = *std::max_element(en.begin(), en.end());
  gemm(true, 1, 0, R_, R_, n_, 1.0, trial, n_, w2_.p, n_, m1_.p, R_);
  const std::vector<double> g = to_host(m1_.p, 2 * R_ * R_);
  double gerr = 0;
  for (long j = 0; j < R_; ++j)
    for (long i = 0; i < R_; ++i)
      gerr = std::max(gerr, std::abs(C(g[2 * (i + j * R_)], g[2 * (i + j * R_) + 1]) - (i == j ? C(1) : C(0))));
  for (double e : en)
    if (!std::isfinite(e)) return false;
  if (!(gerr <= 1e-12) || !(new_max <= tol_ || new_max < .999 * old_max)) return false;
  copy(x_.p, trial, n_ * R_);
  copy(sx_.p, w2_.p, n_ * R_);
  copy(res_.p, errors, n_ * R_);
  values_ = energies;
  return true;
}

// src/strict_residual_polish.hpp:68-123 (no locked bands)
bool Solver::manifold_step() {
  const double old_max = max_res();
  if (old_max <= tol_) return false;
  apply(2, x_.p, R_, w1_.p, w2_.p, true);
  double* h = m1_.p;
  double* s = m1_.p + 2 * R_ * R_;
  double* hh = m1_.p + 4 * R_ * R_;
  double* sh = m1_.p + 6 * R_ * R_;
  check(launch_complex_dots(x_.p, n_, R_, w1_.p, n_, R_, n_, h, R_, s_), "manifold h");
  check(launch_complex_dots(x_.p, n_, R_, w2_.p, n_, R_, n_, s, R_, s_), "manifold s");
  check(launch_hermitise(h, R_, R_, hh, R_, s_), "hermitise");
  check(launch_hermitise(s, R_, R_, sh, R_, s_), "hermitise");
  hegvd(hh, sh, R_, R_);
  double* trial = w3_.p;
  gemm(true, 0, 0, n_, R_, R_, 1.0, x_.p, n_, hh, R_, trial, n_);
  apply(2, trial, R_, w1_.p, w2_.p, true);
  const std::vector<double> norm = real_dots(trial, w2_.p, R_);
  std::vector<double> root(R_);
  for (long r = 0; r < R_; ++r) {
    if (!std::isfinite(norm[r]) || norm[r] <= 0) return false;
    root[r] = std::sqrt(norm[r]);
  }
  upload(small_.p + 8, root.data(), R_);
  check(launch_divide_columns(trial, n_, n_, R_, small_.p + 8, s_), "normalise");
  apply(2, trial, R_, w1_.p, w2_.p, true);
  const std::vector<double> e1 = real_dots(trial, w1_.p, R_);
  const std::vector<double> m = real_dots(trial, w2_.p, R_);
  std::vector<double> energies(R_);
  for (long r = 0; r < R_; ++r) energies[r] = e1[r] / m[r];
  upload(small_.p + 8, energies.data(), R_);
  double* errors = w4_.p;
  check(launch_fma_residual(w1_.p, w2_.p, n_, n_, R_, small_.p + 8, errors, s_), "errors");
  const std::vector<double> en = col_norms(errors, n_, R_);
  const double new_max = *std::max_element(en.begin(), en.end());
  gemm(true, 1, 0, R_, R_, n_, 1.0, trial, n_, w2_.p, n_, h, R_);
  const std::vector<double> g = to_host(h, 2 * R_ * R_);
  double gerr = 0;
  for (long j = 0; j < R_; ++j)
    for (long i = 0; i < R_; ++i)
      gerr = std::max(gerr, std::abs(C(g[2 * (i + j * R_)], g[2 * (i + j * R_) + 1]) - (i == j ? C(1) : C(0))));
  if (!std::isfinite(new_max) || gerr > 1e-12 || !(new_max <= tol_ || new_max < .999 * old_max)) return false;
  copy(x_.p, trial, n_ * R_);
  copy(sx_.p, w2_.p, n_ * R_);
  copy(res_.p, errors, n_ * R_);
  values_ = energies;
  return true;
}

// src/strict_residual_polish.hpp:47-63: [x, B Q_perp] with Q from a Householder
// QR of B^H sx (size x roots).
void Solver::preserved_subspace(long keep, double* retained) {
  copy(retained, x_.p, n_ * R_);
  if (keep <= R_) return;
  const long m = size_;
  double* q = qbuf_.p;
  gemm(true, 1, 0, m, R_, n_, 1.0, basis_.p, n_, sx_.p, n_, q, m);
  auto* tau = reinterpret_cast<rocblas_double_complex*>(small_.p + 8);
  check(rocsolver_zgeqrf(ctx_.blas(), rocblas_int(m), rocblas_int(R_), reinterpret_cast<rocblas_double_complex*>(q),
                         rocblas_int(m), tau),
        "rocsolver_zgeqrf");
  check(rocsolver_zungqr(ctx_.blas(), rocblas_int(m), rocblas_int(m), rocblas_int(R_),
                         reinterpret_cast<rocblas_double_complex*>(q), rocblas_int(m), tau),
        "rocsolver_zungqr");
  gemm(true, 0, 0, n_, keep - R_, m, 1.0, basis_.p, n_, q + 2 * R_ * m, m, retained + 2 * R_ * n_, n_);
}

// Jacobi-Davidson refinement after the hand-off from search arithmetic. The
// block W = [X, G] holds the R wanted Ritz vectors X and the guard Ritz
// vectors G (the next refinement_guards() up), columns 0..k of basis_ with
// their images in hb_ and sb_. Each step:
//   1. FP64 H and S images of W: fresh applications on the first step (W
//      comes from the search subspace), afterwards the FP64 images of [W, T]
//      rotated by the Rayleigh-Ritz vectors as the Davidson rotates its
//      images (two GEMMs instead of a k-column pair application; the
//      certificate applies H and S afresh either way); Rayleigh quotients, S
//      normalisation and residuals of X; when every root meets the
//      tolerance, certify();
//   2. corrections T for the unconverged roots from the correction equations
//      P^H (H - theta_j S) P t_j = r_j, P = I - X (S X)^H, by preconditioned
//      CG in search (FP32) arithmetic (correction_cg);
//   3. T projected against W and S-orthonormalised in FP64, with FP64 images;
//   4. FP64 Rayleigh-Ritz of [W, T]; the lowest k Ritz vectors are the next W.
// The guards enter the Rayleigh-Ritz (roots can move across the block edge)
// but not the projector: guards from the search subspace are far less
// accurate than the wanted roots, and projecting them out distorts the
// correction equations of the roots at the top of the block. Returns true when certified; false (x_ holding the current X, W in
// retained_) when a step could not be formed or the residual stopped
// improving, for the FP64 Davidson iterations to continue from.

// Block preconditioned CG for the active roots (see refine); t: n x |active|.
// The projector is the one of the wanted block X (the first R columns of W);
// the guards enter through the Rayleigh-Ritz only. Applications and GEMMs run
// in FP32 arithmetic whatever the solver's current precision.
void Solver::correction_cg(const std::vector<int>& active, double* t) {
  const long n = n_, R = R_, cols = long(active.size()), k = R;
  const Precision f32 = Precision::FP32;
  const int steps = p_.inner_iterations > 0 ? p_.inner_iterations : 6;
  const double* w = basis_.p;
  const double* sw = sb_.p;
  double* p = w3_.p;
  double* q = p + 2 * n * R;
  double* z = w4_.p;
  double* res = z + 2 * n * R;
  double* coef = m1_.p;  // R x cols
  double* base = small_.p + 8 + 4 * R;
  double* theta = base;
  double* wk = base + R;
  double* rho0 = base + 2 * R;
  double* rho1 = base + 3 * R;
  double* pq = base + 4 * R;
  int* src = reinterpret_cast<int*>(base + 5 * R);
  double* part = base + 6 * R;  // cols x cg_step_parts() partial sums of (r, z)
  if (long(cg_step_parts()) * cols > 8 * R) throw std::logic_error("HIP: CG partial sums exceed their buffer");
  std::vector<double> th(cols);
  for (long c = 0; c < cols; ++c) th[c] = values_[active[c]];
  upload(theta, th.data(), cols);
  upload_bytes(src, active.data(), sizeof(int) * cols);
  check(launch_kinetic_weights(w, n, n, R, op_.kinetic_device(), wk, s_), "kinetic weights");
  check(launch_gather_columns(res_.p, n, src, n, cols, res, n, s_), "gather");
  // P y = y - X (SX)^H y and P^H y = y - SX X^H y in search arithmetic
  auto project = [&](double* y, bool adjoint) {
    ctx_.zgemm(f32, 1, 0, k, cols, n, 1.0, adjoint ? w : sw, n, y, n, coef, k, nullptr, 0, 0.0);
    ctx_.zgemm(f32, 0, 0, n, cols, k, -1.0, adjoint ? sw : w, n, coef, k, y, n, y, n, 1.0);
  };
  project(res, true);
  check(hipMemsetAsync(t, 0, 16 * std::size_t(n) * cols, s_), "zero");
  check(launch_jd_precondition(res, n, src, cols, wk, op_.kinetic_device(), n, z, n, s_), "precondition");
  project(z, false);
  copy(p, z, n * cols);
  check(launch_real_dots(res, n, z, n, n, cols, rho0, s_), "dots");
  for (int it = 0; it < steps; ++it) {
    // q = (H - theta S) p in one application (a single back-projection)
    op_.apply_shifted(p, cols, theta, q, f32, p_.bf16_projectors != 0);
    stats_.h_columns_search += cols;
    project(q, true);
    check(launch_real_dots(p, n, q, n, n, cols, pq, s_), "dots");
    const bool last = it + 1 == steps;
    // the update, the preconditioned residual z = M r and the partial sums of
    // (r, z) in one pass; the last step only updates t and r
    if (last)
      check(launch_cg_update(t, res, p, q, n, n, cols, rho0, pq, s_), "cg update");
    else
      check(launch_cg_step(t, res, p, q, z, n, n, cols, rho0, pq, src, wk, op_.kinetic_device(), part, s_),
            "cg step");
    if (p_.verbose > 1) {
      const std::vector<double> cur = col_norms(res, n, cols);
      const std::vector<double> curv = to_host(pq, cols);
      double worst = 0, neg = 0;
      for (long c = 0; c < cols; ++c) {
        worst = std::max(worst, cur[c] / std::max(res_norms_[active[c]], 1e-300));
        neg += curv[c] <= 0;
      }
      std::fprintf(stderr, "  cg step %d: max relative residual %.3e, nonpositive curvature %g of %ld\n", it, worst, neg,
                   cols);
    }
    if (last) break;
    // the later preconditioned residuals are not projected again (one
    // projection per step instead of two): the first direction starts in the
    // range of P, and step 3 of refine() projects the correction against W
    // whatever the iterates picked up; measured, the solve takes as many
    // steps and 2 % less time
    check(launch_cg_direction_parts(p, z, n, n, cols, part, rho0, rho1, s_), "cg direction");
    std::swap(rho0, rho1);
  }
}

// W (k = R + guards columns, n x k) in retained_ on entry and exit.
bool Solver::refine(long& it, long& iterations) {
  const long n = n_, R = R_, k = R + refinement_guards();
  const double t0 = now();
  double best = std::numeric_limits<double>::infinity();
  int stalls = 0;
  double* wbuf = retained_.p;
  long rotated_at = -1;  // >= 0: W's rotated images wait in hb_/sb_ columns from there
  // Ritz values of the previous step's Rayleigh-Ritz; empty while W still
  // comes from the search subspace
  std::vector<double> ritz;
  for (; it + 1 < p_.max_iterations;) {
    ++it;
    iterations = it + 1;
    // 1. FP64 images of W; X = its first R columns
    copy(basis_.p, wbuf, n * k);
    if (rotated_at >= 0) {
      copy(hb_.p, col(hb_, rotated_at, n), n * k);
      copy(sb_.p, col(sb_, rotated_at, n), n * k);
    } else {
      apply(2, basis_.p, k, hb_.p, sb_.p, true);
    }
    // Rayleigh quotients, S-normalisation of X and its residuals on the
    // device; metrics, energies and residual norms come back in one transfer
    {
      double* dm = small_.p + 8;
      double* de = dm + R;
      double* dv = de + R;
      double* droot = dv + R;
      double* dn = droot + R;
      check(launch_real_dots(basis_.p, n, sb_.p, n, n, R, dm, s_), "dots");
      check(launch_real_dots(basis_.p, n, hb_.p, n, n, R, de, s_), "dots");
      check(launch_rayleigh_quotients(dm, de, R, dv, droot, s_), "Rayleigh quotients");
      for (DevBuf<double>* b : {&basis_, &hb_, &sb_})
        check(launch_divide_columns(b->p, n, n, R, droot, s_), "normalise");
      check(launch_fma_residual(hb_.p, sb_.p, n, n, R, dv, res_.p, s_), "residual");
      check(launch_col_norms(res_.p, n, n, R, dn, s_), "norms");
      const std::vector<double> h = to_host(dm, 5 * R);
      for (long r = 0; r < R; ++r) {
        const double metric = h[r], energy = h[R + r];
        if (!(metric > 0) || !std::isfinite(energy)) throw std::runtime_error("HIP: invalid refinement Ritz metric");
        values_[r] = energy / metric;
        res_norms_[r] = h[4 * R + r];
      }
      converged_ = true;
      for (long r = 0; r < R; ++r) converged_ = converged_ && res_norms_[r] <= tol_;
    }
    stats_.max_residual_history.push_back(max_res());
    if (search_residual_ >= 0) {
      if (max_res() > std::max(8 * tol_, 10 * search_residual_)) ++stats_.search_rejected;
      search_residual_ = -1;
    }
    if (p_.verbose)
      std::fprintf(stderr, "davidson it=%ld refinement max_residual=%.3e t=%.1fms\n", it + 1, max_res(),
                   1e3 * (now() - t0));
    if (converged_) {
      copy(x_.p, basis_.p, n * R);
      certify();
      if (converged_) return true;
    }
    if (max_res() < 0.5 * best) {
      best = max_res();
      stalls = 0;
    } else if (++stalls >= 2) {
      copy(x_.p, basis_.p, n * R);
      return false;
    }
    std::vector<int> active;
    for (long r = 0; r < R; ++r)
      if (res_norms_[r] > .5 * tol_) active.push_back(int(r));
    const long cols = long(active.size());
    ++stats_.refinement_steps;
    // 2. corrections in search arithmetic
    double* t = w4_.p + 4 * n * R;
    correction_cg(active, t);
    // the corrections scale with their roots' residuals, which span many
    // orders of magnitude once some roots converge; unit columns keep the
    // Cholesky pivot test below a test of linear dependence, not of scale
    check(launch_col_norms(t, n, n, cols, small_.p + 8, s_), "norms");
    check(launch_guard_divisors(small_.p + 8, cols, s_), "norms");
    check(launch_divide_columns(t, n, n, cols, small_.p + 8, s_), "unit corrections");
    // 3. T projected against W and S-orthonormalised, with FP64 images
    size_ = k;
    double* v = w3_.p;
    double* metric_v = w3_.p + 2 * n * R;
    double* coef = m1_.p;
    gemm(true, 1, 0, k, cols, n, 1.0, sb_.p, n, t, n, coef, k);
    gemm(true, 0, 0, n, cols, k, -1.0, basis_.p, n, coef, k, v, n, t, n, 1.0);
    // H v and S v in one pair application (one projection B v for both);
    // the Cholesky normalisation T = v U^-1 takes S T = (S v) U^-1 and
    // H T = (H v) U^-1 by the same triangular factor
    double* hv = w3_.p + 4 * n * R;
    apply(2, v, cols, hv, metric_v, true);
    double* gram = m3_.p;
    double* gh = m3_.p + 2 * cols * cols;
    gemm(true, 1, 0, cols, cols, n, 1.0, v, n, metric_v, n, gram, cols);
    check(launch_hermitise(gram, cols, cols, gh, cols, s_), "hermitise");
    long added = cols;
    bool h_images = false;
    if (cholesky_normalise(v, metric_v, gram, gh, cols, true)) {
      gemm(true, 0, 0, n, cols, cols, 1.0, hv, n, gh, cols, col(hb_, k, n), n);
      h_images = true;
    } else {
      // nearly dependent corrections (a correction inside span(W) up to its
      // search-arithmetic rounding): the reference's block CGS2 with its
      // eigen-normalisation and dependency test keeps the independent part
      copy(corr_.p, v, n * cols);
      append_block(corr_.p, cols);
      added = size_ - k;
      if (added <= 0) {
        copy(x_.p, basis_.p, n * R);
        return false;
      }
    }
    if (!h_images) apply(0, col(basis_, k, n), added, col(hb_, k, n), nullptr, true);
    size_ = k + added;
    // 4. Rayleigh-Ritz of [W, T]
    const long m = size_;
    double* A = rr_.p;
    double* B = rrb_.p;
    // W holds Ritz vectors of the previous step's FP64 Rayleigh-Ritz, so its
    // block of the pencil is (diag(theta), I) to rounding and only the columns
    // of T are projected: m x |T| products instead of m x m (the first step's
    // W comes from the search subspace and is projected in full)
    const long first = ritz.empty() ? 0 : k;
    gemm(true, 1, 0, m, m - first, n, 1.0, basis_.p, n, col(hb_, first, n), n, ch_.p + 2 * first * m, m);
    gemm(true, 1, 0, m, m - first, n, 1.0, basis_.p, n, col(sb_, first, n), n, cs_.p + 2 * first * m, m);
    if (first) {
      // X was renormalised in step 1: its Rayleigh quotients, then the guards'
      std::vector<double> theta(values_.begin(), values_.begin() + R);
      theta.insert(theta.end(), ritz.begin() + R, ritz.begin() + k);
      upload(small_.p + 8, theta.data(), k);
      check(launch_ritz_border(ch_.p, m, k, m, small_.p + 8, s_), "Ritz block");
      check(launch_ritz_border(cs_.p, m, k, m, nullptr, s_), "Ritz block");
    }
    check(launch_hermitise(ch_.p, m, m, A, m, s_), "hermitise");
    check(launch_hermitise(cs_.p, m, m, B, m, s_), "hermitise");
    const std::vector<double> w = rr_solve(A, B, m, k);
    if (long(w.size()) < k) throw std::runtime_error("HIP: refinement Rayleigh-Ritz returned too few values");
    ritz.assign(w.begin(), w.begin() + k);
    for (long r = 0; r < R; ++r) values_[r] = w[r];
    gemm(true, 0, 0, n, k, m, 1.0, basis_.p, n, A, m, wbuf, n);
    // the next W's images, past the m columns in use when the buffers have room
    rotated_at = m + k <= cap_ ? m : -1;
    if (rotated_at >= 0) {
      gemm(true, 0, 0, n, k, m, 1.0, hb_.p, n, A, m, col(hb_, m, n), n);
      gemm(true, 0, 0, n, k, m, 1.0, sb_.p, n, A, m, col(sb_, m, n), n);
    }
  }
  copy(x_.p, wbuf, n * R);
  return false;
}

StrictStepOutput Solver::strict_primitive(int kind, const C* x, const C* direction) {
  const long n = n_, R = R_;
  ctx_.upload_bulk(x_.p, x, 16 * std::size_t(n) * R);
  values_.assign(R, 0.);
  res_norms_.assign(R, 0.);
  certify();
  StrictStepOutput out;
  if (kind == 0) {
    ctx_.upload_bulk(corr_.p, direction, 16 * std::size_t(n) * R);
    // the flag is an int in the first word of a zeroed double
    int* flag = reinterpret_cast<int*>(small_.p);
    check(hipMemsetAsync(small_.p, 0, sizeof(double), s_), "memset");
    check(launch_finite_check(corr_.p, n, n, R, flag, s_), "finite");
    if (to_host(small_.p, 1)[0] != 0.0) throw std::invalid_argument("nonfinite strict residual polish state");
    out.accepted = polish_step(corr_.p);
  } else {
    out.accepted = manifold_step();
  }
  refresh_norms();
  out.x.resize(std::size_t(n) * R);
  ctx_.download_bulk(out.x.data(), x_.p, 16 * std::size_t(n) * R);
  out.energies = values_;
  out.residual_norms = res_norms_;
  return out;
}

DavidsonOutput Solver::run(const C* initial) {
  const double t_start = now();
  const long n = n_, R = R_;
  incremental_ = tol_ >= 1e-12;
  bool strict_projection = false;
  unsigned polish_attempts = 0, manifold_attempts = 0;
  double strict_entry = std::numeric_limits<double>::infinity();
  const std::vector<long> order = davidson_order(op_);
  const bool warm = initial != nullptr && 2 * R < n;
  const long nseeds = davidson_seed_columns(n, R, warm);
  if (warm && p_.device_pointers)
    check(hipMemcpyAsync(corr_.p, initial, 16 * std::size_t(n) * R, hipMemcpyDeviceToDevice, s_), "initial block");
  else if (warm)
    ctx_.upload_bulk(corr_.p, initial, 16 * std::size_t(n) * R);
  // the retained block holds cap_ columns; a local subspace of fewer than
  // 3 roots (band groups) takes a block of its own
  DevBuf<double> wide;
  double* block = retained_.p;
  if (nseeds > cap_) block = wide.reserve(2 * std::size_t(n) * nseeds);
  davidson_seed_block(op_, R, tol_, warm ? corr_.p : nullptr, block, s_);
  append_block(block, nseeds);

  values_.assign(R, 0.);
  res_norms_.assign(R, 0.);
  bool certified = false;
  long iterations = 0;
  DevBuf<double>& rrb = rrb_;
  DevBuf<double>& corr = corr_;
  DevBuf<double>& retained = retained_;
  double best_history = std::numeric_limits<double>::infinity();
  long since_halving = 0;
  for (long it = 0; it < p_.max_iterations; ++it) {
    flush_h();
    update_projected(hb_.p, ch_.p);
    update_projected(sb_.p, cs_.p);
    projected_size_ = size_;
    double* A = rr_.p;  // eigenvectors after the solve
    double* B = rrb.p;
    check(launch_hermitise(ch_.p, cap_, size_, A, size_, s_), "hermitise");
    check(launch_hermitise(cs_.p, cap_, size_, B, size_, s_), "hermitise");
    // Ritz vectors wanted: 2 roots for a thick restart, more during an FP32
    // search whose hand-off keeps more
    const long wanted = exact_ ? 2 * R : long(std::max(2, p_.handoff_restart)) * R;
    const std::vector<double> w = rr_solve(A, B, size_, std::min(size_, wanted));
    for (long r = 0; r < R; ++r) values_[r] = w[r];
    // Ritz vectors and residuals (src/pw_core.cpp:1169-1183)
    gemm(exact_, 0, 0, n, R, size_, 1.0, basis_.p, n, A, size_, x_.p, n);
    gemm(exact_, 0, 0, n, R, size_, 1.0, sb_.p, n, A, size_, sx_.p, n);
    gemm(exact_, 0, 0, n, R, size_, 1.0, hb_.p, n, A, size_, hx_.p, n);
    upload(small_.p + 8, values_.data(), R);
    check(launch_ritz_residual(hx_.p, sx_.p, n, n, R, small_.p + 8, res_.p, s_), "ritz residual");
    refresh_norms();
    iterations = it + 1;
    stats_.max_residual_history.push_back(max_res());
    if (exact_ && search_residual_ >= 0) {
      // first Rayleigh-Ritz on FP64 images after the hand-off: the search
      // arithmetic's residual must be confirmed by the original operator
      if (max_res() > std::max(8 * tol_, 10 * search_residual_)) ++stats_.search_rejected;
      search_residual_ = -1;
    }
    if (p_.verbose)
      std::fprintf(stderr, "davidson it=%ld size=%ld exact=%d max_residual=%.3e t=%.1fms\n", it + 1, size_, int(exact_),
                   max_res(), 1e3 * (now() - t_start));
    if (!exact_) {
      // FP32 search hand-off: the reference's 1e-8 threshold, or a stalled
      // FP32 residual (its rounding floor lies far above 1e-8)
      // stagnation counts only near the FP32 rounding floor (below 1e-3), so
      // the slow first iterations of a cold block never trigger it
      if (max_res() < 0.5 * best_history) { best_history = max_res(); since_halving = 0; } else if (max_res() < 1e-3) ++since_halving;
      const bool stalled = since_halving >= p_.stagnation_window;
      if (max_res() < p_.search_handoff || it + 1 == p_.max_iterations || stalled) {
        stats_.handoff_iteration = it + 1;
        search_residual_ = max_res();
        const long jd_block = R + refinement_guards();
        if (p_.refinement && cap_ >= 4 * R && size_ >= jd_block && p_.handoff_restart >= 1 &&
            it + 1 < p_.max_iterations) {
          // the lowest R + guards Ritz vectors of the search subspace in FP64
          // arithmetic, then the Jacobi-Davidson refinement; it returns false
          // to hand them to the FP64 Davidson iterations below
          exact_ = true;
          gemm(true, 0, 0, n, jd_block, size_, 1.0, basis_.p, n, A, size_, retained.p, n);
          if (refine(it, iterations)) {
            certified = true;
            break;
          }
          size_ = 0;
          computed_h_ = 0;
          projected_size_ = 0;
          append_block(retained.p, jd_block);
          continue;
        }
        const long keep = std::min(size_, long(std::max(2, p_.handoff_restart)) * R);
        if (p_.handoff_restart && size_ > keep) {
          // keep the lowest Ritz vectors (FP64 rotation), then the exact block
          // CGS2 and the next flush give them FP64 S and H images
          exact_ = true;
          gemm(true, 0, 0, n, keep, size_, 1.0, basis_.p, n, A, size_, retained.p, n);
          size_ = 0;
          computed_h_ = 0;
          projected_size_ = 0;
          append_block(retained.p, keep);
        } else {
          exact_ = true;
          rebuild_exact();
        }
        continue;
      }
    }
    bool fresh = false;
    if (max_res() <= 8 * tol_) { certify(); fresh = true; }
    if (converged_ && !fresh) { certify(); fresh = true; }
    if (!converged_ && tol_ <= 1e-12 && !strict_projection && max_res() <= 1e-10) {
      if (!fresh) { certify(); fresh = true; }
      if (!converged_) {  // src/pw_core.cpp:1247-1267
        rebuild_exact();
        incremental_ = false;
        strict_projection = true;
        continue;
      }
    }
    bool polished = false;
    // a capped local band group is not a complete invariant manifold: no
    // strict polish there (src/pw_core.cpp:1269-1273)
    if (!(p_.band_group_local && p_.capacity) && !converged_ && tol_ <= 1e-12 && max_res() <= 1e-10 &&
        max_res() < .5 * strict_entry &&
        (polish_attempts < 16 || manifold_attempts < 4)) {
      if (!fresh) certify();
      strict_entry = max_res();
      if (!converged_ && manifold_attempts < 4 && it + 1 < p_.max_iterations) {
        ++manifold_attempts;
        ++it;
        iterations = it + 1;
        const bool accepted = manifold_step();
        ++(accepted ? stats_.manifold_accepted : stats_.manifold_rejected);
        polished |= accepted;
        refresh_norms();
      }
      for (; !converged_ && polish_attempts < 16 && it + 1 < p_.max_iterations; ++polish_attempts) {
        std::vector<int> src(R);
        std::iota(src.begin(), src.end(), 0);
        check(launch_kinetic_weights(x_.p, n, n, R, op_.kinetic_device(), small_.p + 8, s_), "kinetic weights");
        int* srcd = reinterpret_cast<int*>(small_.p + 8 + R);
        upload_bytes(srcd, src.data(), R * sizeof(int));
        check(launch_precondition(res_.p, n, srcd, R, small_.p + 8, op_.kinetic_device(), n, corr.p, n, s_),
              "precondition");
        ++it;
        iterations = it + 1;
        const bool improved = polish_step(corr.p);
        ++(improved ? stats_.polish_accepted : stats_.polish_rejected);
        polished |= improved;
        if (!improved) { ++polish_attempts; break; }
        refresh_norms();
      }
    }
    if (converged_) { certified = true; break; }
    if (polished && it + 1 < p_.max_iterations) {
      const bool at_cap = size_ + R > cap_;
      const long keep = at_cap ? std::min(size_, 2 * R) : size_;
      if (at_cap) {
        copy(retained.p, x_.p, n * R);
        if (keep > R) gemm(exact_, 0, 0, n, keep - R, size_, 1.0, basis_.p, n, A + 2 * R * size_, size_, retained.p + 2 * R * n, n);
      } else {
        preserved_subspace(keep, retained.p);
      }
      size_ = 0;
      computed_h_ = 0;
      projected_size_ = 0;
      append_block(retained.p, keep);
    }
    if (size_ + R > cap_) {  // thick restart (src/pw_core.cpp:1357-1383)
      const long keep = std::min(size_, 2 * R);
      gemm(exact_, 0, 0, n, keep, size_, 1.0, basis_.p, n, A, size_, retained.p, n);
      size_ = 0;
      computed_h_ = 0;
      projected_size_ = 0;
      append_block(retained.p, keep);
    }
    const long previous = size_;
    std::vector<int> src;
    for (long r = 0; r < R && size_ + long(src.size()) < cap_; ++r)
      if (res_norms_[r] > .5 * tol_) src.push_back(int(r));
    // the reference stops collecting once size reaches capacity; it appends
    // at most cap - size columns (append_block truncates the same way)
    if (!src.empty()) {
      check(launch_kinetic_weights(x_.p, n, n, R, op_.kinetic_device(), small_.p + 8, s_), "kinetic weights");
      int* srcd = reinterpret_cast<int*>(small_.p + 8 + R);
      upload_bytes(srcd, src.data(), src.size() * sizeof(int));
      check(launch_precondition(res_.p, n, srcd, long(src.size()), small_.p + 8, op_.kinetic_device(), n, corr.p, n, s_),
            "precondition");
      append_block(corr.p, long(src.size()));
    }
    if (size_ == previous) {
      bool added = false;
      std::vector<double> unit(2 * n, 0.0);
      DevBuf<double> u;
      u.reserve(2 * n);
      for (long j = 0; j < n && !added; ++j) {
        std::fill(unit.begin(), unit.end(), 0.0);
        unit[2 * order[j]] = 1.0;
        upload(u.p, unit.data(), 2 * n);
        added = append(u.p);
      }
      if (!added) break;
    }
  }
  if (!certified) certify();
  check_finite(x_.p, R, "Davidson eigenvectors");
  DavidsonOutput out;
  out.eigenvalues = values_;
  out.residual_norms = res_norms_;
  std::complex<double>* dst = p_.vectors_out;
  if (!dst) {
    out.eigenvectors.resize(std::size_t(n) * R);
    dst = out.eigenvectors.data();
  }
  if (p_.device_pointers)
    check(hipMemcpyAsync(dst, x_.p, 16 * std::size_t(n) * R, hipMemcpyDeviceToDevice, s_), "eigenvectors");
  else
    ctx_.download_bulk(dst, x_.p, 16 * std::size_t(n) * R);
  sync();
  stats_.iterations = iterations;
  stats_.rr_stedc = eig_ ? eig_->stedc_solves() - stedc_base_ : 0;
  stats_.rr_barrier_fallbacks = eig_ ? eig_->barrier_fallbacks() - barrier_base_ : 0;
  stats_.converged = converged_;
  stats_.seconds_total = now() - t_start;
  out.stats = stats_;
  return out;
}
}  // namespace

std::size_t davidson_cache_bytes(DeviceOperator& op) {
  auto ws = std::static_pointer_cast<Workspace>(op.davidson_cache());
  return ws ? workspace_bytes(*ws) : 0;
}

std::vector<long> davidson_order(const DeviceOperator& op) {
  const std::vector<double>& hd = op.h_diagonal();
  const std::vector<double>& sd = op.s_diagonal();
  std::vector<long> order(hd.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](long i, long j) { return hd[i] / sd[i] < hd[j] / sd[j]; });
  return order;
}

long davidson_seed_columns(long n, long roots, bool warm) { return (warm ? roots : 0) + std::min(n, 2 * roots); }

// seeds (src/pw_core.cpp:1027-1112): splitmix64 broad seeds, warm start +
// clamp(100 tol, 1e-10, 1e-6) perturbation, 2 roots canonical guards; the
// block is generated on the device (launch_davidson_seeds) from the same
// stream of uniform() calls, in the host recipe's order
long davidson_seed_block(DeviceOperator& op, long roots, double tol, const double* initial_device, double* block,
                         hipStream_t s) {
  const long n = long(op.n());
  const bool warm = initial_device != nullptr && 2 * roots < n;
  const long guards = std::min(n, 2 * roots);
  const std::vector<long> order = davidson_order(op);
  std::vector<int> guard_rows(order.begin(), order.begin() + guards);
  DevBuf<int> guard_dev;
  guard_dev.reserve(std::size_t(guards));
  check(hipMemcpyAsync(guard_dev.p, guard_rows.data(), sizeof(int) * std::size_t(guards), hipMemcpyHostToDevice, s),
        "guard rows");
  check(launch_davidson_seeds(n, warm ? roots : 0, initial_device, n, std::clamp(100. * tol, 1e-10, 1e-6), guards,
                              guard_dev.p, block, n, s),
        "seeds");
  // guard_dev is released on return
  check(hipStreamSynchronize(s), "seeds");
  return davidson_seed_columns(n, roots, warm);
}

StrictStepOutput strict_step(DeviceOperator& op, int kind, long roots, double tolerance, const std::complex<double>* x,
                             const std::complex<double>* direction) {
  if ((kind != 0 && kind != 1) || !x || (kind == 0 && !direction) || roots <= 0 || long(op.n()) < roots ||
      !std::isfinite(tolerance) || tolerance <= 0)
    throw std::invalid_argument("invalid strict residual polish dimensions");
  DavidsonParams p;
  p.roots = roots;
  p.tolerance = tolerance;
  p.max_iterations = 1;
  Solver s(op, p);
  return s.strict_primitive(kind, x, direction);
}

DavidsonOutput davidson(DeviceOperator& op, const DavidsonParams& params, const std::complex<double>* initial) {
  if (params.device_pointers && !params.vectors_out)
    throw std::invalid_argument("HIP: a device-resident Davidson needs a device eigenvector destination");
  if (params.band_group_size > 0 && !params.band_group_local) return davidson_band_groups(op, params, initial);
  Solver s(op, params);
  return s.run(initial);
}

// src/band_group_solver.cpp:57-211 (no image reuse). Default: two sweeps,
// each group solved to the full iteration limit. Swept (band_group_sweep,
// the reference's SDFT_BAND_GROUP_SWEEP, :79-100 and :197-208): half the
// iteration budget goes to sweeps whose local solves stop after 3 iterations
// (budget / (groups x 3) sweeps; too small a budget solves all roots with the
// Davidson at once), and a block the sweeps leave uncertified is refined by
// the Davidson over all roots with the iterations that remain.
DavidsonOutput davidson_band_groups(DeviceOperator& op, const DavidsonParams& params,
                                    const std::complex<double>* initial) {
  using C = std::complex<double>;
  const long n = long(op.n()), R = params.roots, G = params.band_group_size;
  if (!n || R <= 0 || R > n || params.max_iterations <= 0 || !std::isfinite(params.tolerance) ||
      params.tolerance <= 0)
    throw std::invalid_argument("invalid bounded band-group problem");
  if (G < 1 || G > 64) throw std::invalid_argument("SDFT_BAND_GROUP_SIZE must be in [1,64]");
  const bool swept = params.band_group_sweep != 0;
  const long local_limit = 3, groups = (R + G - 1) / G, local_budget = params.max_iterations / 2;
  // the Davidson over all roots, as a local solve (no strict polish)
  auto whole_block = [&](const std::complex<double>* seed, long iterations, bool seed_on_device) {
    DavidsonParams lp = params;
    lp.band_group_size = 0;
    lp.band_group_sweep = 0;
    lp.band_group_local = 1;
    lp.max_iterations = iterations;
    if (seed && seed_on_device && !params.device_pointers) {
      // the swept block lives on the device; a host caller gets host vectors back
      std::vector<C> host(std::size_t(n) * R);
      check(hipMemcpy(host.data(), seed, 16 * std::size_t(n) * R, hipMemcpyDeviceToHost), "seed");
      return davidson(op, lp, host.data());
    }
    return davidson(op, lp, seed);
  };
  if (swept && groups * local_limit > local_budget) return whole_block(initial, params.max_iterations, false);
  const long sweep_limit = swept ? local_budget / (groups * local_limit) : 2;
  Context& ctx = op.context();
  hipStream_t s = ctx.stream();
  const double t_start = now();
  auto sync = [&] { ctx.sync(); };
  DevBuf<double> x, hx, sx, tmp, small, rot, hh, ss;
  x.reserve(std::size_t(2 * n * R));
  hx.reserve(std::size_t(2 * n * R));
  sx.reserve(std::size_t(2 * n * R));
  tmp.reserve(std::size_t(2 * n * R));
  small.reserve(std::size_t(8 * R + 64));
  rot.reserve(std::size_t(2 * R * R));
  hh.reserve(std::size_t(2 * R * R));
  ss.reserve(std::size_t(2 * R * R));
  const hipMemcpyKind in_kind = params.device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const hipMemcpyKind out_kind = params.device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (initial) {
    check(hipMemcpyAsync(x.p, initial, 16 * std::size_t(n) * R, in_kind, s), "upload");
  } else {
    check(hipMemsetAsync(x.p, 0, 16 * std::size_t(n) * R, s), "zero");
  }
  sync();
  HermitianEigensolver eig(ctx, R);
  const Precision f64 = Precision::FP64;
  auto to_host = [&](const double* d, long count) {
    std::vector<double> h(count);
    check(hipMemcpyAsync(h.data(), d, 8 * std::size_t(count), hipMemcpyDeviceToHost, s), "to host");
    sync();
    return h;
  };
  int* info = reinterpret_cast<int*>(small.p);
  auto potrf_upper = [&](double* a, long m) {
    check(rocsolver_zpotrf(ctx.blas(), rocblas_fill_upper, rocblas_int(m), reinterpret_cast<rocblas_double_complex*>(a),
                           rocblas_int(m), info),
          "rocsolver_zpotrf");
    int pinfo = 0;
    check(hipMemcpyAsync(&pinfo, info, sizeof(int), hipMemcpyDeviceToHost, s), "info");
    sync();
    return pinfo == 0;
  };
  DavidsonOutput out;
  out.eigenvalues.assign(R, 0.);
  out.residual_norms.assign(R, 0.);
  bool converged = false;
  for (long sweep = 0; sweep < sweep_limit && !converged; ++sweep) {
    for (long first = 0; first < R; first += G) {
      const long count = std::min(G, R - first);
      DavidsonParams lp = params;
      lp.roots = count;
      lp.band_group_local = 1;
      if (swept) lp.max_iterations = local_limit;
      lp.vectors_out = nullptr;
      lp.device_pointers = 0;  // the group's seed and result pass through host vectors
      lp.capacity = std::min(n, 5 * G);
      if (first) {
        // S-orthonormalise the completed bands (Cholesky), src/band_group_solver.cpp:100-123
        op.apply(1, x.p, first, nullptr, tmp.p, f64);
        ctx.zgemm(f64, 1, 0, first, first, n, 1.0, x.p, n, tmp.p, n, ss.p, first, nullptr, 0, 0.0);
        check(launch_hermitise(ss.p, first, first, hh.p, first, s), "hermitise");
        if (!potrf_upper(hh.p, first)) throw std::runtime_error("nonpositive completed-band overlap");
        check(rocsolver_ztrtri(ctx.blas(), rocblas_fill_upper, rocblas_diagonal_non_unit, rocblas_int(first),
                               reinterpret_cast<rocblas_double_complex*>(hh.p), rocblas_int(first), info),
              "rocsolver_ztrtri");
        check(launch_zero_lower(hh.p, first, first, s), "zero lower");
        ctx.zgemm(f64, 0, 0, n, first, first, 1.0, x.p, n, hh.p, first, tmp.p, n, nullptr, 0, 0.0);
        check(hipMemcpyAsync(x.p, tmp.p, 16 * std::size_t(n) * first, hipMemcpyDeviceToDevice, s), "locked");
        lp.excluded = x.p;
        lp.excluded_count = first;
      }
      std::vector<C> seed;
      if (sweep || initial) {
        seed.resize(std::size_t(n) * count);
        check(hipMemcpyAsync(seed.data(), x.p + 2 * first * n, 16 * std::size_t(n) * count, hipMemcpyDeviceToHost,
                             s),
              "seed");
        sync();
      }
      const DavidsonOutput r = davidson(op, lp, seed.empty() ? nullptr : seed.data());
      if (params.verbose)
        std::fprintf(stderr, "band_group sweep=%ld first=%ld count=%ld iterations=%ld converged=%d\n", sweep, first,
                     count, r.stats.iterations, int(r.stats.converged));
      out.stats.iterations += r.stats.iterations;
      out.stats.h_columns_exact += r.stats.h_columns_exact;
      out.stats.h_columns_search += r.stats.h_columns_search;
      check(hipMemcpyAsync(x.p + 2 * first * n, r.eigenvectors.data(), 16 * std::size_t(n) * count,
                           hipMemcpyHostToDevice, s),
            "group vectors");
      sync();
    }
    // global roots x roots rotation on fresh H/S, src/band_group_solver.cpp:158-172
    op.apply(2, x.p, R, hx.p, sx.p, f64);
    ctx.zgemm(f64, 1, 0, R, R, n, 1.0, x.p, n, hx.p, n, rot.p, R, nullptr, 0, 0.0);
    check(launch_hermitise(rot.p, R, R, hh.p, R, s), "hermitise");
    ctx.zgemm(f64, 1, 0, R, R, n, 1.0, x.p, n, sx.p, n, rot.p, R, nullptr, 0, 0.0);
    check(launch_hermitise(rot.p, R, R, ss.p, R, s), "hermitise");
    // generalized pencil by Cholesky reduction: S = L L^H, C = L^-1 H L^-H, Y, V = L^-H Y
    const rocblas_double_complex one{1.0, 0.0};
    auto* L = reinterpret_cast<rocblas_double_complex*>(ss.p);
    auto* H = reinterpret_cast<rocblas_double_complex*>(hh.p);
    check(rocsolver_zpotrf(ctx.blas(), rocblas_fill_lower, rocblas_int(R), L, rocblas_int(R), info), "zpotrf");
    {
      int pinfo = 0;
      check(hipMemcpyAsync(&pinfo, info, sizeof(int), hipMemcpyDeviceToHost, s), "info");
      sync();
      if (pinfo) throw std::runtime_error("band-group global rotation failed");
    }
    check(rocblas_ztrsm(ctx.blas(), rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                        rocblas_diagonal_non_unit, rocblas_int(R), rocblas_int(R), &one, L, rocblas_int(R), H,
                        rocblas_int(R)),
          "ztrsm");
    check(rocblas_ztrsm(ctx.blas(), rocblas_side_right, rocblas_fill_lower, rocblas_operation_conjugate_transpose,
                        rocblas_diagonal_non_unit, rocblas_int(R), rocblas_int(R), &one, L, rocblas_int(R), H,
                        rocblas_int(R)),
          "ztrsm");
    check(launch_hermitise(hh.p, R, R, tmp.p, R, s), "hermitise");
    eig.solve(tmp.p, R, R, rot.p);
    check(rocblas_ztrsm(ctx.blas(), rocblas_side_left, rocblas_fill_lower, rocblas_operation_conjugate_transpose,
                        rocblas_diagonal_non_unit, rocblas_int(R), rocblas_int(R), &one, L, rocblas_int(R),
                        reinterpret_cast<rocblas_double_complex*>(rot.p), rocblas_int(R)),
          "ztrsm");
    ctx.zgemm(f64, 0, 0, n, R, R, 1.0, x.p, n, rot.p, R, tmp.p, n, nullptr, 0, 0.0);
    check(hipMemcpyAsync(x.p, tmp.p, 16 * std::size_t(n) * R, hipMemcpyDeviceToDevice, s), "rotate");
    ctx.zgemm(f64, 0, 0, n, R, R, 1.0, sx.p, n, rot.p, R, tmp.p, n, nullptr, 0, 0.0);
    // S-normalise, then certify on fresh FP64 H and S (src/band_group_solver.cpp:173-196)
    double* o = small.p + 8;
    check(launch_real_dots(x.p, n, tmp.p, n, n, R, o, s), "metric");
    std::vector<double> metric = to_host(o, R);
    for (long j = 0; j < R; ++j) {
      if (!(metric[j] > 0)) throw std::runtime_error("band-group invalid final metric");
      metric[j] = std::sqrt(metric[j]);
    }
    check(hipMemcpyAsync(o, metric.data(), 8 * std::size_t(R), hipMemcpyHostToDevice, s), "norms");
    sync();
    check(launch_divide_columns(x.p, n, n, R, o, s), "normalise");
    op.apply(2, x.p, R, hx.p, sx.p, f64);
    check(launch_real_dots(x.p, n, hx.p, n, n, R, o, s), "energy");
    const std::vector<double> num = to_host(o, R);
    check(launch_real_dots(x.p, n, sx.p, n, n, R, o, s), "metric");
    const std::vector<double> den = to_host(o, R);
    for (long j = 0; j < R; ++j) out.eigenvalues[j] = num[j] / den[j];
    check(hipMemcpyAsync(o, out.eigenvalues.data(), 8 * std::size_t(R), hipMemcpyHostToDevice, s), "values");
    sync();
    check(launch_fma_residual(hx.p, sx.p, n, n, R, o, tmp.p, s), "residual");
    check(launch_col_norms(tmp.p, n, n, R, o, s), "norms");
    out.residual_norms = to_host(o, R);
    converged = true;
    for (double r : out.residual_norms) converged = converged && r <= params.tolerance;
  }
  if (swept && !converged) {
    // global refinement from the swept block (src/band_group_solver.cpp:197-208)
    const long search_steps = out.stats.iterations;
    DavidsonOutput refined = whole_block(reinterpret_cast<const C*>(x.p), params.max_iterations - search_steps, true);
    refined.stats.iterations += search_steps;
    return refined;
  }
  std::complex<double>* dst = params.vectors_out;
  if (!dst) {
    out.eigenvectors.resize(std::size_t(n) * R);
    dst = out.eigenvectors.data();
  }
  check(hipMemcpyAsync(dst, x.p, 16 * std::size_t(n) * R, out_kind, s), "vectors");
  sync();
  out.stats.converged = converged;
  out.stats.seconds_total = now() - t_start;
  return out;
}

}  // namespace sdft_hip
