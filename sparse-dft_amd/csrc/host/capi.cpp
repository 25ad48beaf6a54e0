// extern "C" entry points declared in include/sdft_hip.h.
#include "../../../include/sdft_hip.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <exception>
#include <limits>
#include <string>
#include <vector>

#include "../hip/davidson_kernels.hpp"
#include "../hip/scratch.hpp"
#include "../hip/tridiag_eig.hpp"
#include "batch.hpp"
#include "davidson.hpp"
#include "dense_eigen.hpp"
#include "device.hpp"
#include "pw_basis.hpp"
#include "snapshot.hpp"

// A context is reference counted: its creator holds one reference and every
// operator built on it holds another, so sdft_hip_context_destroy of a context
// whose operators are still alive only drops the creator's reference, and the
// HIP streams, handles and buffers go when the last operator is destroyed.
// Operators hold a reference to their context's state (device.hpp), so no
// destruction order of the caller can leave an operator on a freed context.
struct sdft_hip_context {
  std::atomic<long> refs{1};
  sdft_hip::Context impl;
  explicit sdft_hip_context(int device) : impl(device) {}
};

namespace {
void retain(sdft_hip_context* c) { c->refs.fetch_add(1, std::memory_order_relaxed); }
void release(sdft_hip_context* c) {
  if (c->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete c;
}
// the owning reference of an operator: declared before the operator state, so
// the state is destroyed first and the context released after it
struct ContextRef {
  sdft_hip_context* ctx;
  explicit ContextRef(sdft_hip_context* c) : ctx(c) { retain(c); }
  ~ContextRef() { release(ctx); }
  ContextRef(const ContextRef&) = delete;
  ContextRef& operator=(const ContextRef&) = delete;
};
}  // namespace

struct sdft_hip_operator {
  ContextRef owner;
  sdft_hip::DeviceOperator impl;
  sdft_hip_operator(sdft_hip_context* c, const sdft_hip::OperatorDesc& d) : owner(c), impl(c->impl, d) {}
  sdft_hip_operator(sdft_hip_context* c, const sdft_hip::SnapshotArrays& s) : owner(c), impl(c->impl, s) {}
};
struct sdft_hip_batch {
  ContextRef owner;
  sdft_hip::OperatorBatch impl;
  sdft_hip_batch(sdft_hip_context* c, const std::vector<sdft_hip::DeviceOperator*>& ops, long columns,
                 sdft_hip::Precision prec)
      : owner(c), impl(c->impl, ops, columns, prec) {}
};

namespace {
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return SDFT_HIP_OK;
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return SDFT_HIP_INVALID_ARGUMENT;
  } catch (const std::overflow_error& e) {
    g_error = e.what();
    return SDFT_HIP_OVERFLOW;
  } catch (const std::exception& e) {
    g_error = e.what();
    return SDFT_HIP_RUNTIME_ERROR;
  } catch (...) {
    g_error = "HIP: unknown error";
    return SDFT_HIP_RUNTIME_ERROR;
  }
}

using sdft_hip::check;
using sdft_hip::DevBuf;
using sdft_hip::Precision;

Precision to_precision(int p) {
  if (p != 0 && p != 1) throw std::invalid_argument("HIP: precision must be 0 (FP64) or 1 (FP32)");
  return p ? Precision::FP32 : Precision::FP64;
}

void copy_2d_h2d(double* dst, std::size_t rows, const double* src, std::size_t ld, std::size_t cols, hipStream_t s) {
  check(hipMemcpy2DAsync(dst, rows * 16, src, ld * 16, rows * 16, cols, hipMemcpyHostToDevice, s), "copy to device");
}
void copy_2d_d2h(double* dst, std::size_t ld, const double* src, std::size_t rows, std::size_t cols, hipStream_t s) {
  check(hipMemcpy2DAsync(dst, ld * 16, src, rows * 16, rows * 16, cols, hipMemcpyDeviceToHost, s), "copy to host");
}
void check_int_range(std::initializer_list<std::size_t> v) {
  for (auto x : v)
    if (x > std::size_t(std::numeric_limits<int>::max())) throw std::overflow_error("HIP: BLAS dimension overflow");
}
}  // namespace

extern "C" {

const char* sdft_hip_last_error(void) { return g_error.c_str(); }

int sdft_hip_capabilities(char* buffer, size_t size) {
  return guarded([&] {
    const std::string caps =
        "{\"schema_version\":\"sparse-dft-compute-capabilities-1\",\"cpu_fp64\":true,\"metal_compiled\":false,"
        "\"cuda_compiled\":false,\"hip_compiled\":true,\"hip_arch\":\"gfx950\",\"gpu_active\":false,"
        "\"gpu_execution_requires_numerical_dispatch\":true}";
    if (!buffer || size < caps.size() + 1) throw std::invalid_argument("HIP: capability buffer too small");
    std::memcpy(buffer, caps.c_str(), caps.size() + 1);
  });
}

int sdft_hip_gemm_real_products(int precision, int* products) {
  return guarded([&] {
    if (!products) throw std::invalid_argument("HIP: null output");
    *products = sdft_hip::zgemm_real_products(to_precision(precision));
  });
}

int sdft_hip_context_create(int device, sdft_hip_context** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("HIP: null context output");
    *out = new sdft_hip_context(device);
  });
}

int sdft_hip_context_destroy(sdft_hip_context* ctx) {
  return guarded([&] {
    if (ctx) release(ctx);
  });
}

int sdft_hip_inject_fault(sdft_hip_context* ctx, int kind, long count) {
  return guarded([&] {
    if (!ctx) throw std::invalid_argument("HIP: null context");
    ctx->impl.inject_fault(kind, count);
  });
}

int sdft_hip_device_name(sdft_hip_context* ctx, char* buffer, size_t size) {
  return guarded([&] {
    if (!ctx || !buffer || !size) throw std::invalid_argument("HIP: invalid device-name buffer");
    const std::string& n = ctx->impl.name();
    const std::size_t len = std::min(size - 1, n.size());
    std::memcpy(buffer, n.data(), len);
    buffer[len] = 0;
  });
}

int sdft_hip_synchronize(sdft_hip_context* ctx) {
  return guarded([&] { ctx->impl.sync(); });
}

int sdft_hip_kernel_clock(sdft_hip_context* ctx, int enable) {
  return guarded([&] {
    if (!ctx) throw std::invalid_argument("HIP: null context");
    ctx->impl.clock_enable(enable != 0);
  });
}

int sdft_hip_kernel_clock_bytes(sdft_hip_context* ctx, int kind, double* bytes) {
  return guarded([&] {
    if (!ctx || !bytes) throw std::invalid_argument("HIP: null kernel clock argument");
    *bytes = ctx->impl.clock_stat(kind).bytes;
  });
}

int sdft_hip_kernel_clock_stats(sdft_hip_context* ctx, int kind, long* launches, double* milliseconds, double* work) {
  return guarded([&] {
    if (!ctx || !launches || !milliseconds || !work) throw std::invalid_argument("HIP: null kernel clock argument");
    const auto st = ctx->impl.clock_stat(kind);
    *launches = st.launches;
    *milliseconds = st.milliseconds;
    *work = st.work;
  });
}

// src/compute_cuda.cu:186-209 semantics with MFMA kernels underneath
int sdft_hip_zgemm(sdft_hip_context* ctx, int aop, int bop, size_t m, size_t n, size_t k, const double* a, size_t lda,
                   const double* b, size_t ldb, double* c, size_t ldc, int add, uint64_t left_identity,
                   int precision) {
  return guarded([&] {
    if (!ctx || !m || !n || !k || !a || !b || !c || (aop != 0 && aop != 1) || (bop != 0 && bop != 1) ||
        lda < (aop == 0 ? m : k) || ldb < (bop == 0 ? k : n) || ldc < m)
      throw std::invalid_argument("HIP: invalid complex GEMM shape/leading dimension");
    check_int_range({m, n, k, lda, ldb, ldc});
    auto& cx = ctx->impl;
    const Precision prec = to_precision(precision);
    hipStream_t s = cx.stream();
    const std::size_t ar = aop == 0 ? m : k, ac = aop == 0 ? k : m;
    const std::size_t br = bop == 0 ? k : n, bc = bop == 0 ? n : k;
    double* left;
    if (left_identity) {
      auto& res = cx.resident_left[aop];
      const bool reuse = res.identity == left_identity && res.rows == ar && res.cols == ac && res.ld == lda;
      if (!reuse) {
        res.identity = 0;
        res.values.reserve(2 * ar * ac);
        copy_2d_h2d(res.values.p, ar, a, lda, ac, s);
        cx.record("gpu_resident_left_miss", 0);
      } else {
        cx.record("gpu_resident_left_hit", 0);
      }
      left = res.values.p;
      res.identity = left_identity;
      res.rows = ar;
      res.cols = ac;
      res.ld = lda;
    } else {
      cx.ga.reserve(2 * ar * ac);
      copy_2d_h2d(cx.ga.p, ar, a, lda, ac, s);
      left = cx.ga.p;
    }
    cx.gb.reserve(2 * br * bc);
    copy_2d_h2d(cx.gb.p, br, b, ldb, bc, s);
    cx.gc.reserve(2 * m * n);
    if (add) copy_2d_h2d(cx.gc.p, m, c, ldc, n, s);
    cx.zgemm(prec, aop, bop, long(m), long(n), long(k), 1.0, left, long(ar), cx.gb.p, long(br), cx.gc.p, long(m),
             cx.gc.p, long(m), add ? 1.0 : 0.0);
    copy_2d_d2h(c, ldc, cx.gc.p, m, n, s);
    cx.sync();
    cx.record(prec == Precision::FP64 ? "hip_fp64_dispatch" : "hip_fp32_dispatch", 0);
  });
}

int sdft_hip_zgemm_triple_right(sdft_hip_context* ctx, size_t m, size_t n, size_t k, const double* const a[3],
                                size_t lda, const double* b, size_t ldb, double* const c[3], size_t ldc,
                                int precision) {
  return guarded([&] {
    if (!ctx || !m || !n || !k || !b || lda < m || ldb < k || ldc < m || !a || !c)
      throw std::invalid_argument("HIP: invalid joint Ritz rotation dimensions");
    for (int i = 0; i < 3; ++i)
      if (!a[i] || !c[i]) throw std::invalid_argument("HIP: invalid joint Ritz rotation dimensions");
    check_int_range({m, n, k, lda, ldb, ldc});
    auto& cx = ctx->impl;
    const Precision prec = to_precision(precision);
    hipStream_t s = cx.stream();
    cx.gb.reserve(2 * k * n);
    copy_2d_h2d(cx.gb.p, k, b, ldb, n, s);
    cx.ga.reserve(2 * m * k);
    cx.gc.reserve(2 * m * n);
    for (int i = 0; i < 3; ++i) {
      copy_2d_h2d(cx.ga.p, m, a[i], lda, k, s);
      cx.zgemm(prec, 0, 0, long(m), long(n), long(k), 1.0, cx.ga.p, long(m), cx.gb.p, long(k), cx.gc.p, long(m),
               nullptr, 0, 0.0);
      copy_2d_d2h(c[i], ldc, cx.gc.p, m, n, s);
      cx.sync();
    }
  });
}

// src/compute_cuda.cu:141-177 semantics, sphere-restricted FFT pipeline
int sdft_hip_local_hamiltonian(sdft_hip_context* ctx, const int shape[3], const uint64_t* indices, size_t npw,
                               const double* kinetic, const double* potential, const double* input, size_t columns,
                               double* output, uint64_t operator_identity, int precision) {
  return guarded([&] {
    if (!ctx || !shape || !indices || !npw || !kinetic || !potential || !input || !output || !columns ||
        shape[0] <= 0 || shape[1] <= 0 || shape[2] <= 0)
      throw std::invalid_argument("HIP: invalid accelerated FFT dimensions");
    const std::size_t count = std::size_t(shape[0]) * shape[1] * shape[2];
    if (count > std::size_t(std::numeric_limits<int>::max()) || columns > std::size_t(std::numeric_limits<int>::max()))
      throw std::overflow_error("HIP: batched FFT dimensions exceed plan limits");
    for (std::size_t i = 0; i < npw; ++i)
      if (indices[i] >= count) throw std::invalid_argument("HIP: accelerated FFT plane-wave index out of range");
    auto& cx = ctx->impl;
    const Precision prec = to_precision(precision);
    hipStream_t s = cx.stream();
    auto& L = cx.local;
    const std::array<int, 3> shp{shape[0], shape[1], shape[2]};
    const bool fresh = !operator_identity || operator_identity != L.identity || L.shape != shp || L.npw != npw;
    if (fresh) {
      L.identity = 0;
      const auto plan = sdft_hip::build_grid_plan(shp, indices, npw);
      L.grid.upload(plan, s);
      L.vnat.upload(potential, count, s);
      L.v64.reserve(count);
      L.v32.reserve(count);
      L.grid.dev.v64 = L.v64.p;
      L.grid.dev.v32 = L.v32.p;
      check(sdft_hip::launch_potential_to_slots(L.grid.dev, L.vnat.p, L.v64.p, L.v32.p, s), "potential to slots");
      L.shape = shp;
      L.npw = npw;
      cx.record("gpu_fft_operator_upload", 0);
    } else {
      cx.record("gpu_fft_operator_upload_saved", 0, count + npw);
    }
    L.kinetic.upload(kinetic, npw, s);  // kinetic stays dynamic (src/compute_backend.hpp:43-44)
    L.in.upload(input, 2 * npw * columns, s);
    L.out.reserve(2 * npw * columns);
    const std::size_t per_col = std::size_t(L.grid.dev.nsticks) * shp[2] * 16;
    const std::size_t chunk = std::max<std::size_t>(1, std::min<std::size_t>(columns, (std::size_t(512) << 20) / per_col));
    L.t1.reserve(per_col * chunk);
    L.t2.reserve(per_col * chunk);
    for (std::size_t c0 = 0; c0 < columns; c0 += chunk) {
      const long c = long(std::min(chunk, columns - c0));
      const double* x = L.in.p + 2 * c0 * npw;
      double* y = L.out.p + 2 * c0 * npw;
      if (prec == Precision::FP64)
        check(sdft_hip::launch_local_h<double>(L.grid.dev, x, long(npw), c, L.kinetic.p, y, long(npw), L.t1.p, L.t2.p, s),
              "local H");
      else
        check(sdft_hip::launch_local_h<float>(L.grid.dev, x, long(npw), c, L.kinetic.p, y, long(npw), L.t1.p, L.t2.p, s),
              "local H");
    }
    check(hipMemcpyAsync(output, L.out.p, 16 * npw * columns, hipMemcpyDeviceToHost, s), "copy out");
    cx.sync();
    L.identity = operator_identity;
    cx.record(prec == Precision::FP64 ? "hip_fp64_fft_dispatch" : "hip_fp32_fft_dispatch", 0);
  });
}

// src/compute_cuda.cu:95-128 semantics: B resident per projector identity,
// D/Q per operator identity; input uploaded every call
int sdft_hip_paw_pair(sdft_hip_context* ctx, uint64_t identity, size_t npw, size_t np, size_t columns, const double* b,
                      const double* d, const double* q, const double* input, double* h, double* s_out,
                      uint64_t projector_identity, int precision, int* handled) {
  return guarded([&] {
    if (handled) *handled = 0;
    if (!ctx || !identity || !npw || !np || !columns || !b || !d || !q || !input || !h || !s_out)
      throw std::invalid_argument("HIP: invalid PAW chain shape");
    check_int_range({npw, np, columns});
    auto& cx = ctx->impl;
    const Precision prec = to_precision(precision);
    hipStream_t st = cx.stream();
    auto& P = cx.paw;
    if (!projector_identity) projector_identity = identity;
    const bool reuse_b = P.projector_identity == projector_identity && P.npw == npw && P.np == np;
    const bool reuse_d = reuse_b && P.identity == identity;
    if (!reuse_b) {
      P.projector_identity = 0;
      P.b.upload(b, 2 * np * npw, st);
      cx.record("gpu_paw_chain_static_miss", 0);
    } else {
      cx.record("gpu_paw_chain_static_hit", 0);
    }
    if (!reuse_d) {
      P.identity = 0;
      P.d.upload(d, 2 * np * np, st);
      P.q.upload(q, 2 * np * np, st);
      cx.record("gpu_paw_chain_density_miss", 0);
    } else {
      cx.record("gpu_paw_chain_density_hit", 0);
    }
    P.in.upload(input, 2 * npw * columns, st);
    P.p.reserve(2 * np * columns);
    P.dq.reserve(4 * np * columns);
    P.h.reserve(2 * npw * columns);
    P.s.reserve(2 * npw * columns);
    const long N = long(columns), K = long(npw), M = long(np);
    cx.zgemm(prec, 0, 0, M, N, K, 1.0, P.b.p, M, P.in.p, K, P.p.p, M, nullptr, 0, 0.0);
    cx.zgemm(prec, 0, 0, M, N, M, 1.0, P.d.p, M, P.p.p, M, P.dq.p, M, nullptr, 0, 0.0);
    cx.zgemm(prec, 0, 0, M, N, M, 1.0, P.q.p, M, P.p.p, M, P.dq.p + 2 * M * N, M, nullptr, 0, 0.0);
    // one back-projection GEMM over [DP | QP] (2 columns blocks)
    double* hs = P.h.p;
    cx.zgemm(prec, 1, 0, K, N, M, 1.0, P.b.p, M, P.dq.p, M, hs, K, nullptr, 0, 0.0);
    cx.zgemm(prec, 1, 0, K, N, M, 1.0, P.b.p, M, P.dq.p + 2 * M * N, M, P.s.p, K, nullptr, 0, 0.0);
    check(hipMemcpyAsync(h, P.h.p, 16 * npw * columns, hipMemcpyDeviceToHost, st), "copy h");
    check(hipMemcpyAsync(s_out, P.s.p, 16 * npw * columns, hipMemcpyDeviceToHost, st), "copy s");
    cx.sync();
    P.identity = identity;
    P.projector_identity = projector_identity;
    P.npw = npw;
    P.np = np;
    cx.record("gpu_paw_chain_dispatch", 0);
    if (handled) *handled = 1;
  });
}

int sdft_hip_enumerate_g_vectors(const double lattice[9], double cutoff_ev, const double kfrac[3],
                                 double ev_per_hartree, long capacity, int* hkl, double* kinetic_ev,
                                 double* cartesian, long* count) {
  return guarded([&] {
    if (!lattice || !kfrac || !count) throw std::invalid_argument("HIP: invalid G-vector arguments");
    std::array<sdft_hip::V3, 3> a{};
    for (int i = 0; i < 3; ++i) a[i] = {lattice[3 * i], lattice[3 * i + 1], lattice[3 * i + 2]};
    const auto g = sdft_hip::enumerate_g_vectors(sdft_hip::reciprocal_lattice(a), cutoff_ev,
                                                 {kfrac[0], kfrac[1], kfrac[2]}, ev_per_hartree);
    *count = long(g.size());
    if (long(g.size()) > capacity) return;
    for (std::size_t i = 0; i < g.size(); ++i) {
      if (hkl) { hkl[3 * i] = g[i].h; hkl[3 * i + 1] = g[i].k; hkl[3 * i + 2] = g[i].l; }
      if (kinetic_ev) kinetic_ev[i] = g[i].kinetic_ev;
      if (cartesian) {
        cartesian[3 * i] = g[i].cartesian.x;
        cartesian[3 * i + 1] = g[i].cartesian.y;
        cartesian[3 * i + 2] = g[i].cartesian.z;
      }
    }
  });
}

int sdft_hip_grid_indices(const int shape[3], size_t n, const int* hkl, uint64_t* indices) {
  return guarded([&] {
    if (!shape || !hkl || !indices || shape[0] <= 0 || shape[1] <= 0 || shape[2] <= 0)
      throw std::invalid_argument("HIP: invalid grid index arguments");
    const std::array<int, 3> s{shape[0], shape[1], shape[2]};
    for (std::size_t i = 0; i < n; ++i) indices[i] = sdft_hip::grid_index(s, hkl[3 * i], hkl[3 * i + 1], hkl[3 * i + 2]);
  });
}

int sdft_hip_operator_create(sdft_hip_context* ctx, const sdft_hip_operator_desc* desc, sdft_hip_operator** out) {
  return guarded([&] {
    if (!ctx || !desc || !out || !desc->n || !desc->hkl || !desc->kinetic_ev || (desc->nv && (!desc->v_hkl || !desc->v_values)) ||
        (desc->atoms && (!desc->rows || !desc->amplitudes || !desc->d || !desc->q)))
      throw std::invalid_argument("HIP: invalid operator description");
    sdft_hip::OperatorDesc d;
    d.n = desc->n;
    d.hkl = desc->hkl;
    d.kinetic_ev = desc->kinetic_ev;
    d.ev_per_hartree = desc->ev_per_hartree;
    d.nv = desc->nv;
    d.v_hkl = desc->v_hkl;
    d.v_values = desc->v_values;
    d.atoms = desc->atoms;
    d.rows = desc->rows;
    d.amplitudes = desc->amplitudes;
    d.d = desc->d;
    d.q = desc->q;
    d.regularization = desc->regularization;
    d.shape = {desc->shape[0], desc->shape[1], desc->shape[2]};
    *out = new sdft_hip_operator(ctx, d);
  });
}

int sdft_hip_operator_destroy(sdft_hip_operator* op) {
  return guarded([&] { delete op; });
}

int sdft_hip_operator_info(sdft_hip_operator* op, int shape[3], long* count, long* projectors, long* sticks,
                           uint64_t* indices, double* kinetic, double* local_real, double* h_diagonal,
                           double* s_diagonal) {
  return guarded([&] {
    if (!op) throw std::invalid_argument("HIP: null operator");
    auto& o = op->impl;
    if (shape)
      for (int a = 0; a < 3; ++a) shape[a] = o.shape()[a];
    if (count) *count = o.count();
    if (projectors) *projectors = long(o.np());
    if (sticks) *sticks = o.nsticks();
    if (indices) std::copy(o.indices().begin(), o.indices().end(), indices);
    if (kinetic) std::copy(o.kinetic().begin(), o.kinetic().end(), kinetic);
    if (h_diagonal) std::copy(o.h_diagonal().begin(), o.h_diagonal().end(), h_diagonal);
    if (s_diagonal) std::copy(o.s_diagonal().begin(), o.s_diagonal().end(), s_diagonal);
    if (local_real) {
      const auto v = o.local_real_natural();
      std::copy(v.begin(), v.end(), local_real);
    }
  });
}

int sdft_hip_operator_apply(sdft_hip_operator* op, int which, const double* x, size_t columns, double* hy, double* sy,
                            int precision, int device_pointers) {
  return guarded([&] {
    if (!op || !x || !columns || which < 0 || which > 2 || (which != 1 && !hy) || (which != 0 && !sy))
      throw std::invalid_argument("HIP: invalid plane-wave block operator buffers");
    auto& o = op->impl;
    const Precision prec = to_precision(precision);
    if (device_pointers) {
      o.apply(which, x, long(columns), hy, sy, prec);
      o.context().sync();
      return;
    }
    hipStream_t s = o.context().stream();
    const std::size_t total = 2 * o.n() * columns;
    DevBuf<double> dx, dh, ds;
    dx.upload(x, total, s);
    if (which != 1) dh.reserve(total);
    if (which != 0) ds.reserve(total);
    o.apply(which, dx.p, long(columns), dh.p, ds.p, prec);
    if (which != 1) check(hipMemcpyAsync(hy, dh.p, total * 8, hipMemcpyDeviceToHost, s), "copy h");
    if (which != 0) check(hipMemcpyAsync(sy, ds.p, total * 8, hipMemcpyDeviceToHost, s), "copy s");
    o.context().sync();
  });
}

int sdft_hip_operator_apply_shifted(sdft_hip_operator* op, const double* x, size_t columns, const double* theta,
                                    double* y, int precision) {
  return guarded([&] {
    if (!op || !x || !theta || !y || !columns || precision < 0 || precision > 2)
      throw std::invalid_argument("HIP: invalid shifted application arguments");
    auto& o = op->impl;
    hipStream_t s = o.context().stream();
    const std::size_t total = 2 * o.n() * columns;
    DevBuf<double> dx, dy, dt;
    dx.upload(x, total, s);
    dt.upload(theta, columns, s);
    dy.reserve(total);
    o.apply_shifted(dx.p, long(columns), dt.p, dy.p, precision ? Precision::FP32 : Precision::FP64, precision == 2);
    check(hipMemcpyAsync(y, dy.p, total * 8, hipMemcpyDeviceToHost, s), "copy y");
    o.context().sync();
  });
}

int sdft_hip_operator_apply_noalloc(sdft_hip_operator* op, int which, const double* x, size_t columns, double* hy,
                                    double* sy, int precision) {
  return guarded([&] {
    if (!op || !x || !columns || which < 0 || which > 2 || (which != 1 && !hy) || (which != 0 && !sy))
      throw std::invalid_argument("HIP: invalid plane-wave block operator buffers");
    static_assert(SDFT_HIP_SCRATCH_BYTES == sdft_hip::kScratchBytes, "scratch size of the C header");
    auto& o = op->impl;
    o.apply_noalloc(which, x, long(columns), hy, sy, to_precision(precision));
    o.context().sync();
  });
}

int sdft_hip_operator_benchmark(sdft_hip_operator* op, int which, size_t columns, int precision, int repetitions,
                                double* kernel_ms) {
  return guarded([&] {
    if (!op || !columns || repetitions <= 0 || !kernel_ms || which < 0 || which > 2)
      throw std::invalid_argument("HIP: invalid benchmark arguments");
    auto& o = op->impl;
    const Precision prec = to_precision(precision);
    hipStream_t s = o.context().stream();
    const std::size_t total = 2 * o.n() * columns;
    DevBuf<double> dx, dh, ds;
    dx.reserve(total);
    dh.reserve(total);
    ds.reserve(total);
    // deterministic input resident in HBM before timing
    std::vector<double> host(total);
    for (std::size_t i = 0; i < total; ++i) host[i] = std::sin(0.37 * double(i % 9973)) / (1.0 + double(i % 131));
    check(hipMemcpy(dx.p, host.data(), total * 8, hipMemcpyHostToDevice), "bench input");
    o.apply(which, dx.p, long(columns), dh.p, ds.p, prec);  // warm-up
    o.context().sync();
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    check(hipEventRecord(e0, s), "event record");
    for (int r = 0; r < repetitions; ++r) o.apply(which, dx.p, long(columns), dh.p, ds.p, prec);
    check(hipEventRecord(e1, s), "event record");
    check(hipEventSynchronize(e1), "event sync");
    float ms = 0;
    check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    kernel_ms[2] = double(ms) / repetitions;
    // local-H pipeline alone (which = 0 without projectors is not separable:
    // time the S-free path of H minus the projector chain via a local-only run)
    check(hipEventRecord(e0, s), "event record");
    const long chunk = long(o.max_chunk()), n = long(o.n());
    for (int r = 0; r < repetitions; ++r)
      for (long c0 = 0; c0 < long(columns); c0 += chunk) {
        const long c = std::min<long>(chunk, long(columns) - c0);
        void* t1 = o.stick_buffer(0, std::size_t(c), prec);
        void* t2 = o.stick_buffer(1, std::size_t(c), prec);
        if (prec == Precision::FP64)
          check(sdft_hip::launch_local_h<double>(o.grid_plan(), dx.p + 2 * c0 * n, n, c, o.kinetic_device(),
                                                 dh.p + 2 * c0 * n, n, t1, t2, s),
                "local H");
        else
          check(sdft_hip::launch_local_h<float>(o.grid_plan(), dx.p + 2 * c0 * n, n, c, o.kinetic_device(),
                                                dh.p + 2 * c0 * n, n, t1, t2, s),
                "local H");
      }
    check(hipEventRecord(e1, s), "event record");
    check(hipEventSynchronize(e1), "event sync");
    check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    kernel_ms[0] = double(ms) / repetitions;
    kernel_ms[1] = std::max(0.0, kernel_ms[2] - kernel_ms[0]);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

int sdft_hip_davidson(sdft_hip_operator* op, const sdft_hip_davidson_params* params, const double* initial,
                      sdft_hip_davidson_result* result) {
  return guarded([&] {
    if (!op || !params || !result || !result->eigenvalues || !result->residual_norms)
      throw std::invalid_argument("HIP: invalid Davidson arguments");
    sdft_hip::DavidsonParams p;
    p.roots = params->roots;
    p.max_iterations = params->max_iterations;
    p.tolerance = params->tolerance;
    p.subspace_factor = params->subspace_factor ? params->subspace_factor : 8;
    p.search_fp32 = params->search_precision;
    p.search_handoff = params->search_handoff > 0 ? params->search_handoff : 1e-8;
    p.stagnation_window = params->stagnation_window > 0 ? params->stagnation_window : 3;
    p.handoff_restart = params->handoff_restart;
    p.verbose = params->verbose;
    p.timing = params->timing;
    p.band_group_size = params->band_group_size;
    if (params->band_group_sweep != 0 && params->band_group_sweep != 1)
      throw std::invalid_argument("HIP: band_group_sweep must be 0 or 1");
    p.band_group_sweep = params->band_group_sweep;
    p.refinement = params->refinement;
    p.inner_iterations = params->inner_iterations;
    p.refinement_guards = params->refinement_guards;
    if (p.refinement != 0 && p.refinement != 1) throw std::invalid_argument("HIP: refinement must be 0 or 1");
    if (p.inner_iterations < 0) throw std::invalid_argument("HIP: inner iterations must be >= 0");
    if (p.refinement_guards < 0 || p.refinement_guards > params->roots)
      throw std::invalid_argument("HIP: refinement guards must be in 0 .. roots");
    p.vectors_out = reinterpret_cast<std::complex<double>*>(result->eigenvectors);
    if (params->device_pointers != 0 && params->device_pointers != 1)
      throw std::invalid_argument("HIP: device_pointers must be 0 or 1");
    p.device_pointers = params->device_pointers;
    if (params->bf16_projectors != 0 && params->bf16_projectors != 1)
      throw std::invalid_argument("HIP: bf16_projectors must be 0 or 1");
    p.bf16_projectors = params->bf16_projectors;
    if (p.search_fp32 != 0 && p.search_fp32 != 1) throw std::invalid_argument("HIP: search precision must be 0 or 1");
    const auto out = sdft_hip::davidson(op->impl, p, reinterpret_cast<const std::complex<double>*>(initial));
    std::copy(out.eigenvalues.begin(), out.eigenvalues.end(), result->eigenvalues);
    std::copy(out.residual_norms.begin(), out.residual_norms.end(), result->residual_norms);
    // eigenvectors went straight to result->eigenvectors (p.vectors_out)
    result->iterations = out.stats.iterations;
    result->converged = out.stats.converged ? 1 : 0;
    result->h_columns_search = out.stats.h_columns_search;
    result->h_columns_exact = out.stats.h_columns_exact;
    result->handoff_iteration = out.stats.handoff_iteration;
    result->seconds_total = out.stats.seconds_total;
    result->seconds_apply = out.stats.seconds_apply;
    result->seconds_gemm = out.stats.seconds_gemm;
    result->seconds_ritz = out.stats.seconds_ritz;
    result->seconds_ortho = out.stats.seconds_ortho;
    result->cholesky_blocks = out.stats.cholesky_blocks;
    result->rr_cholesky = out.stats.rr_cholesky;
    result->rr_stedc = out.stats.rr_stedc;
    result->rr_barrier_fallbacks = out.stats.rr_barrier_fallbacks;
    result->polish_accepted = out.stats.polish_accepted;
    result->polish_rejected = out.stats.polish_rejected;
    result->manifold_accepted = out.stats.manifold_accepted;
    result->manifold_rejected = out.stats.manifold_rejected;
    result->search_rejected = out.stats.search_rejected;
    result->refinement_steps = out.stats.refinement_steps;
  });
}

int sdft_hip_operator_device_bytes(sdft_hip_operator* op, size_t* bytes) {
  return guarded([&] {
    if (!op || !bytes) throw std::invalid_argument("HIP: invalid device-bytes arguments");
    *bytes = op->impl.device_bytes() + sdft_hip::davidson_cache_bytes(op->impl);
  });
}

int sdft_hip_batch_create(sdft_hip_context* ctx, size_t count, sdft_hip_operator* const* ops, size_t columns,
                          int precision, sdft_hip_batch** out) {
  return guarded([&] {
    if (!ctx || !ops || !count || !columns || !out) throw std::invalid_argument("HIP: invalid operator batch");
    std::vector<sdft_hip::DeviceOperator*> list;
    for (size_t k = 0; k < count; ++k) {
      if (!ops[k] || ops[k]->owner.ctx != ctx) throw std::invalid_argument("HIP: batch operators must share the context");
      list.push_back(&ops[k]->impl);
    }
    check_int_range({columns, count});
    *out = new sdft_hip_batch(ctx, list, long(columns), to_precision(precision));
  });
}

int sdft_hip_batch_destroy(sdft_hip_batch* batch) {
  return guarded([&] { delete batch; });
}

int sdft_hip_batch_offsets(sdft_hip_batch* batch, long* offsets) {
  return guarded([&] {
    if (!batch || !offsets) throw std::invalid_argument("HIP: invalid batch layout output");
    const auto& o = batch->impl.offsets();
    std::copy(o.begin(), o.end(), offsets);
  });
}

int sdft_hip_batch_apply(sdft_hip_batch* batch, const double* x, double* hy, double* sy) {
  return guarded([&] {
    if (!batch || !x || !hy || !sy) throw std::invalid_argument("HIP: invalid batch buffers");
    auto& b = batch->impl;
    auto& cx = batch->owner.ctx->impl;
    const std::size_t bytes = 16 * std::size_t(b.offsets().back());
    cx.upload_bulk(b.input(), x, bytes);
    b.replay();
    cx.download_bulk(hy, b.output_h(), bytes);
    cx.download_bulk(sy, b.output_s(), bytes);
  });
}

int sdft_hip_batch_benchmark(sdft_hip_batch* batch, int replays, double* milliseconds) {
  return guarded([&] {
    if (!batch || !milliseconds) throw std::invalid_argument("HIP: invalid batch benchmark output");
    *milliseconds = batch->impl.benchmark(replays);
  });
}

int sdft_hip_batch_plan(sdft_hip_batch* batch, int* threads, int* planes, int* groups, double* tuned_ms,
                        int* from_cache) {
  return guarded([&] {
    if (!batch) throw std::invalid_argument("HIP: null batch");
    const auto& l = batch->impl.plan();
    if (threads) *threads = l.threads;
    if (planes) *planes = l.planes;
    if (groups) *groups = l.groups;
    if (tuned_ms) *tuned_ms = batch->impl.tuned_ms();
    if (from_cache) *from_cache = batch->impl.plan_from_cache() ? 1 : 0;
  });
}

int sdft_hip_strict_step(sdft_hip_operator* op, int kind, size_t roots, double tolerance, double* x,
                         const double* direction, double* energies, double* residual_norms, int* accepted) {
  return guarded([&] {
    if (!op || !x || !energies || !residual_norms || !accepted)
      throw std::invalid_argument("invalid strict residual polish dimensions");
    using Cx = std::complex<double>;
    const auto r = sdft_hip::strict_step(op->impl, kind, long(roots), tolerance, reinterpret_cast<const Cx*>(x),
                                         reinterpret_cast<const Cx*>(direction));
    std::memcpy(x, r.x.data(), r.x.size() * sizeof(Cx));
    std::copy(r.energies.begin(), r.energies.end(), energies);
    std::copy(r.residual_norms.begin(), r.residual_norms.end(), residual_norms);
    *accepted = r.accepted ? 1 : 0;
  });
}

int sdft_hip_operator_project(sdft_hip_operator* op, const double* x, size_t columns, double* p, int precision) {
  return guarded([&] {
    if (!op || !x || !p || !columns || (precision != 0 && precision != 1))
      throw std::invalid_argument("HIP: invalid PAW projection buffers");
    auto& o = op->impl;
    hipStream_t s = o.context().stream();
    const size_t n = o.n(), np = o.np();
    sdft_hip::DevBuf<double> dx, dp;
    dx.upload(x, 2 * n * columns, s);
    dp.reserve(2 * std::max<size_t>(np, 1) * columns);
    o.project(dx.p, long(columns), dp.p, precision ? sdft_hip::Precision::FP32 : sdft_hip::Precision::FP64);
    sdft_hip::check(hipMemcpyAsync(p, dp.p, 16 * np * columns, hipMemcpyDeviceToHost, s), "download");
    o.context().sync();
  });
}

int sdft_hip_operator_projector_metric(sdft_hip_operator* op, const double* p, size_t columns, double* qp) {
  return guarded([&] {
    if (!op || !p || !qp || !columns) throw std::invalid_argument("HIP: invalid projector metric buffers");
    auto& o = op->impl;
    hipStream_t s = o.context().stream();
    const size_t np = o.np();
    sdft_hip::DevBuf<double> dp, dq;
    dp.upload(p, 2 * np * columns, s);
    dq.reserve(2 * std::max<size_t>(np, 1) * columns);
    o.projector_metric(dp.p, long(columns), dq.p);
    sdft_hip::check(hipMemcpyAsync(qp, dq.p, 16 * np * columns, hipMemcpyDeviceToHost, s), "download");
    o.context().sync();
  });
}

int sdft_hip_operator_real_space_search(sdft_hip_operator* op, double factor, const double* lattice,
                                        const double* fractional, const double* radius, size_t* points,
                                        double* retained) {
  return guarded([&] {
    if (!op) throw std::invalid_argument("HIP: null operator");
    const long kept = op->impl.enable_real_space_search(factor, lattice, fractional, radius, retained);
    if (points) *points = size_t(kept);
  });
}

int sdft_hip_operator_apply_real_space(sdft_hip_operator* op, int which, const double* x, size_t columns, double* y,
                                       int precision) {
  return guarded([&] {
    if (!op || !x || !y || !columns || (which != 0 && which != 1) || (precision != 0 && precision != 1))
      throw std::invalid_argument("HIP: invalid real-space search arguments");
    auto& o = op->impl;
    hipStream_t s = o.context().stream();
    const size_t n = o.n();
    sdft_hip::DevBuf<double> dx, dy;
    dx.upload(x, 2 * n * columns, s);
    dy.reserve(2 * n * columns);
    o.apply_real_space(which, dx.p, long(columns), dy.p,
                       precision ? sdft_hip::Precision::FP32 : sdft_hip::Precision::FP64);
    sdft_hip::check(hipMemcpyAsync(y, dy.p, 16 * n * columns, hipMemcpyDeviceToHost, s), "download");
    o.context().sync();
  });
}

int sdft_hip_operator_real_space_points(sdft_hip_operator* op, long* natural, long* per_atom) {
  return guarded([&] {
    if (!op) throw std::invalid_argument("HIP: null operator");
    std::vector<long> nat, counts;
    op->impl.real_space_points(nat, counts);
    if (natural) std::copy(nat.begin(), nat.end(), natural);
    if (per_atom) std::copy(counts.begin(), counts.end(), per_atom);
  });
}

int sdft_hip_operator_density(sdft_hip_operator* op, const double* coefficients, size_t bands,
                              const double* occupations, double volume, size_t nsupport, const int* support_hkl,
                              double* out) {
  return guarded([&] {
    if (!op || !coefficients || !occupations || !support_hkl || !out || !bands || !nsupport)
      throw std::invalid_argument("invalid plane-wave FFT density dimensions");
    auto& o = op->impl;
    const size_t n = o.n();
    for (size_t i = 0; i < 2 * n * bands; ++i)
      if (!std::isfinite(coefficients[i])) throw std::invalid_argument("nonfinite wavefunction coefficient");
    sdft_hip::DevBuf<double> dx;
    dx.upload(coefficients, 2 * n * bands, o.context().stream());
    const auto r = o.density(dx.p, long(bands), occupations, volume, support_hkl, long(nsupport));
    std::memcpy(out, r.data(), r.size() * sizeof(r[0]));
  });
}

int sdft_hip_snapshot_info_read(const char* path, sdft_hip_snapshot_info* info) {
  return guarded([&] {
    if (!path || !info) throw std::invalid_argument("HIP: invalid snapshot arguments");
    const sdft_hip::SnapshotArrays s = sdft_hip::read_snapshot(path);
    info->format = int(s.form);
    info->n = s.n;
    info->count = s.count;
    for (int a = 0; a < 3; ++a) info->shape[a] = s.shape[a];
    info->projectors = s.projectors;
    info->atoms = s.channels.size();
    info->roots = s.roots;
    info->iterations = s.iterations;
    info->initial_columns = s.initial.size() / s.n;
    info->regularization = s.regularization;
    info->tolerance = s.tolerance;
  });
}

int sdft_hip_snapshot_load(sdft_hip_context* ctx, const char* path, double* initial, sdft_hip_operator** out) {
  return guarded([&] {
    if (!ctx || !path || !out) throw std::invalid_argument("HIP: invalid snapshot arguments");
    const sdft_hip::SnapshotArrays s = sdft_hip::read_snapshot(path);
    auto op = std::make_unique<sdft_hip_operator>(ctx, s);
    if (initial && !s.initial.empty()) std::memcpy(initial, s.initial.data(), s.initial.size() * sizeof(s.initial[0]));
    *out = op.release();
  });
}

int sdft_hip_snapshot_save(sdft_hip_operator* op, const char* path, size_t roots, size_t iterations,
                           double tolerance, const double* initial) {
  return guarded([&] {
    if (!op || !path) throw std::invalid_argument("HIP: invalid snapshot arguments");
    const auto s = op->impl.snapshot(roots, iterations, tolerance, reinterpret_cast<const std::complex<double>*>(initial));
    sdft_hip::write_snapshot(path, s);
  });
}

int sdft_hip_hermitian_eigen(sdft_hip_context* ctx, size_t n, const double* a, size_t k, double* values,
                             double* vectors, int* fused) {
  return guarded([&] {
    if (!ctx || !a || !values || n == 0 || k > n || (k && !vectors))
      throw std::invalid_argument("HIP: invalid Hermitian eigensolve arguments");
    auto& cx = ctx->impl;
    hipStream_t s = cx.stream();
    sdft_hip::DevBuf<double> in, herm, vec;
    in.reserve(2 * n * n);
    herm.reserve(2 * n * n);
    vec.reserve(2 * n * std::max<size_t>(k, 1));
    sdft_hip::check(hipMemcpyAsync(in.p, a, 16 * n * n, hipMemcpyHostToDevice, s), "upload");
    sdft_hip::check(sdft_hip::launch_hermitise(in.p, long(n), long(n), herm.p, long(n), s), "hermitise");
    sdft_hip::HermitianEigensolver eig(cx, long(n));
    const std::vector<double> w = eig.solve(herm.p, long(n), long(k), vec.p);
    std::copy(w.begin(), w.end(), values);
    if (k) sdft_hip::check(hipMemcpyAsync(vectors, vec.p, 16 * n * k, hipMemcpyDeviceToHost, s), "download");
    cx.sync();
    if (fused)
      *fused = eig.launches_fused() > 0 && eig.barrier_fallbacks() == 0 ? (eig.stedc_solves() > 0 ? 2 : 1) : 0;
  });
}

int sdft_hip_davidson_seeds(sdft_hip_operator* op, size_t roots, double tolerance, const double* initial,
                            double* block, size_t* columns) {
  return guarded([&] {
    if (!op || !roots || !block || !columns || !(tolerance > 0)) throw std::invalid_argument("HIP: invalid seed arguments");
    auto& dev = op->impl;
    const long n = long(dev.n()), R = long(roots);
    if (R > n) throw std::invalid_argument("HIP: more roots than basis functions");
    hipStream_t s = dev.context().stream();
    const bool warm = initial != nullptr && 2 * R < n;
    const long cols = sdft_hip::davidson_seed_columns(n, R, warm);
    sdft_hip::DevBuf<double> init, out;
    if (warm) {
      init.reserve(2 * std::size_t(n) * R);
      sdft_hip::check(hipMemcpyAsync(init.p, initial, 16 * std::size_t(n) * R, hipMemcpyHostToDevice, s), "upload");
    }
    out.reserve(2 * std::size_t(n) * cols);
    sdft_hip::davidson_seed_block(dev, R, tolerance, warm ? init.p : nullptr, out.p, s);
    sdft_hip::check(hipMemcpyAsync(block, out.p, 16 * std::size_t(n) * cols, hipMemcpyDeviceToHost, s), "download");
    dev.context().sync();
    *columns = std::size_t(cols);
  });
}

int sdft_hip_tridiagonal_eigen(sdft_hip_context* ctx, size_t n, const double* d, const double* e, size_t k,
                               double* values, double* vectors, int* status) {
  return guarded([&] {
    if (!ctx || !d || (n > 1 && !e) || !values || !status || n == 0 || k > n || (k && !vectors) ||
        n > size_t(sdft_hip::kTridiagEigMaxOrder))
      throw std::invalid_argument("HIP: invalid tridiagonal eigensolve arguments");
    auto& cx = ctx->impl;
    hipStream_t s = cx.stream();
    sdft_hip::DevBuf<double> buf, z;
    buf.reserve(3 * n + 2);
    double* dd = buf.p;
    double* de = dd + n;
    double* dw = de + n;
    int* st = reinterpret_cast<int*>(dw + n);
    z.reserve(2 * n * std::max<size_t>(k, 1));
    sdft_hip::check(hipMemcpyAsync(dd, d, 8 * n, hipMemcpyHostToDevice, s), "upload");
    if (n > 1) sdft_hip::check(hipMemcpyAsync(de, e, 8 * (n - 1), hipMemcpyHostToDevice, s), "upload");
    sdft_hip::check(sdft_hip::launch_tridiag_values(int(n), dd, de, int(n), dw, st, s), "tridiagonal eigenvalues");
    sdft_hip::check(sdft_hip::launch_tridiag_vectors(int(n), dd, de, dw, int(k), z.p, long(n), st, s),
                    "tridiagonal eigenvectors");
    std::vector<double> zh(2 * n * k);
    sdft_hip::check(hipMemcpyAsync(values, dw, 8 * n, hipMemcpyDeviceToHost, s), "download");
    if (k) sdft_hip::check(hipMemcpyAsync(zh.data(), z.p, 16 * n * k, hipMemcpyDeviceToHost, s), "download");
    sdft_hip::check(hipMemcpyAsync(status, st, sizeof(int), hipMemcpyDeviceToHost, s), "download");
    cx.sync();
    for (size_t i = 0; i < n * k; ++i) vectors[i] = zh[2 * i];
  });
}

int sdft_hip_onsite_blocks(sdft_hip_context* ctx, size_t atoms, const long* rows, const double* d, const double* q,
                           const double* p, size_t columns, double* w, int precision) {
  return guarded([&] {
    if (!ctx || !atoms || !rows || !d || !q || !p || !w || !columns)
      throw std::invalid_argument("HIP: invalid onsite block arguments");
    auto& cx = ctx->impl;
    hipStream_t s = cx.stream();
    std::vector<int> offs, rws;
    std::vector<long> boffs;
    std::size_t np = 0, bo = 0;
    for (std::size_t a = 0; a < atoms; ++a) {
      if (rows[a] <= 0) throw std::invalid_argument("HIP: invalid onsite block size");
      offs.push_back(int(np));
      rws.push_back(int(rows[a]));
      boffs.push_back(long(bo));
      np += std::size_t(rows[a]);
      bo += std::size_t(rows[a]) * std::size_t(rows[a]);
    }
    DevBuf<int> doff, drow;
    DevBuf<long> dbo;
    DevBuf<double> dd, dq, dp, dw;
    doff.upload(offs.data(), offs.size(), s);
    drow.upload(rws.data(), rws.size(), s);
    dbo.upload(boffs.data(), boffs.size(), s);
    dd.upload(d, 2 * bo, s);
    dq.upload(q, 2 * bo, s);
    dp.upload(p, 2 * np * columns, s);
    dw.reserve(4 * np * columns);
    check(sdft_hip::launch_atom_blocks(to_precision(precision), int(np), int(columns), int(atoms), doff.p, drow.p, dbo.p,
                                       dd.p, dq.p, dp.p, long(np), dw.p, long(np), 3, s),
          "onsite blocks");
    check(hipMemcpyAsync(w, dw.p, 32 * np * columns, hipMemcpyDeviceToHost, s), "copy");
    cx.sync();
  });
}

}  // extern "C"
