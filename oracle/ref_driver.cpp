// C entry points over the REFERENCE's own compiled sources (pw_core.cpp,
// compute_backend.cpp, band_group_solver.cpp, rmm_trial.cpp,
// indefinite_pencil_davidson.cpp, sdft_core.cpp and the header-only
// paw_compensated_product.hpp), built by oracle/build_ref.sh into
// oracle/_ref/libsdft_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used to pin the CPU oracle and to write golden
// fixtures (tests/golden/make_golden.py). The reference's plane-wave operator
// needs FFTW, which this image lacks, so the Davidson is driven through C
// callbacks that apply the oracle's operator.
#include <complex>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "compute_backend.hpp"
#include "paw_compensated_product.hpp"
#include "pw_core.hpp"

namespace {
thread_local std::string g_error;
using C = std::complex<double>;
}  // namespace

extern "C" {
const char* ref_last_error() { return g_error.c_str(); }

// sdft::enumerate_g_vectors (src/pw_core.cpp:140-171) for given reciprocal rows.
int ref_enumerate_g(const double* reciprocal, double cutoff_ev, const double* kfrac, double ev_per_ha, long capacity,
                    int* hkl, double* kinetic_ev, long* count) {
  try {
    std::array<sdft::Vec3, 3> r{};
    for (int i = 0; i < 3; ++i) r[i] = {reciprocal[3 * i], reciprocal[3 * i + 1], reciprocal[3 * i + 2]};
    const auto g = sdft::enumerate_g_vectors(r, cutoff_ev, {kfrac[0], kfrac[1], kfrac[2]}, ev_per_ha);
    *count = long(g.size());
    if (long(g.size()) > capacity) return 0;
    for (std::size_t i = 0; i < g.size(); ++i) {
      hkl[3 * i] = g[i].h;
      hkl[3 * i + 1] = g[i].k;
      hkl[3 * i + 2] = g[i].l;
      kinetic_ev[i] = g[i].kinetic_ev;
    }
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

double ref_kinetic_preconditioner(double kinetic, double wk) { return sdft::kinetic_preconditioner(kinetic, wk); }

int ref_compensated_product(int aop, long rows, long cols, long inner, const double* a, long lda, const double* b,
                            long ldb, double* c, long ldc, int add) {
  try {
    sdft::compensated_paw_product(aop ? sdft::MatrixOp::Adjoint : sdft::MatrixOp::Normal, rows, cols, inner,
                                  reinterpret_cast<const C*>(a), lda, reinterpret_cast<const C*>(b), ldb,
                                  reinterpret_cast<C*>(c), ldc, add != 0);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

// which: 0 = H block, 1 = S block, 2 = joint pair; column-major n x cols.
typedef void (*ref_block_cb)(void* user, int which, const double* x, long cols, double* hy, double* sy);

// sdft::solve_generalized_hermitian_davidson (src/pw_core.cpp:532-1437) with a
// matrix-free operator whose callbacks are supplied by the caller.
int ref_davidson_callbacks(long n, const double* h_diag, const double* s_diag, const double* kinetic, ref_block_cb cb,
                           void* user, long roots, long max_iterations, double tolerance, const double* initial,
                           double* eigenvalues, double* eigenvectors, double* residuals, long* iterations,
                           int* converged, long group_size) {
  try {
    static int pair_token = 0;
    sdft::HermitianOperator h, s;
    h.dimension = s.dimension = std::size_t(n);
    h.diagonal.assign(h_diag, h_diag + n);
    s.diagonal.assign(s_diag, s_diag + n);
    if (kinetic) h.kinetic.assign(kinetic, kinetic + n);
    auto block = [cb, user](int which, const C* x, std::size_t cols, C* hy, C* sy) {
      cb(user, which, reinterpret_cast<const double*>(x), long(cols), reinterpret_cast<double*>(hy),
         reinterpret_cast<double*>(sy));
    };
    h.apply = [block](const std::vector<C>& x, std::vector<C>& y) { y.resize(x.size()); block(0, x.data(), 1, y.data(), nullptr); };
    s.apply = [block](const std::vector<C>& x, std::vector<C>& y) { y.resize(x.size()); block(1, x.data(), 1, nullptr, y.data()); };
    h.apply_block = [block, n](const std::vector<C>& x, std::size_t c, std::vector<C>& y, bool) {
      y.resize(std::size_t(n) * c);
      block(0, x.data(), c, y.data(), nullptr);
    };
    s.apply_block = [block, n](const std::vector<C>& x, std::size_t c, std::vector<C>& y, bool) {
      y.resize(std::size_t(n) * c);
      block(1, x.data(), c, nullptr, y.data());
    };
    h.apply_block_into = [block](const C* x, std::size_t c, C* y, bool) { block(0, x, c, y, nullptr); };
    s.apply_block_into = [block](const C* x, std::size_t c, C* y, bool) { block(1, x, c, nullptr, y); };
    h.pair_identity = s.pair_identity = &pair_token;
    h.apply_pair_into = [block](const C* x, std::size_t c, C* hy, C* sy, bool) { block(2, x, c, hy, sy); };
    std::vector<C> init;
    if (initial) init.assign(reinterpret_cast<const C*>(initial), reinterpret_cast<const C*>(initial) + n * roots);
    sdft::ComplexGeneralizedEigenResult r;
    if (group_size > 0) {
      // the reference's band-group entry reads its group size from the environment
      const std::string g = std::to_string(group_size);
      setenv("SDFT_BAND_GROUP_SIZE", g.c_str(), 1);
      r = sdft::solve_generalized_hermitian_band_groups(h, s, std::size_t(roots), std::size_t(max_iterations),
                                                        tolerance, init);
      unsetenv("SDFT_BAND_GROUP_SIZE");
    } else {
      r = sdft::solve_generalized_hermitian_davidson(h, s, std::size_t(roots), std::size_t(max_iterations), tolerance,
                                                     init);
    }
    std::memcpy(eigenvalues, r.eigenvalues.data(), sizeof(double) * roots);
    std::memcpy(residuals, r.residual_norms.data(), sizeof(double) * roots);
    if (eigenvectors) std::memcpy(eigenvectors, r.eigenvectors.data(), sizeof(C) * n * roots);
    *iterations = long(r.iterations);
    *converged = r.converged ? 1 : 0;
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}
}
