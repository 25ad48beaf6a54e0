// C entry points of the CPU oracle for ctypes (tests/, smoke, bench CPU leg).
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <cstring>
#include <exception>
#include <string>

#include "oracle.hpp"

namespace {
thread_local std::string g_error;
template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return 2;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}
using orc::C;
}  // namespace

extern "C" {
const char* orc_last_error() { return g_error.c_str(); }

// lattice: 3 real-space vectors (row i = a_i, bohr). Output ordered G-sphere.
// Returns count via *count; if count > capacity nothing but the count is written.
int orc_enumerate_g(const double* lattice, double cutoff_ev, const double* kfrac, double ev_per_ha,
                    long capacity, int* hkl, double* kinetic_ev, double* cartesian, long* count) {
  return guarded([&] {
    std::array<orc::Vec3, 3> a{};
    for (int i = 0; i < 3; ++i) a[i] = {lattice[3 * i], lattice[3 * i + 1], lattice[3 * i + 2]};
    const auto b = orc::reciprocal(a);
    const auto g = orc::enumerate_g_vectors(b, cutoff_ev, {kfrac[0], kfrac[1], kfrac[2]}, ev_per_ha);
    *count = long(g.size());
    if (long(g.size()) > capacity) return;
    for (std::size_t i = 0; i < g.size(); ++i) {
      hkl[3 * i] = g[i].h;
      hkl[3 * i + 1] = g[i].k;
      hkl[3 * i + 2] = g[i].l;
      kinetic_ev[i] = g[i].kinetic_ev;
      cartesian[3 * i] = g[i].cartesian.x;
      cartesian[3 * i + 1] = g[i].cartesian.y;
      cartesian[3 * i + 2] = g[i].cartesian.z;
    }
  });
}

int orc_reciprocal(const double* lattice, double* out) {
  return guarded([&] {
    std::array<orc::Vec3, 3> a{};
    for (int i = 0; i < 3; ++i) a[i] = {lattice[3 * i], lattice[3 * i + 1], lattice[3 * i + 2]};
    const auto b = orc::reciprocal(a);
    for (int i = 0; i < 3; ++i) { out[3 * i] = b[i].x; out[3 * i + 1] = b[i].y; out[3 * i + 2] = b[i].z; }
  });
}

int orc_fast_fft_length(int minimum) { return orc::fast_fft_length(minimum); }

int orc_bind_blas(const char* path) {
  return guarded([&] {
    if (!orc::bind_blas(path)) throw std::runtime_error(std::string("cannot bind BLAS from ") + (path ? path : ""));
  });
}

int orc_set_threads(int n) {
  return guarded([&] { orc::set_threads(n); });
}

// In-place unnormalised 3-D FFT of a row-major grid (sign -1 forward, +1 backward).
int orc_fft3d(const int* shape, double* grid, int sign) {
  return guarded([&] {
    orc::Fft3d f({shape[0], shape[1], shape[2]});
    f.exec(reinterpret_cast<C*>(grid), sign);
  });
}

// Operator from a G basis (hkl, kinetic_ev), a Fourier potential list and
// per-atom blocks (row-major amplitudes/D/Q concatenated atom by atom).
int orc_operator_create(long n, const int* hkl, const double* kinetic_ev, double ev_per_ha, long nv,
                        const int* v_hkl, const double* v_values, long atoms, const long* rows,
                        const double* amplitudes, const double* d, const double* q, double regularization,
                        const int* shape, void** handle) {
  return guarded([&] {
    std::vector<orc::GVector> basis(n);
    for (long i = 0; i < n; ++i) basis[i] = {hkl[3 * i], hkl[3 * i + 1], hkl[3 * i + 2], {}, kinetic_ev[i]};
    orc::FourierDensity v;
    for (long i = 0; i < nv; ++i) v[{v_hkl[3 * i], v_hkl[3 * i + 1], v_hkl[3 * i + 2]}] = C(v_values[2 * i], v_values[2 * i + 1]);
    std::vector<orc::AtomBlock> paw(atoms);
    const C* amp = reinterpret_cast<const C*>(amplitudes);
    const C* dd = reinterpret_cast<const C*>(d);
    const C* qq = reinterpret_cast<const C*>(q);
    std::size_t ao = 0, mo = 0;
    for (long a = 0; a < atoms; ++a) {
      const std::size_t r = rows[a];
      paw[a].projectors = r;
      paw[a].amplitudes.assign(amp + ao, amp + ao + r * n);
      paw[a].d.assign(dd + mo, dd + mo + r * r);
      paw[a].q.assign(qq + mo, qq + mo + r * r);
      ao += r * n;
      mo += r * r;
    }
    *handle = orc::make_operator(basis, v, paw, ev_per_ha, regularization, {shape[0], shape[1], shape[2]});
  });
}

int orc_operator_destroy(void* handle) {
  delete static_cast<orc::Operator*>(handle);
  return 0;
}

// shape[3], count, np; arrays may be null.
int orc_operator_info(void* handle, int* shape, long* count, long* np, unsigned long long* indices,
                      double* kinetic, double* local_real, double* h_diag, double* s_diag, double* all_b) {
  return guarded([&] {
    const auto& op = *static_cast<orc::Operator*>(handle);
    for (int a = 0; a < 3; ++a) shape[a] = op.shape[a];
    *count = long(op.count);
    *np = long(op.np);
    if (indices)
      for (std::size_t i = 0; i < op.n; ++i) indices[i] = op.indices[i];
    if (kinetic) std::memcpy(kinetic, op.kinetic.data(), op.n * sizeof(double));
    if (local_real) std::memcpy(local_real, op.local_real.data(), op.count * sizeof(double));
    if (h_diag) std::memcpy(h_diag, op.h_diagonal.data(), op.n * sizeof(double));
    if (s_diag) std::memcpy(s_diag, op.s_diagonal.data(), op.n * sizeof(double));
    if (all_b) std::memcpy(all_b, op.all_b.data(), op.all_b.size() * sizeof(C));
  });
}

// which: 0 = H, 1 = S, 2 = joint pair (hy and sy)
int orc_operator_apply(void* handle, int which, const double* x, long columns, double* hy, double* sy) {
  return guarded([&] {
    const auto& op = *static_cast<orc::Operator*>(handle);
    const C* in = reinterpret_cast<const C*>(x);
    if (which == 2) orc::apply_pair(op, in, columns, reinterpret_cast<C*>(hy), reinterpret_cast<C*>(sy));
    else orc::apply_block(op, in, columns, reinterpret_cast<C*>(which == 0 ? hy : sy), which == 1);
  });
}

int orc_davidson(void* handle, long roots, long max_iterations, double tolerance, const double* initial,
                 long subspace_factor, double* eigenvalues, double* eigenvectors, double* residual_norms,
                 long* iterations, int* converged, long* h_columns) {
  return guarded([&] {
    const auto& op = *static_cast<orc::Operator*>(handle);
    std::vector<C> init;
    if (initial) init.assign(reinterpret_cast<const C*>(initial), reinterpret_cast<const C*>(initial) + op.n * roots);
    const auto r = orc::davidson(op, roots, max_iterations, tolerance, init, subspace_factor);
    std::memcpy(eigenvalues, r.eigenvalues.data(), roots * sizeof(double));
    std::memcpy(residual_norms, r.residual_norms.data(), roots * sizeof(double));
    if (eigenvectors) std::memcpy(eigenvectors, r.eigenvectors.data(), r.eigenvectors.size() * sizeof(C));
    *iterations = long(r.iterations);
    *converged = r.converged ? 1 : 0;
    if (h_columns) *h_columns = long(r.h_columns);
  });
}

int orc_zgemm(int aop, int bop, long m, long n, long k, const double* a, long lda, const double* b, long ldb,
              double* c, long ldc, int add) {
  return guarded([&] {
    orc::zgemm(aop ? orc::Op::Adjoint : orc::Op::Normal, bop ? orc::Op::Adjoint : orc::Op::Normal, m, n, k,
               reinterpret_cast<const C*>(a), lda, reinterpret_cast<const C*>(b), ldb, reinterpret_cast<C*>(c), ldc,
               add != 0);
  });
}

int orc_compensated_product(int aop, long rows, long cols, long inner, const double* a, long lda, const double* b,
                            long ldb, double* c, long ldc, int add) {
  return guarded([&] {
    orc::compensated_product(aop ? orc::Op::Adjoint : orc::Op::Normal, rows, cols, inner,
                             reinterpret_cast<const C*>(a), lda, reinterpret_cast<const C*>(b), ldb,
                             reinterpret_cast<C*>(c), ldc, add != 0);
  });
}

int orc_kinetic_preconditioner(double kinetic, double wavefunction_kinetic, double* out) {
  return guarded([&] { *out = orc::kinetic_preconditioner(kinetic, wavefunction_kinetic); });
}

int orc_hermitian_generalized_eigen(long n, const double* a, const double* b, double* values, double* vectors) {
  return guarded([&] {
    orc::hermitian_generalized_eigen(n, reinterpret_cast<const C*>(a), reinterpret_cast<const C*>(b), values,
                                     reinterpret_cast<C*>(vectors));
  });
}
}
