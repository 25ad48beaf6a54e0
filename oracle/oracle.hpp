// CPU oracle for the plane-wave H.psi / S.psi path and the block Davidson.
//
// TEST INFRASTRUCTURE ONLY. This is an independent FP64 restatement of the
// reference algorithm, used by tests/ as the parity checker, by
// __graft_entry__.smoke() as the checker of one small device call, and by
// bench.py as the timed CPU baseline. The product library never links it.
//
// Each function cites the reference file:line it restates.
#pragma once
#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

namespace orc {
using C = std::complex<double>;

struct Vec3 { double x{}, y{}, z{}; };
// src/sdft_core.cpp:18-20 (no FMA contraction: built with -ffp-contract=off)
double dot(Vec3 a, Vec3 b);
Vec3 cross(Vec3 a, Vec3 b);
double determinant(const std::array<Vec3, 3>& l);
// src/scf_core.cpp:169-175
std::array<Vec3, 3> reciprocal(const std::array<Vec3, 3>& a);

struct GVector { int h{}, k{}, l{}; Vec3 cartesian{}; double kinetic_ev{}; };
// src/pw_core.cpp:140-171
std::vector<GVector> enumerate_g_vectors(const std::array<Vec3, 3>& reciprocal,
                                         double cutoff_ev, Vec3 kfrac, double hartree_to_ev);
// src/fft_planning.hpp:16-23
int fast_fft_length(int minimum);

// Unnormalised mixed-radix complex FFT of one length (radices 2,3,4,5,7,
// generic DFT for other primes). sign=-1 is FFTW_FORWARD (e^{-i}), +1 is
// FFTW_BACKWARD (e^{+i}).
class Fft1d {
 public:
  explicit Fft1d(int n);
  int size() const { return n_; }
  // in-place on `data` (contiguous), `work` holds n values
  void exec(C* data, C* work, int sign) const;
 private:
  void rec(const C* in, C* out, int n, int stride, int fi, int sign, int tw_stride) const;
  int n_;
  std::vector<int> factors_;
  std::vector<C> tw_fwd_, tw_bwd_;  // e^{-2 pi i j/n}, e^{+2 pi i j/n}
};

// Row-major 3-D grid, index (x*N1+y)*N2+z, z fastest (src/plane_wave_operator.cpp:94-97)
class Fft3d {
 public:
  explicit Fft3d(std::array<int, 3> shape);
  void exec(C* grid, int sign) const;  // thread-safe, allocates its own line buffers
  std::size_t count() const { return count_; }
 private:
  std::array<int, 3> shape_;
  std::size_t count_;
  Fft1d f0_, f1_, f2_;
};

using FourierDensity = std::map<std::tuple<int, int, int>, C>;

struct AtomBlock {            // src/plane_wave_operator.hpp:10-19
  std::size_t projectors{};
  std::vector<C> amplitudes;  // row-major [projector][G]
  std::vector<C> d, q;        // row-major projectors x projectors
};

// The plane-wave operator of src/plane_wave_operator.cpp:29-516 (Impl), FP64.
struct Operator {
  std::size_t n{}, count{};
  std::array<int, 3> shape{};
  double regularization{};
  std::vector<std::size_t> indices;
  std::vector<double> kinetic, local_real, h_diagonal, s_diagonal;
  std::size_t np{};              // total projectors
  std::vector<C> all_b;          // column-major np x n (element (p,g) at p + g*np)
  std::vector<std::size_t> atom_offset, atom_rows;
  std::vector<std::vector<C>> d, q;  // per atom, column-major rows x rows (Hermitised)
  Fft3d* fft{};
  ~Operator();
};

// Constructor restated from src/plane_wave_operator.cpp:393-516 for a given
// grid shape (shape[a] <= 0 selects the reference grid rule).
Operator* make_operator(const std::vector<GVector>& basis, const FourierDensity& potential,
                        const std::vector<AtomBlock>& paw, double hartree_to_ev,
                        double regularization, std::array<int, 3> shape);
// Constructor from precomputed arrays (snapshot style): V(r) given directly.
Operator* make_operator_arrays(std::size_t n, std::array<int, 3> shape, const std::size_t* indices,
                               const double* kinetic, const double* local_real, double regularization,
                               std::size_t atoms, const std::size_t* rows, const C* all_b_colmajor,
                               const C* d_blocks, const C* q_blocks, double v0);

// src/plane_wave_operator.cpp:218-306 with exact=true on the CPU path.
void apply_block(const Operator& op, const C* x, std::size_t columns, C* y, bool overlap,
                 const C* shared_amplitudes = nullptr);
// src/plane_wave_operator.cpp:307-374 (joint H/S sharing P = B X)
void apply_pair(const Operator& op, const C* x, std::size_t columns, C* hy, C* sy);

enum class Op { Normal, Adjoint };
// column-major C = op(A) op(B) (+C) (src/compute_backend.cpp:159-254 semantics)
void zgemm(Op ao, Op bo, std::size_t m, std::size_t n, std::size_t k, const C* a, std::size_t lda,
           const C* b, std::size_t ldb, C* c, std::size_t ldc, bool add = false);
// src/paw_compensated_product.hpp:22-135 (bitwise-specified order)
void compensated_product(Op ao, std::size_t rows, std::size_t cols, std::size_t inner, const C* a,
                         std::size_t lda, const C* b, std::size_t ldb, C* c, std::size_t ldc, bool add);
// src/pw_core.cpp:33-47
double kinetic_preconditioner(double kinetic, double wavefunction_kinetic);

struct DavidsonResult {
  std::vector<double> eigenvalues, residual_norms;
  std::vector<C> eigenvectors;  // column-major n x roots
  bool converged{};
  std::size_t iterations{};
  std::size_t h_columns{};
};
// src/pw_core.cpp:532-1437 with the reference's default controls on the CPU
// backend (exact mode, block CGS2, joint H/S, kinetic preconditioner,
// incremental slabs, strict polish and root-manifold refinement).
DavidsonResult davidson(const Operator& op, std::size_t roots, std::size_t max_iterations,
                        double tolerance, const std::vector<C>& initial, std::size_t subspace_factor = 8);

// Binds CBLAS zgemm / LAPACKE zhegvd, zheevd (ILP64, numpy's OpenBLAS naming)
// from the shared library at `path`; an empty path restores the plain loops.
bool bind_blas(const char* path);
// OpenMP threads of the oracle loops and of the bound BLAS
void set_threads(int n);
bool lapack_hegvd(std::size_t n, C* a, C* b, double* w);  // false when unbound
bool lapack_heevd(std::size_t n, C* a, double* w);

// Dense generalized Hermitian eigenproblem A v = lambda B v (B HPD), all
// eigenpairs ascending; eigenvectors column-major B-orthonormal.
void hermitian_generalized_eigen(std::size_t n, const C* a, const C* b, double* values, C* vectors);
void hermitian_eigen(std::size_t n, const C* a, double* values, C* vectors);

}  // namespace orc
