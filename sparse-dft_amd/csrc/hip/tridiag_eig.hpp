// Eigenpairs of the real symmetric tridiagonal left by the tridiagonalisation
// (tridiag.hpp): every eigenvalue by Sturm-count multisection, the eigenvectors
// of the lowest k by inverse iteration.
#pragma once
#include <hip/hip_runtime_api.h>

namespace sdft_hip {

constexpr int kTridiagEigMaxOrder = 2048;

// Eigenvalues of T closer than this fraction of the Gershgorin bound of |T|
// form one cluster, whose vectors are orthogonalised against each other
// during the iteration.
constexpr double kTridiagClusterGap = 1e-7;

// d (n), e (n - 1): T. w (count): the lowest count eigenvalues, ascending to
// within the bisection tolerance (adjacent values of a degenerate group may
// differ by an ulp in either order). status: one int, set to 0.
hipError_t launch_tridiag_values(int n, const double* d, const double* e, int count, double* w, int* status,
                                 hipStream_t s);

// Eigenvectors of the lowest k eigenvalues w (from launch_tridiag_values on
// the same stream) into the real parts of the complex columns of z (ld ldz,
// imaginary parts zero). Unit 2-norm; vectors of one cluster are orthogonal,
// vectors of different clusters orthogonal to O(eps |T| / gap). *status is set
// to 1 if an iteration did not stay finite (the vectors are then invalid).
hipError_t launch_tridiag_vectors(int n, const double* d, const double* e, const double* w, int k, double* z,
                                  long ldz, int* status, hipStream_t s);

// Largest order of the register-resident back-transformation (16 rows per lane).
constexpr int kBackTransformMaxOrder = 1024;

// z (n x k complex, ld ldz) <- Q z for the Q of the tridiagonalisation
// (reflector tails refl, ld ldr, and tau as launch_hetrd leaves them): the
// product rocSOLVER zunmtr forms with side 'L', uplo 'L', no transpose.
hipError_t launch_back_transform(int n, int k, const double* refl, long ldr, const double* tau, double* z, long ldz,
                                 hipStream_t s);

}  // namespace sdft_hip
