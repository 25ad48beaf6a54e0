// Single-launch Householder tridiagonalisation of a Hermitian matrix (LAPACK
// zhetrd, uplo = 'L', conventions) for the Rayleigh-Ritz eigensolve.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace sdft_hip {

constexpr int kHetrdMaxOrder = 2048;

// Rows per workgroup and workgroup count for order n (one workgroup per CU at
// most, the rows of a workgroup resident in LDS); false when n does not fit.
bool hetrd_plan(int n, int* rows_per_group, int* groups);
size_t hetrd_workspace_bytes(int n);

// a: n x n Hermitian (both triangles, exactly Hermitian), column-major, lda.
// Outputs: d (n), e (n-1) real; tau (n-1 complex); reflector tails in refl
// (column k, rows k+2..n-1; other entries untouched). refl must not alias a.
// `workspace` holds hetrd_workspace_bytes(n); after the launch completes the
// int at hetrd_abort_flag(workspace, n) is nonzero if the grid barrier expired
// (the outputs are then invalid). force_expiry raises that flag before the
// launch (fault injection: the workgroups leave at their first wait).
hipError_t launch_hetrd(const double* a, long lda, int n, double* refl, long ldr, double* d, double* e, double* tau,
                        void* workspace, hipStream_t s, bool force_expiry = false);
const int* hetrd_abort_flag(const void* workspace, int n);

}  // namespace sdft_hip
