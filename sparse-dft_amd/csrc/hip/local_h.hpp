// Device plan for the sphere-restricted FFT pipeline of the local Hamiltonian.
#pragma once
#include <hip/hip_runtime_api.h>

namespace sdft_hip {

struct GridPlanDev {
  int n0 = 0, n1 = 0, n2 = 0;    // grid shape, index (x*n1 + y)*n2 + z
  long count = 0;                 // n0*n1*n2
  int split[3] = {0, 0, 0};       // L1 of each axis (L = L1 * L2)
  int nsticks = 0;                // distinct (x,y) columns touched by the sphere
  const int* stick_g = nullptr;   // nsticks * n2: sphere index at natural z, or -1
  const int* stick_xy = nullptr;  // nsticks: x*n1 + y
  const void* tw64[3] = {nullptr, nullptr, nullptr};  // per axis (cos, sin) FP64
  const void* tw32[3] = {nullptr, nullptr, nullptr};  // per axis FP32
  const double* v64 = nullptr;    // V(r) in slot order [zs][xs][ys]
  const float* v32 = nullptr;
  // x-lines that hold sticks, as two ranges [start, start + lines) of natural
  // x: the y-transforms of the local H run on these lines only
  int xstart[2] = {0, 0}, xlines[2] = {0, 0};
};

// y(:,c) = T x(:,c) + FFT_fwd(V . FFT_bwd(x(:,c))) / N for c < columns.
// t1/t2: stick buffers of columns * nsticks * n2 complex values of T.
template <typename T>
hipError_t launch_local_h(const GridPlanDev& p, const double* x, long ldx, long columns, const double* kinetic,
                          double* y, long ldy, void* t1, void* t2, hipStream_t stream);

// y(:, j) += scale * theta[j] * x(:, j) for j < columns (theta on the device)
hipError_t launch_column_scaled_add(const double* x, long ldx, double* y, long ldy, long rows, long columns,
                                    const double* theta, double scale, hipStream_t stream);
hipError_t launch_scale_copy(const double* x, long ldx, double* y, long ldy, long rows, long columns, double scale,
                             hipStream_t stream);
hipError_t launch_potential_setup(const GridPlanDev& full, const double* grid, void* t1, void* t2, double* v64,
                                  float* v32, double* block_ratio, int ratio_blocks, hipStream_t stream);
hipError_t launch_potential_to_slots(const GridPlanDev& p, const double* v_nat, double* v64, float* v32,
                                     hipStream_t stream);
// Largest dynamic LDS of one workgroup on the current device.
int max_lds_bytes();

// Smooth density pieces: backward transform of columns into full slot-order
// planes (t2: columns x count complex), weighted |psi|^2 accumulation, the
// natural-order complex copy of a slot-order real field, and the gather of
// conj(backward transform) / count at Miller indices (the forward transform of
// a real field).
hipError_t launch_backward_planes(const GridPlanDev& p, const double* x, long ldx, long columns, void* t1, void* t2,
                                  hipStream_t stream);
// Backward transform of columns into full slot-order planes at precision T
// (t2: columns x count), and the forward transform from full slot-order planes
// onto the sphere: y = diag .* x + fwd(planes)[sphere] / N (t_sticks: stick buffer).
template <typename T>
hipError_t launch_backward_full(const GridPlanDev& p, const double* x, long ldx, long columns, void* t1, void* t2,
                                hipStream_t stream);
template <typename T>
hipError_t launch_forward_planes(const GridPlanDev& p, const void* planes, long columns, const double* x, long ldx,
                                 const double* diag, double* y, long ldy, void* t_sticks, hipStream_t stream);
hipError_t launch_density_accumulate(const void* planes, long count, int columns, const double* weights,
                                     double* density, hipStream_t stream);
hipError_t launch_slots_to_natural_complex(const GridPlanDev& p, const double* v_slots, double* grid,
                                           hipStream_t stream);
hipError_t launch_gather_support(const GridPlanDev& p, const void* planes, const int* hkl, long nsupport,
                                 double* out, hipStream_t stream);
hipError_t launch_potential_to_natural(const GridPlanDev& p, const double* v64, double* v_nat, hipStream_t stream);

}  // namespace sdft_hip
