// Fused persistent application of a batch of plane-wave operators that share
// one FFT grid shape (the k-points of one crystal): H psi and S psi of every
// (operator, column) work item in one kernel launch.
#pragma once
#include <hip/hip_runtime_api.h>

namespace sdft_hip {

// One operator of the batch (device pointers; T-precision arrays are FP64 or
// FP32 as the launch precision).
struct BatchOpDev {
  const int* stick_g = nullptr;   // nsticks * n2: sphere index at natural z, or -1
  const int* stick_xy = nullptr;  // nsticks: x * n1 + y
  const int* inverse = nullptr;   // n: stick-buffer position s * ls2 + z of sphere point g
  const double* kinetic = nullptr;
  const void* b = nullptr;        // np x n column-major (element (i, g) at i + g np)
  const void* bt = nullptr;       // n x np column-major (element (g, i) at g + i n), the same values
  const double* dblk = nullptr;   // per-atom D blocks, column-major, Hermitised
  const double* qblk = nullptr;
  const long* block_offset = nullptr;  // per atom, complex elements into dblk / qblk
  const int* atom_first = nullptr;     // per atom: first projector row
  const int* atom_rows = nullptr;      // per atom: projector rows
  const int* proj_atom = nullptr;      // per projector row: its atom
  const void* v = nullptr;        // V(r) in slot order [z slot][x slot][y slot]
  long x_offset = 0;              // complex offset of the operator's n x columns block
  int n = 0, nsticks = 0, np = 0;
  double s_diag = 1.0;            // S local part: 1 + regularisation
};

struct BatchLaunch {
  int n0 = 0, n1 = 0, n2 = 0;
  const void* tw[3] = {nullptr, nullptr, nullptr};  // per-axis twiddles at the launch precision
  int nsticks_max = 0, n_max = 0, np_max = 0;
  int columns = 0, ops = 0;
  int threads = 256, planes = 1, groups = 0;  // tuned launch shape
};

// Cubic 32^3 grids (the k-point meshes of the 8-atom cells) run kernels with
// the line length fixed at compile time (256, 512 or 1024 threads); other
// shapes run the runtime-length dispatch with 256 threads.
bool batch_fixed_length(const BatchLaunch& l);
// Dynamic LDS bytes of one workgroup; 0 when the shape does not fit.
size_t batch_lds_bytes(const BatchLaunch& l, bool fp64);
// hy, sy: same layout as x (every operator's block at its x_offset, ld n).
hipError_t launch_batch_apply(const BatchLaunch& l, const BatchOpDev* ops, bool fp64, const double* x, double* hy,
                              double* sy, hipStream_t s);
// out (n x np column-major) = in (np x n column-major) transposed, complex;
// fp64 selects double or float elements.
hipError_t launch_transpose_projectors(const void* in, int np, int n, bool fp64, void* out, hipStream_t s);

}  // namespace sdft_hip
