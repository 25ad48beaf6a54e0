// Real-space projector search operator kernels (see realspace.hip).
#pragma once
#include <hip/hip_runtime_api.h>

namespace sdft_hip {

// Device view of the retained real-space projector points.
struct RealSpaceDev {
  int atoms = 0;
  long total_points = 0, unique = 0;
  const void* w64 = nullptr;            // per atom rows x points complex (column-major), concatenated
  const void* w32 = nullptr;            // FP32 copy
  const long* w_offset = nullptr;       // atoms: complex offset of W_a
  const long* point_offset = nullptr;   // atoms + 1: offsets into `points`
  const int* atom_offset = nullptr;     // atoms: first projector row of the atom
  const int* atom_rows = nullptr;       // atoms
  const int* points = nullptr;          // total_points slot-order plane indices, atom by atom
  const int* unique_points = nullptr;   // unique slot indices
  const int* entry_start = nullptr;     // unique + 1
  const int* entry_atom = nullptr;      // per entry: atom, sorted by (point, atom)
  const int* entry_local = nullptr;     // per entry: index of the point within its atom
};

template <typename T>
hipError_t launch_rs_gather(const void* planes, long count, const int* points, long total, int columns, void* psi,
                            hipStream_t s);
template <typename T>
hipError_t launch_rs_project(const RealSpaceDev& rs, const void* psi, int columns, double* p, long ldp, hipStream_t s);
// planes <- V planes (H) or 0 (S), then + N W^H amp at the retained points
template <typename T>
hipError_t launch_rs_images(const RealSpaceDev& rs, void* planes, long count, int columns, const T* v_slots,
                            int overlap, const double* amp, long ldamp, hipStream_t s);
// setup helpers (FP64 planes of one atom's rows: rows x count)
hipError_t launch_rs_peak(const void* planes, long count, int rows, double* peak, hipStream_t s);
hipError_t launch_rs_values(const void* planes, long count, int rows, const int* points, long npoints, void* w64,
                            void* w32, hipStream_t s);

}  // namespace sdft_hip
