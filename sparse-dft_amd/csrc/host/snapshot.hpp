// Solver snapshots (SDFTPWO1/2/3): the reference's wire format for a
// plane-wave operator plus a Davidson request, read and written on the host
// (src/plane_wave_operator.cpp:520-653). Little-endian uint64 header, IEEE754
// FP64 arrays, complex arrays interleaved and column-major.
#pragma once
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace sdft_hip {

struct SnapshotArrays {
  enum Form { Folded = 1, Dense = 2, Blocks = 3 };  // PWO1, PWO2, PWO3
  Form form = Blocks;
  std::uint64_t n = 0, count = 0, shape[3]{}, projectors = 0;
  std::uint64_t roots = 0, iterations = 0;
  double regularization = 0, tolerance = 0;
  std::vector<std::uint64_t> indices;                   // n, into the row-major grid
  std::vector<double> kinetic, h_diagonal, s_diagonal;  // n (Hartree)
  std::vector<double> local_real;                       // count, natural grid order
  std::vector<std::complex<double>> all_b;              // projectors x n
  std::vector<std::uint64_t> channels;                  // Blocks: projectors per atom
  // Blocks: per atom D then Q (channels^2 each, column-major), concatenated.
  // Dense: D, Q (projectors^2 each). Folded: D B, Q B (projectors x n each).
  std::vector<std::complex<double>> onsite_d, onsite_q;
  std::vector<std::complex<double>> initial;            // n x roots or empty
};

// Reads and validates a snapshot as the reference loader does (exact length
// before any allocation, dimension limits, finite arrays, Hermitian onsite
// matrices, unique in-range indices, positive overlap diagonal). Throws
// std::runtime_error with the reference's messages.
SnapshotArrays read_snapshot(const std::string& path);
// Writes path.tmp and renames it; refuses to overwrite an existing file.
void write_snapshot(const std::string& path, const SnapshotArrays& s);

}  // namespace sdft_hip
