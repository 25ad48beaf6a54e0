// Host-side G+k sphere (src/pw_core.hpp:11-19) used to build operator inputs.
#pragma once
#include <array>
#include <vector>

namespace sdft_hip {

struct V3 {
  double x{}, y{}, z{};
};
struct GVec {
  int h{}, k{}, l{};
  V3 cartesian{};
  double kinetic_ev{};
};

std::array<V3, 3> reciprocal_lattice(const std::array<V3, 3>& lattice);
std::vector<GVec> enumerate_g_vectors(const std::array<V3, 3>& reciprocal, double cutoff_ev, V3 kfrac,
                                      double ev_per_hartree);

}  // namespace sdft_hip
