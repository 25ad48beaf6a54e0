// G+k sphere enumeration and grid index map on the host (setup work of the
// plane-wave path). Built with -ffp-contract=off so that the FP64 sort key
// rounds as the reference's does on x86-64: the key decides the order inside
// degenerate |G+k| shells and therefore the coefficient order of every block.
#include "pw_basis.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace sdft_hip {
namespace {
constexpr double kPi = 3.1415926535897932384626433832795;
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double determinant(const std::array<V3, 3>& l) { return dot(l[0], cross(l[1], l[2])); }
}  // namespace

// src/scf_core.cpp:169-175
std::array<V3, 3> reciprocal_lattice(const std::array<V3, 3>& a) {
  const double volume = determinant(a);
  if (!std::isfinite(volume) || std::abs(volume) < 1e-12) throw std::invalid_argument("HIP: singular lattice");
  const double f = 2.0 * kPi / volume;
  return {scale(cross(a[1], a[2]), f), scale(cross(a[2], a[0]), f), scale(cross(a[0], a[1]), f)};
}

// src/pw_core.cpp:140-171
std::vector<GVec> enumerate_g_vectors(const std::array<V3, 3>& r, double cutoff_ev, V3 kf, double ev_per_ha) {
  if (cutoff_ev <= 0.0 || !std::isfinite(cutoff_ev))
    throw std::invalid_argument("HIP: plane-wave cutoff must be finite and positive");
  if (ev_per_ha <= 0 || !std::isfinite(ev_per_ha)) throw std::invalid_argument("HIP: invalid Hartree conversion");
  const double gmax = std::sqrt(2.0 * cutoff_ev / ev_per_ha);
  const double det = determinant(r);
  if (std::abs(det) < 1e-12 || !std::isfinite(det)) throw std::invalid_argument("HIP: reciprocal lattice is singular");
  const V3 d0 = scale(cross(r[1], r[2]), 1.0 / det);
  const V3 d1 = scale(cross(r[2], r[0]), 1.0 / det);
  const V3 d2 = scale(cross(r[0], r[1]), 1.0 / det);
  const int hmax = static_cast<int>(std::ceil(gmax * std::sqrt(dot(d0, d0)) + std::abs(kf.x))) + 1;
  const int kmax = static_cast<int>(std::ceil(gmax * std::sqrt(dot(d1, d1)) + std::abs(kf.y))) + 1;
  const int lmax = static_cast<int>(std::ceil(gmax * std::sqrt(dot(d2, d2)) + std::abs(kf.z))) + 1;
  const V3 shift = add(add(scale(r[0], kf.x), scale(r[1], kf.y)), scale(r[2], kf.z));
  std::vector<GVec> out;
  const double limit = cutoff_ev / ev_per_ha + 1e-12;
  for (int h = -hmax; h <= hmax; ++h)
    for (int k = -kmax; k <= kmax; ++k)
      for (int l = -lmax; l <= lmax; ++l) {
        const V3 g = add(add(scale(r[0], h), scale(r[1], k)), scale(r[2], l));
        const V3 s = add(g, shift);
        const double g2 = dot(s, s);
        if (0.5 * g2 <= limit) out.push_back({h, k, l, g, 0.5 * g2 * ev_per_ha});
      }
  std::sort(out.begin(), out.end(), [](const GVec& a, const GVec& b) {
    if (a.kinetic_ev != b.kinetic_ev) return a.kinetic_ev < b.kinetic_ev;
    if (a.h != b.h) return a.h < b.h;
    if (a.k != b.k) return a.k < b.k;
    return a.l < b.l;
  });
  return out;
}

}  // namespace sdft_hip
