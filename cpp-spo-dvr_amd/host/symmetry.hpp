// Surface point groups acting on the (x, y, phi) grid: usable operations, (x, y) orbits with their
// representatives, and the phi index maps that let the potential be stored at representatives only.
#pragma once

#include <string>
#include <vector>

#include "config.hpp"

namespace spo {

// Rotation by beta = 2 pi k / 12 about the normal, or reflection in the line at beta / 2
// (src/symmetry/surface_operation.hpp:13-20).
struct SurfaceOp {
  bool mirror = false;
  int twelfths = 0;
  bool operator==(const SurfaceOp& o) const { return mirror == o.mirror && twelfths == o.twelfths; }
};

// Operations of a named group (identity, rotations by angle, mirrors by angle; src/symmetry/point_group.cpp:31-79).
std::vector<SurfaceOp> point_group_operations(const std::string& name);

// Diffraction order of channel (n, m) on the lattice (src/symmetry/surface_lattice.cpp:53-61).
int diffraction_order(const SystemParameters& system, int n, int m);

// Orbits of the (x, y) grid under the usable operations (src/symmetry/grid_symmetry.cpp:16-108).
struct GridSymmetry {
  int nx = 0, ny = 0, nphi = 0;
  std::string group = "none";
  std::vector<SurfaceOp> ops;
  std::vector<int> representatives;     // full-grid ixy of each representative
  std::vector<int> representative_of;   // per ixy
  std::vector<int> operation_of;        // per ixy: index into ops with op(ixy) = representative
  std::vector<std::vector<int>> phi_index;  // per op: image of phi_i
  int n_representatives() const { return static_cast<int>(representatives.size()); }
};

GridSymmetry make_identity_symmetry(int nx, int ny, int nphi);
GridSymmetry make_grid_symmetry(const Config& config, int nphi);

}  // namespace spo
