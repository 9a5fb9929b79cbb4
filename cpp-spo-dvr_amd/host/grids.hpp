// Coordinate, momentum and angular grids and the rotational (FBR) basis of the propagation.
#pragma once

#include <optional>
#include <vector>

#include "common.hpp"
#include "config.hpp"

namespace spo {

// x_i = start + i * spacing (src/grids/uniform_grid.hpp:12-31).
struct UniformGrid {
  double start = 0.0, spacing = 0.0;
  std::vector<double> points;
  UniformGrid() = default;
  UniformGrid(double s, double d, int n) : start(s), spacing(d), points(n) {
    for (int i = 0; i < n; ++i) points[i] = s + i * d;
  }
  int size() const { return static_cast<int>(points.size()); }
  double operator[](int i) const { return points[i]; }
};

// Kinetic energies 0.5/m (i dk)^2 in FFT order, dk = 2 pi / (n d) (src/grids/momentum_grid.cpp:9-22).
std::vector<double> fft_kinetic_energies(int n, double spacing, double mass);

struct SpatialGrids {
  UniformGrid z, z_specular, r, x, y;
  std::vector<double> ekin_z, ekin_z_specular, ekin_r;
  int iz_analysis = 0;
  double z_analysis = 0.0;
  int nz_potential = 0;         // points with Z < zasymp
  Range z_optical, r_optical;
  Range z_specular_optical;
};

// src/grids/spatial_grids.cpp:66-132 (grid placement decisions with the tolerant comparisons).
SpatialGrids make_spatial_grids(const Config& config);

// Smallest 2^h 3^i 5^j 7^k 11^m (h >= 1, i <= 2, j,k,m <= 1) times `multiple` that is >= n
// (src/fft/fft_sizes.cpp:10-34).
int nearest_fft_size(int n, int multiple = 1);

// Gauss quadrature for associated Legendre weight (src/grids/gauss_legendre.cpp:16-85): roots and weights.
void gauss_associated_legendre(double a, double b, int order, int mj, std::vector<double>& roots,
                               std::vector<double>& weights);
// sqrt(J + 1/2)-normalised P_J^m(cos theta), J = |m| .. j_max (src/grids/associated_legendre.cpp:20-47).
std::vector<double> associated_legendre(int j_max, int m, double theta);

// Half Gauss-Legendre theta grid for a homonuclear molecule and an FFT phi grid
// (src/grids/angular_grid.cpp:14-69).
struct AngularGrid {
  int j_max = 0, j_min = 0, delta_j = 1, mj_max = 0;
  bool homonuclear = true;
  double d_phi = 0.0;
  std::vector<double> thetas, theta_weights, phis;
  int n_theta() const { return static_cast<int>(thetas.size()); }
  int n_phi() const { return static_cast<int>(phis.size()); }
};
AngularGrid make_angular_grid(const Config& config);

// |J mJ> basis ordered mJ = -MJMax..MJMax (outer), J = Jmin..Jmax step dJ (inner), |mJ| <= J
// (src/grids/rotational_basis.cpp:11-26).
struct RotationalBasis {
  int mj_max = 0;
  int initial = -1;
  std::vector<int> j, mj;
  std::vector<Range> blocks;  // per mJ = -mj_max .. mj_max
  int size() const { return static_cast<int>(j.size()); }
};
RotationalBasis make_rotational_basis(const AngularGrid& grid, int j_in, int mj_in);

// The theta part of the angular transform: sqrt(w_t / nphi) P_J^m(theta_t) for every mJ block of the basis, ordered
// (mJ block, theta, J) (src/propagation/angular_transform.cpp:31-39).
std::vector<double> legendre_matrix(const AngularGrid& grid, const RotationalBasis& basis);

}  // namespace spo
