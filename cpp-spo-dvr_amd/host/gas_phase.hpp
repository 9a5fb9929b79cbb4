// Rovibrational states of the free molecule (Fourier grid Hamiltonian on the r grid) and the open /
// analysed channel limits (src/gasphase/gas_phase.cpp:118-187).
#pragma once

#include <vector>

#include "config.hpp"
#include "grids.hpp"

namespace spo {

struct GasPhase {
  int nr = 0, j_max = 0;
  std::vector<double> energies;        // (j, v): j + (j_max + 1) * v
  std::vector<double> wave_functions;  // (ir, j, v)
  int v_max_open = 0, j_max_open = 0;
  int n_v_analysis = 0, j_max_analysis = 0;
  double initial_energy = 0.0;               // E(vin, jin), hartree
  std::vector<double> vibrational_energies;  // E(v,0) - E(0,0), v < n_v_analysis
  std::vector<double> rotational_energies;   // E(0,j) - E(0,0), j <= j_max

  double energy(int v, int j) const { return energies[j + (j_max + 1) * v]; }
  const double* wave_function(int v, int j) const {
    return wave_functions.data() + std::size_t(nr) * (j + (j_max + 1) * std::size_t(v));
  }
};

// Symmetric eigen-decomposition by cyclic Jacobi rotations: eigenvalues ascending, eigenvectors as columns
// (column k at vectors[k * n ...]).
void symmetric_eigen(std::vector<double> matrix, int n, std::vector<double>& values, std::vector<double>& vectors);

// Kinetic energy matrix of the Fourier grid Hamiltonian on n points spaced dx (closed form of the
// real-to-complex / complex-to-real transform of the identity, src/gasphase/fourier_grid_hamiltonian.cpp:15-40).
std::vector<double> fgh_kinetic_matrix(int n, double dx, double mass);

GasPhase compute_gas_phase(const Config& config, const AngularGrid& angular, const std::vector<double>& v_asymptotic);

}  // namespace spo
