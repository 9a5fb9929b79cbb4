// Everything the split-operator step multiplies with, precomputed once in FP64 on the host
// (src/propagation/propagation_tables.hpp:20-59), and the initial wave packet on the specular grid.
#pragma once

#include <vector>

#include "config.hpp"
#include "gas_phase.hpp"
#include "grids.hpp"

namespace spo {

struct StepFactors {
  std::vector<cplx> full, half;
};

struct PropagationTables {
  StepFactors zr_kinetic;        // (ir, iz)
  StepFactors xy_kinetic;        // ixy = iy + ny * ix
  StepFactors specular_kinetic;  // (ir, iz) on the specular grid, with the parallel energy
  StepFactors rotation;          // (irot, ir)
  std::vector<double> xy_kinetic_energies;
  double parallel_energy = 0.0;
  std::vector<double> z_optical, r_optical, z_specular_optical;
  std::vector<cplx> gas_phase;  // exp(-i vasy(r) dt)
  std::vector<cplx> shift_forward, shift_backward;
  std::vector<cplx> specular_to_scattering, scattering_to_specular;
  double fft_scale_z = 0.0, fft_scale_r = 0.0, fft_scale_x = 0.0, fft_scale_y = 0.0;
  bool needs_shift = false;  // some shift_forward factor is not exactly 1
};

// src/propagation/propagation_tables.cpp:141-211.
PropagationTables make_propagation_tables(const Config& config, const SpatialGrids& grids, const RotationalBasis& basis,
                                          const std::vector<double>& v_asymptotic);

// Per-(irot, ir) merged factors of the potential step (src/propagation/potential_operator.cpp:15-26):
// before_interaction = optR rot(dt/2); gas_phase_step = optR rot(dt/2) exp(-i vasy dt) rot(dt/2).
struct PotentialStepFactors {
  std::vector<cplx> before_interaction, gas_phase_step;
};
PotentialStepFactors potential_step_factors(int n_rot, int nr, const PropagationTables& t);

// rot(dt/2) exp(-i vasy dt) rot(dt/2) of the initial channel per ir (src/propagation/specular_operator.cpp:11-19).
std::vector<cplx> specular_gas_phase_factors(int nr, int n_rot, int initial_channel, const PropagationTables& t);

// Gaussian in Z times phi_{vin jin}(r) on the specular grid, layout (ir, iz)
// (src/propagation/initial_wave_packet.cpp:42-92). Throws InputError (TolSp) if too much norm lies outside.
std::vector<cplx> make_initial_specular_wave(const Config& config, const SpatialGrids& grids, const GasPhase& gas);

}  // namespace spo
