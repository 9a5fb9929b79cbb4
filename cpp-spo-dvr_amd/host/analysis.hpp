// Host side of the Balint-Kurti analysis: the analysed channels, the dynamic specular analysis surface, the
// energy grid and the final S-matrix / probabilities from the FP64 time-to-energy sums, and the averages.
#pragma once

#include <optional>
#include <vector>

#include "config.hpp"
#include "gas_phase.hpp"
#include "grids.hpp"

namespace spo {

struct DiffractionChannel {
  int n, m, order, ix, iy;
};

// Analysed states is = irotana * n_dif + idif, layout of coefficients (is, v): is + n_states * v
// (src/analysis/analysis_basis.cpp:58-125).
struct AnalysisBasis {
  std::vector<DiffractionChannel> channels;  // all nx*ny, stable-sorted by order
  int n_diffraction = 0, max_order = 0;
  std::vector<int> rotations;  // propagation-basis index of each analysed rotational state
  int j_max = 0, n_v = 0;
  int initial = -1;
  std::vector<double> diffraction_energies;  // per channel
  std::vector<double> channel_energies;      // (is, v)
  std::vector<int> regular_rotations;
  std::vector<int> state_j, state_mj;  // per is

  int n_rotation() const { return static_cast<int>(rotations.size()); }
  int n_states() const { return n_rotation() * n_diffraction; }
  int state_rotation(int is) const { return is / n_diffraction; }
  int state_diffraction(int is) const { return is % n_diffraction; }
  std::vector<int> regular_state_order() const;
};

AnalysisBasis make_analysis_basis(const Config& config, const AngularGrid& angular, const RotationalBasis& basis,
                                  const GasPhase& gas, const std::vector<double>& xy_kinetic_energies);

// What a checkpoint keeps of the analysis surface (src/propagation/analysis_surface.hpp).
struct AnalysisSurfaceState {
  int index = 0, previous_index = 0;
  bool fixed = false;
  double norm_tolerance = 0.0;
};

// Dynamic analysis surface on the specular grid (src/propagation/analysis_surface.cpp:24-83).
class AnalysisSurface {
 public:
  AnalysisSurfaceState state() const { return {index_, previous_index_, fixed_, norm_tolerance_}; }
  // The state of a checkpoint (src/propagation/analysis_surface.cpp:85-91).
  void restore(const AnalysisSurfaceState& s) {
    index_ = s.index;
    previous_index_ = s.previous_index;
    fixed_ = s.fixed;
    norm_tolerance_ = s.norm_tolerance;
  }
  AnalysisSurface(const Config& config, const SpatialGrids& grids, const std::vector<cplx>& initial_specular);
  int index() const { return index_; }
  bool fixed() const { return fixed_; }
  std::optional<double> fixed_line() const { return fixed_ ? std::optional<double>(z_[index_]) : std::nullopt; }
  // Moves the surface for the specular wave (layout (ir, iz)); returns Z if it became fixed now.
  std::optional<double> update(const std::vector<cplx>& specular);

 private:
  std::vector<double> z_;
  int nr_;
  int iz_analysis_;
  int last_before_absorber_;
  double norm_tolerance_ = 0.0;
  bool fixed_;
  int index_;
  int previous_index_;
};

struct AnalysisEnergies {
  std::vector<double> collision, total, total_hartree;
  int size() const { return static_cast<int>(collision.size()); }
};
AnalysisEnergies make_analysis_energies(const Config& config, const GasPhase& gas);

// The energy-resolved sums' sites and energies (EnergyResolvedSites / EnergyResolvedEnergies,
// src/analysis/energy_resolved.cpp:36-58): each site's grid point, and E_tot = E_coll + E(vin, jin) formed as the
// S matrix's energies are.
struct EnergyResolvedSetup {
  std::vector<std::string> site_names;
  std::vector<int> site_ixy;
  std::vector<double> collision_ev;
  std::vector<double> total_hartree;
  double analysis_time_step = 0.0;
  bool enabled() const { return !site_ixy.empty() && !total_hartree.empty(); }
};
EnergyResolvedSetup make_energy_resolved_setup(const Config& config, const GasPhase& gas);

struct EnergyResult {
  double total_energy = 0.0, collision_energy = 0.0;
  std::vector<cplx> s_matrix;       // (is, v)
  std::vector<double> probability;  // (is, v)
  std::vector<double> phase;
  std::vector<unsigned char> open;
  std::vector<double> p_vibration;
  double p_total = 0.0, p_reaction = 0.0;
};

// S-matrix elements from the accumulated sums (is, v, ien) at one energy, then probabilities and sums
// (src/analysis/s_matrix.cpp:70-176). specular_line: Z of the fixed specular analysis line.
EnergyResult s_matrix_at(const Config& config, const AnalysisBasis& basis, const GasPhase& gas,
                         const AnalysisEnergies& energies, const std::vector<cplx>& sums, int ien,
                         std::optional<double> specular_line);

struct Averages {
  double time = 0.0, norm = 0.0, z_mean = 0.0, r_mean = 0.0, rest_norm = 0.0;
};
// src/propagation/averages.cpp:39-82 from the block densities (n_zr, izr order) and the specular wave.
Averages compute_averages(double time, const std::vector<double>& densities, const std::vector<cplx>& specular,
                          const Config& config, const SpatialGrids& grids, int surface_index);

}  // namespace spo
