// The set-up of a run from its input (grids, potential, gas phase, tables, analysis basis) and the time loop
// that drives a backend through the analysis steps (src/propagation/propagation.cpp:9-95).
#pragma once

#include <memory>
#include <optional>
#include <vector>

#include "backend.hpp"

namespace spo {

struct RunSetup {
  Config config;
  SpatialGrids grids;
  AngularGrid angular;
  RotationalBasis basis;
  PotentialGrid potential;
  GasPhase gas;
  PropagationTables tables;
  AnalysisBasis analysis;
  AnalysisEnergies energies;
  std::vector<cplx> initial_specular;

  BackendInput backend_input() const {
    return {config, grids, angular, basis, tables, potential, gas, analysis, energies, initial_specular};
  }
};

// device_potential: evaluate the potential grid on the GPU (compute_potential_grid_on_device) instead of the host.
std::shared_ptr<RunSetup> make_run_setup(const InputParameters& input, bool device_potential = false);

// File names of the native restart files (src/app/run.cpp:67-68).
inline constexpr const char* kCheckpointFile = "checkpoint.bin";
inline constexpr const char* kBkFile = "bk_coefficients.bin";
// The Fortran program's restart files: written as fort.9 / fort.21, read as fort.8 / fort.22 (src/app/run.cpp:69-72).
inline constexpr const char* kFortranWaveOut = "fort.9";
inline constexpr const char* kFortranWaveIn = "fort.8";
inline constexpr const char* kFortranBkOut = "fort.21";
inline constexpr const char* kFortranBkIn = "fort.22";

class Propagation {
 public:
  Propagation(std::shared_ptr<const RunSetup> setup, const BackendFactory& factory);

  const RunSetup& setup() const { return *setup_; }
  PropagationBackend& backend() { return *backend_; }
  int analysis_steps_total() const;
  int analysis_steps_done() const { return done_; }
  double time() const;

  // One analysis step: propagate, move the analysis surface, project and accumulate; records the coefficients
  // and, every dtAve, the averages.
  void step();
  void run(int max_analysis_steps);
  // Up to max_analysis_steps analysis steps of every run not yet done, in lockstep: the propagations of one step
  // together (propagate_batch), then each run's analysis. Each run's results equal those of run().
  static void run_batch(const std::vector<Propagation*>& runs, int max_analysis_steps);

  // Averages at tStart (before any step) and every dtAve.
  const std::vector<Averages>& averages() const { return averages_; }
  // The BK coefficients of every analysis step: a restart's earlier records first.
  const std::vector<std::vector<cplx>>& bk_records() const { return bk_; }
  int earlier_records() const { return earlier_records_; }
  // The native checkpoint and BK coefficient files of the run so far (checkpoint.bin, bk_coefficients.bin in dir),
  // for a restart with NewCalc = false, RestartDir = dir, tStart = time().
  void write_checkpoint(const std::string& dir);
  void keep_bk_records(bool keep) { keep_bk_ = keep; }
  const AnalysisSurface& surface() const { return surface_; }
  std::optional<double> surface_fixed_time() const { return fixed_time_; }
  std::vector<EnergyResult> results(std::optional<double> specular_line) const;
  std::vector<EnergyResult> results() const { return results(surface_.fixed_line()); }

  // The energy-resolved sums so far (empty when EnergyResolvedSites is not set): [site, energy, Z, r, rot] and
  // [energy, Z_specular, r].
  const EnergyResolvedSetup& energy_resolved_setup() const { return energy_resolved_; }
  void energy_resolved(std::vector<cplx>* sums, std::vector<cplx>* specular_sums);

 private:
  // the analysis part of step(), after the backend has propagated
  void finish_step();
  void record_averages(double time);
  void restore(const std::string& dir);

  std::shared_ptr<const RunSetup> setup_;
  std::unique_ptr<PropagationBackend> backend_;
  AnalysisSurface surface_;
  int done_ = 0;
  bool keep_bk_ = true;
  std::vector<Averages> averages_;
  std::vector<std::vector<cplx>> bk_;
  std::vector<double> bk_times_;
  int earlier_records_ = 0;
  std::optional<double> fixed_time_;
  EnergyResolvedSetup energy_resolved_;
};

}  // namespace spo
