#include "driver.hpp"

#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <string>

#include "checkpoint.hpp"

namespace spo {

std::vector<cplx> exp_potential_table(const PotentialGrid& grid, double dt) {
  std::vector<cplx> table(grid.values.size());
#pragma omp parallel for schedule(static)
  for (std::size_t i = 0; i < grid.values.size(); ++i) {
    const double a = -grid.values[i] * dt;
    table[i] = cplx(std::cos(a), std::sin(a));
  }
  return table;
}

std::shared_ptr<RunSetup> make_run_setup(const InputParameters& input, bool device_potential) {
  auto s = std::make_shared<RunSetup>();
  s->config = build_config(input);
  s->grids = make_spatial_grids(s->config);
  s->angular = make_angular_grid(s->config);
  s->basis = make_rotational_basis(s->angular, s->config.input.j_in, s->config.input.mj_in);
  const auto pes = make_pes(s->config.input);
  s->potential = device_potential ? compute_potential_grid_on_device(*pes, s->config, s->grids, s->angular)
                                  : compute_potential_grid(*pes, s->config, s->grids, s->angular);
  s->gas = compute_gas_phase(s->config, s->angular, s->potential.asymptotic);
  s->tables = make_propagation_tables(s->config, s->grids, s->basis, s->potential.asymptotic);
  s->analysis = make_analysis_basis(s->config, s->angular, s->basis, s->gas, s->tables.xy_kinetic_energies);
  s->energies = make_analysis_energies(s->config, s->gas);
  s->initial_specular = make_initial_specular_wave(s->config, s->grids, s->gas);
  return s;
}

Propagation::Propagation(std::shared_ptr<const RunSetup> setup, const BackendFactory& factory)
    : setup_(std::move(setup)),
      backend_(factory(setup_->backend_input())),
      surface_(setup_->config, setup_->grids, setup_->initial_specular) {
  const InputParameters& p = setup_->config.input;
  if (!p.new_calculation) restore(p.restart_dir);
  record_averages(p.t_end - p.n_steps * p.time_step);
  energy_resolved_ = make_energy_resolved_setup(setup_->config, setup_->gas);
  if (energy_resolved_.enabled())
    backend_->energy_resolved_start(energy_resolved_.site_ixy, static_cast<int>(energy_resolved_.total_hartree.size()),
                                    energy_resolved_.analysis_time_step);
}

// Restart (src/app/run.cpp:139-175, native files): the wave packet and analysis surface of the checkpoint, and the
// earlier run's BK records, added to the S-matrix sums and kept as the first records of this run.
void Propagation::restore(const std::string& dir) {
  const RunSetup& s = *setup_;
  const Config& c = s.config;
  if (!make_energy_resolved_setup(c, s.gas).site_ixy.empty())
    throw NotSupportedError("a restart with EnergyResolvedSites (the earlier run's sums) is not supported");
  const int n_phi = s.angular.n_phi();
  const std::filesystem::path d(dir);
  auto add_record = [&](double t, const std::vector<cplx>& bk) {
    backend_->add_earlier_analysis_step(t, bk);
    bk_.push_back(bk);
    bk_times_.push_back(t);
  };
  CheckpointInfo info;
  if (c.input.restart_format == "native") {
    backend_->read_wave_packet([&](WavePacket& psi) {
      info = read_checkpoint((d / kCheckpointFile).string(), c, n_phi, psi.scattering, psi.n_rot, psi.n_xy, psi.n_zr,
                             psi.specular);
    });
    surface_.restore(info.surface);
    const std::optional<double> last = read_bk_file((d / kBkFile).string(), c, n_phi, s.analysis, add_record);
    if (last && !tol::eq(*last, info.time))
      throw CheckpointError("the BK coefficient file in '" + dir + "' ends at t = " + std::to_string(*last) +
                            ", the checkpoint at t = " + std::to_string(info.time));
  } else {
    // Fortran files: fort.8 (else fort.9) and fort.22 (else fort.21); the wave function file has no time, the
    // BK records give it (src/app/run.cpp:158-170).
    auto first_existing = [&](const char* a, const char* b) {
      if (std::filesystem::exists(d / a)) return (d / a).string();
      if (std::filesystem::exists(d / b)) return (d / b).string();
      throw InputError("RestartDir", std::string("neither ") + a + " nor " + b + " in '" + dir + "'");
    };
    const AnalysisSurfaceState fresh = surface_.state();
    backend_->read_wave_packet([&](WavePacket& psi) {
      info.surface = read_fortran_wave_function(first_existing(kFortranWaveIn, kFortranWaveOut), psi.scattering,
                                                psi.n_rot, psi.n_xy, psi.n_zr, psi.specular, c.input.nr,
                                                c.input.nz_specular, fresh);
    });
    surface_.restore(info.surface);
    read_fortran_bk_file(first_existing(kFortranBkIn, kFortranBkOut), c, s.analysis, s.basis.size(), add_record);
    info.time = bk_times_.empty() ? 0.0 : bk_times_.back();
  }
  if (!tol::eq(info.time, c.input.t_start))
    throw CheckpointError("tStart = " + std::to_string(c.input.t_start) + " but the checkpoint in '" + dir +
                          "' is at t = " + std::to_string(info.time));
  earlier_records_ = static_cast<int>(bk_.size());
}

void Propagation::write_checkpoint(const std::string& dir) {
  const RunSetup& s = *setup_;
  const std::filesystem::path d(dir);
  std::filesystem::create_directories(d);
  const int n_phi = s.angular.n_phi();
  if (bk_.size() != bk_times_.size())
    throw std::logic_error("write_checkpoint needs every BK record (keep_bk_records was switched off)");
  const std::string& format = s.config.input.checkpoint_format;
  const WavePacket& psi = backend_->wave_packet_on_host();
  if (format != "fortran") {
    std::vector<std::pair<double, std::vector<cplx>>> records;
    for (std::size_t k = 0; k < bk_.size(); ++k) records.emplace_back(bk_times_[k], bk_[k]);
    write_bk_file((d / kBkFile).string(), s.config, n_phi, s.analysis, records);
    spo::write_checkpoint((d / kCheckpointFile).string(), s.config, n_phi, time(), psi.scattering, psi.n_rot,
                          psi.n_xy, psi.n_zr, psi.specular, surface_.state());
  }
  if (format != "native") {
    // Record 1 holds the records the run is meant to write, the earlier run's plus this run's analysis steps
    // (src/app/run.cpp:236-240); a reader takes exactly that many.
    FortranBkHeader h;
    h.n_records = earlier_records_ + analysis_steps_total();
    h.n_rotations = s.basis.size();
    h.initial_channel = s.basis.initial;
    h.initial_energy = s.gas.initial_energy;
    write_fortran_bk_file((d / kFortranBkOut).string(), s.config, s.analysis, h, bk_);
    write_fortran_wave_function((d / kFortranWaveOut).string(), psi.scattering, psi.n_rot, psi.n_xy, psi.n_zr,
                                psi.specular, surface_.state());
  }
}

void Propagation::energy_resolved(std::vector<cplx>* sums, std::vector<cplx>* specular_sums) {
  if (!energy_resolved_.enabled()) {
    sums->clear();
    specular_sums->clear();
    return;
  }
  backend_->energy_resolved_read(sums, specular_sums);
}

int Propagation::analysis_steps_total() const {
  return setup_->config.input.n_steps / setup_->config.derived.steps_per_analysis;
}

double Propagation::time() const {
  const InputParameters& p = setup_->config.input;
  const int step = done_ * setup_->config.derived.steps_per_analysis;
  return p.t_end + (step - p.n_steps) * p.time_step;
}

void Propagation::record_averages(double t) {
  averages_.push_back(compute_averages(t, backend_->block_densities(), backend_->specular(), setup_->config,
                                       setup_->grids, surface_.index()));
}

void Propagation::step() {
  backend_->propagate_analysis_step();
  finish_step();
}

void Propagation::run_batch(const std::vector<Propagation*>& runs, int max_steps) {
  // Graph mode (SPO_HIP_GRAPH=1): the runs' steps through one captured hipGraph per set of live runs (recaptured
  // when a run finishes); otherwise their blocks enqueued round robin at every step.
  const char* g = std::getenv("SPO_HIP_GRAPH");
  const bool graph = g && std::string(g) == "1";
  std::unique_ptr<BatchStepper> stepper;
  std::vector<Propagation*> stepper_runs;
  for (int i = 0; i < max_steps; ++i) {
    std::vector<Propagation*> live;
    std::vector<PropagationBackend*> backends;
    for (Propagation* p : runs)
      if (p->done_ < p->analysis_steps_total()) {
        live.push_back(p);
        backends.push_back(p->backend_.get());
      }
    if (live.empty()) return;
    if (graph) {
      if (!stepper || live != stepper_runs) {
        stepper.reset();
        stepper = make_batch_stepper(backends);
        stepper_runs = live;
      }
      stepper->step();
    } else {
      propagate_batch(backends);
    }
    for (Propagation* p : live) p->finish_step();
  }
}

void Propagation::finish_step() {
  done_ += 1;
  const double t = time();
  if (surface_.update(backend_->specular()) && !fixed_time_) fixed_time_ = t;
  std::vector<cplx> bk;
  backend_->analyse(surface_, t, keep_bk_ ? &bk : nullptr);
  if (keep_bk_) {
    bk_.push_back(std::move(bk));
    bk_times_.push_back(t);
  }
  if (energy_resolved_.enabled()) {
    // exp(i E_tot t) per energy, formed as the reference forms it (src/analysis/energy_resolved.cpp:67-68).
    std::vector<cplx> phase(energy_resolved_.total_hartree.size());
    for (std::size_t ien = 0; ien < phase.size(); ++ien)
      phase[ien] = std::exp(cplx(0.0, 1.0) * energy_resolved_.total_hartree[ien] * t);
    backend_->energy_resolved_add(phase);
  }
  if (done_ % setup_->config.derived.analyses_per_average == 0) record_averages(t);
}

void Propagation::run(int max_steps) {
  for (int i = 0; i < max_steps && done_ < analysis_steps_total(); ++i) step();
}

std::vector<EnergyResult> Propagation::results(std::optional<double> specular_line) const {
  const std::vector<cplx> sums = backend_->s_matrix_sums();
  std::vector<EnergyResult> out;
  for (int ien = 0; ien < setup_->energies.size(); ++ien)
    out.push_back(s_matrix_at(setup_->config, setup_->analysis, setup_->gas, setup_->energies, sums, ien,
                              specular_line));
  return out;
}

}  // namespace spo
