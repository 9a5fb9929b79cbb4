// A backend decorator that counts the calls the time loop makes across the seam, after the reference's
// CountingBackend (tests/counting_backend.hpp:35-113, used by tests/test_time_loop_calls.cpp:29-113 to show that the
// loop makes exactly the calls its design lists). Every run is made through it (an increment per call), so the
// counts of a run are always there to read (spo_run_call_counts).
#pragma once

#include <memory>
#include <utility>

#include "backend.hpp"

namespace spo {

// The calls of one run, in the order of spo_run_call_counts.
struct SeamCallCounts {
  long propagate_analysis_step = 0;  // including the steps propagated in a batch
  long kinetic_and_potential = 0;
  long z_plane = 0;
  long block_densities = 0;
  long specular = 0;
  long wave_packet_on_host = 0;
  long read_wave_packet = 0;
  long site_column = 0;
  long analyse = 0;
  long s_matrix_sums = 0;
  long add_earlier_analysis_step = 0;
  long energy_resolved_add = 0;
  static constexpr int kCount = 12;
  void to_array(long* out) const {
    const long v[kCount] = {propagate_analysis_step, kinetic_and_potential, z_plane, block_densities, specular,
                            wave_packet_on_host, read_wave_packet, site_column, analyse, s_matrix_sums,
                            add_earlier_analysis_step, energy_resolved_add};
    for (int i = 0; i < kCount; ++i) out[i] = v[i];
  }
};

class CountingBackend final : public PropagationBackend {
 public:
  CountingBackend(std::unique_ptr<PropagationBackend> inner, SeamCallCounts& counts)
      : inner_(std::move(inner)), counts_(counts) {}

  PropagationBackend* innermost() override { return inner_->innermost(); }
  void note_batched_step() override {
    ++counts_.propagate_analysis_step;
    inner_->note_batched_step();
  }

  void read_wave_packet(const std::function<void(WavePacket&)>& read) override {
    ++counts_.read_wave_packet;
    inner_->read_wave_packet(read);
  }
  const WavePacket& wave_packet_on_host() override {
    ++counts_.wave_packet_on_host;
    return inner_->wave_packet_on_host();
  }
  void propagate_analysis_step() override {
    ++counts_.propagate_analysis_step;
    inner_->propagate_analysis_step();
  }
  void kinetic_and_potential(bool half_step, bool kinetic_only) override {
    ++counts_.kinetic_and_potential;
    inner_->kinetic_and_potential(half_step, kinetic_only);
  }
  const std::vector<cplx>& specular() const override {
    ++counts_.specular;
    return inner_->specular();
  }
  const std::vector<cplx>& z_plane(int iz) const override {
    ++counts_.z_plane;
    return inner_->z_plane(iz);
  }
  std::vector<double> block_densities() const override {
    ++counts_.block_densities;
    return inner_->block_densities();
  }
  void site_column(int ixy, cplx* out) const override {
    ++counts_.site_column;
    inner_->site_column(ixy, out);
  }
  void analyse(const AnalysisSurface& surface, double time, std::vector<cplx>* coefficients) override {
    ++counts_.analyse;
    inner_->analyse(surface, time, coefficients);
  }
  std::vector<cplx> s_matrix_sums() override {
    ++counts_.s_matrix_sums;
    return inner_->s_matrix_sums();
  }
  void add_earlier_analysis_step(double time, const std::vector<cplx>& coefficients) override {
    ++counts_.add_earlier_analysis_step;
    inner_->add_earlier_analysis_step(time, coefficients);
  }
  void energy_resolved_start(const std::vector<int>& site_ixy, int n_energies, double analysis_time_step) override {
    inner_->energy_resolved_start(site_ixy, n_energies, analysis_time_step);
  }
  void energy_resolved_add(const std::vector<cplx>& phase_factors) override {
    ++counts_.energy_resolved_add;
    inner_->energy_resolved_add(phase_factors);
  }
  void energy_resolved_read(std::vector<cplx>* sums, std::vector<cplx>* specular_sums) override {
    inner_->energy_resolved_read(sums, specular_sums);
  }
  void set_kernel_timing(bool on) override { inner_->set_kernel_timing(on); }
  std::vector<KernelTime> kernel_times() override { return inner_->kernel_times(); }

 private:
  std::unique_ptr<PropagationBackend> inner_;
  SeamCallCounts& counts_;
};

}  // namespace spo
