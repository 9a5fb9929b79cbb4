// The HIP backend as a PropagationBackend: packs the set-up into the plain tables of include/spo_hip.h and
// forwards every seam call to libspo_hip.so. Status codes of the C ABI become the exceptions the reference's
// driver expects (device errors -> DeviceError, no device / too little memory -> NotSupportedError).
#include <cstdlib>
#include <cmath>
#include <cstring>

#include "../../include/spo_hip.h"
#include "backend.hpp"

namespace spo {

namespace {

void check(int status) {
  if (status == SPO_HIP_OK) return;
  const std::string message = spo_hip_last_error();
  if (status == SPO_HIP_E_NOT_SUPPORTED) throw NotSupportedError(message);
  if (status == SPO_HIP_E_ARGUMENT) throw std::invalid_argument(message);
  throw DeviceError(message);
}

const double* d(const std::vector<cplx>& v) { return reinterpret_cast<const double*>(v.data()); }

class HipBackend final : public PropagationBackend {
 public:
  explicit HipBackend(const BackendInput& in);
  ~HipBackend() override { spo_hip_destroy(handle_); }

  void read_wave_packet(const std::function<void(WavePacket&)>& read) override {
    read(host_);
    check(spo_hip_upload_wave(handle_, d(host_.scattering), d(host_.specular)));
    specular_ = host_.specular;
  }
  const WavePacket& wave_packet_on_host() override {
    check(spo_hip_download_wave(handle_, reinterpret_cast<double*>(host_.scattering.data()),
                                reinterpret_cast<double*>(host_.specular.data())));
    return host_;
  }
  spo_hip_backend* handle() const { return handle_; }
  void refresh_specular() { check(spo_hip_get_specular(handle_, reinterpret_cast<double*>(specular_.data()))); }
  void propagate_analysis_step() override {
    check(spo_hip_propagate_analysis_step(handle_));
    check(spo_hip_get_specular(handle_, reinterpret_cast<double*>(specular_.data())));
  }
  void kinetic_and_potential(bool half_step, bool kinetic_only) override {
    check(spo_hip_kinetic_and_potential(handle_, half_step ? 1 : 0, kinetic_only ? 1 : 0));
    check(spo_hip_get_specular(handle_, reinterpret_cast<double*>(specular_.data())));
  }
  const std::vector<cplx>& specular() const override { return specular_; }
  const std::vector<cplx>& z_plane(int iz) const override {
    check(spo_hip_get_z_plane(handle_, iz, reinterpret_cast<double*>(plane_.data())));
    return plane_;
  }
  std::vector<double> block_densities() const override {
    std::vector<double> out(std::size_t(host_.n_zr));
    check(spo_hip_block_densities(handle_, out.data()));
    return out;
  }
  void site_column(int ixy, cplx* out) const override {
    check(spo_hip_site_column(handle_, ixy, reinterpret_cast<double*>(out)));
  }
  void energy_resolved_start(const std::vector<int>& site_ixy, int n_energies, double analysis_time_step) override {
    check(spo_hip_energy_resolved_start(handle_, static_cast<int>(site_ixy.size()), site_ixy.data(), n_energies,
                                        analysis_time_step));
    er_sites_ = static_cast<int>(site_ixy.size());
    er_energies_ = n_energies;
  }
  void energy_resolved_add(const std::vector<cplx>& phase_factors) override {
    check(spo_hip_energy_resolved_add(handle_, reinterpret_cast<const double*>(phase_factors.data())));
  }
  void energy_resolved_read(std::vector<cplx>* sums, std::vector<cplx>* specular_sums) override {
    sums->assign(std::size_t(er_sites_) * er_energies_ * host_.n_zr * host_.n_rot, cplx(0.0, 0.0));
    specular_sums->assign(std::size_t(er_energies_) * specular_.size(), cplx(0.0, 0.0));
    check(spo_hip_energy_resolved_read(handle_, reinterpret_cast<double*>(sums->data()),
                                       reinterpret_cast<double*>(specular_sums->data())));
  }
  void analyse(const AnalysisSurface& surface, double time, std::vector<cplx>* coefficients) override {
    double* out = nullptr;
    if (coefficients) {
      coefficients->assign(n_coefficients_, cplx(0.0, 0.0));
      out = reinterpret_cast<double*>(coefficients->data());
    }
    check(spo_hip_analyse(handle_, surface.fixed() ? 1 : 0, surface.index(), time, out));
  }
  void add_earlier_analysis_step(double time, const std::vector<cplx>& coefficients) override {
    if (coefficients.size() != n_coefficients_) throw std::invalid_argument("BK record of the wrong size");
    check(spo_hip_accumulate(handle_, time, reinterpret_cast<const double*>(coefficients.data())));
  }
  std::vector<cplx> s_matrix_sums() override {
    std::vector<cplx> out(n_sums_);
    check(spo_hip_get_s_sums(handle_, reinterpret_cast<double*>(out.data())));
    return out;
  }
  void set_kernel_timing(bool on) override { check(spo_hip_profile_enable(handle_, on ? 1 : 0)); }
  std::vector<KernelTime> kernel_times() override {
    int count = 0;
    check(spo_hip_profile_read(handle_, nullptr, 0, &count));
    std::vector<spo_hip_kernel_time> raw(static_cast<std::size_t>(count));
    check(spo_hip_profile_read(handle_, raw.data(), count, &count));
    std::vector<KernelTime> out;
    for (const auto& r : raw) out.push_back({r.name, r.launches, r.total_ms});
    return out;
  }

 private:
  spo_hip_backend* handle_ = nullptr;
  WavePacket host_;
  std::vector<cplx> specular_;
  mutable std::vector<cplx> plane_;
  std::size_t n_coefficients_ = 0, n_sums_ = 0;
  int er_sites_ = 0, er_energies_ = 0;
};

HipBackend::HipBackend(const BackendInput& in) {
  const InputParameters& p = in.config.input;
  const PropagationTables& t = in.tables;
  const int n_rot = in.basis.size();
  const int n_xy = p.nx * p.ny;
  const int nr = p.nr;
  const int n_phi = in.angular.n_phi();
  const GridSymmetry& sym = in.potential.symmetry;

  spo_hip_shape shape{};
  shape.n_rot = n_rot;
  shape.nx = p.nx;
  shape.ny = p.ny;
  shape.nr = nr;
  shape.nz = p.nz;
  shape.nz_specular = p.nz_specular;
  shape.nz_potential = in.grids.nz_potential;
  shape.n_theta = in.angular.n_theta();
  shape.n_phi = n_phi;
  shape.n_representatives = in.potential.n_rep;
  shape.n_operations = static_cast<int>(sym.ops.size());
  shape.initial_channel = in.basis.initial;
  shape.mj_max = in.basis.mj_max;
  shape.steps_per_analysis = in.config.derived.steps_per_analysis;
  shape.r_absorber_first = in.grids.r_optical.first;
  shape.r_absorber_last = in.grids.r_optical.last;
  shape.needs_shift = t.needs_shift ? 1 : 0;
  if (p.precision == "fp64") shape.precision = SPO_HIP_FP64;
  else if (p.precision == "mixed") shape.precision = SPO_HIP_MIXED;
  else throw InputError("Precision", "expected \"fp64\" or \"mixed\"");

  const PotentialStepFactors f = potential_step_factors(n_rot, nr, t);
  const std::vector<cplx> spec_gas = specular_gas_phase_factors(nr, n_rot, in.basis.initial, t);
  // sqrt(w_t / nphi) P_J^m(theta_t) ordered (mJ block, theta, J) (src/propagation/angular_transform.cpp:31-39).
  const std::vector<double> legendre = legendre_matrix(in.angular, in.basis);
  std::vector<int> block_first, block_size;
  for (const Range& blk : in.basis.blocks) {
    block_first.push_back(blk.first);
    block_size.push_back(blk.size());
  }
  const std::vector<cplx> exp_table = exp_potential_table(in.potential, p.time_step);
  std::vector<int> phi_index(std::size_t(n_phi) * sym.ops.size());
  for (std::size_t op = 0; op < sym.ops.size(); ++op)
    for (int ip = 0; ip < n_phi; ++ip) phi_index[ip + n_phi * op] = sym.phi_index[op][ip];

  spo_hip_tables tb{};
  tb.zr_full = d(t.zr_kinetic.full);
  tb.zr_half = d(t.zr_kinetic.half);
  tb.xy_full = d(t.xy_kinetic.full);
  tb.xy_half = d(t.xy_kinetic.half);
  tb.specular_full = d(t.specular_kinetic.full);
  tb.specular_half = d(t.specular_kinetic.half);
  tb.fft_scale_z = t.fft_scale_z;
  tb.fft_scale_r = t.fft_scale_r;
  tb.fft_scale_x = t.fft_scale_x;
  tb.fft_scale_y = t.fft_scale_y;
  tb.rotation_half = d(t.rotation.half);
  tb.before_interaction = d(f.before_interaction);
  tb.gas_phase_step = d(f.gas_phase_step);
  tb.specular_gas_phase = d(spec_gas);
  tb.z_optical = t.z_optical.data();
  tb.r_optical = t.r_optical.data();
  tb.z_specular_optical = t.z_specular_optical.data();
  tb.shift_forward = d(t.shift_forward);
  tb.shift_backward = d(t.shift_backward);
  tb.specular_to_scattering = d(t.specular_to_scattering);
  tb.scattering_to_specular = d(t.scattering_to_specular);
  tb.legendre = legendre.data();
  tb.block_first = block_first.data();
  tb.block_size = block_size.data();
  tb.exp_potential = d(exp_table);
  tb.representative_of = sym.representative_of.data();
  tb.operation_of = sym.operation_of.data();
  tb.phi_index = phi_index.data();

  const AnalysisBasis& a = in.analysis;
  std::vector<int> slots;
  for (int idif = 0; idif < a.n_diffraction; ++idif) slots.push_back(a.channels[idif].iy + p.ny * a.channels[idif].ix);
  std::vector<double> weights, initial_radial;
  for (int v = 0; v < a.n_v; ++v)
    for (int irotana = 0; irotana < a.n_rotation(); ++irotana) {
      const double* phi = in.gas.wave_function(v, in.basis.j[a.rotations[irotana]]);
      for (int ir = 0; ir < nr; ++ir) weights.push_back(phi[ir] * p.dr);
    }
  for (int v = 0; v < a.n_v; ++v) {
    const double* phi = in.gas.wave_function(v, p.j_in);
    for (int ir = 0; ir < nr; ++ir) initial_radial.push_back(phi[ir]);
  }
  const DerivedParameters& dp = in.config.derived;
  spo_hip_analysis an{};
  an.n_states = a.n_states();
  an.n_v = a.n_v;
  an.n_rotation = a.n_rotation();
  an.n_diffraction = a.n_diffraction;
  an.n_energies = in.energies.size();
  an.iz_analysis = in.grids.iz_analysis;
  an.initial_state = a.initial;
  an.rotation_index = a.rotations.data();
  an.diffraction_slot = slots.data();
  an.radial_weight = weights.data();
  an.initial_radial = initial_radial.data();
  an.dr = p.dr;
  an.xy_scale = std::sqrt(dp.dx * dp.dy * std::sin(dp.system.skew) / (p.nx * p.ny));
  an.analysis_time_step = dp.steps_per_analysis * p.time_step;
  an.total_energy_hartree = in.energies.total_hartree.data();

  check(spo_hip_create(&shape, &tb, &an, d(in.initial_specular), &handle_));
  // SPO_HIP_GRAPH=1: analysis steps replayed from a captured hipGraph (spo_hip_graph_enable)
  if (const char* g = std::getenv("SPO_HIP_GRAPH"); g && std::string(g) == "1") check(spo_hip_graph_enable(handle_, 1));
  host_.n_rot = n_rot;
  host_.n_xy = n_xy;
  host_.n_zr = nr * p.nz;
  host_.nr = nr;
  host_.nz_specular = p.nz_specular;
  host_.scattering.assign(std::size_t(n_rot) * n_xy * host_.n_zr, cplx(0.0, 0.0));
  host_.specular = in.initial_specular;
  specular_ = in.initial_specular;
  plane_.resize(std::size_t(n_rot) * n_xy * nr);
  n_coefficients_ = std::size_t(a.n_states()) * a.n_v;
  n_sums_ = n_coefficients_ * in.energies.size();
}

}  // namespace

void propagate_batch(const std::vector<PropagationBackend*>& backends) {
  std::vector<HipBackend*> hip;
  std::vector<spo_hip_backend*> handles;
  for (PropagationBackend* b : backends) {
    if (auto* h = dynamic_cast<HipBackend*>(b->innermost())) {
      hip.push_back(h);
      handles.push_back(h->handle());
      b->note_batched_step();
    } else {
      b->propagate_analysis_step();
    }
  }
  if (handles.empty()) return;
  check(spo_hip_propagate_batch(handles.data(), static_cast<int>(handles.size())));
  for (HipBackend* h : hip) h->refresh_specular();
}

namespace {

class GraphBatch final : public BatchStepper {
 public:
  explicit GraphBatch(const std::vector<PropagationBackend*>& backends) {
    std::vector<spo_hip_backend*> handles;
    for (PropagationBackend* b : backends) {
      if (auto* h = dynamic_cast<HipBackend*>(b->innermost())) {
        hip_.push_back(h);
        hip_outer_.push_back(b);
        handles.push_back(h->handle());
      } else {
        others_.push_back(b);
      }
    }
    if (!handles.empty()) check(spo_hip_batch_create(handles.data(), static_cast<int>(handles.size()), &batch_));
  }
  ~GraphBatch() override {
    if (batch_) (void)spo_hip_batch_destroy(batch_);
  }
  void step() override {
    if (batch_) {
      check(spo_hip_batch_step(batch_));
      for (HipBackend* h : hip_) h->refresh_specular();
      for (PropagationBackend* b : hip_outer_) b->note_batched_step();
    }
    for (PropagationBackend* b : others_) b->propagate_analysis_step();
  }

 private:
  spo_hip_batch* batch_ = nullptr;
  std::vector<HipBackend*> hip_;
  std::vector<PropagationBackend*> hip_outer_, others_;
};

}  // namespace

std::unique_ptr<BatchStepper> make_batch_stepper(const std::vector<PropagationBackend*>& backends) {
  return std::make_unique<GraphBatch>(backends);
}

std::unique_ptr<PropagationBackend> make_hip_backend(const BackendInput& input) {
  return std::make_unique<HipBackend>(input);
}

}  // namespace spo
