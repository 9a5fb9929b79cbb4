// The reference-side adapter: a spodvr::PropagationBackend of the reference program
// (src/propagation/propagation_backend.hpp:34-59) whose seven calls go to libspo_hip.so through the C ABI of
// include/spo_hip.h. A maintainer adds this file to the reference (as src/hip/hip_backend.cpp), the Device value
// Hip (src/propagation/propagation_backend.hpp:23-27) and one line in backend_factory
// (src/app/backend_factory.cpp:22-29): `if (device == Device::Hip) return hip::make_hip_backend;`, and links
// libspo_hip.so. It is written against the reference's own types (BackendInput, PropagationTables, AngularTransform,
// PotentialGrid, GridSymmetry, WavePacket); its twin for this repository's host types is host/hip_backend.cpp.
//
// The Balint-Kurti projection stays where the reference has it (on the host, from z_plane); the device analysis of
// spo_hip_analyse is optional and not used here, so the analysis tables handed to spo_hip_create are a minimal valid
// set (one state, one energy) that nothing reads.
#include <cmath>
#include <complex>
#include <cstddef>
#include <functional>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/error.hpp"
#include "propagation/angular_transform.hpp"
#include "propagation/potential_operator.hpp"
#include "propagation/propagation_backend.hpp"
#include "spo_hip.h"

namespace spodvr::hip {

namespace {

using cplx = std::complex<double>;

// C ABI status -> the reference's exceptions: no usable device / too little memory -> NotSupportedError (exit 3,
// src/app/run.cpp:602-604), everything else a runtime error (exit 1, as CudaError).
void check(int status) {
  if (status == SPO_HIP_OK) return;
  const std::string message = spo_hip_last_error();
  if (status == SPO_HIP_E_NOT_SUPPORTED) throw NotSupportedError(message);
  throw std::runtime_error("HIP backend: " + message);
}

// read-only (tables, uploads) and written (downloads) views of complex arrays as interleaved doubles
const double* doubles(std::span<const cplx> v) { return reinterpret_cast<const double*>(v.data()); }
double* doubles_out(std::span<cplx> v) { return reinterpret_cast<double*>(v.data()); }

// The host arrays the packed tables point into while spo_hip_create copies them.
struct Packed {
  PotentialStepFactors factors;
  std::vector<cplx> specular_gas, exp_potential;
  std::vector<double> legendre;
  std::vector<int> block_first, block_size, representative_of, operation_of, phi_index;
  // the minimal analysis set (spo_hip_analyse is not called)
  std::vector<int> rotation_index, diffraction_slot;
  std::vector<double> radial_weight, initial_radial, energies;
};

}  // namespace

class HipBackend final : public PropagationBackend {
 public:
  HipBackend(const BackendInput& in, PotentialGrid potential, SpecularWave initial)
      : n_rot_(in.basis.size()),
        n_xy_(in.config.input.nx * in.config.input.ny),
        nr_(in.config.input.nr),
        n_zr_(in.config.input.nr * in.config.input.nz),
        host_{ScatteringWave(n_rot_, n_xy_, n_zr_), SpecularWave(initial.n_r(), initial.n_z())},
        specular_(initial.n_r(), initial.n_z()),
        plane_(std::size_t(n_rot_) * n_xy_ * nr_) {
    const InputParameters& p = in.config.input;
    const PropagationTables& t = in.tables;
    const GridSymmetry& sym = potential.symmetry;
    const AngularTransform transform(in.angular, in.basis, in.fft);
    const int n_phi = in.angular.n_phi();
    Packed k;

    spo_hip_shape shape{};
    shape.n_rot = n_rot_;
    shape.nx = p.nx;
    shape.ny = p.ny;
    shape.nr = nr_;
    shape.nz = p.nz;
    shape.nz_specular = p.nz_specular;
    shape.nz_potential = in.grids.nz_potential;
    shape.n_theta = in.angular.n_theta();
    shape.n_phi = n_phi;
    shape.n_representatives = sym.n_representatives();
    shape.n_operations = static_cast<int>(sym.operations().size());
    shape.initial_channel = in.basis.initial_index();
    shape.mj_max = transform.mj_max();
    shape.steps_per_analysis = in.config.derived.steps_per_analysis;
    shape.r_absorber_first = in.grids.r_optical.first;
    shape.r_absorber_last = in.grids.r_optical.last;
    shape.needs_shift = 0;
    for (const cplx& s : t.shift_forward)
      if (s != cplx(1.0, 0.0)) shape.needs_shift = 1;
    shape.precision = SPO_HIP_FP64;

    // The merged factors of the potential step (src/propagation/potential_operator.hpp:23-26) and the specular
    // grid's gas-phase factor rot exp(-i vasy dt) rot of the initial channel (src/propagation/specular_operator.cpp).
    k.factors = potential_step_factors(ScatteringShape{n_rot_, p.nx, p.ny, nr_, p.nz}, t);
    for (int ir = 0; ir < nr_; ++ir) {
      const cplx rot = t.rotation.half[in.basis.initial_index() + std::size_t(n_rot_) * ir];
      k.specular_gas.push_back(rot * t.gas_phase[ir] * rot);
    }
    // The angular transform's tables: sqrt(w_t / nphi) P_J^m(theta_t) ordered (mJ block, theta, J) and the blocks
    // mJ = -MJMax .. MJMax (src/propagation/angular_transform.hpp:67-75).
    k.legendre = transform.matrix();
    for (const AngularTransform::MjBlock& b : transform.mj_blocks()) {
      k.block_first.push_back(b.fbr.first);
      k.block_size.push_back(b.fbr.size());
    }
    // exp(i A), A = -V dt, on the representatives (src/potential/potential_grid.cpp:164-167), and the maps.
    k.exp_potential.reserve(potential.values.size());
    for (const double v : potential.values) k.exp_potential.push_back(exp_i(-v * p.time_step));
    for (int ixy = 0; ixy < n_xy_; ++ixy) {
      k.representative_of.push_back(sym.representative_of(ixy));
      k.operation_of.push_back(sym.operation_of(ixy));
    }
    for (std::size_t op = 0; op < sym.operations().size(); ++op)
      for (int ip = 0; ip < n_phi; ++ip) k.phi_index.push_back(sym.phi_index(static_cast<int>(op))[ip]);

    spo_hip_tables tb{};
    tb.zr_full = doubles(t.zr_kinetic.full);
    tb.zr_half = doubles(t.zr_kinetic.half);
    tb.xy_full = doubles(t.xy_kinetic.full);
    tb.xy_half = doubles(t.xy_kinetic.half);
    tb.specular_full = doubles(t.specular_kinetic.full);
    tb.specular_half = doubles(t.specular_kinetic.half);
    tb.fft_scale_z = t.fft_scale_z;
    tb.fft_scale_r = t.fft_scale_r;
    tb.fft_scale_x = t.fft_scale_x;
    tb.fft_scale_y = t.fft_scale_y;
    tb.rotation_half = doubles(t.rotation.half);
    tb.before_interaction = doubles(k.factors.before_interaction);
    tb.gas_phase_step = doubles(k.factors.gas_phase_step);
    tb.specular_gas_phase = doubles(k.specular_gas);
    tb.z_optical = t.z_optical.data();
    tb.r_optical = t.r_optical.data();
    tb.z_specular_optical = t.z_specular_optical.data();
    tb.shift_forward = doubles(t.shift_forward);
    tb.shift_backward = doubles(t.shift_backward);
    tb.specular_to_scattering = doubles(t.specular_to_scattering);
    tb.scattering_to_specular = doubles(t.scattering_to_specular);
    tb.legendre = k.legendre.data();
    tb.block_first = k.block_first.data();
    tb.block_size = k.block_size.data();
    tb.exp_potential = doubles(k.exp_potential);
    tb.representative_of = k.representative_of.data();
    tb.operation_of = k.operation_of.data();
    tb.phi_index = k.phi_index.data();

    k.rotation_index = {in.basis.initial_index()};
    k.diffraction_slot = {0};
    k.radial_weight.assign(std::size_t(nr_), 0.0);
    k.initial_radial.assign(std::size_t(nr_), 0.0);
    k.energies = {0.0};
    spo_hip_analysis an{};
    an.n_states = 1;
    an.n_v = 1;
    an.n_rotation = 1;
    an.n_diffraction = 1;
    an.n_energies = 1;
    an.iz_analysis = in.grids.iz_analysis;
    an.initial_state = 0;
    an.rotation_index = k.rotation_index.data();
    an.diffraction_slot = k.diffraction_slot.data();
    an.radial_weight = k.radial_weight.data();
    an.initial_radial = k.initial_radial.data();
    an.dr = p.dr;
    an.xy_scale = 1.0;
    an.analysis_time_step = in.config.derived.steps_per_analysis * p.time_step;
    an.total_energy_hartree = k.energies.data();

    check(spo_hip_create(&shape, &tb, &an, doubles(initial.values), &handle_));
    specular_.values = initial.values;
    host_.specular.values = initial.values;
  }

  HipBackend(const HipBackend&) = delete;
  HipBackend& operator=(const HipBackend&) = delete;
  ~HipBackend() override { spo_hip_destroy(handle_); }

  void read_wave_packet(const std::function<void(WavePacket&)>& read) override {
    read(host_);
    check(spo_hip_upload_wave(handle_, doubles(host_.scattering.values), doubles(host_.specular.values)));
    specular_.values = host_.specular.values;
  }
  const WavePacket& wave_packet_on_host() override {
    check(spo_hip_download_wave(handle_, doubles_out(host_.scattering.values), doubles_out(host_.specular.values)));
    return host_;
  }
  void propagate_analysis_step() override {
    check(spo_hip_propagate_analysis_step(handle_));
    check(spo_hip_get_specular(handle_, doubles_out(specular_.values)));
  }
  const SpecularWave& specular() const override { return specular_; }
  std::span<const cplx> z_plane(int iz) const override {
    check(spo_hip_get_z_plane(handle_, iz, doubles_out(plane_)));
    return plane_;
  }
  std::vector<double> block_densities() const override {
    std::vector<double> densities(static_cast<std::size_t>(n_zr_));
    check(spo_hip_block_densities(handle_, densities.data()));
    return densities;
  }
  void site_column(int ixy, std::span<cplx> column) const override {
    if (column.size() != std::size_t(n_zr_) * n_rot_) throw std::invalid_argument("site_column: wrong buffer size");
    check(spo_hip_site_column(handle_, ixy, doubles_out(column)));
  }

 private:
  int n_rot_, n_xy_, nr_, n_zr_;
  spo_hip_backend* handle_ = nullptr;
  WavePacket host_;
  SpecularWave specular_;
  mutable std::vector<cplx> plane_;
};

// The BackendFactory of Device::Hip (src/propagation/propagation_backend.hpp:73-74).
std::unique_ptr<PropagationBackend> make_hip_backend(const BackendInput& in, PotentialGrid potential,
                                                     SpecularWave initial) {
  return std::make_unique<HipBackend>(in, std::move(potential), std::move(initial));
}

}  // namespace spodvr::hip
