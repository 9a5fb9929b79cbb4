// The propagation seam: the backend owns the wave packet and the split-operator stages that act on it; the
// time loop above talks to it once per analysis step. The first seven methods are the reference's
// PropagationBackend (src/propagation/propagation_backend.hpp:34-59); analyse() and s_matrix_sums() move the
// Balint-Kurti projection and its FP64 time-to-energy accumulation below the seam, so that only the
// coefficients (not the analysis plane) cross it.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "analysis.hpp"
#include "config.hpp"
#include "gas_phase.hpp"
#include "grids.hpp"
#include "potential.hpp"
#include "tables.hpp"

namespace spo {

// Scattering wave in the host layout (irot, ixy, izr), irot fastest, ixy = iy + ny * ix, izr = ir + nr * iz;
// specular wave (ir, iz) (src/propagation/wave_function.hpp:17-59).
struct WavePacket {
  int n_rot = 0, n_xy = 0, n_zr = 0, nr = 0, nz_specular = 0;
  std::vector<cplx> scattering;
  std::vector<cplx> specular;
};

// What a backend is built from; every reference outlives the backend.
struct BackendInput {
  const Config& config;
  const SpatialGrids& grids;
  const AngularGrid& angular;
  const RotationalBasis& basis;
  const PropagationTables& tables;
  const PotentialGrid& potential;  // V in hartree; the backend premultiplies by -dt itself
  const GasPhase& gas;
  const AnalysisBasis& analysis;
  const AnalysisEnergies& energies;
  const std::vector<cplx>& initial_specular;
};

class PropagationBackend {
 public:
  virtual ~PropagationBackend() = default;
  virtual void read_wave_packet(const std::function<void(WavePacket&)>& read) = 0;
  virtual const WavePacket& wave_packet_on_host() = 0;
  // T/2 V, (T V) x (n-1), T/2 (src/propagation/split_operator_schedule.cpp:7-13).
  virtual void propagate_analysis_step() = 0;
  // One T block of that schedule: specular grid, T_Zr, T_xy, then V unless kinetic_only
  // (src/propagation/split_operator.cpp:17-21); for stage-level checks.
  virtual void kinetic_and_potential(bool half_step, bool kinetic_only) = 0;
  virtual const std::vector<cplx>& specular() const = 0;
  // The nr blocks of Z index iz in layout (irot, ixy, ir); valid until the next call.
  virtual const std::vector<cplx>& z_plane(int iz) const = 0;
  // sum |psi|^2 over (irot, ixy) per izr.
  virtual std::vector<double> block_densities() const = 0;
  // n_zr * n_rot values of one (x, y) point, layout (izr, irot).
  virtual void site_column(int ixy, cplx* out) const = 0;
  // Balint-Kurti coefficients (is, v) at the analysis line (the initial channel from the specular grid once
  // the surface is fixed), added to the S-matrix sums with exp(i E t) dtAna (src/analysis/balint_kurti.cpp:42-84,
  // src/analysis/s_matrix.cpp:49-56). coefficients may be null.
  virtual void analyse(const AnalysisSurface& surface, double time, std::vector<cplx>* coefficients) = 0;
  // The FP64 sums (is, v, ien).
  virtual std::vector<cplx> s_matrix_sums() = 0;
  // Adds the coefficients of an analysis step of an earlier run (a restart) to the sums at its time
  // (src/propagation/propagation.hpp:88, SMatrixAccumulator::add).
  virtual void add_earlier_analysis_step(double time, const std::vector<cplx>& coefficients) = 0;

  // Energy-resolved sums at chosen (x, y) sites (src/analysis/energy_resolved.cpp:62-91): start() zeroes the sums of
  // the sites and n_energies energies; add() sums one analysis step, s += psi(site) exp(i E_tot t) dtAna and
  // sigma += chi exp(i E_tot t) dtAna, given exp(i E_tot t) per energy; read() returns them in the reference's
  // layouts, [site, energy, Z, r, rot] (rot fastest) and [energy, Z_specular, r] (r fastest), in FP64.
  virtual void energy_resolved_start(const std::vector<int>& site_ixy, int n_energies, double analysis_time_step) {
    (void)site_ixy, (void)n_energies, (void)analysis_time_step;
    throw NotSupportedError("this backend has no energy-resolved sums");
  }
  virtual void energy_resolved_add(const std::vector<cplx>& phase_factors) {
    (void)phase_factors;
    throw NotSupportedError("this backend has no energy-resolved sums");
  }
  virtual void energy_resolved_read(std::vector<cplx>* sums, std::vector<cplx>* specular_sums) {
    (void)sums, (void)specular_sums;
    throw NotSupportedError("this backend has no energy-resolved sums");
  }

  // Per-kernel device timing (a backend without kernels reports nothing).
  struct KernelTime {
    std::string name;
    long launches = 0;
    double total_ms = 0.0;
  };
  virtual void set_kernel_timing(bool) {}
  virtual std::vector<KernelTime> kernel_times() { return {}; }

  // The backend that does the work (a decorator returns its inner backend's), for propagate_batch; and the
  // notice that propagate_batch propagated this backend's analysis step without calling propagate_analysis_step.
  virtual PropagationBackend* innermost() { return this; }
  virtual void note_batched_step() {}
};

using BackendFactory = std::function<std::unique_ptr<PropagationBackend>(const BackendInput&)>;

// exp(i A) with A = -V dt, (cos A, sin A), over the whole potential grid (src/potential/potential_grid.cpp:164-167,
// src/propagation/potential_operator.cpp:53-64).
std::vector<cplx> exp_potential_table(const PotentialGrid& grid, double dt);

// The backend on the GPU (hip_backend.cpp); throws DeviceError / NotSupportedError when no device is usable.
std::unique_ptr<PropagationBackend> make_hip_backend(const BackendInput& input);

// One analysis step of each backend: the HIP backends' steps enqueued together on their own streams
// (spo_hip_propagate_batch), the others one after another.
void propagate_batch(const std::vector<PropagationBackend*>& backends);

// Analysis steps of a fixed set of backends: the HIP backends' steps captured together into one hipGraph at the
// first step and replayed with one launch per step (spo_hip_batch_*), the others one after another.
class BatchStepper {
 public:
  virtual ~BatchStepper() = default;
  virtual void step() = 0;
};
std::unique_ptr<BatchStepper> make_batch_stepper(const std::vector<PropagationBackend*>& backends);

}  // namespace spo
