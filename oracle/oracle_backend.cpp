// TEST INFRASTRUCTURE — the CPU oracle. A plain FP64 OpenMP restatement of the reference's CPU backend, used by
// tests/, the smoke check and bench.py's CPU-baseline leg as the checker of the HIP backend. Not part of the
// product: the product never links or calls it.
//
// Follows, operator by operator:
//   schedule        src/propagation/split_operator_schedule.cpp:7-13, split_operator.cpp:11-46
//   T_Zr            src/propagation/kinetic_operators.cpp:77-111 (2D FFT in (r, Z), sign -, scale dZ dr / 2 pi;
//                   back sign +, scale 1 / (nZ sZ nR sR))
//   T_xy            src/propagation/kinetic_operators.cpp:134-172 (momentum shift, 2D FFT in (x, y))
//   V               src/propagation/potential_operator.cpp:95-201 (absorbers, rot(dt/2), specular in,
//                   FBR -> DVR, exp(i A) through the symmetry maps, DVR -> FBR, rot(dt/2), projection)
//   angular         src/propagation/angular_transform.cpp:16-114 (Legendre sums per mJ block, phi FFT)
//   specular grid   src/propagation/specular_operator.cpp:40-71
//   BK projection   src/analysis/balint_kurti.cpp:42-84
//   S sums          src/analysis/s_matrix.cpp:49-56
//   densities       src/propagation/averages.cpp:19-30
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "../cpp-spo-dvr_amd/host/backend.hpp"
#include "oracle.hpp"
#include "oracle_fft.hpp"

namespace oracle {

using spo::BackendInput;
using spo::PropagationBackend;
using spo::WavePacket;

namespace {

inline cplx cmul(const cplx& a, const cplx& b) { return a * b; }

// One 2D in-place transform of an (outer x inner) array with the given strides, scale applied afterwards.
struct Fft2d {
  Fft1d outer, inner;
  Fft2d(int n_outer, int n_inner) : outer(n_outer), inner(n_inner) {}
  void run(cplx* x, int outer_stride, int inner_stride, int sign, double scale, cplx* work) const {
    for (int o = 0; o < outer.size(); ++o) inner.run(x + static_cast<long>(o) * outer_stride, inner_stride, sign, work);
    for (int i = 0; i < inner.size(); ++i) outer.run(x + static_cast<long>(i) * inner_stride, outer_stride, sign, work);
    if (scale != 1.0)
      for (int o = 0; o < outer.size(); ++o)
        for (int i = 0; i < inner.size(); ++i) {
          cplx& v = x[static_cast<long>(o) * outer_stride + static_cast<long>(i) * inner_stride];
          v = scale * v;
        }
  }
};

class OracleBackend final : public PropagationBackend {
 public:
  explicit OracleBackend(const BackendInput& in);

  void read_wave_packet(const std::function<void(WavePacket&)>& read) override { read(psi_); }
  const WavePacket& wave_packet_on_host() override { return psi_; }
  void propagate_analysis_step() override;
  const std::vector<cplx>& specular() const override { return psi_.specular; }
  const std::vector<cplx>& z_plane(int iz) const override;
  std::vector<double> block_densities() const override;
  void site_column(int ixy, cplx* out) const override;
  void analyse(const spo::AnalysisSurface& surface, double time, std::vector<cplx>* coefficients) override;
  std::vector<cplx> s_matrix_sums() override { return sums_; }
  void add_earlier_analysis_step(double time, const std::vector<cplx>& coefficients) override {
    accumulate(time, coefficients);
  }

  // Energy-resolved sums as EnergyResolvedAccumulator (src/analysis/energy_resolved.cpp:59-91): each site's column
  // read through site_column, [site, energy, izr, irot] and [energy, iz_specular, ir].
  void energy_resolved_start(const std::vector<int>& site_ixy, int n_energies, double analysis_time_step) override;
  void energy_resolved_add(const std::vector<cplx>& phase_factors) override;
  void energy_resolved_read(std::vector<cplx>* sums, std::vector<cplx>* specular_sums) override {
    *sums = er_sums_;
    *specular_sums = er_specular_;
  }

  // One T block (specular, T_Zr, T_xy, V unless kinetic_only) with full or half kinetic factors.
  void kinetic_and_potential(bool half, bool kinetic_only) override;
  void accumulate(double time, const std::vector<cplx>& bk);

 private:
  void specular_step(bool half, bool kinetic_only);
  void apply_zr(bool half);
  void apply_xy_and_potential(bool half, bool kinetic_only);
  void xy_block(cplx* block, bool half, cplx* work) const;
  void potential_block(int izr, cplx* block, cplx* dvr, cplx* work);

  BackendInput in_;  // a struct of references into the run set-up, which outlives the backend
  int n_rot_, nx_, ny_, n_xy_, nr_, nz_, n_zr_, nzsp_, nzpot_, ichn_, n_theta_, n_phi_, mj_max_;
  int steps_per_analysis_;
  WavePacket psi_;
  Fft2d xy_fft_, spec_fft_;
  Fft1d z_fft_, r_fft_, x_fft_, y_fft_, phi_fft_;
  std::vector<double> legendre_;  // (mJ block, theta, J) as the reference orders it
  std::vector<cplx> exp_table_;   // exp(i A), (itp, irep, izr)
  spo::PotentialStepFactors factors_;
  std::vector<cplx> spec_gas_;
  std::vector<cplx> sums_;
  mutable std::vector<cplx> plane_;
  std::vector<int> er_sites_;
  int er_energies_ = 0;
  double er_dt_ = 0.0;
  std::vector<cplx> er_sums_, er_specular_, er_column_;
};

OracleBackend::OracleBackend(const BackendInput& in)
    : in_(in),
      n_rot_(in.basis.size()),
      nx_(in.config.input.nx),
      ny_(in.config.input.ny),
      n_xy_(nx_ * ny_),
      nr_(in.config.input.nr),
      nz_(in.config.input.nz),
      n_zr_(nr_ * nz_),
      nzsp_(in.config.input.nz_specular),
      nzpot_(in.grids.nz_potential),
      ichn_(in.basis.initial),
      n_theta_(in.angular.n_theta()),
      n_phi_(in.angular.n_phi()),
      mj_max_(in.basis.mj_max),
      steps_per_analysis_(in.config.derived.steps_per_analysis),
      xy_fft_(nx_, ny_),
      spec_fft_(nzsp_, nr_),
      z_fft_(nz_),
      r_fft_(nr_),
      x_fft_(nx_),
      y_fft_(ny_),
      phi_fft_(n_phi_) {
  psi_.n_rot = n_rot_;
  psi_.n_xy = n_xy_;
  psi_.n_zr = n_zr_;
  psi_.nr = nr_;
  psi_.nz_specular = nzsp_;
  psi_.scattering.assign(std::size_t(n_rot_) * n_xy_ * n_zr_, cplx(0.0, 0.0));
  psi_.specular = in.initial_specular;
  // sqrt(w_t / nphi) P_J^m(theta_t), ordered (mJ block, theta, J).
  for (int m = -mj_max_; m <= mj_max_; ++m)
    for (int t = 0; t < n_theta_; ++t) {
      const std::vector<double> p = spo::associated_legendre(in.angular.j_max, m, in.angular.thetas[t]);
      const double f = std::sqrt(in.angular.theta_weights[t] / n_phi_);
      for (int j = in.angular.j_min; j <= in.angular.j_max; j += in.angular.delta_j) {
        if (std::abs(m) > j) continue;
        legendre_.push_back(p[j - std::abs(m)] * f);
      }
    }
  exp_table_.resize(in.potential.values.size());
  for (std::size_t i = 0; i < exp_table_.size(); ++i) {
    const double a = -in.potential.values[i] * in.config.input.time_step;
    exp_table_[i] = cplx(std::cos(a), std::sin(a));
  }
  factors_ = spo::potential_step_factors(n_rot_, nr_, in.tables);
  spec_gas_ = spo::specular_gas_phase_factors(nr_, n_rot_, ichn_, in.tables);
  sums_.assign(std::size_t(in.analysis.n_states()) * in.analysis.n_v * in.energies.size(), cplx(0.0, 0.0));
}

void OracleBackend::propagate_analysis_step() {
  kinetic_and_potential(true, false);
  for (int s = 1; s < steps_per_analysis_; ++s) kinetic_and_potential(false, false);
  kinetic_and_potential(true, true);
}

void OracleBackend::kinetic_and_potential(bool half, bool kinetic_only) {
  specular_step(half, kinetic_only);
  apply_zr(half);
  apply_xy_and_potential(half, kinetic_only);
}

void OracleBackend::specular_step(bool half, bool kinetic_only) {
  const spo::PropagationTables& t = in_.tables;
  std::vector<cplx>& s = psi_.specular;
  std::vector<cplx> work(2 * std::max(nzsp_, nr_));
  const std::vector<cplx>& f = half ? t.specular_kinetic.half : t.specular_kinetic.full;
  spec_fft_.run(s.data(), nr_, 1, -1, t.fft_scale_z * t.fft_scale_r, work.data());
  for (std::size_t i = 0; i < s.size(); ++i) s[i] = cmul(f[i], s[i]);
  spec_fft_.run(s.data(), nr_, 1, +1, (1.0 / (nzsp_ * t.fft_scale_z)) * (1.0 / (nr_ * t.fft_scale_r)), work.data());
  if (kinetic_only) return;
  for (int iz = nzpot_; iz < nzsp_; ++iz)
    for (int ir = 0; ir < nr_; ++ir) s[ir + std::size_t(nr_) * iz] = cmul(spec_gas_[ir], s[ir + std::size_t(nr_) * iz]);
  for (int iz = 0; iz < nzsp_; ++iz)
    for (int ir = 0; ir < nr_; ++ir) s[ir + std::size_t(nr_) * iz] = t.z_specular_optical[iz] * s[ir + std::size_t(nr_) * iz];
  for (int iz = 0; iz < nzsp_; ++iz)
    for (int ir = in_.grids.r_optical.first; ir <= in_.grids.r_optical.last; ++ir)
      s[ir + std::size_t(nr_) * iz] = t.r_optical[ir] * s[ir + std::size_t(nr_) * iz];
}

void OracleBackend::apply_zr(bool half) {
  const spo::PropagationTables& t = in_.tables;
  const std::vector<cplx>& f = half ? t.zr_kinetic.half : t.zr_kinetic.full;
  const std::size_t plane_stride = std::size_t(n_rot_) * n_xy_;  // distance between consecutive izr
  const double s_fwd = t.fft_scale_z * t.fft_scale_r;
  const double s_bwd = (1.0 / (nz_ * t.fft_scale_z)) * (1.0 / (nr_ * t.fft_scale_r));
  // Units of c adjacent (irot, ixy) planes, gathered as (e, ir, iz) with e fastest.
  const int n_planes = n_rot_ * n_xy_;
  const int c = n_planes % 4 == 0 ? 4 : n_planes % 2 == 0 ? 2 : 1;
#pragma omp parallel
  {
    std::vector<cplx> unit(std::size_t(c) * n_zr_);
    std::vector<cplx> work(2 * std::size_t(c) * n_zr_);
    auto transform = [&](int sign, double scale) {
      for (int iz = 0; iz < nz_; ++iz) r_fft_.run_batch(unit.data() + std::size_t(c) * nr_ * iz, c, c, sign, work.data());
      z_fft_.run_batch(unit.data(), long(c) * nr_, c * nr_, sign, work.data());
      for (auto& v : unit) v = scale * v;
    };
#pragma omp for schedule(static)
    for (int u = 0; u < n_planes / c; ++u) {
      const std::size_t p0 = std::size_t(u) * c;
      for (int izr = 0; izr < n_zr_; ++izr)
        for (int e = 0; e < c; ++e) unit[e + std::size_t(c) * izr] = psi_.scattering[p0 + e + plane_stride * izr];
      transform(-1, s_fwd);
      for (int k = 0; k < n_zr_; ++k)
        for (int e = 0; e < c; ++e) unit[e + std::size_t(c) * k] = cmul(f[k], unit[e + std::size_t(c) * k]);
      transform(+1, s_bwd);
      for (int izr = 0; izr < n_zr_; ++izr)
        for (int e = 0; e < c; ++e) psi_.scattering[p0 + e + plane_stride * izr] = unit[e + std::size_t(c) * izr];
    }
  }
}

void OracleBackend::xy_block(cplx* block, bool half, cplx* work) const {
  const spo::PropagationTables& t = in_.tables;
  const std::vector<cplx>& f = half ? t.xy_kinetic.half : t.xy_kinetic.full;
  const double s_fwd = t.fft_scale_x * t.fft_scale_y;
  const double s_bwd = (1.0 / (nx_ * t.fft_scale_x)) * (1.0 / (ny_ * t.fft_scale_y));
  const std::size_t n = std::size_t(n_rot_) * n_xy_;
  // Element (irot, ixy) at block[irot + n_rot ixy]: every irot is one 2D transform over (x, y).
  auto transform = [&](int sign, double scale) {
    for (int ix = 0; ix < nx_; ++ix) y_fft_.run_batch(block + std::size_t(n_rot_) * ny_ * ix, n_rot_, n_rot_, sign, work);
    x_fft_.run_batch(block, long(n_rot_) * ny_, n_rot_ * ny_, sign, work);
    for (std::size_t k = 0; k < n; ++k) block[k] = scale * block[k];
  };
  if (t.needs_shift)
    for (int ixy = 0; ixy < n_xy_; ++ixy)
      for (int irot = 0; irot < n_rot_; ++irot)
        block[irot + std::size_t(n_rot_) * ixy] = block[irot + std::size_t(n_rot_) * ixy] * t.shift_forward[ixy];
  transform(-1, s_fwd);
  for (int ixy = 0; ixy < n_xy_; ++ixy)
    for (int irot = 0; irot < n_rot_; ++irot)
      block[irot + std::size_t(n_rot_) * ixy] = cmul(f[ixy], block[irot + std::size_t(n_rot_) * ixy]);
  transform(+1, s_bwd);
  if (t.needs_shift)
    for (int ixy = 0; ixy < n_xy_; ++ixy)
      for (int irot = 0; irot < n_rot_; ++irot)
        block[irot + std::size_t(n_rot_) * ixy] = block[irot + std::size_t(n_rot_) * ixy] * t.shift_backward[ixy];
}

void OracleBackend::apply_xy_and_potential(bool half, bool kinetic_only) {
  const std::size_t block_size = std::size_t(n_rot_) * n_xy_;
#pragma omp parallel
  {
    const std::size_t dvr_size = std::size_t(n_theta_) * n_phi_ * n_xy_;
    std::vector<cplx> work(std::max(2 * dvr_size, 2 * std::size_t(n_rot_) * n_xy_));
    std::vector<cplx> dvr(dvr_size);
#pragma omp for schedule(dynamic, 1)
    for (int izr = 0; izr < n_zr_; ++izr) {
      cplx* block = psi_.scattering.data() + block_size * izr;
      xy_block(block, half, work.data());
      if (!kinetic_only) potential_block(izr, block, dvr.data(), work.data());
    }
  }
}

void OracleBackend::potential_block(int izr, cplx* block, cplx* dvr, cplx* work) {
  const spo::PropagationTables& t = in_.tables;
  const int ir = izr % nr_, iz = izr / nr_;
  const double opt_z = t.z_optical[iz];
  const cplx* rot_half = t.rotation.half.data() + std::size_t(n_rot_) * ir;
  if (iz >= nzpot_) {
    const cplx* f = factors_.gas_phase_step.data() + std::size_t(n_rot_) * ir;
    for (int ixy = 0; ixy < n_xy_; ++ixy)
      for (int irot = 0; irot < n_rot_; ++irot)
        block[irot + std::size_t(n_rot_) * ixy] = opt_z * cmul(f[irot], block[irot + std::size_t(n_rot_) * ixy]);
    return;
  }
  const cplx* f = factors_.before_interaction.data() + std::size_t(n_rot_) * ir;
  const cplx incoming = psi_.specular[ir + std::size_t(nr_) * iz];
  for (int ixy = 0; ixy < n_xy_; ++ixy) {
    for (int irot = 0; irot < n_rot_; ++irot)
      block[irot + std::size_t(n_rot_) * ixy] = opt_z * cmul(f[irot], block[irot + std::size_t(n_rot_) * ixy]);
    cplx& c = block[ichn_ + std::size_t(n_rot_) * ixy];
    c = c + rot_half[ichn_] * cmul(t.specular_to_scattering[ixy], incoming);
  }

  // Interaction, all (x, y) columns together: FBR -> DVR, x exp(i A) through the symmetry maps, DVR -> FBR.
  // DVR values (phi, column, theta) with theta fastest, so one batched phi FFT covers every column.
  const spo::GridSymmetry& sym = in_.potential.symmetry;
  const int n_tp = n_theta_ * n_phi_;
  const int batch = n_theta_ * n_xy_;
  const cplx* table = exp_table_.data() + std::size_t(n_tp) * in_.potential.n_rep * izr;
  std::fill(dvr, dvr + std::size_t(batch) * n_phi_, cplx(0.0, 0.0));
  for (int ixy = 0; ixy < n_xy_; ++ixy) {
    const cplx* fbr = block + std::size_t(n_rot_) * ixy;
    int index = 0;
    for (int b = 0; b < 2 * mj_max_ + 1; ++b) {
      const int m = b - mj_max_;
      const spo::Range blk = in_.basis.blocks[b];
      const int n_j = blk.size();
      const int slot = m >= 0 ? m : n_phi_ + m;
      for (int th = 0; th < n_theta_; ++th) {
        cplx s(0.0, 0.0);
        for (int j = 0; j < n_j; ++j) s = s + legendre_[index + th * n_j + j] * fbr[blk.first + j];
        dvr[th + std::size_t(n_theta_) * ixy + std::size_t(batch) * slot] = s;
      }
      index += n_theta_ * n_j;
    }
  }
  phi_fft_.run_batch(dvr, batch, batch, +1, work);
  for (int ixy = 0; ixy < n_xy_; ++ixy) {
    const cplx* e = table + std::size_t(n_tp) * sym.representative_of[ixy];
    const std::vector<int>& phi_index = sym.phi_index[sym.operation_of[ixy]];
    for (int ip = 0; ip < n_phi_; ++ip)
      for (int th = 0; th < n_theta_; ++th) {
        cplx& v = dvr[th + std::size_t(n_theta_) * ixy + std::size_t(batch) * ip];
        v = cmul(e[phi_index[ip] + n_phi_ * th], v);
      }
  }
  phi_fft_.run_batch(dvr, batch, batch, -1, work);
  for (int ixy = 0; ixy < n_xy_; ++ixy) {
    cplx* fbr = block + std::size_t(n_rot_) * ixy;
    std::fill(fbr, fbr + n_rot_, cplx(0.0, 0.0));
    int index = 0;
    for (int b = 0; b < 2 * mj_max_ + 1; ++b) {
      const int m = b - mj_max_;
      const spo::Range blk = in_.basis.blocks[b];
      const int n_j = blk.size();
      const int slot = m >= 0 ? m : n_phi_ + m;
      for (int th = 0; th < n_theta_; ++th) {
        const cplx value = dvr[th + std::size_t(n_theta_) * ixy + std::size_t(batch) * slot];
        for (int j = 0; j < n_j; ++j) fbr[blk.first + j] = fbr[blk.first + j] + legendre_[index + j] * value;
        index += n_j;
      }
    }
  }
  for (int ixy = 0; ixy < n_xy_; ++ixy)
    for (int irot = 0; irot < n_rot_; ++irot)
      block[irot + std::size_t(n_rot_) * ixy] = cmul(rot_half[irot], block[irot + std::size_t(n_rot_) * ixy]);
  cplx projection(0.0, 0.0);
  for (int ixy = 0; ixy < n_xy_; ++ixy)
    projection = projection + cmul(t.scattering_to_specular[ixy], block[ichn_ + std::size_t(n_rot_) * ixy]);
  psi_.specular[ir + std::size_t(nr_) * iz] = projection;
  for (int ixy = 0; ixy < n_xy_; ++ixy) {
    cplx& c = block[ichn_ + std::size_t(n_rot_) * ixy];
    c = c - cmul(t.specular_to_scattering[ixy], projection);
  }
}

const std::vector<cplx>& OracleBackend::z_plane(int iz) const {
  const std::size_t plane = std::size_t(n_rot_) * n_xy_ * nr_;
  plane_.assign(psi_.scattering.begin() + plane * iz, psi_.scattering.begin() + plane * (iz + 1));
  return plane_;
}

std::vector<double> OracleBackend::block_densities() const {
  std::vector<double> d(n_zr_, 0.0);
  const std::size_t block_size = std::size_t(n_rot_) * n_xy_;
#pragma omp parallel for schedule(static)
  for (int izr = 0; izr < n_zr_; ++izr) {
    double s = 0.0;
    for (std::size_t k = 0; k < block_size; ++k) {
      const double a = std::abs(psi_.scattering[block_size * izr + k]);
      s = s + a * a;
    }
    d[izr] = s;
  }
  return d;
}

void OracleBackend::energy_resolved_start(const std::vector<int>& site_ixy, int n_energies,
                                          double analysis_time_step) {
  er_sites_ = site_ixy;
  er_energies_ = n_energies;
  er_dt_ = analysis_time_step;
  const std::size_t block = std::size_t(n_zr_) * n_rot_;
  er_sums_.assign(site_ixy.size() * n_energies * block, cplx(0.0, 0.0));
  er_specular_.assign(std::size_t(n_energies) * nzsp_ * nr_, cplx(0.0, 0.0));
  er_column_.assign(block, cplx(0.0, 0.0));
}

void OracleBackend::energy_resolved_add(const std::vector<cplx>& phase) {
  const std::size_t block = std::size_t(n_zr_) * n_rot_;
  for (std::size_t isite = 0; isite < er_sites_.size(); ++isite) {
    site_column(er_sites_[isite], er_column_.data());  // (izr, irot), izr fastest
    cplx* site = er_sums_.data() + isite * er_energies_ * block;
#pragma omp parallel for schedule(static)
    for (int izr = 0; izr < n_zr_; ++izr)
      for (int ien = 0; ien < er_energies_; ++ien) {
        cplx* b = site + ien * block;
        for (int irot = 0; irot < n_rot_; ++irot)
          b[irot + std::size_t(n_rot_) * izr] =
              b[irot + std::size_t(n_rot_) * izr] + er_column_[izr + std::size_t(n_zr_) * irot] * phase[ien] * er_dt_;
      }
  }
  const std::size_t spec_n = std::size_t(nzsp_) * nr_;
  for (int ien = 0; ien < er_energies_; ++ien)
    for (std::size_t k = 0; k < spec_n; ++k)
      er_specular_[ien * spec_n + k] = er_specular_[ien * spec_n + k] + psi_.specular[k] * phase[ien] * er_dt_;
}

void OracleBackend::site_column(int ixy, cplx* out) const {
  for (int irot = 0; irot < n_rot_; ++irot)
    for (int izr = 0; izr < n_zr_; ++izr)
      out[izr + std::size_t(n_zr_) * irot] = psi_.scattering[irot + std::size_t(n_rot_) * (ixy + std::size_t(n_xy_) * izr)];
}

void OracleBackend::analyse(const spo::AnalysisSurface& surface, double time, std::vector<cplx>* coefficients) {
  const spo::AnalysisBasis& a = in_.analysis;
  const spo::PropagationTables& t = in_.tables;
  const spo::DerivedParameters& d = in_.config.derived;
  const double dr = in_.config.input.dr;
  const double xy_scale = std::sqrt(d.dx * d.dy * std::sin(d.system.skew) / (nx_ * ny_));
  const int n_states = a.n_states(), n_v = a.n_v, n_rot_ana = a.n_rotation();
  std::vector<cplx> bk(std::size_t(n_states) * n_v, cplx(0.0, 0.0));
  const std::size_t block_size = std::size_t(n_rot_) * n_xy_;
  const cplx* plane = psi_.scattering.data() + block_size * nr_ * in_.grids.iz_analysis;
#pragma omp parallel
  {
    std::vector<cplx> overlap(n_xy_);
    std::vector<cplx> work(2 * std::max(nx_, ny_));
#pragma omp for schedule(static)
    for (int task = 0; task < n_v * n_rot_ana; ++task) {
      const int v = task / n_rot_ana, irotana = task % n_rot_ana;
      const int irot = a.rotations[irotana];
      const double* phi = in_.gas.wave_function(v, in_.basis.j[irot]);
      std::fill(overlap.begin(), overlap.end(), cplx(0.0, 0.0));
      for (int ir = 0; ir < nr_; ++ir) {
        const double w = phi[ir] * dr;
        const cplx* blk = plane + block_size * ir;
        for (int ixy = 0; ixy < n_xy_; ++ixy)
          overlap[ixy] = overlap[ixy] + t.shift_forward[ixy] * blk[irot + std::size_t(n_rot_) * ixy] * w;
      }
      xy_fft_.run(overlap.data(), ny_, 1, -1, 1.0, work.data());
      for (int idif = 0; idif < a.n_diffraction; ++idif) {
        const spo::DiffractionChannel& ch = a.channels[idif];
        bk[irotana * a.n_diffraction + idif + std::size_t(n_states) * v] = overlap[ch.iy + ny_ * ch.ix] * xy_scale;
      }
    }
  }
  if (surface.fixed()) {
    const int iz = surface.index();
    for (int v = 0; v < n_v; ++v) {
      const double* phi = in_.gas.wave_function(v, in_.config.input.j_in);
      cplx s(0.0, 0.0);
      for (int ir = 0; ir < nr_; ++ir) s = s + psi_.specular[ir + std::size_t(nr_) * iz] * phi[ir];
      bk[a.initial + std::size_t(n_states) * v] = s * dr;
    }
  }
  accumulate(time, bk);
  if (coefficients) *coefficients = std::move(bk);
}

// A(E) += c exp(i E t) dtAna (src/analysis/s_matrix.cpp:49-56).
void OracleBackend::accumulate(double time, const std::vector<cplx>& bk) {
  const double dt_ana = in_.config.derived.steps_per_analysis * in_.config.input.time_step;
  const std::size_t n = bk.size();
  for (int ien = 0; ien < in_.energies.size(); ++ien) {
    const cplx phase = std::exp(cplx(0.0, 1.0) * in_.energies.total_hartree[ien] * time);
    cplx* s = sums_.data() + n * ien;
    for (std::size_t k = 0; k < n; ++k) s[k] = s[k] + bk[k] * phase * dt_ana;
  }
}

}  // namespace

std::unique_ptr<PropagationBackend> make_oracle_backend(const BackendInput& input) {
  return std::make_unique<OracleBackend>(input);
}


}  // namespace oracle
