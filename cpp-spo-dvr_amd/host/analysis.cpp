#include "analysis.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "symmetry.hpp"

namespace spo {

namespace {

struct QRange {
  int min, max;
};
QRange quantum_range(bool dof, int n) { return dof ? QRange{-(n / 2 - 1), n / 2} : QRange{0, 0}; }
int slot(int q, int n) { return q >= 0 ? q : n + q; }

// Highest diffraction order open for EMaxWP on an mx x my momentum grid (src/analysis/analysis_basis.cpp:29-56).
int max_open_order(const Config& c, const GasPhase& gas, int mx, int my) {
  const InputParameters& p = c.input;
  const DerivedParameters& d = c.derived;
  const QRange qx = quantum_range(p.x_dof, mx), qy = quantum_range(p.y_dof, my);
  const double mass = d.total_mass;
  const double cs = std::cos(d.system.skew);
  auto e_par = [&](int i, int j) {
    const double kx = d.kx + i * d.delta_kx, ky = d.ky + j * d.delta_ky;
    return 0.5 * (kx * kx) / mass + -cs * (kx * ky) / mass + 0.5 * (ky * ky) / mass;
  };
  const double e_total = Units::ev_to_hartree(p.e_max_wp) + (gas.energy(p.v_in, p.j_in) - gas.energy(0, 0)) + e_par(0, 0);
  int best = 0;
  for (int j = qy.min; j <= qy.max; ++j)
    for (int i = qx.min; i <= qx.max; ++i)
      if (e_par(i, j) <= e_total) best = std::max(best, diffraction_order(d.system, i, j));
  return best;
}

}  // namespace

std::vector<int> AnalysisBasis::regular_state_order() const {
  std::vector<int> order;
  for (const int irotana : regular_rotations)
    for (int idif = 0; idif < n_diffraction; ++idif) order.push_back(irotana * n_diffraction + idif);
  return order;
}

AnalysisBasis make_analysis_basis(const Config& c, const AngularGrid& angular, const RotationalBasis& basis,
                                  const GasPhase& gas, const std::vector<double>& ekin_xy) {
  const InputParameters& p = c.input;
  AnalysisBasis a;
  a.j_max = gas.j_max_analysis;
  a.n_v = gas.n_v_analysis;
  const QRange qx = quantum_range(p.x_dof, p.nx), qy = quantum_range(p.y_dof, p.ny);
  for (int n = qx.min; n <= qx.max; ++n)
    for (int m = qy.min; m <= qy.max; ++m)
      a.channels.push_back({n, m, diffraction_order(c.derived.system, n, m), slot(n, p.nx), slot(m, p.ny)});
  std::stable_sort(a.channels.begin(), a.channels.end(),
                   [](const DiffractionChannel& x, const DiffractionChannel& y) { return x.order < y.order; });
  a.max_order = p.diffraction_order_analysis ? *p.diffraction_order_analysis : max_open_order(c, gas, p.nx, p.ny);
  for (const auto& ch : a.channels)
    if (ch.order <= a.max_order) ++a.n_diffraction;
  for (const auto& ch : a.channels) a.diffraction_energies.push_back(ekin_xy[ch.iy + p.ny * ch.ix]);
  for (int irot = 0; irot < basis.size(); ++irot)
    if (basis.j[irot] <= a.j_max) a.rotations.push_back(irot);
  for (int irotana = 0; irotana < a.n_rotation(); ++irotana) {
    const int irot = a.rotations[irotana];
    if (basis.j[irot] == p.j_in && basis.mj[irot] == p.mj_in) a.initial = irotana * a.n_diffraction;
  }
  if (a.initial < 0 || p.v_in >= a.n_v) throw InputError("VIn", "the initial state is not in the analysis basis");
  for (int is = 0; is < a.n_states(); ++is) {
    a.state_j.push_back(basis.j[a.rotations[a.state_rotation(is)]]);
    a.state_mj.push_back(basis.mj[a.rotations[a.state_rotation(is)]]);
  }
  a.channel_energies.resize(std::size_t(a.n_states()) * a.n_v);
  for (int v = 0; v < a.n_v; ++v)
    for (int is = 0; is < a.n_states(); ++is)
      a.channel_energies[is + std::size_t(a.n_states()) * v] =
          a.diffraction_energies[a.state_diffraction(is)] + gas.energy(v, a.state_j[is]);
  for (int jj = angular.j_min; jj <= a.j_max; jj += angular.delta_j)
    for (int mm = -basis.mj_max; mm <= basis.mj_max; ++mm) {
      if (std::abs(mm) > jj) continue;
      for (int irotana = 0; irotana < a.n_rotation(); ++irotana)
        if (basis.j[a.rotations[irotana]] == jj && basis.mj[a.rotations[irotana]] == mm)
          a.regular_rotations.push_back(irotana);
    }
  return a;
}

AnalysisSurface::AnalysisSurface(const Config& c, const SpatialGrids& g, const std::vector<cplx>& initial)
    : z_(g.z_specular.points),
      nr_(c.input.nr),
      iz_analysis_(g.iz_analysis),
      last_before_absorber_(g.z_specular_optical.first - 1),
      fixed_(c.derived.fixed_z_analysis_specular),
      index_(last_before_absorber_),
      previous_index_(last_before_absorber_) {
  const InputParameters& p = c.input;
  const double volume = p.dz * p.dr;
  if (fixed_) {
    const bool at_ana_sp = p.z_analysis_specular && *p.z_analysis_specular > 0.0;
    const double z_fix = at_ana_sp ? *p.z_analysis_specular : p.zsp_opt_start;
    for (int iz = 0; iz < static_cast<int>(z_.size()); ++iz)
      if (tol::ge(z_[iz], z_fix)) {
        index_ = at_ana_sp ? iz - 1 : iz;
        break;
      }
    return;
  }
  double norm = 0.0;
  for (int iz = g.z_specular_optical.first; iz < static_cast<int>(z_.size()); ++iz)
    for (int ir = 0; ir < nr_; ++ir) {
      const cplx v = initial[ir + std::size_t(nr_) * iz];
      norm = norm + (v.real() * v.real() + v.imag() * v.imag());
    }
  norm = norm * volume;
  norm_tolerance_ = (norm > p.tol_specular ? norm : p.tol_specular) / volume;
}

std::optional<double> AnalysisSurface::update(const std::vector<cplx>& spec) {
  if (fixed_) return std::nullopt;
  std::optional<double> fixed_at;
  double norm = 0.0;
  for (int iz = last_before_absorber_; iz >= 0; --iz) {
    for (int ir = 0; ir < nr_; ++ir) {
      const cplx v = spec[ir + std::size_t(nr_) * iz];
      norm = norm + (v.real() * v.real() + v.imag() * v.imag());
    }
    if (norm > norm_tolerance_) {
      if (iz + 1 > previous_index_) {
        fixed_ = true;
        index_ = iz + 1;
        if (index_ >= last_before_absorber_) index_ = last_before_absorber_;
        fixed_at = z_[index_];
      }
      index_ = iz + 1;
      previous_index_ = index_;
      break;
    }
    if (iz == iz_analysis_) {
      fixed_ = true;
      index_ = iz_analysis_;
      fixed_at = z_[index_];
    }
  }
  return fixed_at;
}

AnalysisEnergies make_analysis_energies(const Config& c, const GasPhase& gas) {
  AnalysisEnergies e;
  for (int ien = 0; ien < c.input.n_energies; ++ien) {
    const double collision = c.derived.analysis_energy_start + ien * c.derived.analysis_energy_step;
    const double total = collision + Units::hartree_in_ev * gas.initial_energy;
    e.collision.push_back(collision);
    e.total.push_back(total);
    e.total_hartree.push_back(total / Units::hartree_in_ev);
  }
  return e;
}

EnergyResolvedSetup make_energy_resolved_setup(const Config& c, const GasPhase& gas) {
  EnergyResolvedSetup e;
  const InputParameters& p = c.input;
  for (const std::string& name : p.energy_resolved_sites)
    for (const PlotSite& site : p.plot_sites)
      if (site.name == name) {
        e.site_names.push_back(name);
        e.site_ixy.push_back(plot_site_ixy(site, p.nx, p.ny));
        break;
      }
  for (const double collision : p.energy_resolved_energies) {
    const double total = collision + Units::hartree_in_ev * gas.initial_energy;
    e.collision_ev.push_back(collision);
    e.total_hartree.push_back(total / Units::hartree_in_ev);
  }
  e.analysis_time_step = c.derived.steps_per_analysis * p.time_step;
  return e;
}

EnergyResult s_matrix_at(const Config& c, const AnalysisBasis& a, const GasPhase& gas, const AnalysisEnergies& energies,
                         const std::vector<cplx>& sums, int ien, std::optional<double> specular_line) {
  const InputParameters& p = c.input;
  const DerivedParameters& d = c.derived;
  const double chaev = Units::hartree_in_ev;
  const double mass = d.total_mass;
  const int n_states = a.n_states(), n_v = a.n_v;
  EnergyResult e;
  e.total_energy = energies.total[ien];
  e.collision_energy = energies.collision[ien];
  const double kx = d.kx, ky = d.ky;
  const double k_par_sq = kx * kx - (2.0 * std::cos(d.system.skew) * kx * ky) + ky * ky;
  const double k_total = std::sqrt(2.0 * mass * e.collision_energy / chaev);
  const double k0 = std::sqrt(k_total * k_total - k_par_sq);
  const double weight = std::exp(-((d.kz - k0) * (d.kz - k0)) / (2.0 * d.wave_packet_alpha));
  const double prefactor = std::pow(0.5 * d.wave_packet_alpha / kPi, 0.25) / mass;
  const std::size_t n = std::size_t(n_states) * n_v;
  e.s_matrix.assign(n, cplx(0.0, 0.0));
  e.open.assign(n, 0);
  const cplx* s = sums.data() + n * ien;
  for (int v = 0; v < n_v; ++v)
    for (int is = 0; is < n_states; ++is) {
      const std::size_t k = is + std::size_t(n_states) * v;
      const double kz_sq = 2.0 * mass * (e.total_energy / chaev - a.channel_energies[k]);
      if (!(kz_sq > 0.0)) continue;
      const double kz = std::sqrt(kz_sq);
      const double z_measured = (specular_line && is == a.initial) ? *specular_line : p.z_analysis;
      const double arg = -kz * z_measured + (d.kz - k0) * p.z0 - kPi;
      e.s_matrix[k] = s[k] * prefactor * std::exp(cplx(0.0, 1.0) * arg) * std::sqrt(kz * k0 / weight);
      e.open[k] = 1;
    }
  e.probability.assign(n, 0.0);
  e.phase.assign(n, 0.0);
  e.p_vibration.assign(n_v, 0.0);
  double sum = 0.0;
  const std::vector<int> regular = a.regular_state_order();
  for (int v = 0; v < n_v; ++v)
    for (const int is : regular) {
      const std::size_t k = is + std::size_t(n_states) * v;
      const cplx sm = e.s_matrix[k];
      const double prob = sm.real() * sm.real() + sm.imag() * sm.imag();
      e.probability[k] = prob;
      e.phase[k] = sm.real() == 0.0 ? 0.0 : std::atan2(sm.imag(), sm.real());
      e.p_vibration[v] = e.p_vibration[v] + prob;
      sum = sum + prob;
    }
  (void)gas;
  e.p_total = sum;
  e.p_reaction = 1.0 - sum;
  return e;
}

Averages compute_averages(double time, const std::vector<double>& density, const std::vector<cplx>& spec,
                          const Config& c, const SpatialGrids& g, int surface_index) {
  const InputParameters& p = c.input;
  const int nr = p.nr;
  double ns = 0.0, zs = 0.0, rs = 0.0;
  for (int izr = 0; izr < static_cast<int>(density.size()); ++izr) {
    const int ir = izr % nr, iz = izr / nr;
    ns = ns + density[izr];
    zs = zs + g.z[iz] * density[izr];
    rs = rs + g.r[ir] * density[izr];
  }
  const double vol_scat = p.dz * p.dr * c.derived.xy_volume_element;
  ns *= vol_scat;
  zs *= vol_scat;
  rs *= vol_scat;
  auto abs2 = [](const cplx& v) {
    const double a = std::abs(v);
    return a * a;
  };
  double np = 0.0, zp = 0.0, rp = 0.0;
  const int nzsp = g.z_specular.size();
  for (int iz = 0; iz < nzsp; ++iz)
    for (int ir = 0; ir < nr; ++ir) {
      const double dd = abs2(spec[ir + std::size_t(nr) * iz]);
      np = np + dd;
      zp = zp + g.z_specular[iz] * dd;
      rp = rp + g.r[ir] * dd;
    }
  const double vol_spec = p.dz * p.dr;
  np *= vol_spec;
  zp *= vol_spec;
  rp *= vol_spec;
  double rest = 0.0;
  for (int iz = g.iz_analysis + 1; iz <= surface_index; ++iz)
    for (int ir = 0; ir < nr; ++ir) rest = rest + abs2(spec[ir + std::size_t(nr) * iz]);
  Averages av;
  av.time = time;
  av.norm = ns + np;
  av.z_mean = (zs + zp) / av.norm;
  av.r_mean = (rs + rp) / av.norm;
  av.rest_norm = rest * vol_spec;
  return av;
}

}  // namespace spo
