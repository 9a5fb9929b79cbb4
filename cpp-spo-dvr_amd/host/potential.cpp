#include "potential.hpp"

#include <map>
#include <mutex>

#include <cmath>
#include <stdexcept>
#include <string>

#include "../../include/spo_hip.h"

namespace spo {

double HardWallPes::gas_phase(double r_in) {
  const double d = 4.8286, a1 = 2.282, a2 = 1.555, a3 = 0.7533, a4 = 2.23, re = 1.40;
  const double r = r_in - re;
  return -d * (1.0 + a1 * r + a2 * std::pow(r, 2.0) + a3 * std::pow(r, 3.0)) * std::exp(-a4 * r) + d;
}

double HardWallPes::value(double, double, double z, double r, double, double) const {
  const double vr = gas_phase(r);
  const double vz = 1.0 / (z * z * z * z + 0.001);
  const double z_max = 7.0, z_min = 0.1778;
  if (z >= z_max) return vr;
  if (z < z_min) return vr + vz;
  return vr + (1.0 - (0.5 - 0.5 * std::cos(kPi * (z - z_min) / (z_max - z_min)))) * vz;
}

namespace {

// Values at the top, bridge and centre sites, interpolated with the two lowest hexagonal Fourier stars.
struct Sites {
  double top, bridge, centre;
};

constexpr double kCell = 4.84;

double hexagonal_sites(double x, double y, const Sites& v) {
  constexpr double two_pi = 6.28318530717958647688;
  constexpr double four_pi = 12.56637061435917295376;
  const double c0 = v.top / 3.0 + v.centre * 2.0 / 3.0;
  const double c1 = v.top / 4.0 - v.bridge / 4.0;
  const double c2 = v.bridge / 4.0 - v.centre * 2.0 / 9.0 - v.top / 36.0;
  const double s1 = std::cos(two_pi / kCell * x) + std::cos(two_pi / kCell * y) + std::cos(two_pi / kCell * (x + y));
  const double s2 = std::cos(four_pi / kCell * x) + std::cos(four_pi / kCell * y) + std::cos(four_pi / kCell * (x + y));
  return c0 + c1 * s1 + c2 * s2;
}

// LEPS Coulomb and exchange integrals with Sato parameter delta.
double leps_u(double x, double d, double delta, double alpha, double x0) {
  return d / 4.0 / (1.0 + delta) *
         ((3.0 + delta) * std::exp(-2.0 * alpha * (x - x0)) - (2.0 + 6.0 * delta) * std::exp(-alpha * (x - x0)));
}
double leps_a(double x, double d, double delta, double alpha, double x0) {
  return d / 4.0 / (1.0 + delta) *
         ((1.0 + 3.0 * delta) * std::exp(-2.0 * alpha * (x - x0)) - (6.0 + 2.0 * delta) * std::exp(-alpha * (x - x0)));
}

}  // namespace

double DaiLightPes::value(double x, double y, double z, double r, double theta, double phi) const {
  constexpr double m1 = 1.00794, m2 = 1.00794;
  constexpr double d_m = .1744, alpha_m = 1.028, r_m = 1.40, delta_m = -0.15;
  constexpr Sites d_a{.638e-1, .818e-1, .877e-1};
  constexpr Sites alpha_a{.790, .558, .499};
  constexpr Sites z_a{3.02, 2.13, 1.88};
  constexpr Sites delta_a{0.10, 0.10, 0.10};
  constexpr double angle = 1.04719755119659774614;
  const double w1 = m2 / (m1 + m2);
  const double w2 = -m1 / (m1 + m2);
  const double rp = r * std::sin(theta) / std::sin(angle);
  double t = std::sin(angle - phi) * rp;
  const double x1 = x + w1 * t, x2 = x + w2 * t;
  t = std::sin(phi) * rp;
  const double y1 = y + w1 * t, y2 = y + w2 * t;
  t = r * std::cos(theta);
  const double z1 = z + w1 * t, z2 = z + w2 * t;
  const double d1 = hexagonal_sites(x1, y1, d_a), d2 = hexagonal_sites(x2, y2, d_a);
  const double al1 = hexagonal_sites(x1, y1, alpha_a), al2 = hexagonal_sites(x2, y2, alpha_a);
  const double z01 = hexagonal_sites(x1, y1, z_a), z02 = hexagonal_sites(x2, y2, z_a);
  const double de1 = hexagonal_sites(x1, y1, delta_a), de2 = hexagonal_sites(x2, y2, delta_a);
  const double um = leps_u(r, d_m, delta_m, alpha_m, r_m);
  const double am = leps_a(r, d_m, delta_m, alpha_m, r_m);
  const double u1 = leps_u(z1, d1, de1, al1, z01), a1 = leps_a(z1, d1, de1, al1, z01);
  const double u2 = leps_u(z2, d2, de2, al2, z02), a2 = leps_a(z2, d2, de2, al2, z02);
  const double as = a1 + a2;
  double v = um + u1 + u2 - std::sqrt(am * am + as * as - am * as);
  if (tol::ge(z, 15.0)) v = um + am;
  return Units::hartree_to_ev(v);
}

double CrpPes::value(double x, double y, double z, double r, double theta, double phi) const {
  try {
    return surface_->potential_ev(x, y, z, r, theta, phi);
  } catch (const crp::EvaluationError& e) {
    throw InputError("CRPlibDataSet", std::string("the CRP PES cannot be evaluated on this grid: ") + e.what());
  }
}

std::shared_ptr<const crp::QdSurface> load_crp_surface(const std::string& directory) {
  static std::mutex lock;
  static std::map<std::string, std::shared_ptr<const crp::QdSurface>> loaded;
  if (directory.empty()) throw InputError("CRPlibDataSet", "SystemID CRPlibrary needs CRPlibDataSet (the directory of the CRP data set)");
  const std::lock_guard<std::mutex> guard(lock);
  auto& slot = loaded[directory];
  if (!slot) {
    try {
      slot = std::make_shared<const crp::QdSurface>(crp::read_data_set(directory));
    } catch (const crp::DataError& e) {
      loaded.erase(directory);
      throw InputError("CRPlibDataSet", e.what());
    } catch (const crp::EvaluationError& e) {
      loaded.erase(directory);
      throw InputError("CRPlibDataSet", e.what());
    }
  }
  return slot;
}

double ScmPes::value(double x, double y, double z, double r, double theta, double phi) const {
  try {
    return surface_->value_ev(x, y, z, r, theta, phi);
  } catch (const crp::EvaluationError& e) {
    throw InputError("CRPlibDataSet", std::string("the CRP PES cannot be evaluated on this grid: ") + e.what());
  }
}

std::shared_ptr<const scm::ScmSurface> load_scm_surface(const InputParameters& p) {
  static std::mutex lock;
  static std::map<std::string, std::shared_ptr<const scm::ScmSurface>> loaded;
  const std::string key = p.crp_data_set + "\n" + p.scm.parameter_file + "\n" + p.scm.position_file + "\n" +
                          std::to_string(p.scm.configuration) + "\n" + std::to_string(p.scm.rescale) + "\n" +
                          std::to_string(p.scm.cutoff);
  const auto crp = load_crp_surface(p.crp_data_set);
  const std::lock_guard<std::mutex> guard(lock);
  auto& slot = loaded[key];
  if (!slot) {
    try {
      slot = std::make_shared<const scm::ScmSurface>(crp, scm::read_parameters(p.scm.parameter_file),
                                                     scm::read_slab(p.scm.position_file, p.scm.configuration,
                                                                    p.scm.cutoff),
                                                     p.scm.rescale);
    } catch (const crp::DataError& e) {
      loaded.erase(key);
      throw InputError("PosFilepath", e.what());
    }
  }
  return slot;
}

std::unique_ptr<Pes> make_pes(const InputParameters& p) {
  if (p.system_id == "CRPDistrtSurface") return std::make_unique<ScmPes>(load_scm_surface(p));
  return make_pes(p.system_id, p.crp_data_set);
}

std::unique_ptr<Pes> make_pes(const std::string& system_id, const std::string& crp_data_set) {
  const std::string id = canonical_system_id(system_id);
  if (id == "HardWallPES") return std::make_unique<HardWallPes>();
  if (id == "H2Cu111DaiLight") return std::make_unique<DaiLightPes>();
  if (id == "CRPlibrary") return std::make_unique<CrpPes>(load_crp_surface(crp_data_set));
  throw NotSupportedError("SystemID " + id + " is not supported");
}

namespace {

std::vector<double> gas_phase_potential(const Pes& pes, double z, const SpatialGrids& g) {
  std::vector<double> v(g.r.size());
  for (int ir = 0; ir < g.r.size(); ++ir) v[ir] = Units::ev_to_hartree(pes.value(0.0, 0.0, z, g.r[ir], kPi / 2.0, 0.0));
  return v;
}

void clamp(std::vector<double>& v, const Config& c) {
  if (c.input.impose_v_min)
    for (double& x : v)
      if (x < c.derived.v_min) x = c.derived.v_min;
  if (c.input.impose_v_max)
    for (double& x : v)
      if (x > c.derived.v_max) x = c.derived.v_max;
}

}  // namespace

spo_hip_crp_surface crp_surface_view(const crp::Packed& p) {
  spo_hip_crp_surface v{};
  v.c3v = p.c3v;
  v.lattice_constant = p.lattice_constant;
  v.zso_min = p.zso_min;
  v.zso_max = p.zso_max;
  v.r_min = p.r_min;
  v.r_max = p.r_max;
  v.z_min = p.z_min;
  v.h2_asymptote = p.h2_asymptote;
  v.n_corrugation = static_cast<int>(p.corrugation.size() / 3);
  v.corrugation = p.corrugation.data();
  v.n_cluster = static_cast<int>(p.cluster.size() / 4);
  v.cluster = p.cluster.data();
  v.atom_offsets = p.atom_offsets.data();
  v.atom_splines = p.atom_splines.data();
  v.n_cuts = static_cast<int>(p.cut_shape.size() / 3);
  v.cut_shape = p.cut_shape.data();
  v.cut_offsets = p.cut_offsets.data();
  v.cut_data = p.cut_data.data();
  v.n_h2 = static_cast<int>(p.h2.size() / 3);
  v.h2 = p.h2.data();
  v.n_z_dependence = static_cast<int>(p.z_dependence.size() / 3);
  v.z_dependence = p.z_dependence.data();
  v.square = p.square;
  v.cell_nz = p.cell_nz;
  v.cell_spline = p.cell_spline.empty() ? nullptr : p.cell_spline.data();
  return v;
}

spo_hip_scm_correction scm_correction_view(const scm::ScmSurface& s, std::vector<double>& atoms) {
  spo_hip_scm_correction v{};
  const scm::Correction& m = s.correction();
  v.rescale = m.rescaling() ? 1 : 0;
  v.scale = s.crp_scale();
  const auto unit_v = m.to_cartesian(0.0, 1.0);  // (cos, sin) of the skewing angle, as the host computes them
  v.cos_skew = unit_v[0];
  v.sin_skew = unit_v[1];
  for (int l = 0; l < 2; ++l) {
    const scm::Layer& layer = m.layers()[l];
    for (int k = 0; k < 8; ++k) {
      v.layers[l][k] = layer.intercept[k];
      v.layers[l][8 + k] = layer.slope[k];
    }
    v.layers[l][16] = layer.r_clamp_min;
    v.layers[l][17] = layer.r_clamp_max;
  }
  atoms.clear();
  for (const scm::Atom& at : m.atoms())
    atoms.insert(atoms.end(), {at.distorted.x, at.distorted.y, at.distorted.z, at.ideal.x, at.ideal.y, at.ideal.z,
                               static_cast<double>(at.layer)});
  v.n_atoms = static_cast<int>(m.atoms().size());
  v.atoms = atoms.data();
  return v;
}

PotentialGrid compute_potential_grid_on_device(const Pes& pes, const Config& c, const SpatialGrids& g,
                                               const AngularGrid& a) {
  int id = 0;
  if (c.input.system_id == "HardWallPES") id = SPO_HIP_PES_HARDWALL;
  if (c.input.system_id == "H2Cu111DaiLight") id = SPO_HIP_PES_DAILIGHT;
  if (c.input.system_id == "CRPlibrary" || c.input.system_id == "CRPDistrtSurface") id = SPO_HIP_PES_CRP;
  if (id == 0) throw NotSupportedError("SystemID " + c.input.system_id + " has no device surface");
  crp::Packed packed;
  spo_hip_crp_surface crp_view{};
  std::vector<double> scm_atoms;
  spo_hip_scm_correction scm_view{};
  const bool distorted = c.input.system_id == "CRPDistrtSurface";
  if (id == SPO_HIP_PES_CRP) {
    packed = crp::pack(*load_crp_surface(c.input.crp_data_set));
    crp_view = crp_surface_view(packed);
  }
  if (distorted) scm_view = scm_correction_view(*load_scm_surface(c.input), scm_atoms);
  PotentialGrid grid;
  grid.symmetry = make_grid_symmetry(c, a.n_phi());
  const int nr = g.r.size();
  grid.n_tp = a.n_theta() * a.n_phi();
  grid.n_rep = grid.symmetry.n_representatives();
  grid.n_zr = g.nz_potential * nr;
  grid.values.resize(std::size_t(grid.n_tp) * grid.n_rep * grid.n_zr);
  const std::vector<double> v_gas = c.input.use_switching
                                        ? gas_phase_potential(pes, c.derived.z_asymptotic_unswitched, g)
                                        : std::vector<double>(nr, 0.0);
  spo_hip_pes_grid d{};
  d.pes = id;
  d.nx = g.x.size();
  d.ny = g.y.size();
  d.nr = nr;
  d.nz_potential = g.nz_potential;
  d.n_theta = a.n_theta();
  d.n_phi = a.n_phi();
  d.n_representatives = grid.n_rep;
  d.x = g.x.points.data();
  d.y = g.y.points.data();
  d.z = g.z.points.data();
  d.r = g.r.points.data();
  d.theta = a.thetas.data();
  d.phi = a.phis.data();
  d.representatives = grid.symmetry.representatives.data();
  d.use_switching = c.input.use_switching ? 1 : 0;
  d.z_min_switch = c.input.z_min_switch;
  d.z_max_switch = c.input.z_max_switch;
  d.gas_phase = v_gas.data();
  d.impose_v_min = c.input.impose_v_min ? 1 : 0;
  d.impose_v_max = c.input.impose_v_max ? 1 : 0;
  d.v_min = c.derived.v_min;
  d.v_max = c.derived.v_max;
  d.crp = id == SPO_HIP_PES_CRP ? &crp_view : nullptr;
  d.scm = distorted ? &scm_view : nullptr;
  if (spo_hip_potential_grid(&d, grid.values.data()) != SPO_HIP_OK)
    throw std::runtime_error(std::string("device potential grid: ") + spo_hip_last_error());
  grid.asymptotic = gas_phase_potential(pes, c.derived.z_asymptotic, g);
  clamp(grid.asymptotic, c);
  return grid;
}

PotentialGrid compute_potential_grid(const Pes& pes, const Config& c, const SpatialGrids& g, const AngularGrid& a) {
  PotentialGrid grid;
  grid.symmetry = make_grid_symmetry(c, a.n_phi());
  const int n_theta = a.n_theta(), n_phi = a.n_phi();
  const int ny = g.y.size(), nr = g.r.size();
  grid.n_tp = n_theta * n_phi;
  grid.n_rep = grid.symmetry.n_representatives();
  grid.n_zr = g.nz_potential * nr;
  grid.values.resize(std::size_t(grid.n_tp) * grid.n_rep * grid.n_zr);
  const double to_hartree = Units::ev_to_hartree(1.0);  // multiplied, as the reference's 1/chaev factor
#pragma omp parallel for schedule(dynamic, 1)
  for (int izr = 0; izr < grid.n_zr; ++izr) {
    const int iz = izr / nr, ir = izr % nr;
    for (int irep = 0; irep < grid.n_rep; ++irep) {
      const int ixy = grid.symmetry.representatives[irep];
      const int ix = ixy / ny, iy = ixy % ny;
      double* out = grid.values.data() + std::size_t(grid.n_tp) * (irep + std::size_t(grid.n_rep) * izr);
      for (int it = 0; it < n_theta; ++it)
        for (int ip = 0; ip < n_phi; ++ip)
          out[ip + n_phi * it] = pes.value(g.x[ix], g.y[iy], g.z[iz], g.r[ir], a.thetas[it], a.phis[ip]) * to_hartree;
    }
  }
  if (c.input.use_switching) {
    // Cosine blend into the gas phase between ZminSwitch and ZmaxSwitch (src/potential/potential_grid.cpp:35-56).
    const double zmin = c.input.z_min_switch, zmax = c.input.z_max_switch;
    const std::vector<double> v_gas = gas_phase_potential(pes, c.derived.z_asymptotic_unswitched, g);
    const std::size_t block = std::size_t(grid.n_tp) * grid.n_rep;
    for (int iz = 0; iz < g.nz_potential; ++iz) {
      const double z = g.z[iz];
      if (tol::le(z, zmin)) continue;
      const double f = 0.5 + 0.5 * std::cos(kPi * (z - zmin) / (zmax - zmin));
      for (int ir = 0; ir < nr; ++ir) {
        const double va = v_gas[ir];
        double* b = grid.values.data() + (ir + std::size_t(nr) * iz) * block;
        for (std::size_t k = 0; k < block; ++k) b[k] = tol::ge(z, zmax) ? va : f * (b[k] - va) + va;
      }
    }
  }
  clamp(grid.values, c);
  grid.asymptotic = gas_phase_potential(pes, c.derived.z_asymptotic, g);
  clamp(grid.asymptotic, c);
  return grid;
}

}  // namespace spo
