#include "config.hpp"

#include "grids.hpp"
#include "potential.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <numeric>
#include <tuple>
#include <utility>

namespace spo {

double machine_epsilon() {
  static const double eps = [] {
    volatile double e = 1.0;
    while (true) {
      volatile double a = 1.0 - e;
      if (a >= 1.0) break;
      e = e / 2.0;
    }
    double r = std::abs(e);
    return r <= 0.0 ? 1.0e-12 : r;
  }();
  return eps;
}

SystemParameters system_parameters_for(const std::string& id, const std::string& crp_data_set) {
  // src/potential/hard_wall_potential.cpp:15 and src/potential/dai_light_cu111_potential.cpp:38-39, :71-72.
  if (id == "HardWallPES") return {6.239660, 5.239660, 7.0, kPi / 2.0, "C2v"};
  if (id == "H2Cu111DaiLight") return {4.84, 4.84, 15.0, kPi / 3.0, "C6v"};
  if (id == "CRPlibrary") {  // the data set's lattice constant, Zasy and point group (src/potential/crp_potential.cpp:85-87)
    const auto crp = load_crp_surface(crp_data_set);
    const crp::MoleculeSurface& s = crp->surface();
    return {s.lattice_constant(), s.lattice_constant(), s.z_asymptotic(), crp->skewing_angle(),
            s.symmetry() == crp::Symmetry::c3v ? "C3v" : (s.symmetry() == crp::Symmetry::c4v ? "C4v" : "C6v")};
  }
  throw NotSupportedError("SystemID " + id + " is not supported by this backend's set-up");
}

SystemParameters system_parameters_for(const InputParameters& p) {
  if (p.system_id != "CRPDistrtSurface") return system_parameters_for(p.system_id, p.crp_data_set);
  // the CRP's parameters with the slab's lattice constant when it is rescaled (src/potential/scm/
  // scm_crp_potential.cpp:46-54)
  const auto scm = load_scm_surface(p);
  const crp::MoleculeSurface& s = scm->crp().surface();
  return {scm->lattice_constant(), scm->lattice_constant(), s.z_asymptotic(), scm->crp().skewing_angle(),
          s.symmetry() == crp::Symmetry::c3v ? "C3v" : "C6v"};
}

namespace {

void get(const json::Value& d, const char* key, double& out) {
  if (const auto* v = d.find(key)) {
    if (v->kind != json::Value::Kind::Number) throw InputError(key, "expected a number");
    out = v->number;
  }
}
void get(const json::Value& d, const char* key, int& out) {
  if (const auto* v = d.find(key)) {
    if (v->kind != json::Value::Kind::Number || v->number != std::floor(v->number))
      throw InputError(key, "expected an integer");
    out = static_cast<int>(v->number);
  }
}
void get(const json::Value& d, const char* key, bool& out) {
  if (const auto* v = d.find(key)) {
    if (v->kind != json::Value::Kind::Bool) throw InputError(key, "expected true or false");
    out = v->boolean;
  }
}
void get(const json::Value& d, const char* key, std::string& out) {
  if (const auto* v = d.find(key)) {
    if (v->kind != json::Value::Kind::String) throw InputError(key, "expected a string");
    out = v->text;
  }
}
template <typename T>
void get(const json::Value& d, const char* key, std::optional<T>& out) {
  if (d.find(key)) {
    T value{};
    get(d, key, value);
    out = value;
  }
}

// Vibok & Balint-Kurti, J. Phys. Chem. 96, 8712 (1992), table 3: optimal A/ekin against L/lambda.
constexpr std::array<double, 29> kReducedLength = {1.0,  1.5,  2.0,  2.5,  3.0,  3.5,  4.0,  4.5,  5.0,  5.5,
                                                   6.0,  6.5,  7.0,  7.5,  8.0,  8.5,  9.0,  9.5,  10.0, 10.5,
                                                   11.0, 11.5, 12.0, 12.5, 13.0, 13.5, 14.0, 14.5, 15.0};
constexpr std::array<double, 29> kStrengthRatio = {2.750, 2.640, 2.510, 2.380, 2.230, 2.090, 1.940, 1.800,
                                                   1.670, 1.550, 1.440, 1.340, 1.250, 1.170, 1.100, 1.040,
                                                   0.986, 0.949, 0.903, 0.870, 0.840, 0.813, 0.788, 0.764,
                                                   0.741, 0.718, 0.697, 0.678, 0.663};

double energy_strength(double energy_ev, double length, double mass) {
  return absorber_strength(Units::ev_to_hartree(energy_ev), length, mass);
}

double low_strength(const LowAbsorber& a, double mass) {
  if (!a.enabled()) return 0.0;
  if (a.energy && *a.energy > 0.0) return energy_strength(*a.energy, *a.end - *a.start, mass);
  return a.factor.value_or(0.0);
}

}  // namespace

double absorber_strength(double ekin, double length, double mass) {
  const double momentum = std::sqrt(2.0 * mass * ekin);
  const double wavelength = 2.0 * std::acos(-1.0) / momentum;
  const double x = length / wavelength;
  if (x < kReducedLength.front()) return kStrengthRatio.front() * ekin;
  if (x > kReducedLength.back()) return kStrengthRatio.back() * ekin;
  for (std::size_t i = 0; i + 1 < kReducedLength.size(); ++i)
    if (x > kReducedLength[i] && x <= kReducedLength[i + 1])
      return (kStrengthRatio[i] + (x - kReducedLength[i]) * (kStrengthRatio[i + 1] - kStrengthRatio[i]) /
                                      (kReducedLength[i + 1] - kReducedLength[i])) *
             ekin;
  return 0.0;  // L/lambda exactly at the first table point: no interval contains it
}

double wave_packet_alpha(double e_low, double e_high, double mass) {
  const double kmin = std::sqrt(2.0 * mass * e_low);
  const double kmax = std::sqrt(2.0 * mass * e_high);
  const double k0 = (kmin + kmax) / 2.0;
  return std::pow(k0 - kmin, 2) / 4.0;
}

double wave_packet_average_energy(double e_low, double e_high, double mass) {
  const double kmin = -std::sqrt(2.0 * mass * e_low);
  const double kmax = -std::sqrt(2.0 * mass * e_high);
  const double k0 = (kmin + kmax) / 2.0;
  return std::pow(k0, 2) / (2.0 * mass);
}

std::string canonical_system_id(const std::string& id) {
  for (const char* known : {"HardWallPES", "H2Cu111DaiLight", "CRPlibrary", "CRPDistrtSurface"}) {
    const std::string k(known);
    if (k.size() == id.size() && std::equal(k.begin(), k.end(), id.begin(), [](char a, char b) {
          return std::tolower(static_cast<unsigned char>(a)) == std::tolower(static_cast<unsigned char>(b));
        }))
      return k;
  }
  return id;
}

InputParameters read_input(const json::Value& d) {
  if (d.kind != json::Value::Kind::Object) throw InputError("input", "the document is not a JSON object");
  InputParameters p;
  get(d, "Title", p.title);
  get(d, "SystemID", p.system_id);
  p.system_id = canonical_system_id(p.system_id);
  get(d, "CRPlibDataSet", p.crp_data_set);
  get(d, "DistortedPESMode", p.scm.mode);
  get(d, "SCMParFilepath", p.scm.parameter_file);
  get(d, "PosFilepath", p.scm.position_file);
  get(d, "PosFileSlabNum", p.scm.configuration);
  get(d, "EnableSurfExpan", p.scm.rescale);
  get(d, "SCMCutoff", p.scm.cutoff);
  if (p.system_id != "CRPDistrtSurface")
    for (const char* key : {"DistortedPESMode", "SCMParFilepath", "PosFilepath", "PosFileSlabNum", "EnableSurfExpan",
                            "SCMCutoff"})
      if (d.find(key)) throw InputError(key, std::string(key) + " is only used with SystemID CRPDistrtSurface");
  get(d, "XDoF", p.x_dof);
  get(d, "YDoF", p.y_dof);
  get(d, "Mass1", p.mass1);
  get(d, "Mass2", p.mass2);
  get(d, "HomoNuc", p.homonuclear);
  get(d, "ZStart", p.z_start);
  get(d, "dZ", p.dz);
  get(d, "nZ", p.nz);
  get(d, "nZSpecular", p.nz_specular);
  get(d, "RStart", p.r_start);
  get(d, "dR", p.dr);
  get(d, "nR", p.nr);
  get(d, "nX", p.nx);
  get(d, "nY", p.ny);
  get(d, "JMax", p.j_max);
  get(d, "MJMax", p.mj_max);
  if (p.mj_max < 0) p.mj_max = p.j_max;
  get(d, "ProjOp", p.projection_operator);
  get(d, "NewCalc", p.new_calculation);
  get(d, "tStart", p.t_start);
  get(d, "RestartDir", p.restart_dir);
  get(d, "RestartFormat", p.restart_format);
  get(d, "CheckpointFormat", p.checkpoint_format);
  get(d, "tEnd", p.t_end);
  get(d, "dtStep", p.time_step);
  get(d, "nStep", p.n_steps);
  get(d, "dtAna", p.dt_analysis);
  get(d, "dtAve", p.dt_averages);
  get(d, "DynamZSpAna", p.dynamical_zsp_analysis);
  get(d, "Zana", p.z_analysis);
  get(d, "JAna", p.j_analysis);
  get(d, "VAna", p.v_analysis);
  get(d, "DifOrdAna", p.diffraction_order_analysis);
  get(d, "NrOfEs", p.n_energies);
  get(d, "ZAnaSp", p.z_analysis_specular);
  get(d, "EMinWP", p.e_min_wp);
  get(d, "EMaxWP", p.e_max_wp);
  get(d, "Z0", p.z0);
  get(d, "VIn", p.v_in);
  get(d, "JIn", p.j_in);
  get(d, "MJIn", p.mj_in);
  get(d, "EParallel", p.e_parallel);
  get(d, "PhiParallel", p.phi_parallel);
  get(d, "ImposeVMin", p.impose_v_min);
  get(d, "VMin", p.v_min);
  get(d, "ImposeVMax", p.impose_v_max);
  get(d, "VMax", p.v_max);
  get(d, "UseSwitching", p.use_switching);
  get(d, "ZminSwitch", p.z_min_switch);
  get(d, "ZmaxSwitch", p.z_max_switch);
  get(d, "UseSymmetry", p.use_symmetry);
  get(d, "nPhisOptimal", p.phis_optimal);
  get(d, "ZOptStart", p.z_opt_start);
  get(d, "ZOptEnd", p.z_opt_end);
  get(d, "ZOptEnergy", p.z_opt_energy);
  get(d, "ZOptFactor", p.z_opt_factor);
  get(d, "ROptStart", p.r_opt_start);
  get(d, "ROptEnd", p.r_opt_end);
  get(d, "ROptEnergy", p.r_opt_energy);
  get(d, "ROptFactor", p.r_opt_factor);
  get(d, "ZSpOptStart", p.zsp_opt_start);
  get(d, "ZSpOptEnd", p.zsp_opt_end);
  get(d, "ZSpOptEnergy", p.zsp_opt_energy);
  get(d, "ZSpOptFactor", p.zsp_opt_factor);
  get(d, "ZOptLowStart", p.z_opt_low.start);
  get(d, "ZOptLowEnd", p.z_opt_low.end);
  get(d, "ZOptLowEnergy", p.z_opt_low.energy);
  get(d, "ZOptLowFactor", p.z_opt_low.factor);
  get(d, "ZSpOptLowStart", p.zsp_opt_low.start);
  get(d, "ZSpOptLowEnd", p.zsp_opt_low.end);
  get(d, "ZSpOptLowEnergy", p.zsp_opt_low.energy);
  get(d, "ZSpOptLowFactor", p.zsp_opt_low.factor);
  get(d, "TolSp", p.tol_specular);
  get(d, "Precision", p.precision);
  if (const auto* v = d.find("PlotSites")) {
    if (v->kind != json::Value::Kind::Array) throw InputError("PlotSites", "expected a list of sites");
    for (const json::Value& item : v->items) {
      if (item.kind != json::Value::Kind::Object) throw InputError("PlotSites", "each site must be an object");
      for (const auto& [key, value] : item.members)
        if (key != "name" && key != "x" && key != "y") throw InputError("PlotSites", "unknown site field '" + key + "'");
      PlotSite site;
      const json::Value* name = item.find("name");
      const json::Value* x = item.find("x");
      const json::Value* y = item.find("y");
      if (!name || !x || !y) throw InputError("PlotSites", "a site needs a name, x and y");
      if (name->kind != json::Value::Kind::String || x->kind != json::Value::Kind::Number ||
          y->kind != json::Value::Kind::Number)
        throw InputError("PlotSites", "a site's name is a string, x and y are numbers");
      site.name = name->text;
      site.x = x->number;
      site.y = y->number;
      p.plot_sites.push_back(site);
    }
  }
  if (const auto* v = d.find("EnergyResolvedSites")) {
    if (v->kind != json::Value::Kind::Array) throw InputError("EnergyResolvedSites", "expected a list of site names");
    for (const json::Value& item : v->items) {
      if (item.kind != json::Value::Kind::String) throw InputError("EnergyResolvedSites", "expected site names");
      p.energy_resolved_sites.push_back(item.text);
    }
  }
  if (const auto* v = d.find("EnergyResolvedEnergies")) {
    if (v->kind != json::Value::Kind::Array) throw InputError("EnergyResolvedEnergies", "expected a list of energies");
    for (const json::Value& item : v->items) {
      if (item.kind != json::Value::Kind::Number) throw InputError("EnergyResolvedEnergies", "expected numbers (eV)");
      p.energy_resolved_energies.push_back(item.number);
    }
  }
  return p;
}

int plot_site_ixy(const PlotSite& site, int nx, int ny) {
  auto nearest = [](double fraction, int n) { return static_cast<int>(std::lround(fraction * n)) % n; };
  return nearest(site.y, ny) + ny * nearest(site.x, nx);
}

InputParameters read_input_file(const std::string& path) { return read_input(json::parse_file(path)); }

Config build_config(const InputParameters& input) {
  Config c{input, {}};
  const InputParameters& p = c.input;
  DerivedParameters& d = c.derived;
  if (p.system_id == "CRPDistrtSurface") {  // src/input/json_input.cpp:379-405
    if (p.scm.mode != 1) throw InputError("DistortedPESMode", "only mode 1 (SCM) exists");
    if (p.scm.parameter_file.empty()) throw InputError("SCMParFilepath", "SystemID CRPDistrtSurface needs SCMParFilepath");
    if (p.scm.position_file.empty()) throw InputError("PosFilepath", "SystemID CRPDistrtSurface needs PosFilepath");
    if (p.use_symmetry)
      throw InputError("UseSymmetry", "UseSymmetry = true is not supported with SystemID CRPDistrtSurface");
    if (!p.homonuclear) throw InputError("HomoNuc", "HomoNuc = false is not supported with SystemID CRPDistrtSurface");
  }
  d.system = system_parameters_for(p);
  d.z_asymptotic_unswitched = d.system.z_asymptotic;
  d.z_asymptotic = p.use_switching ? p.z_max_switch : d.system.z_asymptotic;

  if (p.projection_operator != 2) throw NotSupportedError("ProjOp other than 2 is not supported");
  if (p.nr < 2) throw InputError("nR", "nR = 1 is not supported");
  if (p.nz_specular < p.nz) throw InputError("nZSpecular", "nZSpecular < nZ");
  if (p.nx < 2 || p.ny < 2) throw NotSupportedError("nX = 1 or nY = 1 is not supported");
  // src/input/validate_outputs.cpp:28-41, :87-110: sites on the grid, energy-resolved keys together and named.
  for (const PlotSite& site : p.plot_sites)
    for (const auto& [axis, fraction, n] : {std::tuple<const char*, double, int>{"x", site.x, p.nx},
                                            std::tuple<const char*, double, int>{"y", site.y, p.ny}}) {
      if (!(fraction >= 0.0 && fraction < 1.0))
        throw InputError("PlotSites", "site '" + site.name + "': " + axis + " is not in [0, 1)");
      const double scaled = fraction * n, nearest = std::round(scaled);
      if (!(std::abs(scaled - nearest) <= 1e-6)) throw InputError("PlotSites", "site '" + site.name + "' is off the grid");
      if (nearest == n) throw InputError("PlotSites", "site '" + site.name + "' rounds to 1: give 0");
    }
  if (p.energy_resolved_sites.empty() != p.energy_resolved_energies.empty())
    throw InputError(p.energy_resolved_sites.empty() ? "EnergyResolvedEnergies" : "EnergyResolvedSites",
                     "EnergyResolvedSites and EnergyResolvedEnergies go together");
  for (std::size_t i = 0; i < p.energy_resolved_sites.size(); ++i) {
    const std::string& name = p.energy_resolved_sites[i];
    bool found = false;
    for (const PlotSite& site : p.plot_sites) found = found || site.name == name;
    if (!found) throw InputError("EnergyResolvedSites", "'" + name + "' is not a site of PlotSites");
    for (std::size_t k = 0; k < i; ++k)
      if (p.energy_resolved_sites[k] == name) throw InputError("EnergyResolvedSites", "'" + name + "' is listed twice");
  }
  for (const double e : p.energy_resolved_energies)
    if (tol::lt(e, p.e_min_wp + p.e_parallel) || tol::gt(e, p.e_max_wp + p.e_parallel))
      throw InputError("EnergyResolvedEnergies", "an energy lies outside the packet's collision energies "
                                                 "(EMinWP, EMaxWP + EParallel)");
  if (p.precision != "fp64" && p.precision != "mixed")
    throw InputError("Precision", "Precision must be \"fp64\" or \"mixed\", not \"" + p.precision + "\"");
  // src/input/validate.cpp:48-66: every FFT grid must have a length 2^h 3^i 5^j 7^k 11^m (h >= 1, i <= 2).
  const std::pair<const char*, int> fft_grids[] = {
      {"nZ", p.nz}, {"nZSpecular", p.nz_specular}, {"nR", p.nr}, {"nX", p.nx}, {"nY", p.ny}};
  for (const auto& [key, n] : fft_grids)
    if (n < 1 || nearest_fft_size(n, 1) != n) throw InputError(key, std::string(key) + " is not a valid FFT size");
  {
    const std::string& g = d.system.point_group;
    if ((g == "C4v" || g == "C6v" || g == "C3v") && p.nx != p.ny) throw InputError("nY", "nX /= nY");
  }

  // src/input/process.cpp:37-94 (ProcessInput).
  d.mj_max_basis = p.mj_max < 1 ? p.j_max : p.mj_max;
  if (p.use_symmetry && p.phis_optimal) {
    // 12 / gcd(12, angles) over the group's operations: C2v 2, C4v 4, C6v 6, C3v 3.
    const std::string& g = d.system.point_group;
    d.phi_points_multiple = g == "C2v" ? 2 : g == "C4v" ? 4 : g == "C6v" ? 6 : g == "C3v" || g == "C3" ? 3 : 1;
  }
  const double total_mass = Units::amu_to_me(p.mass1 + p.mass2);
  const double reduced_mass = Units::amu_to_me((p.mass1 * p.mass2) / (p.mass1 + p.mass2));
  d.optical_power = 2;
  d.z_opt_factor = p.z_opt_factor.value_or(0.0);
  d.r_opt_factor = p.r_opt_factor.value_or(0.0);
  d.zsp_opt_factor = p.zsp_opt_factor.value_or(0.0);
  if (p.z_opt_energy && *p.z_opt_energy > 0.0)
    d.z_opt_factor = energy_strength(*p.z_opt_energy, p.z_opt_end - p.z_opt_start, total_mass);
  d.z_opt_low_factor = low_strength(p.z_opt_low, total_mass);
  if (p.r_opt_energy && *p.r_opt_energy > 0.0)
    d.r_opt_factor = energy_strength(*p.r_opt_energy, p.r_opt_end - p.r_opt_start, reduced_mass);
  if (p.zsp_opt_energy && *p.zsp_opt_energy > 0.0)
    d.zsp_opt_factor = energy_strength(*p.zsp_opt_energy, p.zsp_opt_end - p.zsp_opt_start, total_mass);
  d.zsp_opt_low_factor = low_strength(p.zsp_opt_low, total_mass);
  d.dx = d.system.lattice_x / p.nx;
  d.dy = d.system.lattice_y / p.ny;
  d.translational_energy =
      p.e_parallel + Units::hartree_to_ev(wave_packet_average_energy(
                         Units::ev_to_hartree(p.e_min_wp), Units::ev_to_hartree(p.e_max_wp), total_mass));
  d.wave_packet_alpha =
      wave_packet_alpha(Units::ev_to_hartree(p.e_min_wp), Units::ev_to_hartree(p.e_max_wp), total_mass);
  const double deg = kPi / 180.0;
  const double theta_deg = std::asin(std::sqrt(p.e_parallel / d.translational_energy)) / deg;
  d.analysis_energy_start = p.e_min_wp + p.e_parallel;
  d.analysis_energy_step = (p.e_max_wp - p.e_min_wp) / (p.n_energies - 1);
  d.steps_per_analysis = static_cast<int>(p.dt_analysis / p.time_step);
  d.analyses_per_average = static_cast<int>(p.dt_averages / p.dt_analysis);
  d.fixed_z_analysis_specular = !p.dynamical_zsp_analysis;
  if (d.steps_per_analysis < 1) throw InputError("dtAna", "dtAna < dtStep");
  if (p.n_steps % d.steps_per_analysis != 0) throw InputError("nStep", "nStep is not a multiple of dtAna/dtStep");
  // src/input/validate.cpp:90-103 (the restart and time checks).
  if (tol::lt(p.t_start, 0.0)) throw InputError("tStart", "tStart < 0.");
  if (p.new_calculation && !tol::eq(p.t_start, 0.0))
    throw InputError("tStart", "This is a new calculation but tStart /= 0.");
  if (!p.new_calculation && tol::eq(p.t_start, 0.0)) throw InputError("tStart", "This is a restart but tStart = 0.");
  if (!p.new_calculation && p.restart_dir.empty())
    throw InputError("RestartDir", "This is a restart but RestartDir is not given.");
  if (p.restart_format != "native" && p.restart_format != "fortran")
    throw InputError("RestartFormat", "RestartFormat must be \"native\" or \"fortran\".");
  if (p.checkpoint_format != "native" && p.checkpoint_format != "fortran" && p.checkpoint_format != "both")
    throw InputError("CheckpointFormat", "CheckpointFormat must be \"native\", \"fortran\" or \"both\".");
  if (tol::gt(p.t_start, p.t_end)) throw InputError("tStart", "tStart > tEnd.");
  if (!tol::eq(p.t_end, p.t_start + p.n_steps * p.time_step))
    throw InputError("tEnd", "tEnd /= tStart + nStep*dtStep.");

  // src/input/process.cpp:97-139 (MiscellaneousSetup, scalar part).
  const double skew = d.system.skew;
  d.total_mass = total_mass;
  d.reduced_mass = reduced_mass;
  d.delta_kx = 2.0 * kPi / (d.system.lattice_x * std::sin(skew));
  d.delta_ky = 2.0 * kPi / (d.system.lattice_y * std::sin(skew));
  d.theta_incidence = deg * theta_deg;
  d.phi_incidence = deg * p.phi_parallel;
  d.k_total = std::sqrt(Units::ev_to_hartree(2.0 * total_mass * d.translational_energy));
  const double kx_cart = d.k_total * std::sin(d.theta_incidence) * std::cos(d.phi_incidence);
  const double ky_cart = d.k_total * std::sin(d.theta_incidence) * std::sin(d.phi_incidence);
  d.kx = kx_cart / std::sin(skew);
  d.ky = (std::cos(skew) * kx_cart + std::sin(skew) * ky_cart) / std::sin(skew);
  d.kz = d.k_total * std::cos(d.theta_incidence);
  d.unit_cell_area = d.system.lattice_x * d.system.lattice_y * std::sin(skew);
  d.unit_cell_norm = std::sqrt(1.0 / d.unit_cell_area);
  d.xy_volume_element = d.dx * d.dy * std::sin(skew);
  d.v_min = Units::ev_to_hartree(p.v_min);
  d.v_max = Units::ev_to_hartree(p.v_max);
  const double k_max = d.kz + 2.0 * std::sqrt(d.wave_packet_alpha);
  const double k_min = d.kz - 2.0 * std::sqrt(d.wave_packet_alpha);
  d.confidence_e_max = 0.5 / total_mass * (k_max * k_max) * Units::hartree_to_ev(1.0);
  d.confidence_e_min = 0.5 / total_mass * (k_min * k_min) * Units::hartree_to_ev(1.0);
  return c;
}

}  // namespace spo
