// Run configuration: the input keys the propagation path reads (Fortran DVR.inp names in JSON) and the
// quantities derived from them. Mirrors src/input/config.hpp:63-280 for the keys on the propagation path.
#pragma once

#include <optional>
#include <string>
#include <vector>

#include "common.hpp"
#include "json.hpp"

namespace spo {

// Lattice and asymptote of the molecule-surface system, supplied by the potential
// (src/potential/potential.hpp:14-20).
struct SystemParameters {
  double lattice_x = 0.0;   // bohr (alat)
  double lattice_y = 0.0;   // bohr (blat)
  double z_asymptotic = 0.0;
  double skew = 0.0;        // rad, angle between the lattice vectors
  std::string point_group;  // "C2v", "C6v", ...
};

// A named (x, y) point of the surface cell for the per-site outputs, as fractions of the lattice vectors
// (src/input/config.hpp:53-57).
struct PlotSite {
  std::string name;
  double x = 0.0, y = 0.0;
};

// A low-end absorber, on when both ends are given.
struct LowAbsorber {
  std::optional<double> start, end, energy, factor;
  bool enabled() const { return start.has_value() && end.has_value(); }
};

struct InputParameters {
  std::string title;
  std::string system_id;
  std::string crp_data_set;  // CRPlibDataSet: the CRP data set directory of SystemID CRPlibrary / CRPDistrtSurface
  // SystemID CRPDistrtSurface: the SCM correction's files and choices (src/input/config.hpp:29-36)
  struct Scm {
    int mode = 1;                 // DistortedPESMode: 1 = SCM, the only mode
    std::string parameter_file;   // SCMParFilepath (params.dat)
    std::string position_file;    // PosFilepath (slab snapshots)
    int configuration = 1;        // PosFileSlabNum, 1-based
    bool rescale = true;          // EnableSurfExpan: the slab's lattice constant for the grid, CRP at (u/s, v/s)
    double cutoff = 16.0;         // SCMCutoff (bohr)
  } scm;
  bool x_dof = true, y_dof = true;
  double mass1 = 1.0078, mass2 = 1.0078;
  bool homonuclear = true;
  double z_start = -1.0, dz = 0.25;
  int nz = 56, nz_specular = 80;
  double r_start = 0.4, dr = 0.25;
  int nr = 32, nx = 4, ny = 4;
  int j_max = 6, mj_max = -1;
  int projection_operator = 2;
  // Restart (src/input/config.hpp:67, :96-99, :179-182): NewCalc = false continues the run whose checkpoint is in
  // RestartDir, from tStart: native files by default; RestartFormat / CheckpointFormat "fortran" (or "both" for
  // writing) select the Fortran program's fort.8 / fort.22 (else fort.9 / fort.21) on reading and fort.9 / fort.21 on
  // writing.
  bool new_calculation = true;
  double t_start = 0.0;
  std::string restart_dir;
  std::string restart_format = "native", checkpoint_format = "native";
  double t_end = 1000.0, time_step = 10.0;
  int n_steps = 100;
  double dt_analysis = 50.0, dt_averages = 50.0;
  bool dynamical_zsp_analysis = true;
  double z_analysis = 7.0;
  std::optional<int> j_analysis, v_analysis, diffraction_order_analysis;
  int n_energies = 41;
  std::optional<double> z_analysis_specular;
  double e_min_wp = 0.15, e_max_wp = 0.55, z0 = 10.0;
  int v_in = 0, j_in = 0, mj_in = 0;
  double e_parallel = 0.0, phi_parallel = 0.0;
  bool impose_v_min = false, impose_v_max = false;
  double v_min = -5.5, v_max = 10.0;
  bool use_switching = false;
  double z_min_switch = 6.0, z_max_switch = 8.0;
  bool use_symmetry = false, phis_optimal = false;
  double z_opt_start = 7.0, z_opt_end = 11.25;
  std::optional<double> z_opt_energy, z_opt_factor;
  double r_opt_start = 4.15, r_opt_end = 8.15;
  std::optional<double> r_opt_energy, r_opt_factor;
  double zsp_opt_start = 13.0, zsp_opt_end = 18.0;
  std::optional<double> zsp_opt_energy, zsp_opt_factor;
  LowAbsorber z_opt_low, zsp_opt_low;
  double tol_specular = 0.0;
  // Energy-resolved sums at chosen sites (PlotSites, EnergyResolvedSites, EnergyResolvedEnergies in eV).
  std::vector<PlotSite> plot_sites;
  std::vector<std::string> energy_resolved_sites;
  std::vector<double> energy_resolved_energies;
  // Storage precision of the device backend's scattering wave ("fp64", or "mixed": complex64 with FP64 phases,
  // specular grid and analysis). Not a key of the reference program.
  std::string precision = "fp64";
};

struct DerivedParameters {
  SystemParameters system;
  double z_asymptotic_unswitched = 0.0;
  double z_asymptotic = 0.0;
  int mj_max_basis = 0;
  int phi_points_multiple = 1;
  double total_mass = 0.0, reduced_mass = 0.0;
  double z_opt_factor = 0.0, r_opt_factor = 0.0, zsp_opt_factor = 0.0;
  double z_opt_low_factor = 0.0, zsp_opt_low_factor = 0.0;
  int optical_power = 2;
  double dx = 0.0, dy = 0.0;
  double translational_energy = 0.0;  // eV
  double wave_packet_alpha = 0.0;
  double theta_incidence = 0.0, phi_incidence = 0.0;  // rad
  double analysis_energy_start = 0.0, analysis_energy_step = 0.0;  // eV
  int steps_per_analysis = 1;
  int analyses_per_average = 1;
  bool fixed_z_analysis_specular = false;
  double delta_kx = 0.0, delta_ky = 0.0;
  double k_total = 0.0, kx = 0.0, ky = 0.0, kz = 0.0;
  double unit_cell_area = 0.0, unit_cell_norm = 0.0;
  double xy_volume_element = 0.0;
  double v_min = 0.0, v_max = 0.0;  // hartree
  double confidence_e_min = 0.0, confidence_e_max = 0.0;  // eV
};

struct Config {
  InputParameters input;
  DerivedParameters derived;
};

// The system parameters of a potential by SystemID (HardWallPES, H2Cu111DaiLight; CRPlibrary from its data set).
SystemParameters system_parameters_for(const std::string& system_id, const std::string& crp_data_set = "");
SystemParameters system_parameters_for(const struct InputParameters& input);

// The canonical spelling of a SystemID the set-up knows, matched without case as the reference compares SystemIDs
// (src/input/json_input.cpp:419-420); other names are returned unchanged.
std::string canonical_system_id(const std::string& id);

// Reads the keys from a parsed input document; unknown keys are ignored, keys of other subsystems are not read.
// SystemID is stored in its canonical spelling.
InputParameters read_input(const json::Value& document);
InputParameters read_input_file(const std::string& path);

// Derived parameters (src/input/process.cpp:37-139) and the checks that keep the propagation well defined.
Config build_config(const InputParameters& input);

// The (x, y) grid point of a site, ixy = iy + ny * ix, the nearest grid point to each fraction
// (src/input/validate_outputs.cpp:25, :134-136).
int plot_site_ixy(const PlotSite& site, int nx, int ny);

// Optimal strength of a quadratic absorber of the given length for kinetic energy ekin (hartree), from the
// Vibok / Balint-Kurti table (src/input/wave_packet_params.cpp:14-57).
double absorber_strength(double ekin, double length, double mass);
double wave_packet_alpha(double e_low, double e_high, double mass);
double wave_packet_average_energy(double e_low, double e_high, double mass);

}  // namespace spo
