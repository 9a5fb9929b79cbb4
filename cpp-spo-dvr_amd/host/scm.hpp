// The SCM (static corrugation model) correction: the change dV of the H2-surface potential when the surface atoms
// move from their ideal lattice sites to those of a distorted slab snapshot, a sum of H-metal pair functions over
// the atoms of the top two layers, added to the CRP potential (SystemID CRPDistrtSurface). The reference's
// src/potential/scm/ (scm_correction.cpp, scm_files.cpp, scm_crp_potential.cpp), itself a restatement of the
// Fortran DistortedSurfacePES.dir (SCM_Potential.f90, SCM_IO.f90, SCM_Wrapper.f90). Operand orders follow the
// Fortran: dV agrees with it bit for bit.
#pragma once

#include <array>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "crp.hpp"

namespace spo::scm {

// The pair-function parameters of one layer: at bond length r the eight parameters are slope r_c + intercept,
// r_c = r clamped to [r_clamp_min, r_clamp_max] (params.dat P1-P18; src/potential/scm/scm_files.hpp:13-22).
struct Layer {
  std::array<double, 8> intercept{}, slope{};
  double r_clamp_min = 0.0, r_clamp_max = 0.0;
};

struct Parameters {
  int version = 0;
  std::array<Layer, 2> layers{};  // the top layer, the one below it
};

struct Atom {
  crp::Vec3 distorted, ideal;
  int layer = 0;  // 1 top, 2 below
};

struct Slab {
  std::string tag, structure;
  int nx = 0, ny = 0, nz = 0, configurations = 0, version = 0, max_surface_atoms = 0;
  double lattice_constant = 0.0;  // a_T of the slab (bohr)
  int configuration = 0;
  double cutoff = 0.0;
  std::vector<Atom> atoms;  // layers 1-2; version 3: inside the cut-off sphere around the origin
};

// params.dat: "VERSION 2", then 18 list-directed values per layer (src/potential/scm/scm_files.cpp:222-249).
Parameters read_parameters(const std::string& file);
// A position file's header and snapshot `configuration` (1-based), version 3 or 2
// (src/potential/scm/scm_files.cpp:92-275).
Slab read_slab(const std::string& file, int configuration, double cutoff);

// dV of one snapshot (src/potential/scm/scm_correction.cpp).
class Correction {
 public:
  Correction(const Parameters& parameters, const Slab& slab, std::optional<double> rescaling, double skewing_angle);
  // Cartesian (x, y) of skewed (u, v): (u + cos g v, sin g v)
  std::array<double, 2> to_cartesian(double u, double v) const { return {u + cos_skew_ * v, sin_skew_ * v}; }
  // dV in hartree at the Cartesian centre (x, y), Z, r, theta, phi
  double value_hartree(double x, double y, double z, double r, double theta, double phi) const;
  const std::vector<Atom>& atoms() const { return atoms_; }
  const std::array<Layer, 2>& layers() const { return layers_; }
  std::optional<double> rescaling() const { return rescaling_; }

 private:
  std::array<Layer, 2> layers_;
  std::vector<Atom> atoms_;
  std::optional<double> rescaling_;
  double cos_skew_, sin_skew_;
};

// The pair function's terms at bond length r and its value at an H-metal distance (Setup_3Body_Rydberg_Par,
// Switch_Rydberg_Potential).
struct PairTerms {
  double reference_distance, decay;
  std::array<double, 4> polynomial;
  double switch_start, switch_end, switched_value;
};
PairTerms pair_terms(const Layer& layer, double r);
double pair_energy(const PairTerms& terms, double distance);

// The CRP surface plus dV, V in eV at skewed (u, v) (src/potential/scm/scm_crp_potential.cpp:70-76): the CRP at
// (u / s, v / s) with s = a_T / a_CRP when the slab is rescaled to the CRP lattice, dV in eV with the Fortran's
// Constants_eV2Ha.
class ScmSurface {
 public:
  ScmSurface(std::shared_ptr<const crp::QdSurface> crp, const Parameters& parameters, const Slab& slab, bool rescale);
  double value_ev(double u, double v, double z, double r, double theta, double phi) const;
  const Correction& correction() const { return correction_; }
  const crp::QdSurface& crp() const { return *crp_; }
  double crp_scale() const { return scale_; }
  // the lattice constant of the propagation grid: the slab's a_T when rescaling, else the CRP's
  double lattice_constant() const;

 private:
  std::shared_ptr<const crp::QdSurface> crp_;
  Slab slab_;
  std::optional<double> rescaling_;
  double scale_;
  Correction correction_;
};

}  // namespace spo::scm
