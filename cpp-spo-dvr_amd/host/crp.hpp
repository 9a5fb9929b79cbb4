// The CRP (corrugation-reducing procedure) potential energy surface of a diatomic molecule at a hexagonal (C3v or
// C6v) surface, read from a CRP data set: the reference's CRP library (src/crplib/, a C++ restatement of the
// Fortran library of SPO-DVR) as far as the SystemID "CRPlibrary" of the propagation path uses it.
//
//   data set     src/crplib/dataset.cpp:70-233 (CRPData, CRPData.3D, CRPData.6D, Atom/, Molecule/; list-directed
//                text, src/fortran_text/list_directed_reader.cpp)
//   splines      src/crplib/spline.cpp (DSPLIN, TRIDIA, SPLINT, DPL, DPLP, UBICAR, SPLIN2D, SPLCOFII, SPLINII)
//   atom PES     src/crplib/hexagonal_atom_surface.cpp (MURII cluster sum, FOUR3II basis, ATII)
//   molecule PES src/crplib/molecule_surface.cpp, site_interpolation.cpp, hexagonal_interaction.cpp (CAMBCORD,
//                site functions, FOUR6II / SISII, chain rule, switch to the gas-phase potential)
//   QD wrapper   src/crplib/qd_wrapper.cpp (clamps, skewed coordinates, eV, asymptote shift)
//
// Every arithmetic expression keeps the operand order of the Fortran library, so that results agree with it bit for
// bit where the data set holds the interpolation function (CutsInterpolation T) and to a few ulp otherwise (there the
// Fortran reduces the cuts with a vectorised cosine). The bitwise-neutral speed-ups of the reference (cluster
// pruning, shared grid lookups, value-only evaluation) are left out: they change no result.
#pragma once

#include <array>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace spo::crp {

// The data set cannot be used (src/crplib/errors.hpp:9-14).
struct DataError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// The surface cannot be evaluated at this geometry, where the Fortran stops (src/crplib/errors.hpp:16-21).
struct EvaluationError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

enum class Symmetry { c3v, c4v, c6v };
enum class LengthUnit { bohr = 0, angstrom = 1 };
enum class EnergyUnit { hartree = 2, ev = 3, kcal_per_mol = 4, kj_per_mol = 5, joule = 6 };
struct Units {
  LengthUnit length;
  EnergyUnit energy;
};
inline constexpr Units kInternalUnits{LengthUnit::bohr, EnergyUnit::hartree};

// Unit conversion with the library's own factors (src/crplib/units.cpp:13-79).
double to_length(double x, LengthUnit from, LengthUnit to);
double to_energy(double x, EnergyUnit from, EnergyUnit to);
double to_force(double x, Units from, Units to);

// x -> f table of a two-column file.
struct Table {
  std::vector<double> x, f;
};

// One atom-surface cut: the site (Cartesian, lattice units), end slopes dV/dz and V(z).
struct AtomCut {
  double x = 0.0, y = 0.0;
  std::string file;
  double slope_first = 0.0, slope_last = 0.0;
  Table v;
};

// One molecule-surface cut: the centre (skewed lattice units for C3v/C6v), the orientation in degrees and V on the
// (r, Z) grid, v[ir + nr * iz]. `present` is false for an unused extra cut whose file is missing.
struct MoleculeCut {
  double x = 0.0, y = 0.0, theta = 0.0, phi = 0.0;
  std::string file;
  std::vector<double> r, z, v;
  bool present = true;
};

struct DataSet {
  Symmetry symmetry = Symmetry::c6v;
  Units units = kInternalUnits;
  double lattice_constant = 0.0;
  std::vector<double> interlayer_spacings;
  double displacement_1 = 0.0, displacement_2 = 0.0;
  Table corrugation;
  double corrugation_slope_first = 0.0, corrugation_slope_last = 0.0;
  std::vector<AtomCut> atom_cuts;
  bool cuts_interpolation = false;
  double zso_min = 0.0, zso_max = 0.0;
  Table z_dependence;
  Table h2;
  std::vector<MoleculeCut> molecule_cuts;
};

// Reads and checks the data set in `directory` (DataError naming the file).
DataSet read_data_set(const std::string& directory);

// The molecular cuts the interpolation uses (src/crplib/dataset.cpp:33-40).
int molecule_cuts_used(Symmetry s);

struct Packed;
class QdSurface;
Packed pack(const QdSurface& surface);

// A cubic spline in one variable with F'' at the knots (src/crplib/spline.cpp:81-140).
class Spline1D {
 public:
  enum class End { not_a_knot, second_derivative, first_derivative, natural };
  struct Condition {
    End kind = End::not_a_knot;
    double value = 0.0;
  };
  Spline1D() = default;
  Spline1D(std::vector<double> x, std::vector<double> y, Condition first, Condition last);
  // SPLINT: value and slope, the interval found by bisection (the end cubics continued outside).
  void splint(double t, double& value, double& slope) const;
  // DPL / DPLP: value and slope, linear outside the knots.
  double dpl(double t) const;
  double dplp(double t) const;
  const std::vector<double>& x() const { return x_; }
  const std::vector<double>& y() const { return y_; }

 private:
  friend Packed pack(const QdSurface&);
  std::vector<double> x_, y_, m_;
};

// Bicubic spline on an (x, y) grid, not-a-knot in both directions (src/crplib/spline.cpp:302-411).
class Spline2D {
 public:
  struct Weights {
    double a, b, c, d;
  };
  struct Cell {
    int i, j;
    Weights wx, wy, sx, sy;
  };
  struct Value {
    double v, d_dy, d_dx;
  };
  Spline2D() = default;
  Spline2D(std::vector<double> x, std::vector<double> y, std::vector<double> f);
  std::optional<Cell> locate(double x, double y) const;
  Value at(const Cell& cell) const;
  const std::vector<double>& x() const { return x_; }
  const std::vector<double>& y() const { return y_; }

 private:
  friend Packed pack(const QdSurface&);
  std::vector<double> x_, y_, f_, fxx_, fyy_, fxxyy_;
};

struct Vec3 {
  double x, y, z;
};

struct AtomValue {
  double v;
  Vec3 gradient;
};

// A 3D cubic spline periodic in x and y and clamped in z: the C4v atom surface (src/crplib/periodic_spline_3d.cpp).
class PeriodicSpline3D {
 public:
  PeriodicSpline3D() = default;
  PeriodicSpline3D(std::vector<double> x, std::vector<double> y, std::vector<double> z, std::vector<double> f,
                   const std::vector<double>& dfdz_first, const std::vector<double>& dfdz_last);
  // value, d/dx, d/dy, d/dz; EvaluationError outside the grid
  std::array<double, 4> evaluate(double x, double y, double z) const;
  const std::vector<double>& z() const { return z_; }

 private:
  friend Packed pack(const QdSurface&);
  std::vector<double> x_, y_, z_, f_, xx_, yy_, xxyy_, zz_, zzxx_, zzyy_, zzxxyy_;
};

// The atom-surface PES of a data set (src/crplib/hexagonal_atom_surface.cpp for C3v / C6v,
// src/crplib/square_atom_surface.cpp for C4v).
class AtomSurface {
 public:
  explicit AtomSurface(const DataSet& data);
  AtomValue evaluate(const Vec3& p) const;
  // The corrugation function summed over the MURII cluster at (x, y, z): v, dv/dx, dv/dy, dv/dz.
  std::array<double, 4> repulsion(double x, double y, double z) const;
  double lattice_constant() const { return a_; }

 private:
  friend Packed pack(const QdSurface&);
  struct ClusterAtom {
    double x, y, z;
    int layer;
  };
  bool c3v_;
  bool square_;
  double a_;
  std::vector<double> spacings_;
  Spline1D corrugation_;
  std::vector<ClusterAtom> cluster_;
  std::vector<Spline1D> cuts_;
  const double* table_;
  int n_;
  PeriodicSpline3D spline_;  // C4v
  void basis(double x, double y, double* f, double* dfdx, double* dfdy) const;
  AtomValue evaluate_square(const Vec3& p) const;
  void init_square(const DataSet& data);
};

struct MoleculeValue {
  double v;
  Vec3 gradient_a, gradient_b;
};

// The molecule-surface PES (src/crplib/molecule_surface.cpp, hexagonal_interaction.cpp, site_interpolation.cpp).
class MoleculeSurface {
 public:
  explicit MoleculeSurface(const DataSet& data);
  // V and gradients at atom positions a, b (bohr); nullopt where the Fortran gives up silently (V = 0 there).
  std::optional<MoleculeValue> evaluate(const Vec3& a, const Vec3& b) const;
  struct Switch {
    double f, df_dz;
  };
  Switch switch_function(double z) const;
  struct GasPhase {
    double v, d_z, d_r;
  };
  GasPhase gas_phase(double z, double r) const;
  Symmetry symmetry() const { return symmetry_; }
  double lattice_constant() const { return a_; }
  double r_min() const { return r_min_; }
  double r_max() const { return r_max_; }
  double z_min() const { return z_min_; }
  double zso_min() const { return zso_min_; }
  double zso_max() const { return zso_max_; }
  double z_asymptotic() const { return z_dependence_.x().back(); }
  double h2_asymptote() const { return h2_.y().back(); }
  const AtomSurface& atoms() const { return atom_; }

 private:
  friend Packed pack(const QdSurface&);
  Symmetry symmetry_;
  double a_;
  AtomSurface atom_;
  double zso_min_, zso_max_;
  Spline1D z_dependence_, h2_;
  std::vector<Spline2D> cuts_;
  double r_min_ = 0.0, r_max_ = 100.0, z_min_ = -100.0;
  std::array<double, 7> interaction(const Spline2D::Value* cuts, double theta, double phi, double x, double y) const;
  std::array<double, 7> square_interaction(const Spline2D::Value* cuts, double theta, double phi, double x,
                                           double y) const;
};

// Centre, bond length and orientation of the molecule (CAMBCORD, src/crplib/molecule_surface.cpp:22-41).
struct Geometry {
  double x, y, z, r, theta, phi;
};
Geometry molecule_geometry(const Vec3& a, const Vec3& b);

// V in eV at the centre (u, v skewed, bohr), Z, r, theta, phi with r and Z clamped into the data and the free
// molecule at its last tabulated r as the zero (src/crplib/qd_wrapper.cpp).
class QdSurface {
 public:
  explicit QdSurface(const DataSet& data);
  double potential_ev(double u, double v, double z, double r, double theta, double phi) const;
  const MoleculeSurface& surface() const { return surface_; }
  double skewing_angle() const;

 private:
  MoleculeSurface surface_;
};

// A surface as flat arrays for the device evaluation (csrc/kernels_crp.hpp; the layout of spo_hip_crp_surface).
struct Packed {
  int c3v = 0;
  double lattice_constant = 0.0, zso_min = 0.0, zso_max = 0.0, r_min = 0.0, r_max = 0.0, z_min = 0.0;
  double h2_asymptote = 0.0;                // hartree
  std::vector<double> corrugation;         // x[n] y[n] F''[n]
  std::vector<double> cluster;             // x y z layer per atom
  std::vector<int> atom_offsets;           // site s: knots [off[s], off[s + 1]) of atom_splines' x, y, F'' blocks
  std::vector<double> atom_splines;        // per site at 3 off[s]: x[k] y[k] F''[k]
  std::vector<int> cut_shape;              // per used cut: nr, nz, the first cut with the same (r, Z) grid
  std::vector<long> cut_offsets;           // per used cut: start of r[nr] Z[nz] f fxx fyy fxxyy[nr nz each]
  std::vector<double> cut_data;
  std::vector<double> h2, z_dependence;    // x[n] y[n] F''[n]
  int square = 0;                          // C4v: the cell's periodic 3D spline instead of the atom splines
  int cell_nz = 0;                         // its z points
  std::vector<double> cell_spline;         // x[5] y[5] z[nz], then f xx yy xxyy zz zzxx zzyy zzxxyy (25 nz each)
};

}  // namespace spo::crp
