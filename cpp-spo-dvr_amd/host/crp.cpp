// The CRP potential energy surface of a hexagonal data set. The order of operations of every expression follows the
// Fortran CRP library as the reference's restatement keeps it (file:line of src/crplib/ at each part), so that the
// results agree with the Fortran's bit for bit; integer powers are written as products.
#include "crp.hpp"
#include "crp_tables.hpp"
#include "list_reader.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <sstream>

namespace spo::crp {

namespace {

using crp_tables::kCosPi3;
using crp_tables::kPi;
using crp_tables::kSinPi3;
using crp_tables::kSq3;
using crp_tables::kUsq3;
using crp_tables::kAtomC3vTable;
using crp_tables::kAtomC6vTable;
using crp_tables::kMoleculeTable;

// ---------------------------------------------------------------------------------------------------------------
// Units (src/crplib/units.cpp:13-79)

constexpr double kBohrToAngstrom = 0.529177249;
constexpr double kHartreeToEv = 27.2113961;
constexpr double kHartreeToKcal = 627.50956;
constexpr double kHartreeToKj = 627.50956 * 4.184;
constexpr double kHartreeToJoule = 4.35974434e-18;

double hartree_to(EnergyUnit u) {
  switch (u) {
    case EnergyUnit::ev: return kHartreeToEv;
    case EnergyUnit::kcal_per_mol: return kHartreeToKcal;
    case EnergyUnit::kj_per_mol: return kHartreeToKj;
    case EnergyUnit::joule: return kHartreeToJoule;
    case EnergyUnit::hartree: break;
  }
  return 1.0;
}

double from_hartree_reciprocal(EnergyUnit u) {
  switch (u) {
    case EnergyUnit::ev: return 1.0 / kHartreeToEv;
    case EnergyUnit::kcal_per_mol: return 1.0 / kHartreeToKcal;
    case EnergyUnit::kj_per_mol: return 1.0 / kHartreeToKj;
    case EnergyUnit::joule: return 1.0 / kHartreeToJoule;
    case EnergyUnit::hartree: break;
  }
  return 1.0;
}

}  // namespace

double to_length(double x, LengthUnit from, LengthUnit to) {
  if (from == to) return x;
  double factor = from == LengthUnit::bohr ? 1.0 : 1.0 / kBohrToAngstrom;
  if (to == LengthUnit::angstrom) factor = factor * kBohrToAngstrom;
  return x * factor;
}

double to_energy(double x, EnergyUnit from, EnergyUnit to) {
  if (from == to) return x;
  double factor = from_hartree_reciprocal(from);
  if (to != EnergyUnit::hartree) factor = factor * hartree_to(to);
  return x * factor;
}

double to_force(double x, Units from, Units to) {
  return x * to_energy(1.0, from.energy, to.energy) / to_length(1.0, from.length, to.length);
}

// ---------------------------------------------------------------------------------------------------------------
// Data set (src/crplib/dataset.cpp:70-233; list-directed input, src/fortran_text/list_directed_reader.cpp)

namespace {

Table read_table(ListReader& in, int count) {
  if (count < 0) throw DataError(in.name() + ": negative number of points");
  Table t;
  for (int i = 0; i < count; ++i) {
    const auto s = in.read(2);
    t.x.push_back(in.real(s[0]));
    t.f.push_back(in.real(s[1]));
  }
  return t;
}

void check_grid(const std::vector<double>& x, const std::string& what) {
  if (x.size() < 4) throw DataError(what + ": a cubic spline needs at least 4 points");
  for (std::size_t i = 1; i < x.size(); ++i)
    if (!(x[i] > x[i - 1])) throw DataError(what + ": the grid is not strictly increasing");
}

int atom_cuts_needed(Symmetry s) { return s == Symmetry::c3v ? 10 : (s == Symmetry::c6v ? 7 : 6); }

}  // namespace

int molecule_cuts_used(Symmetry s) {
  switch (s) {
    case Symmetry::c3v: return 29;
    case Symmetry::c4v: return 16;
    case Symmetry::c6v: return 18;
  }
  return 0;
}

DataSet read_data_set(const std::string& directory) {
  namespace fs = std::filesystem;
  const fs::path dir(directory);
  DataSet d;
  const fs::path crpdata = dir / "CRPData";
  if (!fs::exists(crpdata)) throw DataError("CRP data set " + directory + " does not exist (no file CRPData)");
  {
    ListReader in(crpdata);
    const std::string sym = in.read(1)[0];
    if (sym == "C3v") {
      d.symmetry = Symmetry::c3v;
    } else if (sym == "C6v") {
      d.symmetry = Symmetry::c6v;
    } else if (sym == "C4v") {
      d.symmetry = Symmetry::c4v;
    } else {
      throw DataError(crpdata.string() + ": unknown symmetry '" + sym + "'");
    }
    const auto u = in.read(2);
    const int lu = in.integer(u[0]), eu = in.integer(u[1]);
    if (lu != 0 && lu != 1) throw DataError(crpdata.string() + ": unknown length unit code " + u[0]);
    if (eu < 2 || eu > 6) throw DataError(crpdata.string() + ": unknown energy unit code " + u[1]);
    d.units = Units{static_cast<LengthUnit>(lu), static_cast<EnergyUnit>(eu)};
    d.lattice_constant = in.real(in.read(1)[0]);
    const int layers = in.integer(in.read(1)[0]);
    if (layers < 1) throw DataError(crpdata.string() + ": the number of layers must be at least 1");
    for (const std::string& s : in.read(layers - 1)) d.interlayer_spacings.push_back(in.real(s));
    if (auto disp = in.try_read(2); disp && disp->size() == 2) {
      try {
        d.displacement_1 = in.real((*disp)[0]);
        d.displacement_2 = in.real((*disp)[1]);
      } catch (const DataError&) {  // READ(..., ERR=100): unreadable displacements are 0
        d.displacement_1 = d.displacement_2 = 0.0;
      }
    }
  }
  {
    ListReader list(dir / "CRPData.3D");
    const int n = list.integer(list.read(1)[0]);
    for (int i = 0; i < n; ++i) {
      const auto s = list.read(3);
      AtomCut cut;
      cut.x = list.real(s[0]);
      cut.y = list.real(s[1]);
      cut.file = s[2];
      d.atom_cuts.push_back(cut);
    }
    ListReader corr(dir / "Atom" / "intrep.dat");
    d.corrugation_slope_first = corr.real(corr.read(1)[0]);
    d.corrugation_slope_last = corr.real(corr.read(1)[0]);
    d.corrugation = read_table(corr, corr.integer(corr.read(1)[0]));
    for (AtomCut& cut : d.atom_cuts) {
      const fs::path file = dir / "Atom" / cut.file;
      if (!fs::exists(file)) throw DataError(file.string() + ": atom-surface cut file not found");
      ListReader in(file);
      const int count = in.integer(in.read(1)[0]);
      cut.slope_first = in.real(in.read(1)[0]);
      cut.slope_last = in.real(in.read(1)[0]);
      cut.v = read_table(in, count);
    }
  }
  {
    ListReader list(dir / "CRPData.6D");
    const auto header = list.read(2);
    const int n = list.integer(header[0]);
    d.cuts_interpolation = list.logical(header[1]);
    for (int i = 0; i < n; ++i) {
      const auto s = list.read(5);
      MoleculeCut cut;
      cut.x = list.real(s[0]);
      cut.y = list.real(s[1]);
      cut.theta = list.real(s[2]);
      cut.phi = list.real(s[3]);
      cut.file = s[4];
      d.molecule_cuts.push_back(cut);
    }
    ListReader ex(dir / "Molecule" / "extrapol.dat");
    d.zso_min = ex.real(ex.read(1)[0]);
    d.zso_max = ex.real(ex.read(1)[0]);
    d.z_dependence = read_table(ex, ex.integer(ex.read(1)[0]));
    ListReader h2(dir / "Molecule" / "VH2.dat");
    d.h2 = read_table(h2, h2.integer(h2.read(1)[0]));
    const int used = molecule_cuts_used(d.symmetry);
    for (std::size_t k = 0; k < d.molecule_cuts.size(); ++k) {
      MoleculeCut& cut = d.molecule_cuts[k];
      const fs::path file = dir / "Molecule" / cut.file;
      if (!fs::exists(file)) {  // the Fortran skips a missing file; harmless only for an unused extra cut
        if (static_cast<int>(k) < used) throw DataError(file.string() + ": molecule-surface cut file not found");
        cut.present = false;
        continue;
      }
      ListReader in(file);
      const int nr = in.integer(in.read(1)[0]);
      const int nz = in.integer(in.read(1)[0]);
      if (nr < 0 || nz < 0) throw DataError(file.string() + ": negative grid size");
      cut.r.resize(nr);
      cut.z.resize(nz);
      cut.v.resize(static_cast<std::size_t>(nr) * nz);
      for (int j = 0; j < nz; ++j) {  // rows "Z r V", r fastest; the last row read sets the grid values
        for (int i = 0; i < nr; ++i) {
          const auto s = in.read(3);
          cut.z[j] = in.real(s[0]);
          cut.r[i] = in.real(s[1]);
          cut.v[i + static_cast<std::size_t>(nr) * j] = in.real(s[2]);
        }
      }
    }
  }
  // Checks (src/crplib/dataset.cpp:236-283)
  if (!(d.lattice_constant > 0.0)) throw DataError("CRPData: the lattice constant must be positive");
  if (d.symmetry == Symmetry::c3v && d.interlayer_spacings.size() + 1 < 4)
    throw DataError("CRPData: C3v needs at least 4 layers (the atom cluster has 4)");
  if (d.symmetry == Symmetry::c3v && d.displacement_1 != 0.0)
    throw DataError("CRPData: the C3v atom potential does not support displacements");
  check_grid(d.corrugation.x, "Atom/intrep.dat");
  if (d.symmetry == Symmetry::c4v && d.interlayer_spacings.empty())
    throw DataError("CRPData: C4v needs at least 2 layers (the atom cluster uses the first interlayer spacing)");
  const int n_atom = static_cast<int>(d.atom_cuts.size()), need = atom_cuts_needed(d.symmetry);
  if (d.symmetry == Symmetry::c4v ? n_atom < need : n_atom != need)
    throw DataError("CRPData.3D: needs " + std::string(d.symmetry == Symmetry::c4v ? "at least " : "exactly ") +
                    std::to_string(need) + " atom-surface cuts, the data set has " + std::to_string(n_atom));
  for (const AtomCut& cut : d.atom_cuts) {
    check_grid(cut.v.x, "Atom/" + cut.file);
    if (d.symmetry == Symmetry::c4v && cut.v.x != d.atom_cuts.front().v.x)
      throw DataError("Atom/" + cut.file + ": C4v needs the same z values for all sites");
  }
  if (d.zso_min > d.zso_max) throw DataError("Molecule/extrapol.dat: ZsoMin is larger than ZsoMax");
  check_grid(d.z_dependence.x, "Molecule/extrapol.dat");
  check_grid(d.h2.x, "Molecule/VH2.dat");
  if (static_cast<int>(d.molecule_cuts.size()) < molecule_cuts_used(d.symmetry))
    throw DataError("CRPData.6D: needs at least " + std::to_string(molecule_cuts_used(d.symmetry)) +
                    " molecule-surface cuts, the data set has " + std::to_string(d.molecule_cuts.size()));
  for (const MoleculeCut& cut : d.molecule_cuts) {
    if (!cut.present) continue;
    check_grid(cut.r, "Molecule/" + cut.file + " (r)");
    check_grid(cut.z, "Molecule/" + cut.file + " (Z)");
    if (cut.r.size() > 999 || cut.z.size() > 999)
      throw DataError("Molecule/" + cut.file + ": at most 999 points per direction");
  }
  return d;
}

// ---------------------------------------------------------------------------------------------------------------
// Splines (src/crplib/spline.cpp)

namespace {

// Tridiagonal solve: forward elimination, back substitution (TRIDIA, src/crplib/spline.cpp:13-23). sub[i - 1] is
// the sub-diagonal of row i, sup[i] the super-diagonal of row i; diag and rhs are overwritten.
void tridiagonal(std::vector<double>& diag, const std::vector<double>& sub, const std::vector<double>& sup,
                 std::vector<double>& rhs, double* out) {
  const std::size_t n = diag.size();
  for (std::size_t i = 1; i < n; ++i) {
    const double rap = sub[i - 1] / diag[i - 1];
    diag[i] = diag[i] - rap * sup[i - 1];
    rhs[i] = rhs[i] - rap * rhs[i - 1];
  }
  out[n - 1] = rhs[n - 1] / diag[n - 1];
  for (std::size_t i = n - 1; i-- > 0;) out[i] = (rhs[i] - sup[i] * out[i + 1]) / diag[i];
}

// F'' at the knots (DSPLIN, src/crplib/spline.cpp:52-146): the interior unknowns from the tridiagonal system of
// the continuity conditions, the end conditions folded into its first and last rows, then the two end values.
std::vector<double> second_derivatives(const double* x, const double* y, std::size_t n, Spline1D::Condition first,
                                       Spline1D::Condition last) {
  using End = Spline1D::End;
  if (n < 4) throw DataError("a cubic spline needs at least 4 points");
  // (right-hand side, diagonal, off-diagonal) terms of each end
  auto end_terms = [&](Spline1D::Condition c, double h_end, double h_next, double secant, double sign) {
    std::array<double, 3> t{0.0, 0.0, 0.0};
    switch (c.kind) {
      case End::not_a_knot:
        t[1] = h_end * (1.0 + h_end / h_next) / 6.0;
        t[2] = -(h_end * h_end) / (6.0 * h_next);
        break;
      case End::second_derivative: t[0] = -c.value * h_end / 6.0; break;
      case End::first_derivative:
        t[0] = sign * 0.5 * (c.value - secant);
        t[1] = -h_end / 12.0;
        break;
      case End::natural: break;
    }
    return t;
  };
  const auto s = end_terms(first, x[1] - x[0], x[2] - x[1], (y[1] - y[0]) / (x[1] - x[0]), 1.0);
  const auto e = end_terms(last, x[n - 1] - x[n - 2], x[n - 2] - x[n - 3],
                           (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]), -1.0);
  const std::size_t m = n - 2;
  std::vector<double> diag(m), sub(m), sup(m), rhs(m);
  double h_right = x[1] - x[0];
  double slope_right = (y[1] - y[0]) / h_right;
  for (std::size_t i = 2; i < n; ++i) {
    const double h_left = h_right;
    h_right = x[i] - x[i - 1];
    if (h_left * h_right <= 0.0) throw DataError("spline abscissae are not monotonic");
    diag[i - 2] = (h_left + h_right) / 3.0;
    sup[i - 2] = h_right / 6.0;
    sub[i - 2] = sup[i - 2];
    const double slope_left = slope_right;
    slope_right = (y[i] - y[i - 1]) / h_right;
    rhs[i - 2] = slope_right - slope_left;
  }
  rhs[0] = rhs[0] + s[0];
  rhs[m - 1] = rhs[m - 1] + e[0];
  diag[0] = diag[0] + s[1];
  diag[m - 1] = diag[m - 1] + e[1];
  sub[m - 2] = sub[m - 2] + e[2];
  sup[0] = sup[0] + s[2];
  std::vector<double> f2(n);
  tridiagonal(diag, sub, sup, rhs, f2.data() + 1);
  const double h0 = x[1] - x[0];
  switch (first.kind) {
    case End::not_a_knot: f2[0] = f2[1] * (1.0 + h0 / (x[2] - x[1])) - f2[2] * h0 / (x[2] - x[1]); break;
    case End::second_derivative: f2[0] = first.value; break;
    case End::first_derivative: f2[0] = -6.0 * s[0] / h0 - f2[1] / 2.0; break;
    case End::natural: f2[0] = 0.0; break;
  }
  const double hn = x[n - 1] - x[n - 2];
  switch (last.kind) {
    case End::not_a_knot:
      f2[n - 1] = f2[n - 2] * (1.0 + hn / (x[n - 2] - x[n - 3])) - f2[n - 3] * hn / (x[n - 2] - x[n - 3]);
      break;
    case End::second_derivative: f2[n - 1] = last.value; break;
    case End::first_derivative: f2[n - 1] = -6.0 * e[0] / hn - f2[n - 2] / 2.0; break;
    case End::natural: f2[n - 1] = 0.0; break;
  }
  return f2;
}

// The lower knot of the interval holding t by bisection (SPLINT, UBICAR).
int bisect(const std::vector<double>& x, double t) {
  int lo = 0, hi = static_cast<int>(x.size()) - 1;
  while (hi - lo > 1) {
    const int k = (hi + lo) / 2;
    if (x[k] > t) {
      hi = k;
    } else {
      lo = k;
    }
  }
  return lo;
}

// Cyclic tridiagonal solve (TRICYCLE, src/crplib/spline.cpp:25-44): the tridiagonal system plus the corner elements
// bottom_left (row n, column 1) and top_right (row 1, column n) by a rank-one correction; diag and rhs overwritten.
void cyclic_tridiagonal(std::vector<double>& diag, const std::vector<double>& sub, const std::vector<double>& sup,
                        double bottom_left, double top_right, std::vector<double>& rhs, double* out) {
  const std::size_t n = diag.size();
  const double alp = -diag[0];
  diag[0] = diag[0] - alp;
  diag[n - 1] = diag[n - 1] - bottom_left * top_right / alp;
  std::vector<double> z(diag);
  tridiagonal(diag, sub, sup, rhs, out);
  diag = z;
  rhs[0] = alp;
  rhs[n - 1] = bottom_left;
  for (std::size_t i = 1; i + 1 < n; ++i) rhs[i] = 0.0;
  tridiagonal(diag, sub, sup, rhs, z.data());
  const double fact = (out[0] + top_right * out[n - 1] / alp) / (1.0 + z[0] + top_right * z[n - 1] / alp);
  for (std::size_t i = 0; i < n; ++i) out[i] = out[i] - fact * z[i];
}

// F'' of the periodic spline through (x_i, y_i), period x_{n-1} - x_0 (TRISPL, src/crplib/spline.cpp:148-178).
std::vector<double> periodic_second_derivatives(const double* x, const double* y, std::size_t n) {
  if (n < 4) throw DataError("a periodic cubic spline needs at least 4 points");
  const std::size_t m = n - 1;
  std::vector<double> diag(m), sub(m), sup(m), rhs(m);
  double h_right = x[1] - x[0];
  double slope_right = (y[1] - y[0]) / h_right;
  for (std::size_t i = 2; i < n; ++i) {
    const double h_left = h_right;
    h_right = x[i] - x[i - 1];
    if (h_left * h_right <= 0.0) throw DataError("spline abscissae are not monotonic");
    diag[i - 2] = (h_left + h_right) / 3.0;
    sub[i - 2] = h_right / 6.0;
    sup[i - 2] = sub[i - 2];
    const double slope_left = slope_right;
    slope_right = (y[i] - y[i - 1]) / h_right;
    rhs[i - 2] = slope_right - slope_left;
  }
  // the row that closes the period: the last interval followed by the first
  const double h_left = h_right;
  h_right = x[1] - x[0];
  diag[m - 1] = (h_left + h_right) / 3.0;
  sub[m - 1] = h_right / 6.0;
  sup[m - 1] = sub[m - 1];
  const double slope_left = slope_right;
  slope_right = (y[1] - y[0]) / h_right;
  rhs[m - 1] = slope_right - slope_left;
  const double corner = sub[m - 1];
  std::vector<double> f2(n);
  cyclic_tridiagonal(diag, sub, sup, corner, corner, rhs, f2.data() + 1);
  f2[0] = f2[n - 1];
  return f2;
}

}  // namespace

Spline1D::Spline1D(std::vector<double> x, std::vector<double> y, Condition first, Condition last)
    : x_(std::move(x)), y_(std::move(y)) {
  if (x_.size() != y_.size()) throw DataError("spline: different numbers of abscissae and values");
  m_ = second_derivatives(x_.data(), y_.data(), x_.size(), first, last);
}

// SPLINT (src/crplib/spline.cpp:226-254)
void Spline1D::splint(double t, double& value, double& slope) const {
  const int lo = bisect(x_, t), hi = lo + 1;
  const double h = x_[hi] - x_[lo];
  const double a = (x_[hi] - t) / h;
  const double b = (t - x_[lo]) / h;
  value = a * y_[lo] + b * y_[hi] + ((a * a * a - a) * m_[lo] + (b * b * b - b) * m_[hi]) * (h * h) / 6.0;
  slope = (y_[hi] - y_[lo]) / h + h * (m_[hi] * (3.0 * b * b - 1.0) - m_[lo] * (3.0 * a * a - 1.0)) / 6.0;
}

// DPL (src/crplib/spline.cpp:171-190). k - 1 = the number of knots below t, except k = 2 at t = x_0; k = 1 and
// k = n + 1 extrapolate linearly.
double Spline1D::dpl(double t) const {
  const int n = static_cast<int>(x_.size());
  int k = static_cast<int>(std::lower_bound(x_.begin(), x_.end(), t) - x_.begin()) + 1;
  if (t == x_[0]) k = 2;
  if (k <= 1) {
    const double e = x_[1] - x_[0];
    return ((y_[1] - y_[0]) / e - m_[1] * e / 6.0) * (t - x_[0]) + y_[0];
  }
  if (k > n) {
    const double e = x_[n - 1] - x_[n - 2];
    return ((y_[n - 1] - y_[n - 2]) / e + m_[n - 2] * e / 6.0) * (t - x_[n - 1]) + y_[n - 1];
  }
  const int i = k - 1;
  const double f = x_[i] - t;
  const double g = t - x_[i - 1];
  const double e = f + g;
  return (-g * f * (m_[i - 1] * (f + e) + m_[i] * (g + e)) + 6.0 * (g * y_[i] + f * y_[i - 1])) / (6.0 * e);
}

// DPLP (src/crplib/spline.cpp:193-215)
double Spline1D::dplp(double t) const {
  const int n = static_cast<int>(x_.size());
  int k = static_cast<int>(std::lower_bound(x_.begin(), x_.end(), t) - x_.begin()) + 1;
  if (t == x_[0]) k = 2;
  int lo, hi;
  if (k <= 1) {
    lo = 0;
    hi = 1;
  } else if (k > n) {
    lo = n - 2;
    hi = n - 1;
  } else {
    const int i = k - 1;
    const double f = x_[i] - t;
    const double g = t - x_[i - 1];
    const double e = f + g;
    const double e2 = e * e;
    return (-m_[i - 1] * (3.0 * f * f - e2) + m_[i] * (3.0 * g * g - e2) + 6.0 * (y_[i] - y_[i - 1])) / (6.0 * e);
  }
  const double e = x_[hi] - x_[lo];
  const double g = t - x_[lo];
  const double f = x_[hi] - t;
  return (y_[hi] - y_[lo]) / e + (m_[hi] * (2.0 * g - f) + m_[lo] * (g - 2.0 * f)) / 6.0;
}

// SPLIN2D (src/crplib/spline.cpp:302-331): F_xx along x for every y, F_yy along y for every x, then F_xxyy along x
// of F_yy.
Spline2D::Spline2D(std::vector<double> x, std::vector<double> y, std::vector<double> f)
    : x_(std::move(x)), y_(std::move(y)), f_(std::move(f)) {
  const std::size_t nx = x_.size(), ny = y_.size();
  if (f_.size() != nx * ny) throw DataError("2D spline: the number of values does not match the grid");
  check_grid(x_, "2D spline (x)");
  check_grid(y_, "2D spline (y)");
  fxx_.resize(nx * ny);
  fyy_.resize(nx * ny);
  fxxyy_.resize(nx * ny);
  const Spline1D::Condition nak{};
  for (std::size_t j = 0; j < ny; ++j) {
    const auto m = second_derivatives(x_.data(), f_.data() + nx * j, nx, nak, nak);
    std::copy(m.begin(), m.end(), fxx_.begin() + nx * j);
  }
  std::vector<double> column(ny);
  for (std::size_t i = 0; i < nx; ++i) {
    for (std::size_t j = 0; j < ny; ++j) column[j] = f_[i + nx * j];
    const auto m = second_derivatives(y_.data(), column.data(), ny, nak, nak);
    for (std::size_t j = 0; j < ny; ++j) fyy_[i + nx * j] = m[j];
  }
  for (std::size_t j = 0; j < ny; ++j) {
    const auto m = second_derivatives(x_.data(), fyy_.data() + nx * j, nx, nak, nak);
    std::copy(m.begin(), m.end(), fxxyy_.begin() + nx * j);
  }
}

namespace {

// SPLCOFII, one direction (src/crplib/spline.cpp:336-347)
void cell_weights(const std::vector<double>& x, int i, double t, Spline2D::Weights& w, Spline2D::Weights& s) {
  const double h2s6 = (x[i + 1] - x[i]) * (x[i + 1] - x[i]) / 6.0;
  w.a = (x[i + 1] - t) / (x[i + 1] - x[i]);
  w.b = 1.0 - w.a;
  w.c = (w.a * w.a * w.a - w.a) * h2s6;
  w.d = (w.b * w.b * w.b - w.b) * h2s6;
  s.a = -1.0 / (x[i + 1] - x[i]);
  s.b = -s.a;
  s.c = (3.0 * (w.a * w.a) * s.a - s.a) * h2s6;
  s.d = (3.0 * (w.b * w.b) * s.b - s.b) * h2s6;
}

}  // namespace

std::optional<Spline2D::Cell> Spline2D::locate(double x, double y) const {
  if (x < x_.front() || x > x_.back() || y < y_.front() || y > y_.back()) return std::nullopt;
  Cell c{};
  c.i = bisect(x_, x);
  c.j = bisect(y_, y);
  cell_weights(x_, c.i, x, c.wx, c.sx);
  cell_weights(y_, c.j, y, c.wy, c.sy);
  return c;
}

// SPLINII (src/crplib/spline.cpp:366-393): the value and d/dx sum over the x weights outside, d/dy over the y ones.
Spline2D::Value Spline2D::at(const Cell& c) const {
  const std::size_t nx = x_.size();
  const std::size_t p00 = c.i + nx * c.j, p10 = p00 + 1, p01 = p00 + nx, p11 = p01 + 1;
  const auto& v = f_;
  auto x_outer = [&](const Weights& wx, const Weights& wy) {
    return wx.a * (wy.a * v[p00] + wy.b * v[p01] + wy.c * fyy_[p00] + wy.d * fyy_[p01]) +
           wx.b * (wy.a * v[p10] + wy.b * v[p11] + wy.c * fyy_[p10] + wy.d * fyy_[p11]) +
           wx.c * (wy.a * fxx_[p00] + wy.b * fxx_[p01] + wy.c * fxxyy_[p00] + wy.d * fxxyy_[p01]) +
           wx.d * (wy.a * fxx_[p10] + wy.b * fxx_[p11] + wy.c * fxxyy_[p10] + wy.d * fxxyy_[p11]);
  };
  auto y_outer = [&](const Weights& wx, const Weights& wy) {
    return wy.a * (wx.a * v[p00] + wx.b * v[p10] + wx.c * fxx_[p00] + wx.d * fxx_[p10]) +
           wy.b * (wx.a * v[p01] + wx.b * v[p11] + wx.c * fxx_[p01] + wx.d * fxx_[p11]) +
           wy.c * (wx.a * fyy_[p00] + wx.b * fyy_[p10] + wx.c * fxxyy_[p00] + wx.d * fxxyy_[p10]) +
           wy.d * (wx.a * fyy_[p01] + wx.b * fyy_[p11] + wx.c * fxxyy_[p01] + wx.d * fxxyy_[p11]);
  };
  return {x_outer(c.wx, c.wy), y_outer(c.wx, c.sy), x_outer(c.sx, c.wy)};
}

// SPLIN3D periodic (src/crplib/periodic_spline_3d.cpp:48-105): clamped splines along z for every (x, y), then the
// periodic 2D second derivatives (along y, along x, along x of the y ones) of F and of F_zz in every z plane.
PeriodicSpline3D::PeriodicSpline3D(std::vector<double> x, std::vector<double> y, std::vector<double> z,
                                   std::vector<double> f, const std::vector<double>& dfdz_first,
                                   const std::vector<double>& dfdz_last)
    : x_(std::move(x)), y_(std::move(y)), z_(std::move(z)), f_(std::move(f)) {
  const std::size_t nx = x_.size(), ny = y_.size(), nz = z_.size(), n = f_.size();
  if (n != nx * ny * nz || dfdz_first.size() != nx * ny || dfdz_last.size() != nx * ny)
    throw DataError("3D spline: the number of values does not match the grid");
  auto at = [&](std::size_t i, std::size_t j, std::size_t k) { return i + nx * (j + ny * k); };
  zz_.resize(n);
  std::vector<double> line(nz);
  for (std::size_t i = 0; i < nx; ++i)
    for (std::size_t j = 0; j < ny; ++j) {
      for (std::size_t k = 0; k < nz; ++k) line[k] = f_[at(i, j, k)];
      const auto m = second_derivatives(z_.data(), line.data(), nz,
                                        {Spline1D::End::first_derivative, dfdz_first[i + nx * j]},
                                        {Spline1D::End::first_derivative, dfdz_last[i + nx * j]});
      for (std::size_t k = 0; k < nz; ++k) zz_[at(i, j, k)] = m[k];
    }
  for (auto* v : {&xx_, &yy_, &xxyy_, &zzxx_, &zzyy_, &zzxxyy_}) v->resize(n);
  auto periodic_2d = [&](const std::vector<double>& g, std::vector<double>& gxx, std::vector<double>& gyy,
                         std::vector<double>& gxxyy) {
    std::vector<double> col(ny);
    for (std::size_t i = 0; i < nx; ++i) {
      for (std::size_t j = 0; j < ny; ++j) col[j] = g[i + nx * j];
      const auto m = periodic_second_derivatives(y_.data(), col.data(), ny);
      for (std::size_t j = 0; j < ny; ++j) gyy[i + nx * j] = m[j];
    }
    for (std::size_t j = 0; j < ny; ++j) {
      const auto m = periodic_second_derivatives(x_.data(), g.data() + nx * j, nx);
      for (std::size_t i = 0; i < nx; ++i) gxx[i + nx * j] = m[i];
    }
    for (std::size_t j = 0; j < ny; ++j) {
      const auto m = periodic_second_derivatives(x_.data(), gyy.data() + nx * j, nx);
      for (std::size_t i = 0; i < nx; ++i) gxxyy[i + nx * j] = m[i];
    }
  };
  std::vector<double> plane(nx * ny), plane_zz(nx * ny), xx(nx * ny), yy(nx * ny), xxyy(nx * ny), zx(nx * ny),
      zy(nx * ny), zxy(nx * ny);
  for (std::size_t k = 0; k < nz; ++k) {
    for (std::size_t j = 0; j < ny; ++j)
      for (std::size_t i = 0; i < nx; ++i) {
        plane[i + nx * j] = f_[at(i, j, k)];
        plane_zz[i + nx * j] = zz_[at(i, j, k)];
      }
    periodic_2d(plane, xx, yy, xxyy);
    periodic_2d(plane_zz, zx, zy, zxy);
    for (std::size_t j = 0; j < ny; ++j)
      for (std::size_t i = 0; i < nx; ++i) {
        const std::size_t q = at(i, j, k), l = i + nx * j;
        xx_[q] = xx[l];
        yy_[q] = yy[l];
        xxyy_[q] = xxyy[l];
        zzxx_[q] = zx[l];
        zzyy_[q] = zy[l];
        zzxxyy_[q] = zxy[l];
      }
  }
}

namespace {

// SPLINT3D weights in one direction (src/crplib/periodic_spline_3d.cpp:33-46): x**3.0 is a call of pow there.
void weights_3d(const std::vector<double>& x, int i, double t, Spline2D::Weights& w, Spline2D::Weights& s) {
  const double h = x[i + 1] - x[i];
  w.a = (x[i + 1] - t) / h;
  w.b = 1.0 - w.a;
  w.c = (std::pow(w.a, 3.0) - w.a) / 6.0 * (h * h);
  w.d = (std::pow(w.b, 3.0) - w.b) / 6.0 * (h * h);
  s.a = -1.0 / h;
  s.b = -s.a;
  s.c = (3.0 * (w.a * w.a) * s.a - s.a) / 6.0 * (h * h);
  s.d = (3.0 * (w.b * w.b) * s.b - s.b) / 6.0 * (h * h);
}

}  // namespace

// SPLINT3D (src/crplib/periodic_spline_3d.cpp:108-165)
std::array<double, 4> PeriodicSpline3D::evaluate(double x, double y, double z) const {
  auto inside = [](const std::vector<double>& g, double t) { return !(t < g.front() || t > g.back()); };
  if (!inside(x_, x) || !inside(y_, y) || !inside(z_, z))
    throw EvaluationError("C4v atom-surface spline: point outside the interpolation grid");
  const int i = bisect(x_, x), j = bisect(y_, y), k = bisect(z_, z);
  Spline2D::Weights wx, sx, wy, sy, wz, sz;
  weights_3d(x_, i, x, wx, sx);
  weights_3d(y_, j, y, wy, sy);
  weights_3d(z_, k, z, wz, sz);
  const std::size_t nx = x_.size(), ny = y_.size();
  auto plane_sum = [&](const Spline2D::Weights& ex, const Spline2D::Weights& ey, const std::vector<double>& g,
                       const std::vector<double>& gxx, const std::vector<double>& gyy,
                       const std::vector<double>& gxxyy, int kk) {
    const std::size_t p00 = i + nx * (j + ny * kk), p10 = p00 + 1, p01 = p00 + nx, p11 = p01 + 1;
    return ey.a * (ex.a * g[p00] + ex.b * g[p10] + ex.c * gxx[p00] + ex.d * gxx[p10]) +
           ey.b * (ex.a * g[p01] + ex.b * g[p11] + ex.c * gxx[p01] + ex.d * gxx[p11]) +
           ey.c * (ex.a * gyy[p00] + ex.b * gyy[p10] + ex.c * gxxyy[p00] + ex.d * gxxyy[p10]) +
           ey.d * (ex.a * gyy[p01] + ex.b * gyy[p11] + ex.c * gxxyy[p01] + ex.d * gxxyy[p11]);
  };
  auto terms = [&](const Spline2D::Weights& ex, const Spline2D::Weights& ey) {
    return std::array<double, 4>{plane_sum(ex, ey, f_, xx_, yy_, xxyy_, k), plane_sum(ex, ey, f_, xx_, yy_, xxyy_, k + 1),
                                 plane_sum(ex, ey, zz_, zzxx_, zzyy_, zzxxyy_, k),
                                 plane_sum(ex, ey, zz_, zzxx_, zzyy_, zzxxyy_, k + 1)};
  };
  auto combine = [](const Spline2D::Weights& e, const std::array<double, 4>& t) {
    return e.a * t[0] + e.b * t[1] + e.c * t[2] + e.d * t[3];
  };
  const std::array<double, 4> base = terms(wx, wy);
  return {combine(wz, base), combine(wz, terms(sx, wy)), combine(wz, terms(wx, sy)), combine(sz, base)};
}

// ---------------------------------------------------------------------------------------------------------------
// Atom-surface PES (src/crplib/hexagonal_atom_surface.cpp, atom_surface.cpp)

namespace {

// The C6v cluster: crystal coordinates of 30 top-layer atoms (Fortran DATA ivec of MURII).
constexpr std::array<int, 30> kC6vI = {0,  1,  2, 3, -1, 0, 1,  2,  3, -2, -1, 0,  1,  2, 3,
                                       -2, -1, 0, 1, 2,  3, -2, -1, 0, 1,  2,  -2, -1, 0, 1};
constexpr std::array<int, 30> kC6vJ = {-2, -2, -2, -2, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0,
                                       1,  1,  1,  1,  1,  1,  2,  2,  2,  2, 2, 3, 3, 3, 3};

// INTEGER = REAL of the cell index, one further for negative values (ORIGEN2, src/crplib/atom_surface.cpp:58-70).
int cell_index(double ori) {
  if (!(std::abs(ori) < 2.0e9))
    throw EvaluationError("CRP atom-surface PES: coordinate " + std::to_string(ori) +
                          " lattice constants from the origin");
  return ori >= 0.0 ? static_cast<int>(ori) : static_cast<int>(ori - 1.0);
}

// (x, y) folded into the square unit cell at the origin (ORIGEN, src/crplib/atom_surface.cpp:84-89).
std::pair<double, double> fold_square(double x, double y, double delta) {
  const int i1 = cell_index(x / delta);
  const int i2 = cell_index(y / delta);
  return {x - i1 * delta, y - i2 * delta};
}

// A point of the square cell mapped into the triangle 0 <= y <= x <= delta/2 by one of the 8 operations; returns
// the operation's number 1..8 (REGION, src/crplib/atom_surface.cpp:92-129; the comparisons as written there).
int fold_into_triangle(double& x, double& y, double delta) {
  if (y <= delta / 2.0) {
    if (x <= delta / 2.0) {
      if (x >= y) return 1;
      const double t = x;
      x = y;
      y = t;
      return 3;
    }
    if (y >= (delta - x)) {
      const double t = x;
      x = y;
      y = delta - t;
      return 4;
    }
    x = delta - x;
    return 2;
  }
  if (x <= delta / 2.0) {
    if (y > (delta - x)) {
      y = delta - y;
      return 7;
    }
    const double t = x;
    x = delta - y;
    y = t;
    return 5;
  }
  if (x > y) {
    const double t = x;
    x = delta - y;
    y = delta - t;
    return 6;
  }
  x = delta - x;
  y = delta - y;
  return 8;
}

// The gradient of operation `region` from the triangle's (TRANSF, src/crplib/atom_surface.cpp:132-163).
void unfold_gradient(int region, double& gx, double& gy) {
  double t = 0.0;
  switch (region) {
    case 1: return;
    case 2: gx = -gx; return;
    case 3: t = gx; gx = gy; gy = t; return;
    case 4: t = gx; gx = -gy; gy = t; return;
    case 5: t = gx; gx = gy; gy = -t; return;
    case 6: t = gx; gx = -gy; gy = -t; return;
    case 7: gy = -gy; return;
    case 8: gx = -gx; gy = -gy; return;
    default: throw EvaluationError("CRP C4v atom-surface PES: unknown region");
  }
}

// MURO: 16 atoms of the top layer and 16 of the second (layer -1), lattice units (src/crplib/square_atom_surface.cpp
// :18-28, Fortran DATA VEC)
struct MuroAtom {
  double x, y, layer;
};
constexpr std::array<MuroAtom, 32> kMuro = {{
    {0., -2., 0.},    {-1., -1., 0.},    {0., -1., 0.},     {1., -1., 0.},    {2., -1., 0.},     {-2., 0., 0.},
    {-1., 0., 0.},    {0., 0., 0.},      {1., 0., 0.},      {2., 0., 0.},     {-1., 1., 0.},     {0., 1., 0.},
    {1., 1., 0.},     {2., 1., 0.},      {0., 2., 0.},      {1., 2., 0.},     {-0.5, -1.5, -1.}, {0.5, -1.5, -1.},
    {1.5, -1.5, -1.}, {-1.5, -0.5, -1.}, {-0.5, -0.5, -1.}, {0.5, -0.5, -1.}, {1.5, -0.5, -1.},  {-1.5, 0.5, -1.},
    {-0.5, 0.5, -1.}, {0.5, 0.5, -1.},   {1.5, 0.5, -1.},   {2.5, 0.5, -1.},  {-0.5, 1.5, -1.},  {0.5, 1.5, -1.},
    {1.5, 1.5, -1.},  {0.5, 2.5, -1.},
}};

// The (i, j) points (1-based) of the 5 x 5 grid each C4v site cut fills, sites in CRPData.3D order: top, hollow,
// bridge, t2h, b2h, t2b (src/crplib/square_atom_surface.cpp:31-40)
const std::array<std::vector<std::pair<int, int>>, 6> kSquareSitePoints = {{
    {{1, 1}, {1, 5}, {5, 1}, {5, 5}},
    {{3, 3}},
    {{3, 1}, {1, 3}, {5, 3}, {3, 5}},
    {{2, 2}, {4, 4}, {4, 2}, {2, 4}},
    {{3, 2}, {2, 3}, {4, 3}, {3, 4}},
    {{2, 1}, {4, 1}, {1, 2}, {5, 2}, {1, 4}, {5, 4}, {2, 5}, {4, 5}},
}};

// (x, y) folded into the skewed unit cell at the origin (ORIGEN2, src/crplib/atom_surface.cpp:73-82).
std::pair<double, double> fold(double x, double y, double delta) {
  double bigx = x - y * kUsq3;
  double bigy = 2.00 * y * kUsq3;
  const int i1 = cell_index(bigx / delta);
  const int i2 = cell_index(bigy / delta);
  bigx = bigx - i1 * delta;
  bigy = bigy - i2 * delta;
  return {bigx + bigy * 0.500, kSq3 * bigy * 0.500};
}

}  // namespace

AtomSurface::AtomSurface(const DataSet& data)
    : c3v_(data.symmetry == Symmetry::c3v),
      square_(data.symmetry == Symmetry::c4v),
      a_(to_length(data.lattice_constant, data.units.length, LengthUnit::bohr)),
      table_(data.symmetry == Symmetry::c3v ? kAtomC3vTable.data() : kAtomC6vTable.data()),
      n_(data.symmetry == Symmetry::c3v ? 10 : 7) {
  for (double s : data.interlayer_spacings) spacings_.push_back(to_length(s, data.units.length, LengthUnit::bohr));
  {  // the corrugation function, clamped at both ends (src/crplib/atom_surface.cpp:41-51)
    const double s1 = to_force(data.corrugation_slope_first, data.units, kInternalUnits);
    const double s2 = to_force(data.corrugation_slope_last, data.units, kInternalUnits);
    std::vector<double> x, f;
    for (std::size_t i = 0; i < data.corrugation.x.size(); ++i) {
      x.push_back(to_length(data.corrugation.x[i], data.units.length, LengthUnit::bohr));
      f.push_back(to_energy(data.corrugation.f[i], data.units.energy, EnergyUnit::hartree));
    }
    corrugation_ = Spline1D(x, f, {Spline1D::End::first_derivative, s1}, {Spline1D::End::first_derivative, s2});
  }
  if (square_) {  // MURO cluster, reduced cuts on the 5 x 5 grid of the cell, periodic 3D spline
    init_square(data);
    return;
  }
  // MURII cluster in its loop order (src/crplib/hexagonal_atom_surface.cpp:29-60)
  if (!c3v_) {
    for (std::size_t k = 0; k < kC6vI.size(); ++k) {
      const double bx = kC6vI[k] * a_;
      const double by = kC6vJ[k] * a_;
      cluster_.push_back({bx + by * 0.500, by * kSq3 * 0.500, 0.0, 0});
    }
  } else {
    for (int i = -3; i <= 4; ++i) {
      for (int j = -3; j <= 4; ++j) {
        for (int k = 0; k < 4; ++k) {  // fcc(111) ABC stacking: layer k shifted by (k mod 3)/3 of both vectors
          const double u = (static_cast<double>(i) + (k % 3) * (1.0 / 3.0)) * a_;
          const double v = (static_cast<double>(j) + (k % 3) * (1.0 / 3.0)) * a_;
          double z = 0.0;
          for (int l = 0; l < k; ++l) z = z - spacings_[l];
          cluster_.push_back({u + v * kCosPi3, v * kSinPi3, z, k});
        }
      }
    }
  }
  // The cuts reduced by the repulsive sum at their sites (src/crplib/hexagonal_atom_surface.cpp:80-106)
  for (const AtomCut& cut : data.atom_cuts) {
    const double s1 = to_force(cut.slope_first, data.units, kInternalUnits);
    const double s2 = to_force(cut.slope_last, data.units, kInternalUnits);
    const std::size_t n = cut.v.x.size();
    std::vector<double> z(n), reduced(n);
    double d1 = 0.0, d2 = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
      z[i] = to_length(cut.v.x[i], data.units.length, LengthUnit::bohr);
      const double pot = to_energy(cut.v.f[i], data.units.energy, EnergyUnit::hartree);
      std::array<double, 4> rep{};
      try {
        rep = repulsion(cut.x * a_, cut.y * a_, z[i]);
      } catch (const EvaluationError& e) {
        throw DataError("Atom/" + cut.file + ": " + e.what());
      }
      reduced[i] = pot - rep[0];
      if (i == 0) d1 = s1 - rep[3];
      if (i == n - 1) d2 = s2 - rep[3];
    }
    cuts_.emplace_back(z, reduced, Spline1D::Condition{Spline1D::End::first_derivative, d1},
                       Spline1D::Condition{Spline1D::End::first_derivative, d2});
  }
}

// MURO positions, the reduced cuts (V minus the cluster sum at the site) and the periodic 3D spline of the cell
// (src/crplib/square_atom_surface.cpp:112-146). A negative first spacing is in units of a / sqrt(2) (kept).
void AtomSurface::init_square(const DataSet& data) {
  const double spacing = spacings_.front();
  for (const MuroAtom& m : kMuro) {
    const double z = spacing < 0.0 ? -m.layer * spacing * a_ / std::sqrt(2.0) : m.layer * spacing;
    cluster_.push_back({m.x * a_, m.y * a_, z, m.layer == 0.0 ? 0 : 1});
  }
  const int nz = static_cast<int>(data.atom_cuts.front().v.x.size());
  std::vector<double> z(nz), reduced(6 * static_cast<std::size_t>(nz)), first(6), last(6);
  for (int site = 0; site < 6; ++site) {
    const AtomCut& cut = data.atom_cuts[site];
    const double s1 = to_force(cut.slope_first, data.units, kInternalUnits);
    const double s2 = to_force(cut.slope_last, data.units, kInternalUnits);
    for (int i = 0; i < nz; ++i) {
      z[i] = to_length(cut.v.x[i], data.units.length, LengthUnit::bohr);
      const double pot = to_energy(cut.v.f[i], data.units.energy, EnergyUnit::hartree);
      const std::array<double, 4> rep = repulsion(cut.x * a_, cut.y * a_, z[i]);
      reduced[site * nz + i] = pot - rep[0];
      if (i == 0) first[site] = s1 - rep[3];
      if (i == nz - 1) last[site] = s2 - rep[3];
    }
  }
  std::vector<double> grid(5);
  for (int i = 0; i < 5; ++i) grid[i] = static_cast<double>(i) * a_ * 0.25;
  std::vector<double> f(25 * static_cast<std::size_t>(nz)), d1(25), d2(25);
  for (int site = 0; site < 6; ++site)
    for (const auto& [gi, gj] : kSquareSitePoints[site]) {
      const int ij = (gi - 1) + 5 * (gj - 1);
      for (int k = 0; k < nz; ++k) f[ij + 25 * k] = reduced[site * nz + k];
      d1[ij] = first[site];
      d2[ij] = last[site];
    }
  spline_ = PeriodicSpline3D(grid, grid, z, f, d1, d2);
}

// int_PES6CuHATOMCu100 (src/crplib/square_atom_surface.cpp:148-158): 0 above the data; the point folded into the
// triangle, the cluster sum and the spline there, the gradient mapped back.
AtomValue AtomSurface::evaluate_square(const Vec3& p) const {
  if (p.z >= spline_.z().back()) return {0.0, {0.0, 0.0, 0.0}};
  auto [x, y] = fold_square(p.x, p.y, a_);
  const int region = fold_into_triangle(x, y, a_);
  const std::array<double, 4> rep = repulsion(x, y, p.z);
  const std::array<double, 4> sp = spline_.evaluate(x, y, p.z);
  double gx = sp[1] + rep[1], gy = sp[2] + rep[2];
  unfold_gradient(region, gx, gy);
  return {sp[0] + rep[0], {gx, gy, sp[3] + rep[3]}};
}

// MURII: the corrugation function of every cluster atom within its reach, summed in cluster order
// (src/crplib/hexagonal_atom_surface.cpp:109-154).
std::array<double, 4> AtomSurface::repulsion(double x, double y, double z) const {
  std::array<double, 4> s{0.0, 0.0, 0.0, 0.0};
  const double reach = corrugation_.x().back(), nearest = corrugation_.x().front();
  if (square_) {  // MURO / VREP1D: no lower limit (SPLINT continues the first cubic), dV/dr (x - x_a) / r
    for (const ClusterAtom& at : cluster_) {
      const double r = std::sqrt((x - at.x) * (x - at.x) + (y - at.y) * (y - at.y) + (z - at.z) * (z - at.z));
      double vr = 0.0, dvdr = 0.0;
      if (!(r > reach)) corrugation_.splint(r, vr, dvdr);
      s[0] = s[0] + vr;
      if (r > 0.0) {
        s[1] = s[1] + dvdr * (x - at.x) / r;
        s[2] = s[2] + dvdr * (y - at.y) / r;
        s[3] = s[3] + dvdr * (z - at.z) / r;
      }
    }
    return s;
  }
  for (const ClusterAtom& at : cluster_) {
    const double dx = x - at.x, dy = y - at.y, dz = z - at.z;
    const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    double vr = 0.0, dvdr = 0.0;
    if (!(r > reach)) {
      if (r < nearest)
        throw EvaluationError("CRP atom-surface PES: atom " + std::to_string(r) +
                              " bohr from a surface atom, closer than the corrugation function");
      corrugation_.splint(r, vr, dvdr);
    }
    s[0] = s[0] + vr;
    if (!(r > 0.0) && at.layer != 0) throw EvaluationError("CRP atom-surface PES: atom on a sub-surface atom");
    if (r > 0.0) {
      const double q = dvdr / r;
      s[1] = s[1] + q * dx;
      s[2] = s[2] + q * dy;
      s[3] = s[3] + q * dz;
    } else {
      s[3] = s[3] + dvdr;
    }
  }
  return s;
}

// FOUR3II: the star functions of the hexagonal cell and their derivatives, in crystal coordinates, the y
// derivative turned Cartesian at the end (src/crplib/hexagonal_atom_surface.cpp:187-248).
void AtomSurface::basis(double xa, double ya, double* f, double* dfdx, double* dfdy) const {
  const double fac = 2.00 * kPi / a_;
  const double bigx = xa - ya * kUsq3;
  const double bigy = 2.00 * ya * kUsq3;
  const double p = fac * bigx, q = fac * bigy;
  const double c2a = std::cos(p), c2b = std::cos(q), c2c = std::cos(p + q);
  const double s2a = std::sin(p), s2b = std::sin(q), s2c = std::sin(p + q);
  const double c3a = std::cos(2.00 * p + q), c3b = std::cos(p + 2.00 * q), c3c = std::cos(p - q);
  const double s3a = std::sin(2.00 * p + q), s3b = std::sin(p + 2.00 * q), s3c = std::sin(p - q);
  const double c4a = std::cos(2.00 * p), c4b = std::cos(2.00 * q), c4c = std::cos(2.00 * (p + q));
  const double s4a = std::sin(2.00 * p), s4b = std::sin(2.00 * q), s4c = std::sin(2.00 * (p + q));
  const double c5a = std::cos(3.00 * p + q), c5b = std::cos(p + 3.00 * q), c5c = std::cos(2.00 * p + 3.00 * q);
  const double c5d = std::cos(3.00 * p + 2.00 * q), c5e = std::cos(-p + 2.00 * q), c5f = std::cos(2.00 * p - q);
  const double s5a = std::sin(3.00 * p + q), s5b = std::sin(p + 3.00 * q), s5c = std::sin(2.00 * p + 3.00 * q);
  const double s5d = std::sin(3.00 * p + 2.00 * q), s5e = std::sin(-p + 2.00 * q), s5f = std::sin(2.00 * p - q);
  const double c6a = std::cos(3.00 * p), c6b = std::cos(3.00 * q), c6c = std::cos(3.00 * (p + q));
  const double s6a = std::sin(3.00 * p), s6b = std::sin(3.00 * q), s6c = std::sin(3.00 * (p + q));
  const double c7a = std::cos(2.00 * (2.00 * p + q)), c7b = std::cos(2.00 * (p + 2.00 * q));
  const double c7c = std::cos(2.00 * (p - q));
  const double s7a = std::sin(2.00 * (2.00 * p + q)), s7b = std::sin(2.00 * (p + 2.00 * q));
  const double s7c = std::sin(2.00 * (p - q));
  int k = 0;
  auto put = [&](double v, double dx, double dy) {
    f[k] = v;
    dfdx[k] = dx;
    dfdy[k] = dy;
    ++k;
  };
  put(1.00, 0.00, 0.00);
  put(c2a + c2b + c2c, -fac * (s2a + s2c), -fac * (s2b + s2c));
  if (c3v_) put(-s2a - s2b + s2c, -fac * (c2a - c2c), -fac * (c2b - c2c));
  put(c3a + c3b + c3c, -fac * (2.00 * s3a + s3b + s3c), -fac * (s3a + 2.00 * s3b - s3c));
  put(c4a + c4b + c4c, -2.00 * fac * (s4a + s4c), -2.00 * fac * (s4b + s4c));
  if (c3v_) put(-s4a - s4b + s4c, -2.00 * fac * (c4a - c4c), -2.00 * fac * (c4b - c4c));
  put(c5a + c5b + c5c + c5d + c5e + c5f, -fac * (3.00 * s5a + s5b + 2.00 * s5c + 3.00 * s5d - s5e + 2.00 * s5f),
      -fac * (s5a + 3.00 * s5b + 3.00 * s5c + 2.00 * s5d + 2.00 * s5e - s5f));
  if (c3v_)
    put(-s5a - s5b + s5c + s5d - s5e - s5f, fac * (-3.00 * c5a - c5b + 2.00 * c5c + 3.00 * c5d + c5e - 2.00 * c5f),
        fac * (-c5a - 3.00 * c5b + 3.00 * c5c + 2.00 * c5d - 2.00 * c5e + c5f));
  put(c6a + c6b + c6c, -3.00 * fac * (s6a + s6c), -3.00 * fac * (s6b + s6c));
  put(c7a + c7b + c7c, -2.00 * fac * (2.00 * s7a + s7b + s7c), -2.00 * fac * (s7a + 2.00 * s7b - s7c));
  for (int i = 0; i < n_; ++i) dfdy[i] = (-dfdx[i] + 2.00 * dfdy[i]) * kUsq3;
}

// ATII (src/crplib/hexagonal_atom_surface.cpp:253-283): the repulsive sum at the folded position plus the Fourier
// interpolation of the reduced cuts.
AtomValue AtomSurface::evaluate(const Vec3& p) const {
  if (square_) return evaluate_square(p);
  const auto [xf, yf] = fold(p.x, p.y, a_);
  const std::array<double, 4> rep = repulsion(xf, yf, p.z);
  const std::vector<double>& z = cuts_.front().x();
  if (p.z < z.front() || p.z > z.back())
    throw EvaluationError("CRP atom-surface PES: z = " + std::to_string(p.z) + " bohr outside the data range");
  std::array<double, 10> vs{}, dvs{}, f{}, dfdx{}, dfdy{};
  for (int i = 0; i < n_; ++i) {
    vs[i] = cuts_[i].dpl(p.z);
    dvs[i] = cuts_[i].dplp(p.z);
  }
  basis(p.x, p.y, f.data(), dfdx.data(), dfdy.data());
  double v = 0.00, gx = 0.00, gy = 0.00, gz = 0.00;
  for (int j = 0; j < n_; ++j) {
    double a = 0.00, b = 0.00;
    for (int s = 0; s < n_; ++s) {
      a = a + table_[s + n_ * j] * vs[s];
      b = b + table_[s + n_ * j] * dvs[s];
    }
    v = v + a * f[j];
    gx = gx + a * dfdx[j];
    gy = gy + a * dfdy[j];
    gz = gz + b * f[j];
  }
  return {v + rep[0], {gx + rep[1], gy + rep[2], gz + rep[3]}};
}

// ---------------------------------------------------------------------------------------------------------------
// Molecule-surface PES

// CAMBCORD with the centre at the midpoint (src/crplib/molecule_surface.cpp:22-41)
Geometry molecule_geometry(const Vec3& a, const Vec3& b) {
  constexpr double cc = 0.500;
  const double r = std::sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y) + (b.z - a.z) * (b.z - a.z));
  if (r == 0.0) throw EvaluationError("CRP molecule-surface PES: the two atoms coincide");
  Geometry g{};
  g.r = r;
  g.x = a.x * (1.0 - cc) + b.x * cc;
  g.y = a.y * (1.0 - cc) + b.y * cc;
  g.z = a.z * (1.0 - cc) + b.z * cc;
  double ct = (b.z - a.z) / r;
  if (ct > 1.0) ct = 1.0;
  g.theta = std::acos(ct);
  const double dx = b.x - a.x, dy = b.y - a.y;
  g.phi = (dx == 0.0 && dy == 0.0) ? 0.0 : std::atan2(dy, dx);
  return g;
}

namespace {

// The site functions: angular interpolation at one site from its cuts at fixed orientations
// (src/crplib/site_interpolation.cpp:60-222). Each returns (value, d/dZ, d/dr, d/dphi, d/dtheta); the first three
// apply the same formula to the cut values, their Z and their r derivatives.
struct Site {
  double v[3];
  double d_phi, d_theta;
};
struct Trig {
  double c2, s2, c4, s4;  // of 2 theta, 4 theta
};

double cutc(const Spline2D::Value* cuts, int comp, int k) {
  const Spline2D::Value& c = cuts[k];
  return comp == 0 ? c.v : (comp == 1 ? c.d_dy : c.d_dx);
}

// theta = 0, 45, 90, 135 degrees combined (the hollow and t2 sites)
double four(double v0, double v45, double v90, double v135, const Trig& t) {
  return 0.2500 * ((v0 + v45 + v90 + v135) + (v0 + v90 - v45 - v135) * t.c4) +
         0.500 * ((v0 - v90) * t.c2 + (v45 - v135) * t.s2);
}
double four_dtheta(double v0, double v45, double v90, double v135, const Trig& t) {
  return -(v0 - v90) * t.s2 + (v45 - v135) * t.c2 - (v0 + v90 - v45 - v135) * t.s4;
}
double four_dphi(double d45, double d90, double d135, const Trig& t) {
  return 0.2500 * ((d45 + d90 + d135) + (d90 - d45 - d135) * t.c4) + 0.500 * ((-d90) * t.c2 + (d45 - d135) * t.s2);
}

// top: theta 0; 90/0; 90/30
Site top_site(const Spline2D::Value* c, const Trig& t, double phi) {
  const double c6 = std::cos(6.00 * phi), s6 = std::sin(6.00 * phi);
  Site s{};
  for (int i = 0; i < 3; ++i) {
    const double v0 = cutc(c, i, 0);
    const double v90 = 0.500 * ((cutc(c, i, 1) + cutc(c, i, 2)) + (cutc(c, i, 1) - cutc(c, i, 2)) * c6);
    s.v[i] = 0.500 * ((v0 + v90) + (v0 - v90) * t.c2);
    if (i == 0) {
      const double d90 = -3.00 * (cutc(c, 0, 1) - cutc(c, 0, 2)) * s6;
      s.d_phi = 0.500 * d90 * (1.00 - t.c2);
      s.d_theta = -(v0 - v90) * t.s2;
    }
  }
  return s;
}

// bridge: theta 0; 90/0; 90/60; 90/90
Site bridge_site(const Spline2D::Value* c, const Trig& t, double phi) {
  constexpr double u3 = 1.00 / 3.00;
  const double c2 = std::cos(2.00 * phi), c4 = std::cos(4.00 * phi);
  const double s2 = std::sin(2.00 * phi), s4 = std::sin(4.00 * phi);
  Site s{};
  for (int i = 0; i < 3; ++i) {
    const double v0 = cutc(c, i, 0), a = cutc(c, i, 1), b = cutc(c, i, 2), d = cutc(c, i, 3);
    const double v90 = (a + 2.00 * b) * u3 - 0.500 * (d - a) * c2 + (0.500 * (d - a) - 2.00 * (b - a) * u3) * c4;
    s.v[i] = 0.500 * (v0 * (1.00 + t.c2) + v90 * (1.00 - t.c2));
    if (i == 0) {
      const double d90 = (d - a) * s2 - (2.00 * (d - a) - 8.00 * (b - a) * u3) * s4;
      s.d_phi = 0.500 * d90 * (1.00 - t.c2);
      s.d_theta = (v90 - v0) * t.s2;
    }
  }
  return s;
}

// fcc / hcp hollow: theta 0; 90/0; 90/30; 45/30; 45/210
Site hollow_site(const Spline2D::Value* c, const Trig& t, double phi) {
  const double s3 = std::sin(3.00 * phi), c6 = std::cos(6.00 * phi);
  const double c3 = std::cos(3.00 * phi), s6 = std::sin(6.00 * phi);
  Site s{};
  for (int i = 0; i < 3; ++i) {
    const double v0 = cutc(c, i, 0), a = cutc(c, i, 1), b = cutc(c, i, 2), d = cutc(c, i, 3), e = cutc(c, i, 4);
    const double v45 = 0.500 * ((d + e) + (d - e) * s3);
    const double v135 = 0.500 * ((d + e) - (d - e) * s3);
    const double v90 = 0.500 * ((a + b) + (a - b) * c6);
    s.v[i] = four(v0, v45, v90, v135, t);
    if (i == 0) {
      const double d45 = 1.500 * (d - e) * c3;
      const double d90 = -3.00 * (a - b) * s6;
      s.d_phi = d90 * (0.2500 * (1.00 + t.c4) - 0.500 * t.c2) + d45 * t.s2;
      s.d_theta = four_dtheta(v0, v45, v90, v135, t);
    }
  }
  return s;
}

// The t2 sites: theta 0, two 90 cuts (p, q), three 45 cuts (g, h, k) at the azimuth phi' = phi - shift. `mid`
// selects which 45 cut is the middle one of the cos 2 phi' term and `sign` the orientation of the cos phi' terms
// (C6v t2h: 45 cuts 30, 210, 120 with mid 5; C3v t2h: 30, 120, 210 with mid 4; C3v t2f: 150, 240, 330 with mid 4
// and the odd terms reversed; src/crplib/site_interpolation.cpp:120-194).
Site t2_site(const Spline2D::Value* c, const Trig& t, double phi, double shift, int mid, bool reversed) {
  const double ph = phi - shift;
  const double cp = std::cos(ph), sp = std::sin(ph), c2p = std::cos(2.00 * ph), s2p = std::sin(2.00 * ph);
  const int first45 = 3, last45 = mid == 5 ? 4 : 5;
  Site s{};
  for (int i = 0; i < 3; ++i) {
    const double v0 = cutc(c, i, 0), p = cutc(c, i, 1), q = cutc(c, i, 2);
    const double g = cutc(c, i, first45), k = cutc(c, i, last45), m = cutc(c, i, mid);
    const double odd = reversed ? (-g + k) : (g - k);
    const double v45 = 0.2500 * ((g + 2.00 * m + k) + (g - 2.00 * m + k) * c2p) + 0.500 * odd * cp;
    const double v135 = 0.2500 * ((g + 2.00 * m + k) + (g - 2.00 * m + k) * c2p) - 0.500 * odd * cp;
    const double v90 = reversed ? 0.500 * ((p + q) + (-p + q) * c2p) : 0.500 * ((p + q) + (p - q) * c2p);
    s.v[i] = four(v0, v45, v90, v135, t);
    if (i == 0) {
      const double d45 = -0.500 * (odd * sp + (g - 2.00 * m + k) * s2p);
      const double d135 = 0.500 * (odd * sp - (g - 2.00 * m + k) * s2p);
      const double d90 = reversed ? -(-p + q) * s2p : -(p - q) * s2p;
      s.d_phi = four_dphi(d45, d90, d135, t);
      s.d_theta = four_dtheta(v0, v45, v90, v135, t);
    }
  }
  return s;
}

// FOUR6II: the 12 basis functions of the molecular interpolation and their Cartesian x, y derivatives
// (src/crplib/hexagonal_interaction.cpp:16-72).
void molecule_basis(double x, double y, double a, double* f, double* dfdx, double* dfdy) {
  const double fac = 2.00 * kPi / a;
  const double bigx = x - y * kUsq3;
  const double bigy = 2.00 * y * kUsq3;
  const double p = fac * bigx, q = fac * bigy;
  const double c2x = std::cos(2.00 * p), c2y = std::cos(2.00 * q), c2xy = std::cos(2.00 * (p + q));
  const double s2x = std::sin(2.00 * p), s2y = std::sin(2.00 * q), s2xy = std::sin(2.00 * (p + q));
  f[0] = 1.00;
  f[1] = std::cos(p);
  f[2] = std::cos(q);
  f[3] = std::cos(p + q);
  f[4] = std::sin(p);
  f[5] = std::sin(q);
  f[6] = std::sin(p + q);
  f[7] = std::cos(2.00 * p + q);
  f[8] = std::cos(p + 2.00 * q);
  f[9] = std::cos(p - q);
  f[10] = c2x + c2y + c2xy;
  f[11] = -s2x - s2y + s2xy;
  dfdx[0] = 0.00;
  dfdx[1] = -fac * f[4];
  dfdx[2] = 0.00;
  dfdx[3] = -fac * f[6];
  dfdx[4] = fac * f[1];
  dfdx[5] = 0.00;
  dfdx[6] = fac * f[3];
  dfdx[7] = -2.00 * fac * std::sin(2.00 * p + q);
  dfdx[8] = -fac * std::sin(p + 2.00 * q);
  dfdx[9] = -fac * std::sin(p - q);
  dfdx[10] = -2.0 * fac * (s2x + s2xy);
  dfdx[11] = 2.0 * fac * (-c2x + c2xy);
  dfdy[0] = 0.00;
  dfdy[1] = 0.00;
  dfdy[2] = -fac * f[5];
  dfdy[3] = -fac * f[6];
  dfdy[4] = 0.00;
  dfdy[5] = fac * f[2];
  dfdy[6] = fac * f[3];
  dfdy[7] = 0.500 * dfdx[7];
  dfdy[8] = 2.00 * dfdx[8];
  dfdy[9] = -dfdx[9];
  dfdy[10] = -2.0 * fac * (s2y + s2xy);
  dfdy[11] = 2.0 * fac * (-c2y + c2xy);
  for (int i = 0; i < 12; ++i) dfdy[i] = (-dfdx[i] + 2.00 * dfdy[i]) * kUsq3;
}

// The C4v site functions (src/crplib/site_interpolation.cpp:196-270): top and hollow (theta 0; 90/0; 90/45), bridge
// (0; 90/0; 90/45; 90/90), t2h (0; 45/225; 45/45; 90/45; 90/135; 45/135).
Site square_top_site(const Spline2D::Value* c, const Trig& t, double phi) {
  Site s{};
  for (int i = 0; i < 3; ++i) {
    const double v0 = cutc(c, i, 0), c1 = cutc(c, i, 1), c2 = cutc(c, i, 2);
    const double v90 = 0.5 * (c1 + c2) + 0.5 * (c1 - c2) * std::cos(4.0 * phi);
    s.v[i] = 0.5 * (v0 + v90) + 0.5 * (v0 - v90) * t.c2;
    if (i == 0) {
      const double d0 = 0.0;
      const double d90 = -4.0 * 0.5 * (c1 - c2) * std::sin(4.0 * phi);
      s.d_phi = 0.5 * (d0 + d90) + 0.5 * (d0 - d90) * t.c2;
      s.d_theta = -2.0 * 0.5 * (v0 - v90) * t.s2;
    }
  }
  return s;
}

Site square_bridge_site(const Spline2D::Value* c, const Trig& t, double phi) {
  Site s{};
  for (int i = 0; i < 3; ++i) {
    const double v0 = cutc(c, i, 0), c1 = cutc(c, i, 1), c2 = cutc(c, i, 2), c3 = cutc(c, i, 3);
    const double v90 = (0.25 * c1 + 0.5 * c2 + 0.25 * c3) + 0.5 * (c1 - c3) * std::cos(2.0 * phi) +
                       (0.25 * c1 - 0.5 * c2 + 0.25 * c3) * std::cos(4.0 * phi);
    s.v[i] = 0.5 * (v0 + v90) + 0.5 * (v0 - v90) * t.c2;
    if (i == 0) {
      const double d0 = 0.0;
      const double d90 =
          -2.0 * 0.5 * (c1 - c3) * std::sin(2.0 * phi) - 4.0 * (0.25 * c1 - 0.5 * c2 + 0.25 * c3) * std::sin(4.0 * phi);
      s.d_phi = 0.5 * (d0 + d90) + 0.5 * (d0 - d90) * t.c2;
      s.d_theta = -2.0 * 0.5 * (v0 - v90) * t.s2;
    }
  }
  return s;
}

Site square_t2h_site(const Spline2D::Value* c, const Trig& t, double phi) {
  const double phi45 = phi - kPi / 4.0;
  const double phi135 = phi + kPi - kPi / 4.0;
  Site s{};
  for (int i = 0; i < 3; ++i) {
    const double v0 = cutc(c, i, 0), c1 = cutc(c, i, 1), c2 = cutc(c, i, 2);
    const double c3 = cutc(c, i, 3), c4 = cutc(c, i, 4), c5 = cutc(c, i, 5);
    // V(45, phi) = a0 + b0 cos(phi - 45) + c0 cos(2 (phi - 45)) through the three 45-degree cuts
    const double cc0 = 0.25 * (c1 + c2 - 2.0 * c5);
    const double b0 = -c5 + c2 - 2.0 * cc0;
    const double a0 = c2 - b0 - cc0;
    const double v45 = a0 + b0 * std::cos(phi45) + cc0 * std::cos(2.0 * phi45);
    const double v135 = a0 + b0 * std::cos(phi135) + cc0 * std::cos(2.0 * phi135);
    const double v90 = 0.5 * (c3 + c4) + 0.5 * (c3 - c4) * std::cos(2.0 * phi45);
    s.v[i] = 0.25 * (v0 + v45 + v90 + v135) + 0.5 * (v0 - v90) * t.c2 + 0.5 * (v45 - v135) * t.s2 +
             0.25 * (v0 + v90 - v45 - v135) * t.c4;
    if (i == 0) {
      const double d0 = 0.0;
      const double d45 = -b0 * std::sin(phi45) - 2.0 * cc0 * std::sin(2.0 * phi45);
      const double d135 = -b0 * std::sin(phi135) - 2.0 * cc0 * std::sin(2.0 * phi135);
      const double d90 = -2.0 * 0.5 * (c3 - c4) * std::sin(2.0 * phi45);
      s.d_phi = 0.25 * (d0 + d45 + d90 + d135) + 0.5 * (d0 - d90) * t.c2 + 0.5 * (d45 - d135) * t.s2 +
                0.25 * (d0 + d90 - d45 - d135) * t.c4;
      s.d_theta = -2.0 * 0.5 * (v0 - v90) * t.s2 + 2.0 * 0.5 * (v45 - v135) * t.c2 -
                  4.0 * 0.25 * (v0 + v90 - v45 - v135) * t.s4;
    }
  }
  return s;
}

std::vector<double> lengths(const std::vector<double>& x, const Units& u) {
  std::vector<double> y;
  for (double v : x) y.push_back(to_length(v, u.length, LengthUnit::bohr));
  return y;
}
std::vector<double> energies(const std::vector<double>& x, const Units& u) {
  std::vector<double> y;
  for (double v : x) y.push_back(to_energy(v, u.energy, EnergyUnit::hartree));
  return y;
}

}  // namespace

// Read_MoleculeSurface_Data (src/crplib/molecule_surface.cpp:119-192)
MoleculeSurface::MoleculeSurface(const DataSet& data)
    : symmetry_(data.symmetry),
      a_(to_length(data.lattice_constant, data.units.length, LengthUnit::bohr)),
      atom_(data) {
  const Units& in = data.units;
  zso_min_ = to_length(data.zso_min, in.length, LengthUnit::bohr);
  zso_max_ = to_length(data.zso_max, in.length, LengthUnit::bohr);
  z_dependence_ = Spline1D(lengths(data.z_dependence.x, in), energies(data.z_dependence.f, in), {},
                           {Spline1D::End::first_derivative, 0.00});
  h2_ = Spline1D(lengths(data.h2.x, in), energies(data.h2.f, in), {}, {});
  for (const MoleculeCut& cut : data.molecule_cuts) {
    if (!cut.present) continue;
    const std::vector<double> r = lengths(cut.r, in), z = lengths(cut.z, in);
    std::vector<double> v = energies(cut.v, in);
    if (!data.cuts_interpolation) {  // total energies: subtract both atom potentials (corrugation reduction)
      const double ucm = cut.x * a_, vcm = cut.y * a_;  // skewed for C3v / C6v, Cartesian for C4v
      const bool sq = symmetry_ == Symmetry::c4v;
      const double xcm = sq ? ucm : ucm + vcm / 2.0, ycm = sq ? vcm : std::sqrt(3.0) * vcm / 2.0;
      const double theta = cut.theta * kPi / 180.0, phi = cut.phi * kPi / 180.0;
      for (std::size_t j = 0; j < z.size(); ++j) {
        for (std::size_t i = 0; i < r.size(); ++i) {
          const Vec3 pa{xcm - 0.5 * r[i] * std::sin(theta) * std::cos(phi),
                        ycm - 0.5 * r[i] * std::sin(theta) * std::sin(phi), z[j] - 0.5 * r[i] * std::cos(theta)};
          const Vec3 pb{xcm + 0.5 * r[i] * std::sin(theta) * std::cos(phi),
                        ycm + 0.5 * r[i] * std::sin(theta) * std::sin(phi), z[j] + 0.5 * r[i] * std::cos(theta)};
          double va = 0.0, vb = 0.0;
          try {
            va = atom_.evaluate(pa).v;
            vb = atom_.evaluate(pb).v;
          } catch (const EvaluationError& e) {
            throw DataError("Molecule/" + cut.file + ": " + e.what());
          }
          double& value = v[i + r.size() * j];
          value = value - va - vb;
        }
      }
    }
    if (r_min_ < r.front()) r_min_ = r.front();
    if (r_max_ > r.back()) r_max_ = r.back();
    if (z_min_ < z.front()) z_min_ = z.front();
    cuts_.emplace_back(r, z, v);
  }
}

MoleculeSurface::Switch MoleculeSurface::switch_function(double z) const {
  if (z <= zso_min_) return {1.00, 0.00};
  if (z >= zso_max_) return {0.00, 0.00};
  const double f = 0.500 + 0.500 * std::cos((z - zso_min_) * kPi / (zso_max_ - zso_min_));
  const double df = -0.500 * kPi / (zso_max_ - zso_min_) * std::sin((z - zso_min_) * kPi / (zso_max_ - zso_min_));
  return {f, df};
}

MoleculeSurface::GasPhase MoleculeSurface::gas_phase(double z, double r) const {
  double hv = 0.0, hs = 0.0;
  h2_.splint(r, hv, hs);
  double v1 = 0.0, dz = 0.0;
  if (z >= zso_min_ && z <= z_dependence_.x().back()) {
    z_dependence_.splint(z, v1, dz);
  } else if (z > z_dependence_.x().back()) {
    v1 = z_dependence_.y().back();
    dz = 0.00;
  } else {
    throw EvaluationError("CRP gas-phase potential: Z = " + std::to_string(z) + " below ZsoMin");
  }
  return {hv + v1, dz, hs};
}

// RRINTII's site part and SISII (src/crplib/hexagonal_interaction.cpp:77-160): the 12 Fourier sites from their
// cuts (rotated azimuths for the symmetry images), then the Fourier interpolation over the centre. Returns
// (I, dI/dZ, dI/dr, dI/dphi, dI/dtheta, dI/dX, dI/dY).
std::array<double, 7> MoleculeSurface::interaction(const Spline2D::Value* c, double theta, double phi, double x,
                                                   double y) const {
  constexpr double pis3 = kPi / 3.00;
  constexpr double pi2s3 = 2.00 * pis3;
  constexpr double pis6 = kPi / 6.00;
  const Trig t{std::cos(2.00 * theta), std::sin(2.00 * theta), std::cos(4.00 * theta), std::sin(4.00 * theta)};
  std::array<Site, 12> s{};
  s[0] = top_site(c, t, phi);
  s[1] = bridge_site(c + 3, t, phi);
  s[4] = bridge_site(c + 3, t, phi + pis3);
  s[5] = bridge_site(c + 3, t, phi - pis3);
  if (symmetry_ == Symmetry::c6v) {
    s[2] = hollow_site(c + 7, t, phi);
    s[8] = hollow_site(c + 7, t, phi + kPi);
    const Spline2D::Value* t2 = c + 12;
    s[3] = t2_site(t2, t, phi, pis6, 5, false);
    s[6] = t2_site(t2, t, phi - pi2s3, pis6, 5, false);
    s[7] = t2_site(t2, t, phi + pi2s3, pis6, 5, false);
    s[9] = t2_site(t2, t, phi + kPi, pis6, 5, false);
    s[10] = t2_site(t2, t, phi + pis3, pis6, 5, false);
    s[11] = t2_site(t2, t, phi - pis3, pis6, 5, false);
  } else {
    s[2] = hollow_site(c + 7, t, phi);
    const Spline2D::Value* t2h = c + 12;
    s[3] = t2_site(t2h, t, phi, pis6, 4, false);
    s[6] = t2_site(t2h, t, phi - pi2s3, pis6, 4, false);
    s[7] = t2_site(t2h, t, phi + pi2s3, pis6, 4, false);
    const Spline2D::Value* t2f = c + 18;
    s[10] = t2_site(t2f, t, phi, 11.00 * pis6, 4, true);
    s[9] = t2_site(t2f, t, phi + pi2s3, 11.00 * pis6, 4, true);
    s[11] = t2_site(t2f, t, phi - pi2s3, 11.00 * pis6, 4, true);
    s[8] = hollow_site(c + 24, t, phi);
  }
  std::array<double, 12> f{}, dfdx{}, dfdy{};
  molecule_basis(x, y, a_, f.data(), dfdx.data(), dfdy.data());
  std::array<double, 7> out{};
  for (int j = 0; j < 12; ++j) {
    double a = 0.00;
    for (int k = 0; k < 12; ++k) a = a + kMoleculeTable[k + 12 * j] * s[k].v[0];
    out[0] = out[0] + a * f[j];
    out[5] = out[5] + a * dfdx[j];
    out[6] = out[6] + a * dfdy[j];
  }
  for (int i = 1; i < 5; ++i) {
    for (int j = 0; j < 12; ++j) {
      double a = 0.00;
      for (int k = 0; k < 12; ++k) {
        const double comp = i == 1 ? s[k].v[1] : (i == 2 ? s[k].v[2] : (i == 3 ? s[k].d_phi : s[k].d_theta));
        a = a + kMoleculeTable[k + 12 * j] * comp;
      }
      out[i] = out[i] + a * f[j];
    }
  }
  return out;
}

// RRINT's site part and SIS8x8 (src/crplib/square_interaction.cpp): 8 Fourier sites (top, the bridges, hollow, the
// t2h sites) and the Fourier interpolation over the centre (x, y Cartesian).
std::array<double, 7> MoleculeSurface::square_interaction(const Spline2D::Value* c, double theta, double phi,
                                                          double x, double y) const {
  const Trig t{std::cos(2.00 * theta), std::sin(2.00 * theta), std::cos(4.00 * theta), std::sin(4.00 * theta)};
  std::array<Site, 8> s{};
  s[0] = square_top_site(c, t, phi);
  s[1] = square_bridge_site(c + 3, t, phi);
  s[4] = square_bridge_site(c + 3, t, phi + kPi / 2.0);
  s[2] = square_top_site(c + 7, t, phi);  // the hollow has the top's cut pattern
  s[3] = square_t2h_site(c + 10, t, phi);
  s[5] = square_t2h_site(c + 10, t, phi - kPi / 2.0);
  s[6] = square_t2h_site(c + 10, t, phi - kPi);
  s[7] = square_t2h_site(c + 10, t, phi + kPi / 2.0);
  auto comp = [&](int k, int i) {
    return i == 0 ? s[k].v[0] : (i == 1 ? s[k].v[1] : (i == 2 ? s[k].v[2] : (i == 3 ? s[k].d_phi : s[k].d_theta)));
  };
  // the coefficients of the 8 basis functions from the site values (the inverse of the basis at the sites)
  auto coefficients = [&](int i) {
    const double p1 = comp(0, i), p2 = comp(1, i), p3 = comp(2, i), p4 = comp(3, i);
    const double p5 = comp(4, i), p6 = comp(5, i), p7 = comp(6, i), p8 = comp(7, i);
    return std::array<double, 8>{1.0 / 8.0 * (p1 + p2 + p5 + p3 + p4 + p6 + p7 + p8),
                                 1.0 / 4.0 * (p1 - p2 + p5 - p3),
                                 1.0 / 4.0 * (p1 + p2 - p5 - p3),
                                 1.0 / 4.0 * (p4 - p6 - p7 + p8),
                                 1.0 / 4.0 * (p4 + p6 - p7 - p8),
                                 1.0 / 8.0 * (p1 - p2 - p5 + p3 - p4 + p6 - p7 + p8),
                                 1.0 / 8.0 * (p1 - p2 - p5 + p3 + p4 - p6 + p7 - p8),
                                 1.0 / 16.0 * (p1 + p2 + p5 + p3 - p4 - p6 - p7 - p8)};
  };
  const double fac = 2.0 * kPi / a_;
  const double f[8] = {1.0,
                       std::cos(fac * x),
                       std::cos(fac * y),
                       std::sin(fac * x),
                       std::sin(fac * y),
                       std::cos(fac * (x + y)),
                       std::cos(fac * (x - y)),
                       std::cos(fac * (2.0 * x)) + std::cos(fac * (2.0 * y))};
  const double fx[8] = {0.0,
                        -fac * std::sin(fac * x),
                        0.0,
                        fac * std::cos(fac * x),
                        0.0,
                        -fac * std::sin(fac * (x + y)),
                        -fac * std::sin(fac * (x - y)),
                        -2.0 * fac * std::sin(fac * (2.0 * x))};
  const double fy[8] = {0.0,
                        0.0,
                        -fac * std::sin(fac * y),
                        0.0,
                        fac * std::cos(fac * y),
                        -fac * std::sin(fac * (x + y)),
                        fac * std::sin(fac * (x - y)),
                        -2.0 * fac * std::sin(fac * (2.0 * y))};
  auto dot = [](const std::array<double, 8>& a, const double* g) {
    return a[0] * g[0] + a[1] * g[1] + a[2] * g[2] + a[3] * g[3] + a[4] * g[4] + a[5] * g[5] + a[6] * g[6] +
           a[7] * g[7];
  };
  std::array<double, 7> out{};
  const std::array<double, 8> a0 = coefficients(0);
  out[0] = dot(a0, f);
  out[5] = dot(a0, fx);
  out[6] = dot(a0, fy);
  for (int i = 1; i < 5; ++i) out[i] = dot(coefficients(i), f);
  return out;
}

// POT6D -> CAMBCORD -> RRINTII (src/crplib/molecule_surface.cpp:249-315)
std::optional<MoleculeValue> MoleculeSurface::evaluate(const Vec3& a, const Vec3& b) const {
  const Geometry g = molecule_geometry(a, b);
  if (!(g.z < zso_max_)) {  // far from the surface: the gas-phase potential only
    if (g.r < h2_.x().front() || g.r > h2_.x().back()) return std::nullopt;
    const GasPhase gp = gas_phase(g.z, g.r);
    const double rxa = -(b.x - a.x) / g.r, rya = -(b.y - a.y) / g.r, rza = -(b.z - a.z) / g.r;
    return MoleculeValue{gp.v,
                         {rxa * gp.d_r, rya * gp.d_r, 0.500 * gp.d_z + rza * gp.d_r},
                         {-rxa * gp.d_r, -rya * gp.d_r, 0.500 * gp.d_z + -rza * gp.d_r}};
  }
  const int n = molecule_cuts_used(symmetry_);
  std::array<Spline2D::Value, 29> cv{};
  for (int k = 0; k < n; ++k) {
    const auto cell = cuts_[k].locate(g.r, g.z);
    if (!cell) return std::nullopt;  // outside a cut's grid the Fortran gives up silently
    cv[k] = cuts_[k].at(*cell);
  }
  const bool square = symmetry_ == Symmetry::c4v;
  const std::array<double, 7> vi = square ? square_interaction(cv.data(), g.theta, g.phi, g.x, g.y)
                                          : interaction(cv.data(), g.theta, g.phi, g.x, g.y);
  // derivatives of r, theta, phi with respect to atom a (RRINTII; RRINT for C4v in its own operand order)
  const double rxa = -(b.x - a.x) / g.r, rya = -(b.y - a.y) / g.r, rza = -(b.z - a.z) / g.r;
  double txa = 0.0, tya = 0.0, pxa = 0.0, pya = 0.0;
  const double rho2 = (b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y);
  const double rho = std::sqrt(rho2);
  if (rho2 > 0.00) {
    txa = square ? -(b.z - a.z) * (b.x - a.x) / (g.r * g.r) / std::sqrt(rho2)
                 : -(b.z - a.z) * (b.x - a.x) / (g.r * g.r * rho);
    tya = square ? -(b.z - a.z) * (b.y - a.y) / (g.r * g.r) / std::sqrt(rho2)
                 : -(b.z - a.z) * (b.y - a.y) / (g.r * g.r * rho);
    pxa = (b.y - a.y) / rho2;
    pya = -(b.x - a.x) / rho2;
  }
  const double tza = rho / (g.r * g.r);
  const AtomValue va = atom_.evaluate(a);
  const AtomValue vb = atom_.evaluate(b);
  const double iv = vi[0], iz = vi[1], ir = vi[2], iphi = vi[3], itheta = vi[4], ix = vi[5], iy = vi[6];
  const Vec3 fa{va.gradient.x + 0.500 * ix + rxa * ir + txa * itheta + pxa * iphi,
                va.gradient.y + 0.500 * iy + rya * ir + tya * itheta + pya * iphi,
                va.gradient.z + 0.500 * iz + rza * ir + tza * itheta};
  const Vec3 fb{vb.gradient.x + 0.500 * ix + (-rxa) * ir + (-txa) * itheta + (-pxa) * iphi,
                vb.gradient.y + 0.500 * iy + (-rya) * ir + (-tya) * itheta + (-pya) * iphi,
                vb.gradient.z + 0.500 * iz + (-rza) * ir + (-tza) * itheta};
  if (g.z <= zso_min_) return MoleculeValue{va.v + vb.v + iv, fa, fb};
  const Switch sw = switch_function(g.z);  // switching region: (Va + Vb + I) F + (1 - F) V2D
  const GasPhase gp = gas_phase(g.z, g.r);
  const double dfz = 0.500 * sw.df_dz;
  const double near = va.v + vb.v + iv;
  return MoleculeValue{near * sw.f + (1.00 - sw.f) * gp.v,
                       {fa.x * sw.f + (1.00 - sw.f) * gp.d_r * rxa, fa.y * sw.f + (1.00 - sw.f) * gp.d_r * rya,
                        fa.z * sw.f + near * dfz - dfz * gp.v + (1.00 - sw.f) * (gp.d_z * 0.500 + rza * gp.d_r)},
                       {fb.x * sw.f + (1.00 - sw.f) * gp.d_r * (-rxa), fb.y * sw.f + (1.00 - sw.f) * gp.d_r * (-rya),
                        fb.z * sw.f + near * dfz - dfz * gp.v +
                            (1.00 - sw.f) * (gp.d_z * 0.500 + (-rza) * gp.d_r)}};
}

// ---------------------------------------------------------------------------------------------------------------
// Flat arrays for the device

Packed pack(const QdSurface& q) {
  const MoleculeSurface& m = q.surface();
  const AtomSurface& at = m.atom_;
  Packed p;
  auto spline = [](const Spline1D& s, std::vector<double>& out) {
    out.insert(out.end(), s.x_.begin(), s.x_.end());
    out.insert(out.end(), s.y_.begin(), s.y_.end());
    out.insert(out.end(), s.m_.begin(), s.m_.end());
  };
  p.c3v = at.c3v_ ? 1 : 0;
  p.lattice_constant = m.a_;
  p.zso_min = m.zso_min_;
  p.zso_max = m.zso_max_;
  p.r_min = m.r_min_;
  p.r_max = m.r_max_;
  p.z_min = m.z_min_;
  p.h2_asymptote = m.h2_asymptote();
  spline(at.corrugation_, p.corrugation);
  for (const auto& c : at.cluster_) p.cluster.insert(p.cluster.end(), {c.x, c.y, c.z, static_cast<double>(c.layer)});
  if (at.square_) {
    const PeriodicSpline3D& c = at.spline_;
    p.square = 1;
    p.cell_nz = static_cast<int>(c.z_.size());
    for (const auto* v : {&c.x_, &c.y_, &c.z_, &c.f_, &c.xx_, &c.yy_, &c.xxyy_, &c.zz_, &c.zzxx_, &c.zzyy_, &c.zzxxyy_})
      p.cell_spline.insert(p.cell_spline.end(), v->begin(), v->end());
  }
  p.atom_offsets.push_back(0);
  for (const Spline1D& c : at.cuts_) {
    spline(c, p.atom_splines);
    p.atom_offsets.push_back(p.atom_offsets.back() + static_cast<int>(c.x_.size()));
  }
  const int used = molecule_cuts_used(m.symmetry_);
  for (int k = 0; k < used; ++k) {
    const Spline2D& c = m.cuts_[k];
    int group = k;
    for (int g = 0; g < k && group == k; ++g)
      if (m.cuts_[g].x_ == c.x_ && m.cuts_[g].y_ == c.y_) group = g;
    p.cut_shape.insert(p.cut_shape.end(), {static_cast<int>(c.x_.size()), static_cast<int>(c.y_.size()), group});
    p.cut_offsets.push_back(static_cast<long>(p.cut_data.size()));
    for (const auto* v : {&c.x_, &c.y_, &c.f_, &c.fxx_, &c.fyy_, &c.fxxyy_})
      p.cut_data.insert(p.cut_data.end(), v->begin(), v->end());
  }
  p.cut_offsets.push_back(static_cast<long>(p.cut_data.size()));
  spline(m.h2_, p.h2);
  spline(m.z_dependence_, p.z_dependence);
  return p;
}

// ---------------------------------------------------------------------------------------------------------------
// QD wrapper (src/crplib/qd_wrapper.cpp:20-66)

QdSurface::QdSurface(const DataSet& data) : surface_(data) {}

double QdSurface::skewing_angle() const {
  return surface_.symmetry() == Symmetry::c4v ? kPi / 2.0 : kPi / 3.0;
}

double QdSurface::potential_ev(double u, double v, double z, double r, double theta, double phi) const {
  constexpr double delta = 1.0e-12;  // keeps the clamped values inside the data
  double rr = r, zz = z;
  if (rr < surface_.r_min() + delta) {
    rr = surface_.r_min() + delta;
  } else if (rr > surface_.r_max() - delta) {
    rr = surface_.r_max() - delta;
  }
  if (zz < surface_.z_min() + delta) zz = surface_.z_min() + delta;
  // [[1, cos g], [0, sin g]] (u, v); cos(pi/2) as gfortran folds it, bits 0x3C91A62633145C07
  const bool sq = surface_.symmetry() == Symmetry::c4v;
  const double cg = sq ? 0x1.1a62633145c07p-54 : kCosPi3, sg = sq ? 1.0 : kSinPi3;
  const double x = 1.0 * u + cg * v;
  const double y = 0.0 * u + sg * v;
  const Vec3 a{x - 0.5 * rr * std::sin(theta) * std::cos(phi), y - 0.5 * rr * std::sin(theta) * std::sin(phi),
               zz - 0.5 * rr * std::cos(theta)};
  const Vec3 b{x + 0.5 * rr * std::sin(theta) * std::cos(phi), y + 0.5 * rr * std::sin(theta) * std::sin(phi),
               zz + 0.5 * rr * std::cos(theta)};
  const std::optional<MoleculeValue> m = surface_.evaluate(a, b);
  const double value = to_energy(m ? m->v : 0.00, EnergyUnit::hartree, EnergyUnit::ev);
  return value - to_energy(surface_.h2_asymptote(), EnergyUnit::hartree, EnergyUnit::ev);
}

}  // namespace spo::crp
