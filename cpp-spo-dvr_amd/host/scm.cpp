// The SCM correction and the CRP + SCM surface (src/potential/scm/scm_correction.cpp, scm_files.cpp,
// scm_crp_potential.cpp). dV keeps the Fortran's operand order step by step (SCM_Potential.f90,
// SCM_Wrapper.f90): its results are the Fortran's bit for bit.
#include "scm.hpp"

#include <cmath>
#include <stdexcept>

#include "crp_tables.hpp"
#include "list_reader.hpp"

namespace spo::scm {

namespace {

[[noreturn]] void fail(const std::string& file, const std::string& what) { throw crp::DataError(file + ": " + what); }

// The unswitched pair function -(p3 + q (p4 + q (p5 + q p6))) exp(-p2 q) (Switch_Rydberg_Potential)
double unswitched(const PairTerms& t, double q) {
  const double cubic = t.polynomial[0] + q * (t.polynomial[1] + q * (t.polynomial[2] + q * t.polynomial[3]));
  return -(cubic * std::exp(-(t.decay * q)));
}

// |a - b| as SQRT(SUM((a - b)**2)): sqrt((dx^2 + dy^2) + dz^2)
double distance(const crp::Vec3& a, const crp::Vec3& b) {
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return std::sqrt((dx * dx + dy * dy) + dz * dz);
}

using Pair = std::array<crp::Vec3, 2>;

// The two H atoms at the centre (x, y, Z) with the unclamped r, offsets ((r/2 sin t) cos p, (r/2 sin t) sin p,
// r/2 cos t) (ConvertCoMToXYZ)
Pair hydrogens(double x, double y, double z, double r, double theta, double phi) {
  const double st = std::sin(theta), ct = std::cos(theta), sp = std::sin(phi), cp = std::cos(phi);
  const double h = 0.5 * r;
  const crp::Vec3 o{(h * st) * cp, (h * st) * sp, h * ct};
  return {crp::Vec3{x - o.x, y - o.y, z - o.z}, crp::Vec3{x + o.x, y + o.y, z + o.z}};
}

// The H atoms in the ideal (CRP) lattice's frame: the centre's x and y divided by s, the bond kept
// (SCM_SurfInteract_3Body with EnableExpansion)
Pair ideal_frame(const Pair& h, double s) {
  const crp::Vec3 bond{h[1].x - h[0].x, h[1].y - h[0].y, h[1].z - h[0].z};
  const crp::Vec3 c{((h[0].x + h[1].x) / 2.0) / s, ((h[0].y + h[1].y) / 2.0) / s, (h[0].z + h[1].z) / 2.0};
  return {crp::Vec3{c.x - 0.5 * bond.x, c.y - 0.5 * bond.y, c.z - 0.5 * bond.z},
          crp::Vec3{c.x + 0.5 * bond.x, c.y + 0.5 * bond.y, c.z + 0.5 * bond.z}};
}

}  // namespace

PairTerms pair_terms(const Layer& layer, double r) {
  double t = r;
  if (r < layer.r_clamp_min) {
    t = layer.r_clamp_min;
  } else if (r > layer.r_clamp_max) {
    t = layer.r_clamp_max;
  }
  std::array<double, 8> p{};
  for (int k = 0; k < 8; ++k) p[k] = layer.slope[k] * t + layer.intercept[k];
  PairTerms terms{p[0], p[1], {p[2], p[3], p[4], p[5]}, p[6], p[7], 0.0};
  terms.switched_value = unswitched(terms, terms.switch_end - terms.reference_distance);
  return terms;
}

double pair_energy(const PairTerms& t, double d) {
  if (d > t.switch_end) return t.switched_value;
  const double f = unswitched(t, d - t.reference_distance);
  if (!(d > t.switch_start)) return f;
  const double phase = ((d - t.switch_end) * crp_tables::kPi) / (t.switch_end - t.switch_start);
  const double rho = 0.5 * std::cos(phase) + 0.5;
  return (1.0 - rho) * f + rho * t.switched_value;
}

Parameters read_parameters(const std::string& file) {
  if (!std::filesystem::is_regular_file(file)) fail(file, "no such file (SCMParFilepath)");
  crp::ListReader in(file);
  Parameters p;
  const std::optional<std::string> first = in.read_line();
  if (!first || first->compare(0, 7, "VERSION") != 0) fail(file, "line 1 must be 'VERSION 2'");
  const std::string rest = first->substr(7);
  const std::size_t b = rest.find_first_not_of(" \t");
  if (b == std::string::npos) fail(file, "line 1: 'VERSION' without a number");
  p.version = in.integer(rest.substr(b, rest.find_first_of(" \t,/", b) - b));
  if (p.version != 2) fail(file, "only parameter version 2 is supported");
  int number = 1;
  for (Layer& layer : p.layers) {
    const auto v = in.read(18);
    for (int k = 0; k < 8; ++k) {
      layer.intercept[k] = in.real(v[k]);
      layer.slope[k] = in.real(v[8 + k]);
    }
    layer.r_clamp_min = in.real(v[16]);
    layer.r_clamp_max = in.real(v[17]);
    if (!(layer.r_clamp_min < layer.r_clamp_max))
      fail(file, "layer " + std::to_string(number) + ": P17 must be smaller than P18");
    ++number;
  }
  return p;
}

Slab read_slab(const std::string& file, int configuration, double cutoff) {
  if (!(cutoff > 0.0)) throw crp::DataError("SCMCutoff: the cut-off must be positive");
  if (configuration < 1) throw crp::DataError("PosFileSlabNum: the configuration number must be at least 1");
  if (!std::filesystem::is_regular_file(file)) fail(file, "no such file (PosFilepath)");
  Slab s;
  s.configuration = configuration;
  s.cutoff = cutoff;
  {
    crp::ListReader in(file);
    const auto names = in.read(2);
    s.tag = names[0];
    s.structure = names[1];
    const auto sizes = in.read(6);
    s.nx = in.integer(sizes[0]);
    s.ny = in.integer(sizes[1]);
    s.nz = in.integer(sizes[2]);
    s.configurations = in.integer(sizes[3]);
    s.version = in.integer(sizes[4]);
    s.max_surface_atoms = in.integer(sizes[5]);
    s.lattice_constant = in.real(in.read(1)[0]);
  }
  if (s.structure != "FCC111") fail(file, "SlabSurfStruct '" + s.structure + "': only FCC111 is supported");
  if (configuration > s.configurations) fail(file, "configuration " + std::to_string(configuration) + " asked for");
  if (s.nz < 2) fail(file, "nz must be at least 2");
  if (s.version != 2 && s.version != 3) fail(file, "PosFileVersion " + std::to_string(s.version) + " (3 and 2 only)");
  if (s.version == 3 && (s.nx < 1 || s.ny < 1)) fail(file, "nx and ny must be at least 1");
  if (s.version == 2 && s.max_surface_atoms < 1) fail(file, "MaxNumSurfAtoms must be at least 1");
  const long per = s.version == 2 ? s.max_surface_atoms : static_cast<long>(s.nx) * s.ny * s.nz;
  crp::ListReader in(file);
  const long skip = 3 + (configuration - 1) * per;  // the header's three records, then the earlier snapshots
  for (long i = 0; i < skip; ++i)
    if (!in.read_line()) fail(file, "the file ends before configuration " + std::to_string(configuration));
  for (long i = 0; i < per; ++i) {
    const auto v = in.read(7);
    Atom a;
    a.distorted = {in.real(v[0]), in.real(v[1]), in.real(v[2])};
    a.ideal = {in.real(v[3]), in.real(v[4]), in.real(v[5])};
    a.layer = in.integer(v[6]);
    if (a.layer < 1) fail(file, "atom " + std::to_string(i + 1) + ": layers are numbered from 1");
    if (a.layer > 2) continue;
    if (s.version == 3) {  // strictly inside the cut-off sphere, compared squared in this order
      const double d2 = (a.distorted.x * a.distorted.x + a.distorted.y * a.distorted.y) + a.distorted.z * a.distorted.z;
      if (!(d2 < cutoff * cutoff)) continue;
    }
    s.atoms.push_back(a);
  }
  if (s.atoms.empty()) fail(file, "configuration " + std::to_string(configuration) + " has no atom of layers 1-2 in use");
  return s;
}

Correction::Correction(const Parameters& parameters, const Slab& slab, std::optional<double> rescaling,
                       double skewing_angle)
    : layers_(parameters.layers),
      atoms_(slab.atoms),
      rescaling_(rescaling),
      cos_skew_(std::cos(skewing_angle)),
      sin_skew_(std::sin(skewing_angle)) {
  if (rescaling_ && !(*rescaling_ > 0.0)) throw crp::DataError("SCM: the rescaling factor must be positive");
}

// GetSCMPotential (src/potential/scm/scm_correction.cpp:105-125): every atom in file order, H atom 1 then 2,
// dV = (dV - f(ideal)) + f(distorted) starting from +0.
double Correction::value_hartree(double x, double y, double z, double r, double theta, double phi) const {
  const Pair hd = hydrogens(x, y, z, r, theta, phi);
  const Pair hi = rescaling_ ? ideal_frame(hd, *rescaling_) : hd;
  const std::array<PairTerms, 2> terms = {pair_terms(layers_[0], r), pair_terms(layers_[1], r)};
  double dv = +0.0;
  for (const Atom& atom : atoms_) {
    const PairTerms& t = terms[atom.layer - 1];
    for (int a = 0; a < 2; ++a) {
      const double f_ideal = pair_energy(t, distance(hi[a], atom.ideal));
      const double f_distorted = pair_energy(t, distance(hd[a], atom.distorted));
      dv = (dv - f_ideal) + f_distorted;
    }
  }
  return dv;
}

namespace {
std::optional<double> rescaling_of(const crp::QdSurface& crp, const Slab& slab, bool rescale) {
  if (!rescale) return std::nullopt;
  if (!(slab.lattice_constant > 0.0)) throw crp::DataError("PosFilepath: the slab's lattice constant must be positive");
  return slab.lattice_constant / crp.surface().lattice_constant();
}
}  // namespace

ScmSurface::ScmSurface(std::shared_ptr<const crp::QdSurface> crp, const Parameters& parameters, const Slab& slab,
                       bool rescale)
    : crp_(std::move(crp)),
      slab_(slab),
      rescaling_(rescaling_of(*crp_, slab, rescale)),
      scale_(rescaling_.value_or(1.0)),
      correction_(parameters, slab, rescaling_, crp_->skewing_angle()) {
  if (crp_->surface().symmetry() == crp::Symmetry::c4v)
    throw crp::DataError("the SCM slab is an FCC111 surface: the CRP data set must be C3v or C6v");
}

double ScmSurface::lattice_constant() const {
  return rescaling_ ? slab_.lattice_constant : crp_->surface().lattice_constant();
}

// CRP at (u / s, v / s) in eV plus dV / Constants_eV2Ha (src/potential/scm/scm_crp_potential.cpp:70-76)
double ScmSurface::value_ev(double u, double v, double z, double r, double theta, double phi) const {
  const std::array<double, 2> xy = correction_.to_cartesian(u, v);
  const double v_crp = crp_->potential_ev(u / scale_, v / scale_, z, r, theta, phi);
  return v_crp + correction_.value_hartree(xy[0], xy[1], z, r, theta, phi) / 0.036749308286;
}

}  // namespace spo::scm
