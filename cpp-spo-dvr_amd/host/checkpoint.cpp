#include "checkpoint.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <type_traits>
#include <utility>

#include "scm.hpp"

namespace spo {

namespace {

using Magic = std::array<char, 16>;
constexpr Magic kCheckpointMagic = {'S', 'P', 'O', 'D', 'V', 'R', '-', 'C', 'H', 'E', 'C', 'K', 'P', 'N', 'T', '\0'};
constexpr Magic kBkMagic = {'S', 'P', 'O', 'D', 'V', 'R', '-', 'B', 'K', 'C', 'O', 'E', 'F', 'F', 'S', '\0'};
constexpr std::uint32_t kVersion = 2;
constexpr std::uint64_t kEndMarker = 0x454e444f464350ULL;
constexpr std::uint64_t kMaxParameters = 1000;

// Raw little-endian output; any failure is a CheckpointError naming the file.
class Out {
 public:
  explicit Out(const std::string& path) : path_(path), f_(path, std::ios::binary | std::ios::trunc) {
    if (!f_) throw CheckpointError("cannot open '" + path + "' for writing");
  }
  template <typename T>
  void put(const T& v) {
    static_assert(std::is_trivially_copyable_v<T>);
    bytes(&v, sizeof(T));
  }
  void put_string(const std::string& s) {
    put(static_cast<std::uint64_t>(s.size()));
    bytes(s.data(), s.size());
  }
  void put_array(const std::vector<cplx>& a) { bytes(a.data(), a.size() * sizeof(cplx)); }
  void put_bytes(const void* p, std::size_t n) { bytes(p, n); }
  void close() {
    f_.close();
    if (!f_) throw CheckpointError("error writing '" + path_ + "'");
  }

 private:
  void bytes(const void* p, std::size_t n) {
    f_.write(static_cast<const char*>(p), static_cast<std::streamsize>(n));
    if (!f_) throw CheckpointError("error writing '" + path_ + "'");
  }
  std::string path_;
  std::ofstream f_;
};

class In {
 public:
  explicit In(const std::string& path) : path_(path), f_(path, std::ios::binary) {
    if (!f_) throw CheckpointError("cannot open '" + path + "'");
  }
  template <typename T>
  T get() {
    static_assert(std::is_trivially_copyable_v<T>);
    T v;
    bytes(&v, sizeof(T));
    return v;
  }
  std::string get_string() {
    const auto n = get<std::uint64_t>();
    if (n > (1u << 20)) throw CheckpointError("corrupt file '" + path_ + "' (string length)");
    std::string s(n, '\0');
    bytes(s.data(), n);
    return s;
  }
  void get_array(std::vector<cplx>& a) { bytes(a.data(), a.size() * sizeof(cplx)); }
  void get_bytes(void* p, std::size_t n) { bytes(p, n); }
  bool at_end() { return f_.peek() == std::ifstream::traits_type::eof(); }
  const std::string& path() const { return path_; }

 private:
  void bytes(void* p, std::size_t n) {
    f_.read(static_cast<char*>(p), static_cast<std::streamsize>(n));
    if (static_cast<std::size_t>(f_.gcount()) != n) throw CheckpointError("'" + path_ + "' ends too early");
  }
  std::string path_;
  std::ifstream f_;
};

using Parameters = std::vector<std::pair<std::string, double>>;

// The numeric identity of a run on a distorted surface (SystemID CRPDistrtSurface; src/io/checkpoint.cpp:87-106,
// statistics as src/potential/scm/scm_diagnostics.cpp:286-311): the snapshot, the cut-off and the rescaling, the
// slab's lattice constant a_T, the number of slab atoms the correction uses, their rms displacement, the mean
// displacement (x, y, z) of the top layer, the highest distorted z, and the 2 x 18 values of the parameter file. A
// restart on another slab, cut-off or parameter set is then refused, and the list is the reference program's.
void add_scm_identity(Parameters& list, const InputParameters& p) {
  const scm::Slab slab = scm::read_slab(p.scm.position_file, p.scm.configuration, p.scm.cutoff);
  const scm::Parameters parameters = scm::read_parameters(p.scm.parameter_file);
  double squares = 0.0, highest = 0.0;
  double shift[3] = {0.0, 0.0, 0.0};
  int top = 0;
  for (std::size_t k = 0; k < slab.atoms.size(); ++k) {
    const scm::Atom& a = slab.atoms[k];
    const double dx = a.distorted.x - a.ideal.x, dy = a.distorted.y - a.ideal.y, dz = a.distorted.z - a.ideal.z;
    const double d = std::sqrt((dx * dx + dy * dy) + dz * dz);
    squares = squares + d * d;
    if (k == 0 || std::isnan(a.distorted.z) || (!std::isnan(highest) && a.distorted.z > highest)) highest = a.distorted.z;
    if (a.layer != 1) continue;
    ++top;
    shift[0] = shift[0] + dx;
    shift[1] = shift[1] + dy;
    shift[2] = shift[2] + dz;
  }
  const int n = static_cast<int>(slab.atoms.size());
  const double rms = n > 0 ? std::sqrt(squares / n) : 0.0;
  for (double& v : shift) v = top > 0 ? v / top : 0.0;
  list.emplace_back("PosFileSlabNum", double(p.scm.configuration));
  list.emplace_back("SCMCutoff", p.scm.cutoff);
  list.emplace_back("EnableSurfExpan", p.scm.rescale ? 1.0 : 0.0);
  list.emplace_back("PosFilepath (a_T)", slab.lattice_constant);
  list.emplace_back("PosFilepath (atoms used)", double(n));
  list.emplace_back("PosFilepath (rms displacement)", rms);
  list.emplace_back("PosFilepath (layer 1 mean shift x)", shift[0]);
  list.emplace_back("PosFilepath (layer 1 mean shift y)", shift[1]);
  list.emplace_back("PosFilepath (layer 1 mean shift z)", shift[2]);
  list.emplace_back("PosFilepath (highest atom)", highest);
  for (int layer = 0; layer < 2; ++layer) {
    const scm::Layer& l = parameters.layers[layer];
    for (int k = 0; k < 18; ++k) {
      const double v = k < 8 ? l.intercept[k] : k < 16 ? l.slope[k - 8] : k == 16 ? l.r_clamp_min : l.r_clamp_max;
      list.emplace_back("SCMParFilepath (layer " + std::to_string(layer + 1) + ", P" + std::to_string(k + 1) + ")", v);
    }
  }
}

// The input parameters a restart must keep, under the reference's names and in its order (format version 2,
// src/io/checkpoint.cpp:35-119).
Parameters restart_parameters(const Config& c, int n_phi) {
  const InputParameters& p = c.input;
  const DerivedParameters& d = c.derived;
  Parameters list = {{"ZStart", p.z_start},
                     {"dZ", p.dz},
                     {"nZ", double(p.nz)},
                     {"nZSpecular", double(p.nz_specular)},
                     {"RStart", p.r_start},
                     {"dR", p.dr},
                     {"nR", double(p.nr)},
                     {"nX", double(p.nx)},
                     {"nY", double(p.ny)},
                     {"JMax", double(p.j_max)},
                     {"MJMax", double(d.mj_max_basis)},
                     {"number of phi points (MJMax, nPhisOptimal)", double(n_phi)},
                     {"Mass1", p.mass1},
                     {"Mass2", p.mass2},
                     {"HomoNuc", p.homonuclear ? 1.0 : 0.0},
                     {"dtStep", p.time_step},
                     {"dtAna", p.dt_analysis},
                     {"Zana", p.z_analysis},
                     {"Z0", p.z0},
                     {"EMinWP", p.e_min_wp},
                     {"EMaxWP", p.e_max_wp},
                     {"EParallel", p.e_parallel},
                     {"PhiParallel", p.phi_parallel},
                     {"VIn", double(p.v_in)},
                     {"JIn", double(p.j_in)},
                     {"MJIn", double(p.mj_in)},
                     {"ZOptStart", p.z_opt_start},
                     {"ZOptEnd", p.z_opt_end},
                     {"ROptStart", p.r_opt_start},
                     {"ROptEnd", p.r_opt_end},
                     {"ZSpOptStart", p.zsp_opt_start},
                     {"ZSpOptEnd", p.zsp_opt_end},
                     {"ZOptFactor (effective)", d.z_opt_factor},
                     {"ROptFactor (effective)", d.r_opt_factor},
                     {"ZSpOptFactor (effective)", d.zsp_opt_factor}};
  // Low-Z absorbers only when on, so that files without them keep the shorter list.
  auto low = [&](const std::string& prefix, const LowAbsorber& a, double strength) {
    if (!a.enabled()) return;
    list.emplace_back(prefix + "Start", *a.start);
    list.emplace_back(prefix + "End", *a.end);
    list.emplace_back(prefix + "Factor (effective)", strength);
  };
  low("ZOptLow", p.z_opt_low, d.z_opt_low_factor);
  low("ZSpOptLow", p.zsp_opt_low, d.zsp_opt_low_factor);
  if (p.system_id == "CRPDistrtSurface") add_scm_identity(list, p);
  return list;
}

bool same_ignoring_case(const std::string& a, const std::string& b) {
  return a.size() == b.size() && std::equal(a.begin(), a.end(), b.begin(), [](char x, char y) {
           return std::tolower(static_cast<unsigned char>(x)) == std::tolower(static_cast<unsigned char>(y));
         });
}

std::string exact(double v) {
  std::ostringstream s;
  s.precision(17);
  s << v;
  return s.str();
}

void put_header(Out& f, const Magic& magic, const Config& c, int n_phi) {
  f.put(magic);
  f.put(kVersion);
  f.put_string(c.input.system_id);
  const Parameters list = restart_parameters(c, n_phi);
  f.put(static_cast<std::uint64_t>(list.size()));
  for (const auto& [name, value] : list) {
    f.put_string(name);
    f.put(value);
  }
}

void check_header(In& f, const Magic& magic, const std::string& kind, const Config& c, int n_phi) {
  const std::string where = " in " + kind + " '" + f.path() + "'";
  if (f.get<Magic>() != magic) throw CheckpointError("not a " + kind + ": '" + f.path() + "'");
  if (const auto v = f.get<std::uint32_t>(); v != kVersion)
    throw CheckpointError("unsupported format version " + std::to_string(v) + where);
  if (const std::string id = f.get_string(); !same_ignoring_case(id, c.input.system_id))
    throw CheckpointError("SystemID is '" + c.input.system_id + "' in the input but '" + id + "'" + where);
  const auto count = f.get<std::uint64_t>();
  if (count > kMaxParameters) throw CheckpointError("unexpected parameter list" + where);
  Parameters stored;
  for (std::uint64_t k = 0; k < count; ++k) {
    std::string name = f.get_string();
    const double value = f.get<double>();
    stored.emplace_back(std::move(name), value);
  }
  const Parameters expected = restart_parameters(c, n_phi);
  const bool same_names = expected.size() == stored.size() &&
                          std::equal(expected.begin(), expected.end(), stored.begin(),
                                     [](const auto& a, const auto& b) { return a.first == b.first; });
  if (!same_names) {
    for (const char* prefix : {"ZOptLow", "ZSpOptLow"}) {
      auto has = [&](const Parameters& l) {
        return std::any_of(l.begin(), l.end(), [&](const auto& e) { return e.first == std::string(prefix) + "Start"; });
      };
      if (has(expected) != has(stored))
        throw CheckpointError(std::string("the low-Z absorber ") + prefix + " is " + (has(expected) ? "on" : "off") +
                              " in the input but " + (has(stored) ? "on" : "off") + where);
    }
    throw CheckpointError("unexpected parameter list" + where + "; a restart must use the same parameters");
  }
  for (std::size_t k = 0; k < expected.size(); ++k)
    if (!(stored[k].second == expected[k].second))  // NaN never matches
      throw CheckpointError(expected[k].first + " is " + exact(expected[k].second) + " in the input but " +
                            exact(stored[k].second) + where + "; a restart must use the same parameters");
}

void check_count(In& f, const std::string& what, int expected) {
  if (const auto v = f.get<std::int32_t>(); v != expected)
    throw CheckpointError(what + " is " + std::to_string(expected) + " but " + std::to_string(v) + " in '" + f.path() +
                          "'");
}

}  // namespace

void write_checkpoint(const std::string& path, const Config& c, int n_phi, double time,
                      const std::vector<cplx>& scattering, int n_rot, int n_xy, int n_zr,
                      const std::vector<cplx>& specular, const AnalysisSurfaceState& surface) {
  const std::string temporary = path + ".tmp";
  {
    Out f(temporary);
    put_header(f, kCheckpointMagic, c, n_phi);
    f.put(time);
    for (const int n : {n_rot, n_xy, n_zr, c.input.nr, c.input.nz_specular}) f.put(static_cast<std::int32_t>(n));
    f.put(static_cast<std::int32_t>(surface.index));
    f.put(static_cast<std::int32_t>(surface.previous_index));
    f.put(static_cast<std::uint8_t>(surface.fixed ? 1 : 0));
    f.put(surface.norm_tolerance);
    f.put_array(specular);
    f.put_array(scattering);
    f.put(kEndMarker);
    f.close();
  }
  std::error_code ec;
  std::filesystem::rename(temporary, path, ec);
  if (ec) throw CheckpointError("cannot rename '" + temporary + "' to '" + path + "': " + ec.message());
}

CheckpointInfo read_checkpoint(const std::string& path, const Config& c, int n_phi, std::vector<cplx>& scattering,
                               int n_rot, int n_xy, int n_zr, std::vector<cplx>& specular) {
  In f(path);
  check_header(f, kCheckpointMagic, "checkpoint", c, n_phi);
  CheckpointInfo info;
  info.time = f.get<double>();
  check_count(f, "the number of rotational states", n_rot);
  check_count(f, "the number of (x, y) points", n_xy);
  check_count(f, "the number of (Z, r) points", n_zr);
  check_count(f, "nR", c.input.nr);
  check_count(f, "nZSpecular", c.input.nz_specular);
  info.surface.index = f.get<std::int32_t>();
  info.surface.previous_index = f.get<std::int32_t>();
  info.surface.fixed = f.get<std::uint8_t>() != 0;
  info.surface.norm_tolerance = f.get<double>();
  for (const int index : {info.surface.index, info.surface.previous_index})
    if (index < 0 || index >= c.input.nz_specular)
      throw CheckpointError("checkpoint '" + path + "' is corrupt (analysis surface index " + std::to_string(index) +
                            " is outside the specular grid, 0 .. " + std::to_string(c.input.nz_specular - 1) + ")");
  specular.assign(std::size_t(c.input.nr) * c.input.nz_specular, cplx(0.0, 0.0));
  scattering.assign(std::size_t(n_rot) * n_xy * n_zr, cplx(0.0, 0.0));
  f.get_array(specular);
  f.get_array(scattering);
  if (f.get<std::uint64_t>() != kEndMarker || !f.at_end())
    throw CheckpointError("checkpoint '" + path + "' is corrupt (bad end marker)");
  return info;
}

void write_bk_file(const std::string& path, const Config& c, int n_phi, const AnalysisBasis& basis,
                   const std::vector<std::pair<double, std::vector<cplx>>>& records) {
  Out f(path);
  put_header(f, kBkMagic, c, n_phi);
  const int n_states = basis.n_states();
  f.put(static_cast<std::int32_t>(n_states));
  f.put(static_cast<std::int32_t>(basis.n_v));
  for (int is = 0; is < n_states; ++is) {
    const DiffractionChannel& ch = basis.channels[basis.state_diffraction(is)];
    for (const int label : {basis.state_j[is], basis.state_mj[is], ch.n, ch.m, ch.order})
      f.put(static_cast<std::int32_t>(label));
  }
  for (int v = 0; v < basis.n_v; ++v)
    for (int is = 0; is < n_states; ++is) f.put(basis.channel_energies[is + std::size_t(n_states) * v]);
  for (const auto& [time, bk] : records) {
    if (bk.size() != std::size_t(n_states) * basis.n_v) throw CheckpointError("BK record of the wrong size");
    f.put(time);
    f.put_array(bk);
  }
  f.close();
}

std::optional<double> read_bk_file(const std::string& path, const Config& c, int n_phi, const AnalysisBasis& basis,
                                   const std::function<void(double, const std::vector<cplx>&)>& on_record) {
  In f(path);
  check_header(f, kBkMagic, "BK coefficient file", c, n_phi);
  const int n_states = basis.n_states();
  check_count(f, "the number of analysed states (nstata)", n_states);
  check_count(f, "the number of analysed vibrational levels (nviba)", basis.n_v);
  for (int is = 0; is < n_states; ++is) {
    const DiffractionChannel& ch = basis.channels[basis.state_diffraction(is)];
    for (const int label : {basis.state_j[is], basis.state_mj[is], ch.n, ch.m, ch.order})
      check_count(f, "a state label of the analysis basis", label);
  }
  for (int k = 0; k < n_states * basis.n_v; ++k) (void)f.get<double>();
  std::optional<double> last;
  std::vector<cplx> bk(std::size_t(n_states) * basis.n_v);
  while (!f.at_end()) {
    const double time = f.get<double>();
    f.get_array(bk);
    on_record(time, bk);
    last = time;
  }
  return last;
}

namespace {

// gfortran sequential unformatted records: int32 length, the bytes, int32 length. Records above the largest
// subrecord (2^31 - 9 bytes) would be split by gfortran; none of these files has one.
class FortranOut {
 public:
  explicit FortranOut(const std::string& path) : out_(path) {}
  void record(const void* p, std::size_t n) {
    if (n > 2147483639u) throw CheckpointError("a Fortran record of " + std::to_string(n) + " bytes needs subrecords");
    const auto len = static_cast<std::int32_t>(n);
    out_.put(len);
    out_.put_bytes(p, n);
    out_.put(len);
  }
  void record(const std::vector<unsigned char>& b) { record(b.data(), b.size()); }
  void close() { out_.close(); }

 private:
  Out out_;
};

class FortranIn {
 public:
  explicit FortranIn(const std::string& path) : in_(path) {}
  // The bytes of the next record (what names it in the error message).
  std::vector<unsigned char> record(const std::string& what) {
    if (in_.at_end()) throw CheckpointError("'" + in_.path() + "' ends before " + what);
    const auto n = in_.get<std::int32_t>();
    if (n < 0) throw CheckpointError("'" + in_.path() + "': " + what + " is split into subrecords");
    std::vector<unsigned char> b(static_cast<std::size_t>(n));
    in_.get_bytes(b.data(), b.size());
    if (in_.get<std::int32_t>() != n) throw CheckpointError("'" + in_.path() + "': bad record marker after " + what);
    return b;
  }
  const std::string& path() const { return in_.path(); }

 private:
  In in_;
};

struct Bytes {
  std::vector<unsigned char> b;
  template <typename T>
  Bytes& add(const T& v) {
    const auto* p = reinterpret_cast<const unsigned char*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
    return *this;
  }
  template <typename T>
  Bytes& add_all(const std::vector<T>& v) {
    for (const T& x : v) add(x);
    return *this;
  }
};

// Values of one record in order.
struct Fields {
  const std::vector<unsigned char>& b;
  std::string what, path;
  std::size_t at = 0;
  template <typename T>
  T next() {
    if (at + sizeof(T) > b.size()) throw CheckpointError("'" + path + "': " + what + " is too short");
    T v;
    std::memcpy(&v, b.data() + at, sizeof(T));
    at += sizeof(T);
    return v;
  }
  void end() const {
    if (at != b.size()) throw CheckpointError("'" + path + "': " + what + " is longer than expected");
  }
};

std::array<std::vector<std::int32_t>, 6> state_labels(const AnalysisBasis& basis) {
  std::array<std::vector<std::int32_t>, 6> labels;
  for (int is = 0; is < basis.n_states(); ++is) {
    const DiffractionChannel& ch = basis.channels[basis.state_diffraction(is)];
    labels[0].push_back(basis.state_rotation(is) + 1);
    labels[1].push_back(basis.state_diffraction(is) + 1);
    labels[2].push_back(ch.n);
    labels[3].push_back(ch.m);
    labels[4].push_back(basis.state_j[is]);
    labels[5].push_back(basis.state_mj[is]);
  }
  return labels;
}

double analysis_step(const Config& c) { return c.input.time_step * c.derived.steps_per_analysis; }

void expect(std::int32_t stored, int expected, const std::string& what, const std::string& path) {
  if (stored != expected)
    throw CheckpointError(what + " is " + std::to_string(expected) + " but " + std::to_string(stored) + " in '" + path +
                          "'");
}

}  // namespace

void write_fortran_wave_function(const std::string& path, const std::vector<cplx>& scattering, int n_rot, int n_xy,
                                 int n_zr, const std::vector<cplx>& specular, const AnalysisSurfaceState& surface) {
  FortranOut f(path);
  f.record(Bytes().add<std::int32_t>(2).b);
  f.record(specular.data(), specular.size() * sizeof(cplx));
  f.record(Bytes().add<std::int32_t>(surface.fixed ? 1 : 0).add<std::int32_t>(surface.index + 1).b);
  const std::size_t block = std::size_t(n_rot) * n_xy;
  for (int izr = 0; izr < n_zr; ++izr) f.record(scattering.data() + block * izr, block * sizeof(cplx));
  f.close();
}

AnalysisSurfaceState read_fortran_wave_function(const std::string& path, std::vector<cplx>& scattering, int n_rot,
                                                int n_xy, int n_zr, std::vector<cplx>& specular, int nr, int n_z_specular,
                                                const AnalysisSurfaceState& fresh) {
  FortranIn f(path);
  {
    const auto b = f.record("record 1 (nproj)");
    Fields r{b, "record 1 (nproj)", path};
    expect(r.next<std::int32_t>(), 2, "ProjOp", path);
    r.end();
  }
  specular.assign(std::size_t(nr) * n_z_specular, cplx(0.0, 0.0));
  {
    const auto b = f.record("record 2 (SpecWave)");
    if (b.size() != specular.size() * sizeof(cplx))
      throw CheckpointError("'" + path + "': the specular wave (record 2) does not have nR x nZSpecular values");
    std::memcpy(specular.data(), b.data(), b.size());
  }
  AnalysisSurfaceState state = fresh;
  {
    const auto b = f.record("record 3 (FixedZAnalysisSpec, izanlySpecChan)");
    Fields r{b, "record 3 (FixedZAnalysisSpec, izanlySpecChan)", path};
    state.fixed = r.next<std::int32_t>() != 0;
    const std::int32_t iz = r.next<std::int32_t>();
    r.end();
    if (iz < 1 || iz > n_z_specular)
      throw CheckpointError("'" + path + "': izanlySpecChan " + std::to_string(iz) +
                            " is outside the specular grid, 1 .. " + std::to_string(n_z_specular));
    state.index = iz - 1;
    state.previous_index = state.index;
  }
  const std::size_t block = std::size_t(n_rot) * n_xy;
  scattering.assign(block * n_zr, cplx(0.0, 0.0));
  for (int izr = 0; izr < n_zr; ++izr) {
    const std::string what = "ScatWave record " + std::to_string(izr + 1);
    const auto b = f.record(what);
    if (b.size() != block * sizeof(cplx))
      throw CheckpointError("'" + path + "': " + what + " does not hold n_rot x nX x nY values");
    std::memcpy(scattering.data() + block * izr, b.data(), b.size());
  }
  return state;
}

void write_fortran_bk_file(const std::string& path, const Config& c, const AnalysisBasis& basis,
                           const FortranBkHeader& h, const std::vector<std::vector<cplx>>& records) {
  const InputParameters& p = c.input;
  const DerivedParameters& d = c.derived;
  const int n_dif_total = static_cast<int>(basis.channels.size());
  FortranOut f(path);
  f.record(Bytes()
               .add<std::int32_t>(h.n_records)
               .add<std::int32_t>(n_dif_total * h.n_rotations)
               .add<std::int32_t>(n_dif_total)
               .add<std::int32_t>(h.n_rotations)
               .b);
  f.record(Bytes().add(analysis_step(c)).add(d.kz).add(p.z0).add(d.wave_packet_alpha).add(p.z_analysis)
               .add(h.initial_energy).b);
  f.record(Bytes()
               .add(d.translational_energy)
               .add(d.theta_incidence)
               .add(d.phi_incidence)
               .add<std::int32_t>(p.j_in)
               .add<std::int32_t>(p.mj_in)
               .add<std::int32_t>(h.initial_channel + 1)
               .b);
  for (const auto& labels : state_labels(basis)) f.record(Bytes().add_all(labels).b);
  std::vector<double> energies;
  for (int v = 0; v < basis.n_v; ++v)
    for (int is = 0; is < basis.n_states(); ++is)
      energies.push_back(basis.channel_energies[is + std::size_t(basis.n_states()) * v]);
  f.record(Bytes().add_all(energies).b);
  for (const auto& bk : records) f.record(bk.data(), bk.size() * sizeof(cplx));
  f.close();
}

int read_fortran_bk_file(const std::string& path, const Config& c, const AnalysisBasis& basis, int n_rotations,
                         const std::function<void(double, const std::vector<cplx>&)>& on_record) {
  FortranIn f(path);
  const int n_dif_total = static_cast<int>(basis.channels.size());
  int n_records = 0;
  {
    const auto b = f.record("record 1 (sizes)");
    Fields r{b, "record 1 (sizes)", path};
    n_records = r.next<std::int32_t>();  // the records the file holds: exactly these are read
    if (n_records < 0) throw CheckpointError("'" + path + "': negative record count in record 1");
    expect(r.next<std::int32_t>(), n_dif_total * n_rotations, "nstatw", path);
    expect(r.next<std::int32_t>(), n_dif_total, "ndftow", path);
    expect(r.next<std::int32_t>(), n_rotations, "nrotw", path);
    r.end();
  }
  const double deltat = analysis_step(c);
  {
    const auto b = f.record("record 2 (deltat, ...)");
    Fields r{b, "record 2 (deltat, ...)", path};
    if (r.next<double>() != deltat)
      throw CheckpointError("the analysis time step (dtAna) differs from the one in '" + path + "'");
  }
  (void)f.record("record 3");
  const auto expected = state_labels(basis);
  for (int k = 0; k < 6; ++k) {
    const std::string what = "record " + std::to_string(4 + k);
    const auto b = f.record(what);
    if (b.size() != expected[k].size() * sizeof(std::int32_t))
      throw CheckpointError("the analysis basis (nstata) differs from the one in '" + path + "'");
    Fields r{b, what, path};
    for (const std::int32_t label : expected[k])
      if (r.next<std::int32_t>() != label)
        throw CheckpointError("the analysis basis differs from the one in '" + path + "'");
  }
  (void)f.record("record 10 (channel energies)");
  const std::size_t n = std::size_t(basis.n_states()) * basis.n_v;
  std::vector<cplx> bk(n);
  // Exactly the records record 1 announces (src/io/fortran_files.cpp:177-183): a file cut short by a killed run
  // is refused.
  for (int k = 1; k <= n_records; ++k) {
    const std::vector<unsigned char> b = f.record("BK record " + std::to_string(k) + " of " + std::to_string(n_records));
    if (b.size() != n * sizeof(cplx))
      throw CheckpointError("'" + path + "': BK record " + std::to_string(k) + " has the wrong length");
    std::memcpy(bk.data(), b.data(), b.size());
    on_record(k * deltat, bk);
  }
  return n_records;
}

}  // namespace spo
