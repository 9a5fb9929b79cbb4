#include "checkpoint.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <type_traits>
#include <utility>

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

// The input parameters a restart must keep, under the reference's names and in its order (format version 2,
// src/io/checkpoint.cpp:35-119); the SCM identity of its distorted-surface runs does not arise here.
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

}  // namespace spo
