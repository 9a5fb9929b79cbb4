#include "symmetry.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <optional>

namespace spo {

namespace {

int wrap(int a, int n) { return ((a % n) + n) % n; }

SurfaceOp compose(const SurfaceOp& a, const SurfaceOp& b) {  // first b, then a
  return {a.mirror != b.mirror, wrap(a.twelfths + (a.mirror ? -1 : 1) * b.twelfths, 12)};
}
SurfaceOp inverse(const SurfaceOp& g) { return g.mirror ? g : SurfaceOp{false, wrap(-g.twelfths, 12)}; }

struct Vec2 {
  double x, y;
};

Vec2 act(const SurfaceOp& g, Vec2 v) {
  const double beta = 2.0 * kPi * g.twelfths / 12.0;
  const double c = std::cos(beta), s = std::sin(beta);
  const double y = g.mirror ? -v.y : v.y;
  return {c * v.x - s * y, s * v.x + c * y};
}

struct IntMatrix {
  int a11, a12, a21, a22;
};

// The operation's matrix in lattice (fractional) coordinates, or nothing if it is not integral.
std::optional<IntMatrix> lattice_matrix(const SystemParameters& s, const SurfaceOp& g) {
  const double a = s.lattice_x, b = s.lattice_y, skew = s.skew;
  auto to_cart = [&](double x, double y) { return Vec2{x + y * std::cos(skew), y * std::sin(skew)}; };
  auto to_skewed = [&](Vec2 v) {
    const double y = v.y / std::sin(skew);
    return Vec2{v.x - y * std::cos(skew), y};
  };
  const Vec2 i1 = to_skewed(act(g, to_cart(a, 0.0)));
  const Vec2 i2 = to_skewed(act(g, to_cart(0.0, b)));
  const double e[4] = {i1.x / a, i2.x / a, i1.y / b, i2.y / b};
  int r[4];
  for (int k = 0; k < 4; ++k) {
    r[k] = static_cast<int>(std::lround(e[k]));
    if (std::abs(e[k] - r[k]) > 1e-10) return std::nullopt;
  }
  return IntMatrix{r[0], r[1], r[2], r[3]};
}

double reciprocal_length_squared(const SystemParameters& s, int n, int m) {
  const double sn = std::sin(s.skew);
  const double u = n / s.lattice_x, w = m / s.lattice_y;
  return (u * u + w * w - 2.0 * std::cos(s.skew) * u * w) / (sn * sn);
}

}  // namespace

std::vector<SurfaceOp> point_group_operations(const std::string& name) {
  std::vector<SurfaceOp> gens;
  if (name == "C2v") gens = {{false, 6}, {true, 0}};
  else if (name == "C4v") gens = {{false, 3}, {true, 0}};
  else if (name == "C6v") gens = {{false, 2}, {true, 0}};
  else if (name == "C3v") gens = {{false, 4}, {true, 6}};
  else if (name == "C3") gens = {{false, 4}};
  else if (name != "none") throw InputError("SystemID", "unknown point group " + name);
  std::vector<SurfaceOp> group = {{false, 0}};
  for (std::size_t i = 0; i < group.size(); ++i)
    for (const SurfaceOp& g : gens) {
      const SurfaceOp p = compose(g, group[i]);
      if (std::find(group.begin(), group.end(), p) == group.end()) group.push_back(p);
    }
  std::sort(group.begin(), group.end(), [](const SurfaceOp& a, const SurfaceOp& b) {
    if (a.mirror != b.mirror) return !a.mirror;
    return a.twelfths < b.twelfths;
  });
  return group;
}

int diffraction_order(const SystemParameters& s, int n, int m) {
  const double longest = std::max(reciprocal_length_squared(s, 1, 0), reciprocal_length_squared(s, 0, 1));
  const double limit = longest * (1.0 + 1e-12);
  const bool short_sum = reciprocal_length_squared(s, 1, 1) <= limit;
  const bool short_difference = reciprocal_length_squared(s, 1, -1) <= limit;
  if ((short_sum && n * m > 0) || (short_difference && n * m < 0)) return std::max(std::abs(n), std::abs(m));
  return std::abs(n) + std::abs(m);
}

namespace {

void build_orbits(GridSymmetry& s, const std::vector<IntMatrix>& mats) {
  const int n_ops = static_cast<int>(s.ops.size());
  const int nxy = s.nx * s.ny;
  s.representative_of.assign(nxy, -1);
  s.operation_of.assign(nxy, -1);
  std::vector<int> inv(n_ops, -1);
  for (int i = 0; i < n_ops; ++i)
    for (int j = 0; j < n_ops; ++j)
      if (s.ops[j] == inverse(s.ops[i])) inv[i] = j;
  for (int ix = 0; ix < s.nx; ++ix)
    for (int iy = 0; iy < s.ny; ++iy) {
      const int ixy = iy + s.ny * ix;
      if (s.representative_of[ixy] >= 0) continue;
      const int irep = s.n_representatives();
      s.representatives.push_back(ixy);
      for (int iop = 0; iop < n_ops; ++iop) {
        const IntMatrix& a = mats[iop];
        const int jx = wrap(a.a11 * ix + (a.a12 * s.nx / s.ny) * iy, s.nx);
        const int jy = wrap((a.a21 * s.ny / s.nx) * ix + a.a22 * iy, s.ny);
        const int image = jy + s.ny * jx;
        if (s.representative_of[image] >= 0) continue;
        s.representative_of[image] = irep;
        s.operation_of[image] = inv[iop];
      }
    }
  s.phi_index.clear();
  for (const SurfaceOp& g : s.ops) {
    const int product = g.twelfths * s.nphi;
    const int shift = product % 12 == 0 ? product / 12 : -1;
    std::vector<int> index(s.nphi);
    for (int i = 0; i < s.nphi; ++i) index[i] = wrap((g.mirror ? -i : i) + shift, s.nphi);
    s.phi_index.push_back(std::move(index));
  }
}

}  // namespace

GridSymmetry make_identity_symmetry(int nx, int ny, int nphi) {
  GridSymmetry s;
  s.nx = nx;
  s.ny = ny;
  s.nphi = nphi;
  s.ops = {{false, 0}};
  build_orbits(s, {{1, 0, 0, 1}});
  return s;
}

GridSymmetry make_grid_symmetry(const Config& c, int nphi) {
  const InputParameters& p = c.input;
  if (!p.use_symmetry) return make_identity_symmetry(p.nx, p.ny, nphi);
  GridSymmetry s;
  s.nx = p.nx;
  s.ny = p.ny;
  s.nphi = nphi;
  s.group = c.derived.system.point_group;
  std::vector<IntMatrix> mats;
  for (const SurfaceOp& g : point_group_operations(s.group)) {
    const auto a = lattice_matrix(c.derived.system, g);
    if (!a) throw InputError("UseSymmetry", "point group " + s.group + " does not fit the surface cell");
    const bool phi_misfit = (g.twelfths * nphi) % 12 != 0;
    const bool xy_misfit = (a->a12 * p.nx) % p.ny != 0 || (a->a21 * p.ny) % p.nx != 0;
    if (phi_misfit || xy_misfit) continue;
    s.ops.push_back(g);
    mats.push_back(*a);
  }
  if (s.ops.size() == 1) throw InputError("UseSymmetry", "no operation of point group " + s.group + " is usable");
  build_orbits(s, mats);
  return s;
}

}  // namespace spo
