#include "grids.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>

namespace spo {

std::vector<double> fft_kinetic_energies(int n, double spacing, double mass) {
  const double dk = 2.0 * kPi / (n * spacing);
  std::vector<double> e(n, 0.0);
  const int half = n / 2;
  e[half] = 0.5 / mass * std::pow(half * dk, 2);
  for (int i = 1; i < half; ++i) {
    e[i] = 0.5 / mass * std::pow(i * dk, 2);
    e[n - i] = e[i];
  }
  return e;
}

namespace {

Range absorber_points(const UniformGrid& g, double start, double end) {
  int first = 0;
  while (first < g.size() && tol::lt(g[first], start)) ++first;
  int last = g.size() - 1;
  while (last >= 0 && tol::gt(g[last], end)) --last;
  return {first, last};
}

}  // namespace

SpatialGrids make_spatial_grids(const Config& c) {
  const InputParameters& p = c.input;
  const DerivedParameters& d = c.derived;
  SpatialGrids g;
  g.z = UniformGrid(p.z_start, p.dz, p.nz);
  g.z_specular = UniformGrid(p.z_start, p.dz, p.nz_specular);
  g.r = UniformGrid(p.r_start, p.dr, p.nr);
  g.x = UniformGrid(0.0, d.dx, p.nx);
  g.y = UniformGrid(0.0, d.dy, p.ny);
  g.ekin_z = fft_kinetic_energies(p.nz, p.dz, d.total_mass);
  g.ekin_z_specular = fft_kinetic_energies(p.nz_specular, p.dz, d.total_mass);
  g.ekin_r = fft_kinetic_energies(p.nr, p.dr, d.reduced_mass);

  int iz = 0;
  while (iz < g.z.size() && !tol::ge(g.z[iz], p.z_analysis)) ++iz;
  if (iz == g.z.size()) throw InputError("Zana", "analysis value ZAnalysis lies outside grid");
  g.iz_analysis = iz;
  g.z_analysis = g.z[iz];
  if (tol::gt(g.z_analysis, p.z_opt_start)) throw InputError("Zana", "ZAnalysis lies inside the optical potential");

  iz = 0;
  while (iz < g.z.size() && !tol::ge(g.z[iz], d.z_asymptotic)) ++iz;
  if (iz == g.z.size()) throw InputError("nZ", "the scattering grid in Z does not reach ZAsymptotic");
  g.nz_potential = iz;
  g.z_optical = absorber_points(g.z, p.z_opt_start, p.z_opt_end);
  g.z_specular_optical = absorber_points(g.z_specular, p.zsp_opt_start, p.zsp_opt_end);
  g.r_optical = absorber_points(g.r, p.r_opt_start, p.r_opt_end);
  return g;
}

int nearest_fft_size(int n, int multiple) {
  constexpr std::int64_t limit = std::int64_t{1} << 30;
  std::int64_t base = multiple;
  if (base % 2 == 0) base /= 2;
  std::int64_t best = limit;
  for (int h = 1; h <= 25; ++h)
    for (int i = 0; i <= 2; ++i)
      for (int j = 0; j <= 1; ++j)
        for (int k = 0; k <= 1; ++k)
          for (int m = 0; m <= 1; ++m) {
            std::int64_t size = base << h;
            for (int t = 0; t < i; ++t) size *= 3;
            if (j) size *= 5;
            if (k) size *= 7;
            if (m) size *= 11;
            if (size <= 0 || size >= limit) continue;
            if (size >= n && size <= best) best = size;
          }
  return static_cast<int>(best);
}

namespace {

std::vector<double> log_factorials(int n) {
  std::vector<double> lf(n + 1);
  lf[0] = 0.0;
  for (int i = 1; i <= n; ++i) lf[i] = std::log(i) + lf[i - 1];
  return lf;
}

}  // namespace

// Newton iteration on the Jacobi-type recursion of the order-th polynomial, initial guesses from the
// asymptotic root formulas; the symmetric half of the roots is mirrored.
void gauss_associated_legendre(double a, double b, int order, int mj, std::vector<double>& roots,
                               std::vector<double>& weights) {
  const int amj = std::abs(mj);
  const double eps = machine_epsilon();
  const std::vector<double> lf = log_factorials(2 * order + 2 * amj + 5);
  const double ord = order;
  roots.assign(order, 0.0);
  weights.assign(order, 0.0);
  const int m = (order + 1) / 2;
  const double sb = 0.5 * (b + a);
  const double sa = 0.5 * (b - a);
  const double ratio = amj / ord;
  const double fw = amj != 0 ? std::pow(2.0, 2 * amj + 1) * (order + amj) * sa *
                                   std::exp(2.0 * lf[order + amj - 1] - lf[order] - lf[order + 2 * amj])
                             : 2.0 * sa / ord;
  double z = 0.0;
  for (int i = 1; i <= m; ++i) {
    const double di = i;
    if (i == (m - order / 2) * m) {
      z = 0.0;
    } else if (amj == 0) {
      z = std::cos(kPi * (di - 0.25) / (ord + 0.5));
    } else if (i == 1) {
      const double dm = amj;
      z = 1.0 - ((1.0 + dm) * (2.78 / (4.0 + ord * ord) + 0.768 * ratio / ord)) / (1.0 + ratio * (2.44 + 1.282 * ratio));
    } else if (i == 2) {
      const double dm = amj;
      z = z - (1.0 - z) * (1.0 + 0.012 * dm * (1.0 + 0.25 * dm) / ord) * (1.0 + 0.06 * (order - 8) * (1.0 + 0.12 * dm) / ord) *
                  (4.1 + dm) / ((1.0 + dm) * (1.0 + 0.156 * dm));
    } else {
      z = (roots[i - 3] - 2.0 * roots[i - 2] + sb) / sa;
    }
    double p1 = 0.0, p2 = 0.0, dp = 0.0;
    for (int it = 1;; ++it) {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= order; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = (j + amj) * ((2 * (j + amj) - 1) * z * p2 - (j + amj - 1) * p3) / (j * (j + 2 * amj));
      }
      dp = (ord * z * p1 - (order + amj) * p2) / (z * z - 1.0);
      const double previous = z;
      z = previous - p1 / dp;
      if (std::abs(z - previous) <= eps || it >= 500) break;
    }
    roots[i - 1] = sb - sa * z;
    weights[i - 1] = fw / (dp * p2);
    roots[order - i] = sb + sa * z;
    weights[order - i] = weights[i - 1];
  }
}

std::vector<double> associated_legendre(int j_max, int m, double theta) {
  const int am = std::abs(m);
  const std::vector<double> lf = log_factorials(j_max + am + 1);
  std::vector<double> p(j_max - am + 1);
  double p1 = 1.0;
  if (am > 0) p1 = std::exp(lf[2 * am] / 2 - lf[am]) * ipow(std::sin(theta) / 2, am);
  if (m > 0 && am % 2 == 1) p1 = -p1;  // Condon-Shortley phase
  p[0] = std::sqrt(am + 0.5) * p1;
  const double x = std::cos(theta);
  double p2 = 0.0;
  for (int l = am; l < j_max; ++l) {
    const double p3 = p2;
    p2 = p1;
    p1 = ((2 * l + 1) * x * p2 - (l + am) * p3) / (l - am + 1);
    p[l - am + 1] = std::sqrt(l + 1.5) * p1;
  }
  if (am == 0) return p;
  for (int i = 1; i < static_cast<int>(p.size()); ++i) p[i] = std::exp((lf[i] + lf[2 * am] - lf[2 * am + i]) / 2) * p[i];
  return p;
}

AngularGrid make_angular_grid(const Config& c) {
  const InputParameters& p = c.input;
  AngularGrid a;
  a.j_max = p.j_max;
  a.j_min = 0;
  a.delta_j = 1;
  a.mj_max = c.derived.mj_max_basis;
  a.homonuclear = p.homonuclear;
  if (p.j_in < 0 || p.j_max < 0) throw InputError("JMax", "JIn and JMax must not be negative");
  if (std::abs(p.mj_in) > p.j_in) throw InputError("MJIn", "|MJIn| > JIn");
  if (p.homonuclear && (p.j_in % 2) != (p.j_max % 2))
    throw InputError("JMax", "JMax and JIn must have the same parity for a homonuclear molecule");
  if (a.mj_max > p.j_max) throw InputError("MJMax", "MJMax > JMax");
  if (p.j_max < p.j_in) throw InputError("JIn", "JIn > JMax");

  const int n_phi = nearest_fft_size(2 * a.mj_max + 1, c.derived.phi_points_multiple);
  a.d_phi = 2.0 * kPi / n_phi;
  if (n_phi < 2 * a.mj_max + 1) throw InputError("MJMax", "nphi < 2 MJMax + 1");
  int n_theta = 0;
  if (a.homonuclear) {
    if (p.j_in % 2 == 0) {
      if (a.j_max % 2 != 0) a.j_max -= 1;
      if (a.j_min % 2 != 0) a.j_min += 1;
    } else {
      if (a.j_max % 2 == 0) a.j_max -= 1;
      if (a.j_min % 2 == 0) a.j_min += 1;
    }
    n_theta = (a.j_max - a.j_min) / 2 + 1;
    a.delta_j = 2;
  } else {
    n_theta = a.j_max - a.j_min + 1;
  }
  a.phis.resize(n_phi);
  for (int i = 0; i < n_phi; ++i) a.phis[i] = i * a.d_phi;
  std::vector<double> roots, weights;
  gauss_associated_legendre(1.0, -1.0, a.j_max + 1, 0, roots, weights);
  a.thetas.resize(n_theta);
  a.theta_weights.assign(weights.begin(), weights.begin() + n_theta);
  double norm = -1.0;
  if (a.homonuclear) {
    norm = -2.0;
    if (a.j_max % 2 == 0) a.theta_weights.back() = 0.5 * a.theta_weights.back();  // theta = pi/2 counted once
  }
  for (int t = 0; t < n_theta; ++t) {
    a.theta_weights[t] = a.theta_weights[t] * norm;
    a.thetas[t] = std::acos(roots[t]);
  }
  return a;
}

std::vector<double> legendre_matrix(const AngularGrid& g, const RotationalBasis& basis) {
  std::vector<double> out;
  for (int m = -basis.mj_max; m <= basis.mj_max; ++m)
    for (int th = 0; th < g.n_theta(); ++th) {
      const std::vector<double> pl = associated_legendre(g.j_max, m, g.thetas[th]);
      const double w = std::sqrt(g.theta_weights[th] / g.n_phi());
      for (int j = g.j_min; j <= g.j_max; j += g.delta_j)
        if (std::abs(m) <= j) out.push_back(pl[j - std::abs(m)] * w);
    }
  return out;
}

RotationalBasis make_rotational_basis(const AngularGrid& g, int j_in, int mj_in) {
  RotationalBasis b;
  b.mj_max = std::min(g.mj_max, g.j_max);
  for (int m = -b.mj_max; m <= b.mj_max; ++m) {
    const int first = b.size();
    for (int j = g.j_min; j <= g.j_max; j += g.delta_j) {
      if (std::abs(m) > j) continue;
      if (j == j_in && m == mj_in) b.initial = b.size();
      b.j.push_back(j);
      b.mj.push_back(m);
    }
    b.blocks.push_back({first, b.size() - 1});
  }
  if (b.initial < 0) throw InputError("JIn", "the initial state (JIn, MJIn) is not in the rotational basis");
  return b;
}

}  // namespace spo
