#include "gas_phase.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace spo {

void symmetric_eigen(std::vector<double> a, int n, std::vector<double>& values, std::vector<double>& vectors) {
  std::vector<double> v(std::size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) v[i + std::size_t(n) * i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[i + std::size_t(n) * j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int j = 0; j < n; ++j) {
      diag += A(j, j) * A(j, j);
      for (int i = 0; i < j; ++i) off += A(i, j) * A(i, j);
    }
    if (off <= 1e-34 * diag) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A(p, q);
        if (apq == 0.0) continue;
        const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0);
        const double s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        A(p, q) = 0.0;
        A(q, p) = 0.0;
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k + std::size_t(n) * p], vkq = v[k + std::size_t(n) * q];
          v[k + std::size_t(n) * p] = c * vkp - s * vkq;
          v[k + std::size_t(n) * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return A(x, x) < A(y, y); });
  values.resize(n);
  vectors.resize(std::size_t(n) * n);
  for (int k = 0; k < n; ++k) {
    values[k] = A(order[k], order[k]);
    std::copy_n(v.begin() + std::size_t(n) * order[k], n, vectors.begin() + std::size_t(n) * k);
  }
}

std::vector<double> fgh_kinetic_matrix(int n, double dx, double mass) {
  const int n_half = n / 2 + 1;
  std::vector<double> ek(n_half);
  for (int l = 0; l < n_half; ++l) ek[l] = 2.0 / mass * std::pow(l * kPi / n / dx, 2.0);
  // T[j, l] = (1/n) sum_k E_k cos(2 pi k (l - j) / n) over the full (Hermitian-completed) spectrum.
  std::vector<double> row(n);
  for (int d = 0; d < n; ++d) {
    long double sum = ek[0];
    for (int k = 1; k < n_half; ++k) {
      const bool nyquist = (n % 2 == 0) && k == n / 2;
      const long double c = std::cos(2.0L * 3.14159265358979323846264338327950288L * k * d / n);
      sum += (nyquist ? 1.0L : 2.0L) * ek[k] * c;
    }
    row[d] = static_cast<double>(sum / n);
  }
  std::vector<double> t(std::size_t(n) * n);
  for (int j = 0; j < n; ++j)
    for (int l = 0; l < n; ++l) t[j + std::size_t(n) * l] = row[((l - j) % n + n) % n];
  return t;
}

GasPhase compute_gas_phase(const Config& c, const AngularGrid& angular, const std::vector<double>& vasy) {
  const InputParameters& p = c.input;
  const int nr = p.nr;
  const double mu = c.derived.reduced_mass;
  if (nearest_fft_size(nr, 1) != nr) throw InputError("nR", "nR must be a valid FFT size");
  const double lower = p.r_start, upper = p.r_start + nr * p.dr;
  const double dx = (upper - lower) / nr;
  std::vector<double> grid(nr);
  for (int l = 0; l < nr; ++l) grid[l] = l * dx + lower;
  const std::vector<double> kinetic = fgh_kinetic_matrix(nr, dx, mu);

  GasPhase gas;
  gas.nr = nr;
  gas.j_max = p.j_max;
  gas.energies.assign(std::size_t(nr) * (p.j_max + 1), 0.0);
  gas.wave_functions.assign(std::size_t(nr) * nr * (p.j_max + 1), 0.0);
  std::vector<double> values, vectors;
  for (int j = 0; j <= p.j_max; ++j) {
    std::vector<double> h = kinetic;
    for (int i = 0; i < nr; ++i) {
      double v = vasy[i];
      if (j > 0) {
        double r = grid[i];
        if (r == 0.0) r = r + machine_epsilon();
        v = vasy[i] + (j * (j + 1)) / 2.0 / mu / (r * r);
      }
      h[i + std::size_t(nr) * i] = h[i + std::size_t(nr) * i] + v;
    }
    symmetric_eigen(std::move(h), nr, values, vectors);
    for (int v = 0; v < nr; ++v) {
      double* psi = vectors.data() + std::size_t(nr) * v;
      // Sign: the first local maximum of |psi| seen from the last point is positive.
      int i = nr - 1;
      while (i > 0 && std::abs(psi[i - 1]) >= std::abs(psi[i])) --i;
      if (psi[i] < 0.0)
        for (int k = 0; k < nr; ++k) psi[k] = -1.0 * psi[k];
      double norm = 0.0;
      for (int k = 0; k < nr; ++k) norm = norm + psi[k] * psi[k];
      norm = 1.0 / std::sqrt(norm * dx);
      for (int k = 0; k < nr; ++k) psi[k] = psi[k] * norm;
      gas.energies[j + (p.j_max + 1) * std::size_t(v)] = values[v];
      std::copy_n(psi, nr, gas.wave_functions.begin() + std::size_t(nr) * (j + (p.j_max + 1) * std::size_t(v)));
    }
  }

  const double e00 = gas.energy(0, 0);
  const double e_in = gas.energy(p.v_in, p.j_in);
  const double available = Units::ev_to_hartree(p.e_max_wp) + (e_in - e00) + Units::ev_to_hartree(p.e_parallel);
  int iv = 1;
  while (gas.energy(iv - 1, 0) - e00 <= available) {
    iv = iv + 1;
    if (iv == nr) break;
  }
  gas.v_max_open = iv - 1;
  int jj = p.j_in;
  while (gas.energy(0, jj) - e00 <= available) {
    jj = jj + angular.delta_j;
    if (jj > p.j_max) break;
  }
  gas.j_max_open = jj - angular.delta_j;
  gas.n_v_analysis = (p.v_analysis && *p.v_analysis >= 1) ? *p.v_analysis : gas.v_max_open;
  gas.j_max_analysis = (p.j_analysis && *p.j_analysis >= 0) ? *p.j_analysis : gas.j_max_open;
  if (gas.n_v_analysis <= p.v_in) throw InputError("VIn", "VIn is not among the analysed vibrational levels");
  if (gas.j_max_analysis < p.j_in) throw InputError("JIn", "JIn is above the highest analysed j");
  gas.initial_energy = e_in;
  for (int j = 0; j <= p.j_max; ++j) gas.rotational_energies.push_back(gas.energy(0, j) - e00);
  for (int v = 0; v < gas.n_v_analysis; ++v) gas.vibrational_energies.push_back(gas.energy(v, 0) - e00);
  return gas;
}

}  // namespace spo
