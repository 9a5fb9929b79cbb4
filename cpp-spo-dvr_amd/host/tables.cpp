#include "tables.hpp"

#include <cmath>

namespace spo {

namespace {

inline cplx exp_minus_i(double phase) { return std::exp(cplx(0.0, -phase)); }

// Full step exp(-i E dt); the half step as exp(-i E dt / 2) or as the principal root of the full step.
StepFactors step_factors(const std::vector<double>& energies, double dt, bool half_by_root) {
  StepFactors f;
  for (const double e : energies) {
    const double phase = e * dt;
    f.full.push_back(exp_minus_i(phase));
    f.half.push_back(half_by_root ? std::sqrt(f.full.back()) : exp_minus_i(phase / 2.0));
  }
  return f;
}

// k0, k0 + dk, ..., k0 + (n/2) dk, ..., k0 - dk with the offset accumulated step by step.
std::vector<double> shifted_momenta(int n, double k0, double dk) {
  std::vector<double> k(n);
  k[0] = k0;
  k[n / 2] = k0 + (n / 2) * dk;
  double offset = dk;
  for (int i = 1; i < n / 2; ++i) {
    k[i] = k0 + offset;
    k[n - i] = k0 - offset;
    offset = offset + dk;
  }
  return k;
}

// exp(-V_opt dt), V_opt = A (n+1)/2 x^n on [start, end] (x from 0 at the weak end to 1 at the strong end).
void set_optical(std::vector<double>& f, const UniformGrid& g, double start, double end, double strength, bool high,
                 int power, double dt) {
  const double length = end - start;
  for (int i = 0; i < g.size(); ++i) {
    const double q = g[i];
    if (!(q >= start && q <= end)) continue;
    const double x = high ? (q - start) / length : (end - q) / length;
    const double v = ((strength * 0.5) * (power + 1)) * ipow(x, power);
    f[i] = std::exp(-(v * dt));
  }
}

std::vector<double> optical(const UniformGrid& g, double start, double end, double strength, const LowAbsorber* low,
                            double low_strength, int power, double dt) {
  std::vector<double> f(g.size(), 1.0);
  set_optical(f, g, start, end, strength, true, power, dt);
  if (low && low->enabled()) set_optical(f, g, *low->start, *low->end, low_strength, false, power, dt);
  return f;
}

}  // namespace

PropagationTables make_propagation_tables(const Config& c, const SpatialGrids& g, const RotationalBasis& basis,
                                          const std::vector<double>& vasy) {
  const InputParameters& p = c.input;
  const DerivedParameters& d = c.derived;
  const double dt = p.time_step;
  PropagationTables t;
  const StepFactors z = step_factors(g.ekin_z, dt, true);
  const StepFactors r = step_factors(g.ekin_r, dt, true);
  for (std::size_t iz = 0; iz < z.full.size(); ++iz)
    for (std::size_t ir = 0; ir < r.full.size(); ++ir) {
      t.zr_kinetic.full.push_back(z.full[iz] * r.full[ir]);
      t.zr_kinetic.half.push_back(z.half[iz] * r.half[ir]);
    }

  const std::vector<double> kx = shifted_momenta(p.nx, d.kx, d.delta_kx);
  const std::vector<double> ky = shifted_momenta(p.ny, d.ky, d.delta_ky);
  const double ecnvrt = 0.5 / d.total_mass;
  const double prefac = -2.0 * std::cos(d.system.skew);
  for (int ix = 0; ix < p.nx; ++ix)
    for (int iy = 0; iy < p.ny; ++iy) {
      const double ekx = ecnvrt * (kx[ix] * kx[ix]);
      const double eky = ecnvrt * (ky[iy] * ky[iy]);
      const double ekxy = ecnvrt * prefac * kx[ix] * ky[iy];
      t.xy_kinetic_energies.push_back(ekx + ekxy + eky);
    }
  t.parallel_energy = t.xy_kinetic_energies[0];
  t.xy_kinetic = step_factors(t.xy_kinetic_energies, dt, false);

  std::vector<double> spec;
  for (const double ez : g.ekin_z_specular) {
    const double total = ez + t.parallel_energy;
    for (const double er : g.ekin_r) spec.push_back(er + total);
  }
  t.specular_kinetic = step_factors(spec, dt, true);

  std::vector<double> erot;
  for (int ir = 0; ir < g.r.size(); ++ir) {
    const double rr = g.r[ir];
    for (int irot = 0; irot < basis.size(); ++irot) {
      const int j = basis.j[irot];
      erot.push_back((j * (j + 1)) / (2.0 * d.reduced_mass * (rr * rr)));
    }
  }
  t.rotation = step_factors(erot, dt, false);

  const int power = d.optical_power;
  t.z_optical = optical(g.z, p.z_opt_start, p.z_opt_end, d.z_opt_factor, &p.z_opt_low, d.z_opt_low_factor, power, dt);
  t.r_optical = optical(g.r, p.r_opt_start, p.r_opt_end, d.r_opt_factor, nullptr, 0.0, power, dt);
  t.z_specular_optical = optical(g.z_specular, p.zsp_opt_start, p.zsp_opt_end, d.zsp_opt_factor, &p.zsp_opt_low,
                                 d.zsp_opt_low_factor, power, dt);
  for (const double v : vasy) t.gas_phase.push_back(exp_minus_i(v * dt));

  const double sin_skew = std::sin(d.system.skew);
  for (int ix = 0; ix < p.nx; ++ix) {
    const double conx = sin_skew * d.kx * ix * d.dx;
    for (int iy = 0; iy < p.ny; ++iy) {
      const double phase = conx + sin_skew * d.ky * iy * d.dy;
      const cplx forward = exp_minus_i(phase);
      const cplx backward = std::exp(cplx(0.0, phase));
      t.shift_forward.push_back(forward);
      t.shift_backward.push_back(backward);
      t.specular_to_scattering.push_back(backward * d.unit_cell_norm);
      t.scattering_to_specular.push_back(forward * d.unit_cell_norm * d.xy_volume_element);
      if (forward.real() != 1.0 || forward.imag() != 0.0) t.needs_shift = true;
    }
  }
  const double sqrt_two_pi = std::sqrt(2.0 * kPi);
  t.fft_scale_z = p.dz / sqrt_two_pi;
  t.fft_scale_r = p.dr / sqrt_two_pi;
  t.fft_scale_x = std::sqrt(d.dx / p.nx);
  t.fft_scale_y = std::sqrt(d.dy / p.ny);
  return t;
}

PotentialStepFactors potential_step_factors(int n_rot, int nr, const PropagationTables& t) {
  PotentialStepFactors f;
  for (int ir = 0; ir < nr; ++ir)
    for (int irot = 0; irot < n_rot; ++irot) {
      const cplx rot = t.rotation.half[irot + std::size_t(n_rot) * ir];
      const double opt_r = t.r_optical[ir];
      f.before_interaction.push_back(opt_r * rot);
      f.gas_phase_step.push_back(opt_r * rot * t.gas_phase[ir] * rot);
    }
  return f;
}

std::vector<cplx> specular_gas_phase_factors(int nr, int n_rot, int ichn, const PropagationTables& t) {
  std::vector<cplx> f;
  for (int ir = 0; ir < nr; ++ir) {
    const cplx rot = t.rotation.half[ichn + std::size_t(n_rot) * ir];
    f.push_back(rot * t.gas_phase[ir] * rot);
  }
  return f;
}

std::vector<cplx> make_initial_specular_wave(const Config& c, const SpatialGrids& g, const GasPhase& gas) {
  const InputParameters& p = c.input;
  const DerivedParameters& d = c.derived;
  const int nzsp = g.z_specular.size();
  const double alpha = d.wave_packet_alpha;
  const double amplitude = std::pow(2.0 * alpha / kPi, 0.25);
  std::vector<cplx> gauss(nzsp);
  for (int iz = 0; iz < nzsp; ++iz) {
    const double dz = g.z_specular[iz] - p.z0;
    gauss[iz] = amplitude * std::exp(cplx(-(alpha * (dz * dz)), -(d.kz * g.z_specular[iz])));
  }
  int i_low = 0;
  for (int iz = 0; iz < nzsp; ++iz)
    if (tol::ge(g.z_specular[iz], g.z_analysis)) {
      i_low = iz;
      break;
    }
  int i_high = g.z_specular_optical.first - 1;
  if (d.fixed_z_analysis_specular && p.z_analysis_specular && *p.z_analysis_specular > 0.0)
    for (int iz = 0; iz < nzsp; ++iz)
      if (tol::ge(g.z_specular[iz], *p.z_analysis_specular)) {
        i_high = iz - 1;
        break;
      }
  double inside = 0.0;
  for (int i = i_low; i <= i_high; ++i) inside = inside + (gauss[i].real() * gauss[i].real() + gauss[i].imag() * gauss[i].imag());
  inside = inside * p.dz;
  if ((1.0 - inside) > p.tol_specular)
    throw InputError("TolSp", "too much norm of the initial wave packet outside the allowed region");
  const double* phi = gas.wave_function(p.v_in, p.j_in);
  std::vector<cplx> spec(std::size_t(p.nr) * nzsp);
  for (int iz = 0; iz < nzsp; ++iz)
    for (int ir = 0; ir < p.nr; ++ir) spec[ir + std::size_t(p.nr) * iz] = phi[ir] * gauss[iz];
  return spec;
}

}  // namespace spo
