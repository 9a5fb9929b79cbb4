// The potential energy surfaces on the device, and the potential grid built from them: V(itp, irep, izr) in hartree
// at the symmetry representatives, with the cosine switch into the gas phase and the VMin / VMax limits
// (src/potential/potential_grid.cpp:35-56, :65-127). The formulas are the published ones of the two surfaces
// (src/potential/hard_wall_potential.cpp, src/potential/dai_light_cu111_potential.cpp); every comparison that
// places a point uses the absolute tolerance 1e-12 of src/core/real_compare.hpp:14-33.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_crp.hpp"

namespace spo_dev {

constexpr double kPesPi = 3.1415926535897932384626433832795028841971;
constexpr double kHartreeInEv = 27.2114;
constexpr double kTolerance = 1.0e-12;

__device__ __forceinline__ bool tol_eq(double a, double b) { return fabs(a - b) < kTolerance; }
__device__ __forceinline__ bool tol_ge(double a, double b) { return tol_eq(a, b) || (a - b) > kTolerance; }
__device__ __forceinline__ bool tol_gt(double a, double b) { return (a - b) > kTolerance; }

enum PesId : int { kPesHardWall = 1, kPesDaiLight = 2 };

// H2 gas phase (extended Rydberg) plus a 1/Z^4 wall switched off between Z = 0.1778 and 7 bohr; eV.
__device__ __forceinline__ double hard_wall_gas_phase(double r_in) {
  const double d = 4.8286, a1 = 2.282, a2 = 1.555, a3 = 0.7533, a4 = 2.23, re = 1.40;
  const double r = r_in - re;
  return -d * (1.0 + a1 * r + a2 * pow(r, 2.0) + a3 * pow(r, 3.0)) * exp(-a4 * r) + d;
}

__device__ __forceinline__ double hard_wall_value(double z, double r) {
  const double vr = hard_wall_gas_phase(r);
  const double vz = 1.0 / (z * z * z * z + 0.001);
  const double z_max = 7.0, z_min = 0.1778;
  if (z >= z_max) return vr;
  if (z < z_min) return vr + vz;
  return vr + (1.0 - (0.5 - 0.5 * cos(kPesPi * (z - z_min) / (z_max - z_min)))) * vz;
}

// Site values (top, bridge, centre) interpolated over the hexagonal cell with the two lowest Fourier stars.
struct PesSites {
  double top, bridge, centre;
};

__device__ __forceinline__ double pes_hexagonal(double s1, double s2, PesSites v) {
  const double c0 = v.top / 3.0 + v.centre * 2.0 / 3.0;
  const double c1 = v.top / 4.0 - v.bridge / 4.0;
  const double c2 = v.bridge / 4.0 - v.centre * 2.0 / 9.0 - v.top / 36.0;
  return c0 + c1 * s1 + c2 * s2;
}

// LEPS Coulomb (u) and exchange (a) integrals with Sato parameter delta; e1 = exp(-alpha (x - x0)).
__device__ __forceinline__ double leps_coulomb(double e1, double d, double delta) {
  return d / 4.0 / (1.0 + delta) * ((3.0 + delta) * (e1 * e1) - (2.0 + 6.0 * delta) * e1);
}
__device__ __forceinline__ double leps_exchange(double e1, double d, double delta) {
  return d / 4.0 / (1.0 + delta) * ((1.0 + 3.0 * delta) * (e1 * e1) - (6.0 + 2.0 * delta) * e1);
}

// H2 / Cu(111) LEPS of Dai and Light; eV.
__device__ __forceinline__ double dai_light_value(double x, double y, double z, double r, double theta, double phi) {
  constexpr double m1 = 1.00794, m2 = 1.00794;
  constexpr double d_m = .1744, alpha_m = 1.028, r_m = 1.40, delta_m = -0.15;
  constexpr PesSites d_a{.638e-1, .818e-1, .877e-1};
  constexpr PesSites alpha_a{.790, .558, .499};
  constexpr PesSites z_a{3.02, 2.13, 1.88};
  constexpr PesSites delta_a{0.10, 0.10, 0.10};
  constexpr double angle = 1.04719755119659774614;
  constexpr double cell = 4.84;
  const double w1 = m2 / (m1 + m2);
  const double w2 = -m1 / (m1 + m2);
  const double rp = r * sin(theta) / sin(angle);
  double t = sin(angle - phi) * rp;
  const double x1 = x + w1 * t, x2 = x + w2 * t;
  t = sin(phi) * rp;
  const double y1 = y + w1 * t, y2 = y + w2 * t;
  t = r * cos(theta);
  const double z1 = z + w1 * t, z2 = z + w2 * t;
  // The two Fourier stars at each atom, shared by the four site-dependent parameters.
  const double k1 = 2.0 * kPesPi / cell, k2 = 4.0 * kPesPi / cell;
  const double s11 = cos(k1 * x1) + cos(k1 * y1) + cos(k1 * (x1 + y1));
  const double s12 = cos(k2 * x1) + cos(k2 * y1) + cos(k2 * (x1 + y1));
  const double s21 = cos(k1 * x2) + cos(k1 * y2) + cos(k1 * (x2 + y2));
  const double s22 = cos(k2 * x2) + cos(k2 * y2) + cos(k2 * (x2 + y2));
  const double d1 = pes_hexagonal(s11, s12, d_a), d2 = pes_hexagonal(s21, s22, d_a);
  const double al1 = pes_hexagonal(s11, s12, alpha_a), al2 = pes_hexagonal(s21, s22, alpha_a);
  const double z01 = pes_hexagonal(s11, s12, z_a), z02 = pes_hexagonal(s21, s22, z_a);
  const double de1 = pes_hexagonal(s11, s12, delta_a), de2 = pes_hexagonal(s21, s22, delta_a);
  const double em = exp(-alpha_m * (r - r_m));
  const double e1 = exp(-al1 * (z1 - z01)), e2 = exp(-al2 * (z2 - z02));
  const double um = leps_coulomb(em, d_m, delta_m), am = leps_exchange(em, d_m, delta_m);
  const double u1 = leps_coulomb(e1, d1, de1), a1 = leps_exchange(e1, d1, de1);
  const double u2 = leps_coulomb(e2, d2, de2), a2 = leps_exchange(e2, d2, de2);
  const double as = a1 + a2;
  double v = um + u1 + u2 - sqrt(am * am + as * as - am * as);
  if (tol_ge(z, 15.0)) v = um + am;
  return v * kHartreeInEv;
}

__device__ __forceinline__ double pes_value(int pes, double x, double y, double z, double r, double theta, double phi) {
  return pes == kPesDaiLight ? dai_light_value(x, y, z, r, theta, phi) : hard_wall_value(z, r);
}

// PES (eV) at n arbitrary points (x, y, z, r, theta, phi).
__global__ __launch_bounds__(256) void pes_points_kernel(int pes, long n, const double* points, double* out) {
  const long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double* p = points + 6 * i;
  out[i] = pes_value(pes, p[0], p[1], p[2], p[3], p[4], p[5]);
}

struct PesGrid {
  int pes;
  int ny, nr, n_theta, n_phi, n_rep, n_zr;
  const double *x, *y, *z, *r, *theta, *phi;
  const int* representatives;  // full-grid ixy (iy + ny ix) of each representative
  int use_switching;
  double z_min_switch, z_max_switch;
  const double* gas_phase;  // nr values, hartree: the switch's asymptote
  int impose_v_min, impose_v_max;
  double v_min, v_max;
  CrpDev crp;  // pes == 3 (CRP) only
  ScmDev scm;  // pes == 3 with a distorted slab (scm.on)
};

// One value per thread, in the table's own order: index = iphi + n_phi (itheta + n_theta (irep + n_rep izr)).
__global__ __launch_bounds__(256) void potential_grid_kernel(PesGrid g, double* values) {
  const long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  const long n_tp = static_cast<long>(g.n_theta) * g.n_phi;
  if (i >= n_tp * g.n_rep * g.n_zr) return;
  const int ip = static_cast<int>(i % g.n_phi);
  const int it = static_cast<int>((i / g.n_phi) % g.n_theta);
  const int irep = static_cast<int>((i / n_tp) % g.n_rep);
  const int izr = static_cast<int>(i / (n_tp * g.n_rep));
  const int iz = izr / g.nr, ir = izr % g.nr;
  const int ixy = g.representatives[irep];
  const int ix = ixy / g.ny, iy = ixy % g.ny;
  const double z = g.z[iz];
  // eV -> hartree by a multiplication with 1 / 27.2114, as the reference's grid builder does
  const double ev = g.pes == 3 ? crp_scm_qd(g.crp, g.scm, g.x[ix], g.y[iy], z, g.r[ir], g.theta[it], g.phi[ip])
                               : pes_value(g.pes, g.x[ix], g.y[iy], z, g.r[ir], g.theta[it], g.phi[ip]);
  double v = ev * (1.0 / kHartreeInEv);
  // A point the surface cannot evaluate (NaN) stays NaN through the switch and the limits, so that the entry point
  // refuses the grid where the host build stops, also above ZmaxSwitch.
  if (isnan(v)) {
    values[i] = v;
    return;
  }
  if (g.use_switching && tol_gt(z, g.z_min_switch)) {
    const double va = g.gas_phase[ir];
    if (tol_ge(z, g.z_max_switch)) {
      v = va;
    } else {
      const double f = 0.5 + 0.5 * cos(kPesPi * (z - g.z_min_switch) / (g.z_max_switch - g.z_min_switch));
      v = f * (v - va) + va;
    }
  }
  if (g.impose_v_min && v < g.v_min) v = g.v_min;
  if (g.impose_v_max && v > g.v_max) v = g.v_max;
  values[i] = v;
}

}  // namespace spo_dev
