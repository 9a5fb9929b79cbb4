// The CRP surface (SystemID CRPlibrary) on the device: V alone at a molecular geometry, one thread per point, FP64.
// The statements are the value statements of the host restatement (host/crp.cpp) in the same operand order
// (src/crplib/qd_wrapper.cpp:46-66, molecule_surface.cpp:249-315, hexagonal_interaction.cpp, site_interpolation.cpp,
// hexagonal_atom_surface.cpp:109-283, spline.cpp); FMA contraction is off so that the only differences from the
// host are those of the device's sin / cos / acos / atan2. A geometry where the host stops with an error (an atom
// on a surface atom, z outside the atom-surface data) gives NaN; the caller turns it into an error.
#pragma once

#include <hip/hip_runtime.h>

#include <array>

#include "../host/crp_tables.hpp"

#pragma clang fp contract(off)

namespace spo_dev {

__constant__ std::array<double, 100> kCrpAtomC3v = spo::crp_tables::kAtomC3vTable;
__constant__ std::array<double, 49> kCrpAtomC6v = spo::crp_tables::kAtomC6vTable;
__constant__ std::array<double, 144> kCrpMolecule = spo::crp_tables::kMoleculeTable;

// Device pointers of a packed surface (the layout of spo_hip_crp_surface).
struct CrpDev {
  int c3v, n_basis, n_cuts, n_corrugation, n_cluster, n_h2, n_zdep;
  double a, zso_min, zso_max, r_min, r_max, z_min, h2_asymptote;
  const double* corrugation;
  const double* cluster;
  const int* atom_offsets;
  const double* atom_splines;
  const int* cut_shape;
  const long* cut_offsets;
  const double* cut_data;
  const double* h2;
  const double* zdep;
  int square, cell_nz;
  const double* cell;  // C4v: x[5] y[5] z[nz], f xx yy xxyy zz zzxx zzyy zzxxyy (25 nz each)
};

// SPLINT / UBICAR bisection: the lower knot of the interval holding t.
__device__ inline int crp_bisect(const double* x, int n, double t) {
  int lo = 0, hi = n - 1;
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

// SPLINT value; s = x[n] y[n] F''[n]
__device__ inline double crp_splint(const double* s, int n, double t) {
  const double* x = s;
  const double* y = s + n;
  const double* m = s + 2 * n;
  const int lo = crp_bisect(x, n, t), hi = lo + 1;
  const double h = x[hi] - x[lo];
  const double a = (x[hi] - t) / h;
  const double b = (t - x[lo]) / h;
  return a * y[lo] + b * y[hi] + ((a * a * a - a) * m[lo] + (b * b * b - b) * m[hi]) * (h * h) / 6.0;
}

// DPL: linear outside the knots
__device__ inline double crp_dpl(const double* s, int n, double t) {
  const double* x = s;
  const double* y = s + n;
  const double* m = s + 2 * n;
  int lo = 0, hi = n;
  while (lo < hi) {  // the number of knots below t
    const int mid = (lo + hi) / 2;
    if (x[mid] < t) {
      lo = mid + 1;
    } else {
      hi = mid;
    }
  }
  int k = lo + 1;
  if (t == x[0]) k = 2;
  if (k <= 1) {
    const double e = x[1] - x[0];
    return ((y[1] - y[0]) / e - m[1] * e / 6.0) * (t - x[0]) + y[0];
  }
  if (k > n) {
    const double e = x[n - 1] - x[n - 2];
    return ((y[n - 1] - y[n - 2]) / e + m[n - 2] * e / 6.0) * (t - x[n - 1]) + y[n - 1];
  }
  const int i = k - 1;
  const double f = x[i] - t;
  const double g = t - x[i - 1];
  const double e = f + g;
  return (-g * f * (m[i - 1] * (f + e) + m[i] * (g + e)) + 6.0 * (g * y[i] + f * y[i - 1])) / (6.0 * e);
}


// SPLINT3D value weights in one direction (x**3.0 a call of pow, as there)
__device__ inline void crp_w3(const double* x, int i, double t, double& a, double& b, double& c, double& d) {
  const double h = x[i + 1] - x[i];
  a = (x[i + 1] - t) / h;
  b = 1.0 - a;
  c = (pow(a, 3.0) - a) / 6.0 * (h * h);
  d = (pow(b, 3.0) - b) / 6.0 * (h * h);
}

// The C4v atom-surface V (int_PES6CuHATOMCu100 value): 0 above the data; the point folded into the triangle of the
// square cell (ORIGEN, REGION), the MURO cluster sum and the cell's periodic 3D spline there.
__device__ inline double crp_atom_square(const CrpDev& s, double px, double py, double pz, bool& bad) {
  const int nz = s.cell_nz;
  const double* gx = s.cell;
  const double* gz = s.cell + 10;
  if (pz >= gz[nz - 1]) return 0.0;
  const double a = s.a;
  const double o1 = px / a, o2 = py / a;
  if (!(fabs(o1) < 2.0e9) || !(fabs(o2) < 2.0e9)) {
    bad = true;
    return 0.0;
  }
  const int i1 = o1 >= 0.0 ? static_cast<int>(o1) : static_cast<int>(o1 - 1.0);
  const int i2 = o2 >= 0.0 ? static_cast<int>(o2) : static_cast<int>(o2 - 1.0);
  double x = px - i1 * a, y = py - i2 * a;
  // REGION (the gradient's region is not needed for V)
  if (y <= a / 2.0) {
    if (x <= a / 2.0) {
      if (!(x >= y)) {
        const double t = x;
        x = y;
        y = t;
      }
    } else if (y >= (a - x)) {
      const double t = x;
      x = y;
      y = a - t;
    } else {
      x = a - x;
    }
  } else if (x <= a / 2.0) {
    if (y > (a - x)) {
      y = a - y;
    } else {
      const double t = x;
      x = a - y;
      y = t;
    }
  } else if (x > y) {
    const double t = x;
    x = a - y;
    y = a - t;
  } else {
    x = a - x;
    y = a - y;
  }
  // MURO / VREP1D value (no lower limit)
  const int nc = s.n_corrugation;
  const double reach = s.corrugation[nc - 1];
  double rep = 0.0;
  for (int k = 0; k < s.n_cluster; ++k) {
    const double* at = s.cluster + 4 * k;
    const double r = sqrt((x - at[0]) * (x - at[0]) + (y - at[1]) * (y - at[1]) + (pz - at[2]) * (pz - at[2]));
    double vr = 0.0;
    if (!(r > reach)) vr = crp_splint(s.corrugation, nc, r);
    rep = rep + vr;
  }
  // SPLINT3D value
  if (x < gx[0] || x > gx[4] || y < gx[5] || y > gx[9] || pz < gz[0]) {
    bad = true;
    return 0.0;
  }
  const int i = crp_bisect(gx, 5, x), j = crp_bisect(gx + 5, 5, y), k = crp_bisect(gz, nz, pz);
  double xa, xb, xc, xd, ya, yb, yc, yd, za, zb, zc, zd;
  crp_w3(gx, i, x, xa, xb, xc, xd);
  crp_w3(gx + 5, j, y, ya, yb, yc, yd);
  crp_w3(gz, k, pz, za, zb, zc, zd);
  const long plane = 25L * nz;
  const double* f = s.cell + 10 + nz;
  auto plane_sum = [&](const double* g, const double* gxx, const double* gyy, const double* gxxyy, int kk) {
    const long p00 = i + 5L * (j + 5L * kk), p10 = p00 + 1, p01 = p00 + 5, p11 = p01 + 1;
    return ya * (xa * g[p00] + xb * g[p10] + xc * gxx[p00] + xd * gxx[p10]) +
           yb * (xa * g[p01] + xb * g[p11] + xc * gxx[p01] + xd * gxx[p11]) +
           yc * (xa * gyy[p00] + xb * gyy[p10] + xc * gxxyy[p00] + xd * gxxyy[p10]) +
           yd * (xa * gyy[p01] + xb * gyy[p11] + xc * gxxyy[p01] + xd * gxxyy[p11]);
  };
  const double *fxx = f + plane, *fyy = f + 2 * plane, *fxxyy = f + 3 * plane, *fzz = f + 4 * plane,
               *fzzxx = f + 5 * plane, *fzzyy = f + 6 * plane, *fzzxxyy = f + 7 * plane;
  const double t0 = plane_sum(f, fxx, fyy, fxxyy, k), t1 = plane_sum(f, fxx, fyy, fxxyy, k + 1);
  const double t2 = plane_sum(fzz, fzzxx, fzzyy, fzzxxyy, k), t3 = plane_sum(fzz, fzzxx, fzzyy, fzzxxyy, k + 1);
  return za * t0 + zb * t1 + zc * t2 + zd * t3 + rep;
}

// The atom-surface V (ATII value): the corrugation function over the cluster at the folded position plus the
// Fourier interpolation of the reduced cuts. `bad` is set where the host stops.
__device__ inline double crp_atom(const CrpDev& s, double px, double py, double pz, bool& bad) {
  using namespace spo::crp_tables;
  if (s.square) return crp_atom_square(s, px, py, pz, bad);
  const double delta = s.a;
  // ORIGEN2
  double bigx = px - py * kUsq3;
  double bigy = 2.00 * py * kUsq3;
  const double o1 = bigx / delta, o2 = bigy / delta;
  if (!(fabs(o1) < 2.0e9) || !(fabs(o2) < 2.0e9)) {
    bad = true;
    return 0.0;
  }
  const int i1 = o1 >= 0.0 ? static_cast<int>(o1) : static_cast<int>(o1 - 1.0);
  const int i2 = o2 >= 0.0 ? static_cast<int>(o2) : static_cast<int>(o2 - 1.0);
  bigx = bigx - i1 * delta;
  bigy = bigy - i2 * delta;
  const double xf = bigx + bigy * 0.500, yf = kSq3 * bigy * 0.500;
  // MURII value; an atom beyond the corrugation function's reach adds +0 (skipped on the squared distance, with
  // a margin that keeps every atom whose rounded distance could be within reach)
  const int nc = s.n_corrugation;
  const double reach = s.corrugation[nc - 1], nearest = s.corrugation[0];
  const double reach2 = reach * reach * (1.0 + 1.0e-12);
  double rep = 0.0;
  for (int k = 0; k < s.n_cluster; ++k) {
    const double* at = s.cluster + 4 * k;
    const double dx = xf - at[0], dy = yf - at[1], dz = pz - at[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 > reach2) continue;
    const double r = sqrt(r2);
    double vr = 0.0;
    if (!(r > reach)) {
      if (r < nearest) bad = true;
      vr = crp_splint(s.corrugation, nc, r);
    }
    rep = rep + vr;
    if (!(r > 0.0) && at[3] != 0.0) bad = true;
  }
  // z within the first site's data
  const int n0 = s.atom_offsets[1] - s.atom_offsets[0];
  if (pz < s.atom_splines[0] || pz > s.atom_splines[n0 - 1]) bad = true;
  double vs[10];
  for (int i = 0; i < s.n_basis; ++i) {
    const int off = s.atom_offsets[i], n = s.atom_offsets[i + 1] - off;
    vs[i] = crp_dpl(s.atom_splines + 3 * off, n, pz);
  }
  // FOUR3II values
  const double fac = 2.00 * kPi / s.a;
  const double bx = px - py * kUsq3;
  const double by = 2.00 * py * kUsq3;
  const double p = fac * bx, q = fac * by;
  double f[10];
  int k = 0;
  f[k++] = 1.00;
  f[k++] = cos(p) + cos(q) + cos(p + q);
  if (s.c3v) f[k++] = -sin(p) - sin(q) + sin(p + q);
  f[k++] = cos(2.00 * p + q) + cos(p + 2.00 * q) + cos(p - q);
  f[k++] = cos(2.00 * p) + cos(2.00 * q) + cos(2.00 * (p + q));
  if (s.c3v) f[k++] = -sin(2.00 * p) - sin(2.00 * q) + sin(2.00 * (p + q));
  f[k++] = cos(3.00 * p + q) + cos(p + 3.00 * q) + cos(2.00 * p + 3.00 * q) + cos(3.00 * p + 2.00 * q) +
           cos(-p + 2.00 * q) + cos(2.00 * p - q);
  if (s.c3v)
    f[k++] = -sin(3.00 * p + q) - sin(p + 3.00 * q) + sin(2.00 * p + 3.00 * q) + sin(3.00 * p + 2.00 * q) -
             sin(-p + 2.00 * q) - sin(2.00 * p - q);
  f[k++] = cos(3.00 * p) + cos(3.00 * q) + cos(3.00 * (p + q));
  f[k++] = cos(2.00 * (2.00 * p + q)) + cos(2.00 * (p + 2.00 * q)) + cos(2.00 * (p - q));
  const double* table = s.c3v ? kCrpAtomC3v.data() : kCrpAtomC6v.data();
  double v = 0.00;
  for (int j = 0; j < s.n_basis; ++j) {
    double a = 0.00;
    for (int i = 0; i < s.n_basis; ++i) a = a + table[i + s.n_basis * j] * vs[i];
    v = v + a * f[j];
  }
  return v + rep;
}

// theta = 0, 45, 90, 135 degrees combined
__device__ inline double crp_four(double v0, double v45, double v90, double v135, double c2, double s2, double c4) {
  return 0.2500 * ((v0 + v45 + v90 + v135) + (v0 + v90 - v45 - v135) * c4) +
         0.500 * ((v0 - v90) * c2 + (v45 - v135) * s2);
}

// The t2 sites (value): 45 cuts g, m, k around the middle one, 90 cuts p, q, azimuth phi - shift
__device__ inline double crp_t2(const double* c, double c2, double s2, double c4, double phi, double shift, int mid,
                                bool reversed) {
  const double ph = phi - shift;
  const double cp = cos(ph), c2p = cos(2.00 * ph);
  const double g = c[3], k = c[mid == 5 ? 4 : 5], m = c[mid];
  const double odd = reversed ? (-g + k) : (g - k);
  const double v45 = 0.2500 * ((g + 2.00 * m + k) + (g - 2.00 * m + k) * c2p) + 0.500 * odd * cp;
  const double v135 = 0.2500 * ((g + 2.00 * m + k) + (g - 2.00 * m + k) * c2p) - 0.500 * odd * cp;
  const double v90 = reversed ? 0.500 * ((c[1] + c[2]) + (-c[1] + c[2]) * c2p)
                              : 0.500 * ((c[1] + c[2]) + (c[1] - c[2]) * c2p);
  return crp_four(c[0], v45, v90, v135, c2, s2, c4);
}

__device__ inline double crp_hollow(const double* c, double c2, double s2, double c4, double phi) {
  const double s3 = sin(3.00 * phi), c6 = cos(6.00 * phi);
  const double v45 = 0.500 * ((c[3] + c[4]) + (c[3] - c[4]) * s3);
  const double v135 = 0.500 * ((c[3] + c[4]) - (c[3] - c[4]) * s3);
  const double v90 = 0.500 * ((c[1] + c[2]) + (c[1] - c[2]) * c6);
  return crp_four(c[0], v45, v90, v135, c2, s2, c4);
}

__device__ inline double crp_bridge(const double* c, double c2t, double phi) {
  constexpr double u3 = 1.00 / 3.00;
  const double c2 = cos(2.00 * phi), c4 = cos(4.00 * phi);
  const double v90 =
      (c[1] + 2.00 * c[2]) * u3 - 0.500 * (c[3] - c[1]) * c2 + (0.500 * (c[3] - c[1]) - 2.00 * (c[2] - c[1]) * u3) * c4;
  return 0.500 * (c[0] * (1.00 + c2t) + v90 * (1.00 - c2t));
}

// The molecule-surface V at atom positions a, b (POT6D value statements). `silent`: outside the data, where the
// library returns V = 0 without an error.
__device__ inline double crp_molecule(const CrpDev& s, const double* a, const double* b, bool& silent, bool& bad) {
  using namespace spo::crp_tables;
  const double r = sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]) + (b[2] - a[2]) * (b[2] - a[2]));
  if (r == 0.0) {
    bad = true;
    return 0.0;
  }
  const double gx = a[0] * (1.0 - 0.500) + b[0] * 0.500;
  const double gy = a[1] * (1.0 - 0.500) + b[1] * 0.500;
  const double gz = a[2] * (1.0 - 0.500) + b[2] * 0.500;
  const int nh = s.n_h2, nz = s.n_zdep;
  if (!(gz < s.zso_max)) {  // the gas-phase potential alone
    if (r < s.h2[0] || r > s.h2[nh - 1]) {
      silent = true;
      return 0.0;
    }
    const double v1 = gz <= s.zdep[nz - 1] ? crp_splint(s.zdep, nz, gz) : s.zdep[2 * nz - 1];
    return crp_splint(s.h2, nh, r) + v1;
  }
  double ct = (b[2] - a[2]) / r;
  if (ct > 1.0) ct = 1.0;
  const double theta = acos(ct);
  const double ddx = b[0] - a[0], ddy = b[1] - a[1];
  const double phi = (ddx == 0.0 && ddy == 0.0) ? 0.0 : atan2(ddy, ddx);
  // the cuts at (r, Z); one cell lookup per group of cuts on the same grid
  double cv[29];
  int group = -1, ci = 0, cj = 0;
  double wxa = 0, wxb = 0, wxc = 0, wxd = 0, wya = 0, wyb = 0, wyc = 0, wyd = 0;
  for (int k = 0; k < s.n_cuts; ++k) {
    const int nr = s.cut_shape[3 * k], nzc = s.cut_shape[3 * k + 1], g = s.cut_shape[3 * k + 2];
    const double* blk = s.cut_data + s.cut_offsets[k];
    const double* xr = blk;
    const double* yz = blk + nr;
    if (g != group) {
      if (r < xr[0] || r > xr[nr - 1] || gz < yz[0] || gz > yz[nzc - 1]) {
        silent = true;
        return 0.0;
      }
      group = g;
      ci = crp_bisect(xr, nr, r);
      cj = crp_bisect(yz, nzc, gz);
      const double hx = (xr[ci + 1] - xr[ci]) * (xr[ci + 1] - xr[ci]) / 6.0;
      wxa = (xr[ci + 1] - r) / (xr[ci + 1] - xr[ci]);
      wxb = 1.0 - wxa;
      wxc = (wxa * wxa * wxa - wxa) * hx;
      wxd = (wxb * wxb * wxb - wxb) * hx;
      const double hy = (yz[cj + 1] - yz[cj]) * (yz[cj + 1] - yz[cj]) / 6.0;
      wya = (yz[cj + 1] - gz) / (yz[cj + 1] - yz[cj]);
      wyb = 1.0 - wya;
      wyc = (wya * wya * wya - wya) * hy;
      wyd = (wyb * wyb * wyb - wyb) * hy;
    }
    const long plane = static_cast<long>(nr) * nzc;
    const double* f = blk + nr + nzc;
    const double* fxx = f + plane;
    const double* fyy = fxx + plane;
    const double* fxxyy = fyy + plane;
    const long p00 = ci + static_cast<long>(nr) * cj, p10 = p00 + 1, p01 = p00 + nr, p11 = p01 + 1;
    cv[k] = wxa * (wya * f[p00] + wyb * f[p01] + wyc * fyy[p00] + wyd * fyy[p01]) +
            wxb * (wya * f[p10] + wyb * f[p11] + wyc * fyy[p10] + wyd * fyy[p11]) +
            wxc * (wya * fxx[p00] + wyb * fxx[p01] + wyc * fxxyy[p00] + wyd * fxxyy[p01]) +
            wxd * (wya * fxx[p10] + wyb * fxx[p11] + wyc * fxxyy[p10] + wyd * fxxyy[p11]);
  }
  const double c2t = cos(2.00 * theta), s2t = sin(2.00 * theta), c4t = cos(4.00 * theta);
  double iv = 0.00;
  if (s.square) {  // the 8 sites of C4v (RRINT) and SIS8x8, values only
    auto top = [&](const double* c) {
      const double v90 = 0.5 * (c[1] + c[2]) + 0.5 * (c[1] - c[2]) * cos(4.0 * phi);
      return 0.5 * (c[0] + v90) + 0.5 * (c[0] - v90) * c2t;
    };
    auto bridge = [&](const double* c, double ph) {
      const double v90 = (0.25 * c[1] + 0.5 * c[2] + 0.25 * c[3]) + 0.5 * (c[1] - c[3]) * cos(2.0 * ph) +
                         (0.25 * c[1] - 0.5 * c[2] + 0.25 * c[3]) * cos(4.0 * ph);
      return 0.5 * (c[0] + v90) + 0.5 * (c[0] - v90) * c2t;
    };
    auto t2h = [&](const double* c, double ph) {
      const double p45 = ph - kPi / 4.0, p135 = ph + kPi - kPi / 4.0;
      const double cc0 = 0.25 * (c[1] + c[2] - 2.0 * c[5]);
      const double b0 = -c[5] + c[2] - 2.0 * cc0;
      const double a0 = c[2] - b0 - cc0;
      const double v45 = a0 + b0 * cos(p45) + cc0 * cos(2.0 * p45);
      const double v135 = a0 + b0 * cos(p135) + cc0 * cos(2.0 * p135);
      const double v90 = 0.5 * (c[3] + c[4]) + 0.5 * (c[3] - c[4]) * cos(2.0 * p45);
      return 0.25 * (c[0] + v45 + v90 + v135) + 0.5 * (c[0] - v90) * c2t + 0.5 * (v45 - v135) * s2t +
             0.25 * (c[0] + v90 - v45 - v135) * c4t;
    };
    const double p1 = top(cv), p2 = bridge(cv + 3, phi), p5 = bridge(cv + 3, phi + kPi / 2.0);
    const double p3 = top(cv + 7), p4 = t2h(cv + 10, phi), p6 = t2h(cv + 10, phi - kPi / 2.0);
    const double p7 = t2h(cv + 10, phi - kPi), p8 = t2h(cv + 10, phi + kPi / 2.0);
    const double fac = 2.0 * kPi / s.a;
    iv = 1.0 / 8.0 * (p1 + p2 + p5 + p3 + p4 + p6 + p7 + p8) * 1.0 + 1.0 / 4.0 * (p1 - p2 + p5 - p3) * cos(fac * gx) +
         1.0 / 4.0 * (p1 + p2 - p5 - p3) * cos(fac * gy) + 1.0 / 4.0 * (p4 - p6 - p7 + p8) * sin(fac * gx) +
         1.0 / 4.0 * (p4 + p6 - p7 - p8) * sin(fac * gy) +
         1.0 / 8.0 * (p1 - p2 - p5 + p3 - p4 + p6 - p7 + p8) * cos(fac * (gx + gy)) +
         1.0 / 8.0 * (p1 - p2 - p5 + p3 + p4 - p6 + p7 - p8) * cos(fac * (gx - gy)) +
         1.0 / 16.0 * (p1 + p2 + p5 + p3 - p4 - p6 - p7 - p8) * (cos(fac * (2.0 * gx)) + cos(fac * (2.0 * gy)));
  } else {
  // the 12 Fourier sites (RRINTII) and the Fourier interpolation (SISII, FOUR6II)
  constexpr double pis3 = kPi / 3.00, pi2s3 = 2.00 * pis3, pis6 = kPi / 6.00;
  double site[12];
  {
    const double c6 = cos(6.00 * phi);
    const double v90 = 0.500 * ((cv[1] + cv[2]) + (cv[1] - cv[2]) * c6);
    site[0] = 0.500 * ((cv[0] + v90) + (cv[0] - v90) * c2t);
  }
  site[1] = crp_bridge(cv + 3, c2t, phi);
  site[4] = crp_bridge(cv + 3, c2t, phi + pis3);
  site[5] = crp_bridge(cv + 3, c2t, phi - pis3);
  if (!s.c3v) {
    site[2] = crp_hollow(cv + 7, c2t, s2t, c4t, phi);
    site[8] = crp_hollow(cv + 7, c2t, s2t, c4t, phi + kPi);
    site[3] = crp_t2(cv + 12, c2t, s2t, c4t, phi, pis6, 5, false);
    site[6] = crp_t2(cv + 12, c2t, s2t, c4t, phi - pi2s3, pis6, 5, false);
    site[7] = crp_t2(cv + 12, c2t, s2t, c4t, phi + pi2s3, pis6, 5, false);
    site[9] = crp_t2(cv + 12, c2t, s2t, c4t, phi + kPi, pis6, 5, false);
    site[10] = crp_t2(cv + 12, c2t, s2t, c4t, phi + pis3, pis6, 5, false);
    site[11] = crp_t2(cv + 12, c2t, s2t, c4t, phi - pis3, pis6, 5, false);
  } else {
    site[2] = crp_hollow(cv + 7, c2t, s2t, c4t, phi);
    site[3] = crp_t2(cv + 12, c2t, s2t, c4t, phi, pis6, 4, false);
    site[6] = crp_t2(cv + 12, c2t, s2t, c4t, phi - pi2s3, pis6, 4, false);
    site[7] = crp_t2(cv + 12, c2t, s2t, c4t, phi + pi2s3, pis6, 4, false);
    site[10] = crp_t2(cv + 18, c2t, s2t, c4t, phi, 11.00 * pis6, 4, true);
    site[9] = crp_t2(cv + 18, c2t, s2t, c4t, phi + pi2s3, 11.00 * pis6, 4, true);
    site[11] = crp_t2(cv + 18, c2t, s2t, c4t, phi - pi2s3, 11.00 * pis6, 4, true);
    site[8] = crp_hollow(cv + 24, c2t, s2t, c4t, phi);
  }
  const double fac = 2.00 * kPi / s.a;
  const double p = fac * (gx - gy * kUsq3), q = fac * (2.00 * gy * kUsq3);
  double fb[12];
  fb[0] = 1.00;
  fb[1] = cos(p);
  fb[2] = cos(q);
  fb[3] = cos(p + q);
  fb[4] = sin(p);
  fb[5] = sin(q);
  fb[6] = sin(p + q);
  fb[7] = cos(2.00 * p + q);
  fb[8] = cos(p + 2.00 * q);
  fb[9] = cos(p - q);
  fb[10] = cos(2.00 * p) + cos(2.00 * q) + cos(2.00 * (p + q));
  fb[11] = -sin(2.00 * p) - sin(2.00 * q) + sin(2.00 * (p + q));
  for (int j = 0; j < 12; ++j) {
    double c = 0.00;
    for (int k = 0; k < 12; ++k) c = c + kCrpMolecule[k + 12 * j] * site[k];
    iv = iv + c * fb[j];
  }
  }
  const double va = crp_atom(s, a[0], a[1], a[2], bad);
  const double vb = crp_atom(s, b[0], b[1], b[2], bad);
  if (gz <= s.zso_min) return va + vb + iv;
  const double sw = 0.500 + 0.500 * cos((gz - s.zso_min) * kPi / (s.zso_max - s.zso_min));
  const double v1 = gz <= s.zdep[nz - 1] ? crp_splint(s.zdep, nz, gz) : s.zdep[2 * nz - 1];
  if (gz < s.zso_min) bad = true;
  const double gas = crp_splint(s.h2, nh, r) + v1;
  return (va + vb + iv) * sw + (1.00 - sw) * gas;
}

// The QD wrapper: V in eV at the centre (u, v skewed), Z, r, theta, phi; NaN where the host stops.
__device__ inline double crp_qd(const CrpDev& s, double u, double v, double z, double r, double theta, double phi) {
  using namespace spo::crp_tables;
  constexpr double delta = 1.0e-12;
  constexpr double ev = 27.2113961;  // the CRP library's hartree in eV (to_energy: x * (1.0 * 27.2113961))
  double rr = r, zz = z;
  if (rr < s.r_min + delta) {
    rr = s.r_min + delta;
  } else if (rr > s.r_max - delta) {
    rr = s.r_max - delta;
  }
  if (zz < s.z_min + delta) zz = s.z_min + delta;
  // [[1, cos g], [0, sin g]] (u, v); cos(pi/2) as gfortran folds it for C4v
  const double cg = s.square ? 0x1.1a62633145c07p-54 : kCosPi3, sg = s.square ? 1.0 : kSinPi3;
  const double x = 1.0 * u + cg * v;
  const double y = 0.0 * u + sg * v;
  const double st = sin(theta), ctt = cos(theta), cph = cos(phi), sph = sin(phi);
  const double a[3] = {x - 0.5 * rr * st * cph, y - 0.5 * rr * st * sph, zz - 0.5 * rr * ctt};
  const double b[3] = {x + 0.5 * rr * st * cph, y + 0.5 * rr * st * sph, zz + 0.5 * rr * ctt};
  bool silent = false, bad = false;
  double m = crp_molecule(s, a, b, silent, bad);
  if (bad) return __builtin_nan("");
  if (silent) m = 0.00;
  return m * (1.0 * ev) - s.h2_asymptote * (1.0 * ev);
}

// The SCM correction (spo_hip_scm_correction): the pair functions of the slab's top two layers
// (src/potential/scm/scm_correction.cpp; Fortran SCM_Potential.f90), the same operand order as host/scm.cpp.
struct ScmDev {
  int on, rescale;
  double scale, cos_skew, sin_skew;
  double layers[2][18];
  int n_atoms;
  const double* atoms;
};

// -(p3 + q (p4 + q (p5 + q p6))) exp(-p2 q); t = p1..p8, C
__device__ inline double scm_unswitched(const double* t, double q) {
  const double cubic = t[2] + q * (t[3] + q * (t[4] + q * t[5]));
  return -(cubic * exp(-(t[1] * q)));
}

__device__ inline void scm_terms(const double* layer, double r, double* t) {
  double c = r;
  if (r < layer[16]) {
    c = layer[16];
  } else if (r > layer[17]) {
    c = layer[17];
  }
  for (int k = 0; k < 8; ++k) t[k] = layer[8 + k] * c + layer[k];
  t[8] = scm_unswitched(t, t[7] - t[0]);
}

__device__ inline double scm_pair(const double* t, double d) {
  if (d > t[7]) return t[8];
  const double f = scm_unswitched(t, d - t[0]);
  if (!(d > t[6])) return f;
  const double phase = ((d - t[7]) * spo::crp_tables::kPi) / (t[7] - t[6]);
  const double rho = 0.5 * cos(phase) + 0.5;
  return (1.0 - rho) * f + rho * t[8];
}

__device__ inline double scm_distance(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt((dx * dx + dy * dy) + dz * dz);
}

// dV (hartree) at the Cartesian centre (x, y), Z, r, theta, phi
__device__ inline double scm_delta(const ScmDev& m, double x, double y, double z, double r, double theta, double phi) {
  const double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
  const double h = 0.5 * r;
  const double o[3] = {(h * st) * cp, (h * st) * sp, h * ct};
  double hd[2][3] = {{x - o[0], y - o[1], z - o[2]}, {x + o[0], y + o[1], z + o[2]}};
  double hi[2][3];
  if (m.rescale) {
    const double bond[3] = {hd[1][0] - hd[0][0], hd[1][1] - hd[0][1], hd[1][2] - hd[0][2]};
    const double c[3] = {((hd[0][0] + hd[1][0]) / 2.0) / m.scale, ((hd[0][1] + hd[1][1]) / 2.0) / m.scale,
                         (hd[0][2] + hd[1][2]) / 2.0};
    for (int k = 0; k < 3; ++k) {
      hi[0][k] = c[k] - 0.5 * bond[k];
      hi[1][k] = c[k] + 0.5 * bond[k];
    }
  } else {
    for (int k = 0; k < 3; ++k) {
      hi[0][k] = hd[0][k];
      hi[1][k] = hd[1][k];
    }
  }
  double t[2][9];
  scm_terms(m.layers[0], r, t[0]);
  scm_terms(m.layers[1], r, t[1]);
  double dv = +0.0;
  for (int i = 0; i < m.n_atoms; ++i) {
    const double* at = m.atoms + 7 * i;
    const double* tl = t[at[6] == 1.0 ? 0 : 1];
    for (int a = 0; a < 2; ++a) {
      const double f_ideal = scm_pair(tl, scm_distance(hi[a], at + 3));
      const double f_distorted = scm_pair(tl, scm_distance(hd[a], at));
      dv = (dv - f_ideal) + f_distorted;
    }
  }
  return dv;
}

// The CRP surface, with the SCM correction when m.on, in eV at skewed (u, v)
__device__ inline double crp_scm_qd(const CrpDev& s, const ScmDev& m, double u, double v, double z, double r,
                                    double theta, double phi) {
  if (!m.on) return crp_qd(s, u, v, z, r, theta, phi);
  const double x = u + m.cos_skew * v, y = m.sin_skew * v;
  const double v_crp = crp_qd(s, u / m.scale, v / m.scale, z, r, theta, phi);
  return v_crp + scm_delta(m, x, y, z, r, theta, phi) / 0.036749308286;
}

__global__ __launch_bounds__(256) void crp_points_kernel(CrpDev s, ScmDev m, long n, const double* points,
                                                         double* values) {
  const long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double* p = points + 6 * i;
  values[i] = crp_scm_qd(s, m, p[0], p[1], p[2], p[3], p[4], p[5]);
}

}  // namespace spo_dev

#pragma clang fp contract(on)
