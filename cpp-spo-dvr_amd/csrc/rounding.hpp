// Rounding of complex factors to complex64 that keeps their modulus.
//
// A factor of the split-operator step that is applied every time step (a kinetic or rotational phase, exp(i A),
// an FFT twiddle) multiplies the wave by the same complex64 value thousands of times. Rounding its two components
// to the nearest floats independently changes its modulus by up to about 6e-8 relative, always in the same
// direction for that factor, so the error grows linearly with the number of steps instead of as a random walk
// (2e4 steps: about 1e-3). Here the two floats are chosen among the neighbours of the nearest ones so that
// x^2 + y^2 matches the FP64 modulus as closely as possible, at a phase error of a few units in the last place
// (a phase error is unitary: it does not change the norm, and it moves an energy by about 1e-7 / dt).
#pragma once

#include <hip/hip_runtime.h>

#include <bit>
#include <cmath>
#include <cstdint>

namespace spo_dev {

// The float `k` representable values above f (below for k < 0); f finite and nonzero, no zero crossing.
__host__ __device__ constexpr float float_step(float f, int k) {
  const std::int32_t b = std::bit_cast<std::int32_t>(f);
  return std::bit_cast<float>(f > 0.0f ? b + k : b - k);
}

// (fx, fy) close to (x, y) with fx^2 + fy^2 as close as the search finds to x^2 + y^2 (the sum of two exactly
// representable squares, evaluated in FP64). One coordinate is stepped up to 8 values around its nearest float; the
// other is then solved for and rounded, and its two neighbours are tried; both roles are tried. Candidates farther
// than 1e-6 |z| from (x, y) are not taken. Measured over uniformly random phases: | |z'|^2 / |z|^2 - 1 | has median
// 6e-10 and 90th percentile 4e-9 (plain rounding: 3e-8 and 5e-8); phases within a few tenths of a degree of a
// multiple of 45 degrees (about 2 %) keep errors of plain size, since near them no nearby pair does better.
__host__ __device__ inline void round_keep_modulus(double x, double y, float& fx, float& fy) {
  const float x0 = static_cast<float>(x), y0 = static_cast<float>(y);
  fx = x0;
  fy = y0;
  const double m2 = x * x + y * y;
  if (!(m2 > 0.0) || !(m2 < 1e300)) return;
  const double tol = 1e-6 * sqrt(m2);
  double best = fabs(static_cast<double>(x0) * x0 + static_cast<double>(y0) * y0 - m2);
  auto consider = [&](float cx, float cy) {
    if (fabs(cx - x) + fabs(cy - y) > tol) return;
    const double e = fabs(static_cast<double>(cx) * cx + static_cast<double>(cy) * cy - m2);
    if (e < best) {
      best = e;
      fx = cx;
      fy = cy;
    }
  };
  for (int role = 0; role < 2; ++role) {
    const float a0 = role == 0 ? x0 : y0;          // stepped coordinate
    const double other = role == 0 ? y : x;        // solved coordinate (its sign is kept)
    if (a0 == 0.0f) continue;
    for (int i = -8; i <= 8; ++i) {
      const float a = float_step(a0, i);
      const double r = m2 - static_cast<double>(a) * a;
      if (r < 0.0) continue;
      const float b0 = static_cast<float>(copysign(sqrt(r), other));
      if (b0 == 0.0f) continue;
      for (int j = -1; j <= 1; ++j) {
        const float b = float_step(b0, j);
        if (role == 0)
          consider(a, b);
        else
          consider(b, a);
      }
    }
  }
}

// round_keep_modulus over a table of n complex values (interleaved re, im; times a real factor, in FP64 first),
// one value per thread.
__global__ inline void round_keep_modulus_kernel(const double* in, float* out, long n, double factor) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  round_keep_modulus(in[2 * i] * factor, in[2 * i + 1] * factor, out[2 * i], out[2 * i + 1]);
}

// A unit complex constant (c, s) as two floats whose squares sum to 1 as closely as a search over a few units in
// the last place finds (compile time), for the fixed twiddles inside the register DFTs.
struct UnitF32 {
  float c, s;
};

constexpr double ct_sqrt(double v) {
  if (v <= 0.0) return 0.0;
  double g = v > 1.0 ? v : 1.0;
  for (int i = 0; i < 80; ++i) g = 0.5 * (g + v / g);
  return g;
}

constexpr double ct_abs(double v) { return v < 0 ? -v : v; }

constexpr UnitF32 unit_f32(double c, double s) {
  const float c0 = static_cast<float>(c), s0 = static_cast<float>(s);
  UnitF32 best{c0, s0};
  double err = ct_abs(static_cast<double>(c0) * c0 + static_cast<double>(s0) * s0 - 1.0);
  if (c0 == 0.0f || s0 == 0.0f) return best;  // 1, i and their negatives are exact
  for (int role = 0; role < 2; ++role)
    for (int i = -6; i <= 6; ++i) {
      const float a = float_step(role == 0 ? c0 : s0, i);
      const double r = 1.0 - static_cast<double>(a) * a;
      if (r < 0.0) continue;
      const double other = role == 0 ? s : c;
      const float b0 = static_cast<float>(other < 0 ? -ct_sqrt(r) : ct_sqrt(r));
      for (int j = -1; j <= 1; ++j) {
        const float b = float_step(b0, j);
        const float cc = role == 0 ? a : b, ss = role == 0 ? b : a;
        const double e = ct_abs(static_cast<double>(cc) * cc + static_cast<double>(ss) * ss - 1.0);
        if (e < err) {
          err = e;
          best = {cc, ss};
        }
      }
    }
  return best;
}

// c x for a constant c: in FP32 as c_hi x + c_lo x (c_lo = c - c_hi), so that c takes part with about twice the
// float precision; a single rounded c would scale a DFT's outputs by the same wrong factor at every step.
template <typename T>
__device__ __forceinline__ T mul_const(double c, T x) {
  if constexpr (sizeof(T) == 8) {
    return static_cast<T>(c) * x;
  } else {
    const float hi = static_cast<float>(c);
    const float lo = static_cast<float>(c - static_cast<double>(hi));
    return fmaf(hi, x, lo * x);
  }
}

// acc + c x with the same treatment of c.
template <typename T>
__device__ __forceinline__ T fma_const(double c, T x, T acc) {
  if constexpr (sizeof(T) == 8) {
    return fma(static_cast<T>(c), x, acc);
  } else {
    const float hi = static_cast<float>(c);
    const float lo = static_cast<float>(c - static_cast<double>(hi));
    return fmaf(hi, x, fmaf(lo, x, acc));
  }
}

}  // namespace spo_dev
