// Complex64 forms of the constants and factors of the split-operator step.
//
// A factor that multiplies the wave at every time step (a kinetic or rotational phase, exp(i A), an FFT twiddle)
// is applied thousands of times with the same complex64 value, so its rounding errors add up linearly with the
// number of steps instead of as a random walk. Two kinds of error matter:
//   modulus: rounding the two components to the nearest floats changes |z| by up to about 6e-8 relative, always in
//     the same direction for that entry; a table of such factors carries a per-entry correction
//     eps = |z| / |fl(z)| - 1 in FP32 that the kernels apply to the product (mul_factor, twiddle);
//   phase: the nearest floats are also the smallest phase error (half a unit in the last place per component);
//     a phase error is an energy shift of about 1e-7 / dt that accumulates coherently. Choosing other nearby floats
//     whose squared sum matches |z|^2 better trades a few units of phase for the modulus: on a 3000-step
//     CRP + SCM run that moved the reaction probabilities by 1.4e-4, against 1.4e-5 for nearest rounding with
//     the eps correction. So the factor tables are rounded to nearest, and only the fixed twiddles inside the
//     register DFTs (unit_f32) trade phase for modulus: a DFT whose twiddles have unit modulus is exactly unitary
//     and its inverse applies the conjugates, so their phase errors cancel over the forward and backward pair.
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

// A table of n complex values (interleaved re, im) times a real factor (in FP64 first), rounded to the nearest
// complex64 values; one value per thread.
__global__ inline void round_to_complex64_kernel(const double* in, float* out, long n, double factor) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = static_cast<float>(in[2 * i] * factor);
  out[2 * i + 1] = static_cast<float>(in[2 * i + 1] * factor);
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
