// Compile-time-shaped Stockham FFTs over sequences staged in LDS (the fast path of the hot kernels).
//
// The length N, its factorisation into radices {2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 16} (fewest stages, larger
// radices first), the sequence layout and the thread count are template parameters, so every index computation
// divides by constants and every butterfly's values live in registers. A stage is one read-all / barrier /
// write-all round over whole sequences (in place). Twiddles between stages come from the table
// W[m] = exp(-2 pi i m / N) (host-built in long double); twiddles inside composite radices are compile-time
// constants. The exponent sign is a run-time argument and only conjugates.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "fft_lds.hpp"

namespace spo_dev {

// ---- compile-time unit roots (cos, sin)(2 pi m / R), exact at the quarter turns --------------------------------
struct CS {
  double c, s;
};

constexpr double ct_sin_series(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term = -term * x * x / ((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}
constexpr double ct_cos_series(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term = -term * x * x / ((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}
constexpr CS unit_root(int m, int R) {
  m %= R;
  if (m < 0) m += R;
  const int q = (4 * m) / R;  // quarter turns
  const int r = 4 * m - q * R;  // angle within the quarter: (pi/2) r / R
  constexpr double half_pi = 1.57079632679489661923;
  double c = 1.0, s = 0.0;
  if (r == 0) {
    c = 1.0;
    s = 0.0;
  } else if (2 * r <= R) {
    const double t = half_pi * r / R;
    c = ct_cos_series(t);
    s = ct_sin_series(t);
  } else {
    const double t = half_pi * (R - r) / R;
    c = ct_sin_series(t);
    s = ct_cos_series(t);
  }
  switch (q) {
    case 1: return {-s, c};
    case 2: return {-c, -s};
    case 3: return {s, -c};
    default: return {c, s};
  }
}

// ---- compile-time factorisation -------------------------------------------------------------------------------
struct Factorisation {
  int count = 0;
  int radix[8] = {};
};

constexpr bool is_radix(int r) {
  return r == 2 || r == 3 || r == 4 || r == 5 || r == 6 || r == 7 || r == 8 || r == 10 || r == 11 || r == 12 ||
         r == 16;
}

// Fewest stages; among those, the lexicographically largest radix sequence.
constexpr Factorisation factorise(int n) {
  Factorisation best;
  best.count = 99;
  if (n == 1) {
    best.count = 0;
    return best;
  }
  for (int r = 16; r >= 2; --r) {
    if (!is_radix(r) || n % r != 0) continue;
    const Factorisation rest = factorise(n / r);
    if (rest.count + 1 < best.count) {
      best.count = rest.count + 1;
      best.radix[0] = r;
      for (int i = 0; i < rest.count; ++i) best.radix[i + 1] = rest.radix[i];
    }
  }
  return best;
}

template <int N>
struct CtPlan {
  static constexpr Factorisation f = factorise(N);
  static_assert(f.count > 0 && f.count < 8, "FFT length not supported by the compile-time engine");
  static constexpr int stages = f.count;
  static constexpr int radix(int i) { return f.radix[i]; }
  static constexpr int span(int i) {
    int s = 1;
    for (int k = 0; k < i; ++k) s *= f.radix[k];
    return s;
  }
};

// ---- register DFTs --------------------------------------------------------------------------------------------
template <int... I, class F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F&& f) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

template <int R, typename T>
__device__ __forceinline__ void rdft(cx<T>* v, int sign);

// a * w_R^m (w_R = exp(sign 2 pi i / R)) for a compile-time m: the quarter turns are sign swaps, the odd eighth
// turns (1 +- i) / sqrt 2 two adds and two multiplications, the rest a full complex product.
template <int M, int R, typename T>
__device__ __forceinline__ cx<T> mul_unit_root(cx<T> a, int sign) {
  constexpr int m = ((M % R) + R) % R;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (4 * m == R) {
    return times_i(a, sign);
  } else if constexpr (2 * m == R) {
    return {-a.x, -a.y};
  } else if constexpr (4 * m == 3 * R) {
    return times_i(a, -sign);
  } else if constexpr ((8 * m) % R == 0) {
    // odd eighth turn: w = (cx + i sign sx) with |cx| = |sx| = sqrt(1/2); in FP32 sqrt(1/2) takes part with split
    // precision (mul_const): one rounded float h gives |w|^2 = 2 h^2 = 1 - 3.4e-8, the same shrink at every
    // transform, and no pair of floats near (h, h) has a modulus much closer to 1
    constexpr CS w = unit_root(m, R);
    constexpr double h = 0.70710678118654752440;
    const T sx = T(sign) * T(w.s > 0 ? 1 : -1);
    const T cxs = T(w.c > 0 ? 1 : -1);
    // (a.x + i a.y)(cxs + i sx) h
    return {mul_const(h, cxs * a.x - sx * a.y), mul_const(h, cxs * a.y + sx * a.x)};
  } else if constexpr (sizeof(T) == 8) {
    constexpr CS w = unit_root(m, R);
    return a * cx<T>{T(w.c), T(sign) * T(w.s)};
  } else {
    // FP32: a pair of floats of unit modulus (unit_f32), so that the twiddle does not scale its modes at every
    // transform
    constexpr CS w = unit_root(m, R);
    constexpr UnitF32 u = unit_f32(w.c, w.s);
    return a * cx<T>{u.c, T(sign) * u.s};
  }
}

// Length R = R1 R2 from DFTs of lengths R1 and R2.
// Coprime R1, R2 (Good-Thomas): x[(R2 n1 + R1 n2) mod R] -> DFT over n1 -> DFT over n2 -> X[k], k = k1 mod R1 and
// k2 mod R2 (Chinese remainder), with no twiddles at all.
// Otherwise (Cooley-Tukey): y[k1 + R1 k2] = sum_{n2} w_R^{n2 k1} (sum_{n1} x[R2 n1 + n2] w_R1^{n1 k1}) w_R2^{n2 k2}.
constexpr int ct_gcd(int a, int b) { return b == 0 ? a : ct_gcd(b, a % b); }
constexpr int ct_crt(int k1, int r1, int k2, int r2) {
  for (int k = 0; k < r1 * r2; ++k)
    if (k % r1 == k1 && k % r2 == k2) return k;
  return -1;
}

template <int R1, int R2, typename T>
__device__ __forceinline__ void rdft_composite(cx<T>* v, int sign) {
  constexpr int R = R1 * R2;
  cx<T> a[R2][R1];
  if constexpr (ct_gcd(R1, R2) == 1) {
    static_for<R2 * R1>([&](auto idx) {
      constexpr int n2 = decltype(idx)::value / R1;
      constexpr int n1 = decltype(idx)::value % R1;
      a[n2][n1] = v[(R2 * n1 + R1 * n2) % R];
    });
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) rdft<R1, T>(a[n2], sign);
    static_for<R1>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      cx<T> b[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) b[n2] = a[n2][k1];
      rdft<R2, T>(b, sign);
      static_for<R2>([&](auto k2c) {
        constexpr int k2 = decltype(k2c)::value;
        v[ct_crt(k1, R1, k2, R2)] = b[k2];
      });
    });
  } else {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) a[n2][n1] = v[R2 * n1 + n2];
      rdft<R1, T>(a[n2], sign);
    }
    // Internal twiddles w_R^{n2 k1}, n2, k1 >= 1, as compile-time constants.
    static_for<R2 * R1>([&](auto idx) {
      constexpr int n2 = decltype(idx)::value / R1;
      constexpr int k1 = decltype(idx)::value % R1;
      if constexpr (n2 > 0 && k1 > 0) a[n2][k1] = mul_unit_root<n2 * k1, R, T>(a[n2][k1], sign);
    });
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      cx<T> b[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) b[n2] = a[n2][k1];
      rdft<R2, T>(b, sign);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
    }
  }
}

template <int R, typename T>
__device__ __forceinline__ void rdft(cx<T>* v, int sign) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 7 || R == 8 || R == 11) {
    dft<R, T>(v, sign);
  } else if constexpr (R == 6) {
    rdft_composite<3, 2, T>(v, sign);
  } else if constexpr (R == 10) {
    rdft_composite<5, 2, T>(v, sign);
  } else if constexpr (R == 12) {
    rdft_composite<4, 3, T>(v, sign);
  } else if constexpr (R == 16) {
    rdft_composite<4, 4, T>(v, sign);
  }
}

// ---- one stage, whole sequences per round ---------------------------------------------------------------------
// Sequence s starts at (s % INNER) * IS + (s / INNER) * OS (INNER = 0: at s * IS); elements ES apart.
template <int INNER, int IS, int OS, int ES>
struct CtLayout {
  static constexpr int es = ES;
  static __device__ __forceinline__ int base(int s) {
    if constexpr (INNER == 0) {
      return s * IS;
    } else {
      return (s % INNER) * IS + (s / INNER) * OS;
    }
  }
};

template <typename T>
constexpr int ct_max_values() {
  return sizeof(T) == 8 ? 16 : 32;
}

// eps (complex64): the twiddle table's modulus corrections, 1 / |table[m]| - 1 (applied through mul_hilo).
template <typename T, int N, int STAGE, int THREADS, class Layout>
__device__ __forceinline__ void ct_stage(cx<T>* buf, int count, const cx<T>* table, const T* eps, int sign) {
  using P = CtPlan<N>;
  constexpr int R = P::radix(STAGE);
  constexpr int SPAN = P::span(STAGE);
  constexpr int M = N / R;
  constexpr int TW = N / (SPAN * R);
  constexpr int ES = Layout::es;
  constexpr int NB_MAX = ct_max_values<T>() / R > 0 ? ct_max_values<T>() / R : 1;
  constexpr int SEQ_PER_ROUND = (THREADS * NB_MAX) / M > 0 ? (THREADS * NB_MAX) / M : 1;
  const int tid = threadIdx.x;
  for (int s0 = 0; s0 < count; s0 += SEQ_PER_ROUND) {
    const int n_seq = min(SEQ_PER_ROUND, count - s0);
    const int total = n_seq * M;
    constexpr int NB = (SEQ_PER_ROUND * M + THREADS - 1) / THREADS;
    cx<T> v[NB][R];
    int out[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int b = tid + q * THREADS;
      out[q] = -1;
      if (b < total) {
        const int s = s0 + b / M;
        const int j = b % M;
        const int k = j % SPAN;
        const int base = Layout::base(s);
#pragma unroll
        for (int r = 0; r < R; ++r) v[q][r] = buf[base + (j + r * M) * ES];
        if constexpr (SPAN > 1) {
          if (k > 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
              cx<T> w = table[r * k * TW];
              if (sign > 0) w = conj(w);
              if constexpr (sizeof(T) == 4) {
                const T e = eps[r * k * TW];
                v[q][r] = mul_hilo(w, {e * w.x, e * w.y}, v[q][r]);
              } else {
                v[q][r] = v[q][r] * w;
              }
            }
          }
        }
        rdft<R, T>(v[q], sign);
        out[q] = base + ((j / SPAN) * SPAN * R + k) * ES;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q)
      if (out[q] >= 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) buf[out[q] + r * SPAN * ES] = v[q][r];
      }
    __syncthreads();
  }
}

template <typename T, int N, int THREADS, class Layout, int STAGE = 0>
__device__ __forceinline__ void ct_fft(cx<T>* buf, int count, const cx<T>* table, const T* eps, int sign) {
  if constexpr (STAGE < CtPlan<N>::stages) {
    ct_stage<T, N, STAGE, THREADS, Layout>(buf, count, table, eps, sign);
    ct_fft<T, N, THREADS, Layout, STAGE + 1>(buf, count, table, eps, sign);
  }
}

}  // namespace spo_dev
