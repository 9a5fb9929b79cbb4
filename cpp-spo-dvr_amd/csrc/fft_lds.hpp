// Mixed-radix Stockham FFTs over sequences staged in LDS, executed by a whole workgroup.
//
// A batch of `count` sequences of length n lives in LDS; element i of sequence s sits at
//   base(s) + i * elem_stride,  base(s) = (s % inner) * inner_stride + (s / inner) * outer_stride,
// so one call covers the rows or the columns of a 2D tile, or the phi rows of a DVR tile. Each stage of radix R
// is a read-all / barrier / write-all round (in place): a thread holds up to 16 values in registers, whole
// sequences are assigned to a round, so no write of a round touches a value another thread still has to read.
// Twiddles come from one table W[m] = exp(-2 pi i m / n) (built on the host in long double); the exponent sign
// only conjugates them. No scaling.
#pragma once

#include <hip/hip_runtime.h>

#include "rounding.hpp"

namespace spo_dev {

// Aligned to its size, so that a complex value moves as one 16-byte (FP64) or 8-byte (FP32) access: in LDS one
// ds_read_b128 / ds_write_b128 instead of a pair of 8-byte accesses, which use only half of the banks per pass.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T>
__device__ __forceinline__ cx<T> scale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// a * (sign * i)
template <typename T>
__device__ __forceinline__ cx<T> times_i(cx<T> a, int sign) { return sign < 0 ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x}; }

// w v for a complex64 factor held as hi + lo: hi the nearest complex64 value of the FP64 factor, lo the remainder
// in FP32. The lo products are formed first and the hi products fused on top of them, so the sum is rounded once
// more at most and the factor takes part with about twice the float precision, modulus and phase. A correction
// added after the product is rounded (p + eps p) does not do this: a relative change below half a unit in the last
// place of p rounds back to p. FP64: hi v.
template <typename T>
__device__ __forceinline__ cx<T> mul_hilo(cx<T> hi, cx<T> lo, cx<T> v) {
  if constexpr (sizeof(T) == 8) {
    return hi * v;
  } else {
    const float re = fmaf(hi.x, v.x, fmaf(-hi.y, v.y, fmaf(lo.x, v.x, -lo.y * v.y)));
    const float im = fmaf(hi.x, v.y, fmaf(hi.y, v.x, fmaf(lo.x, v.y, lo.y * v.x)));
    return {re, im};
  }
}

// Plan of one length: radices in the order of the stages and the twiddle table.
struct FftPlan {
  int n;
  int n_stages;
  int radix[12];
  const void* table;  // cx<T>[n], exp(-2 pi i m / n)
  const void* eps;    // complex64: T[n], 1 / |table[m]| - 1 (applied through mul_hilo); FP64: unused
};

struct SeqLayout {
  int count;         // number of sequences
  int inner;         // sequences per outer step
  int inner_stride;  // LDS elements between consecutive inner sequences
  int outer_stride;
  int elem_stride;   // LDS elements between consecutive elements of a sequence
  __device__ __forceinline__ int base(int s) const { return (s % inner) * inner_stride + (s / inner) * outer_stride; }
};

// Odd-prime DFT by the symmetric pairs (q, R - q): c_q = cos(2 pi q / R), s_q = sin(2 pi q / R).
template <int R, typename T>
__device__ __forceinline__ void odd_dft(cx<T>* v, int sign, const double* c, const double* s) {
  constexpr int H = (R - 1) / 2;
  cx<T> a[H], b[H];
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    a[q - 1] = v[q] + v[R - q];
    b[q - 1] = v[q] - v[R - q];
  }
  cx<T> out[R];
  out[0] = v[0];
#pragma unroll
  for (int q = 0; q < H; ++q) out[0] = out[0] + a[q];
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    cx<T> re = v[0];
    cx<T> im = {T(0), T(0)};
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      const int idx = (k * q) % R;
      const int qq = idx <= H ? idx : R - idx;       // cos(2 pi idx / R) = c[qq]
      const double sgn = idx <= H ? 1.0 : -1.0;      // sin(2 pi idx / R) = sgn * s[qq]
      // the constants take part with split precision in FP32 (fma_const), so the DFT stays unitary to ~1e-14
      re = {fma_const(c[qq - 1], a[q - 1].x, re.x), fma_const(c[qq - 1], a[q - 1].y, re.y)};
      im = {fma_const(sgn * s[qq - 1], b[q - 1].x, im.x), fma_const(sgn * s[qq - 1], b[q - 1].y, im.y)};
    }
    // y_k = re + sign i im, y_{R-k} = re - sign i im
    const cx<T> jim = times_i(im, sign);
    out[k] = re + jim;
    out[R - k] = re - jim;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = out[k];
}

__device__ __constant__ const double kCos5[2] = {0.30901699437494745, -0.8090169943749473};
__device__ __constant__ const double kSin5[2] = {0.9510565162951535, 0.5877852522924732};
__device__ __constant__ const double kCos7[3] = {0.6234898018587336, -0.22252093395631434, -0.900968867902419};
__device__ __constant__ const double kSin7[3] = {0.7818314824680298, 0.9749279121818236, 0.43388373911755823};
__device__ __constant__ const double kCos11[5] = {0.8412535328311812, 0.41541501300188644, -0.142314838273285,
                                                  -0.654860733945285, -0.9594929736144974};
__device__ __constant__ const double kSin11[5] = {0.5406408174555976, 0.9096319953545183, 0.9898214418809328,
                                                  0.7557495743542583, 0.28173255684142967};

template <int R, typename T>
__device__ __forceinline__ void dft(cx<T>* v, int sign) {
  if constexpr (R == 2) {
    const cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (R == 3) {
    // sqrt(3)/2 with split precision in FP32: a rounded float would scale the difference modes by 1 - 1.8e-8 at
    // every transform.
    const double s = 0.8660254037844386 * sign;
    const cx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
    const cx<T> m = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    const cx<T> js = {mul_const(-s, t2.y), mul_const(s, t2.x)};
    v[0] = v[0] + t1;
    v[1] = m + js;
    v[2] = m - js;
  } else if constexpr (R == 4) {
    const cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
    const cx<T> b0 = v[1] + v[3], b1 = times_i(v[1] - v[3], sign);
    v[0] = a0 + b0;
    v[2] = a0 - b0;
    v[1] = a1 + b1;
    v[3] = a1 - b1;
  } else if constexpr (R == 8) {
    // Two radix-4 transforms of the even and odd elements, then the radix-2 combination.
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft<4, T>(e, sign);
    dft<4, T>(o, sign);
    // w^k = exp(sign i pi k / 4): w^1 = h (1 + sign i), w^3 = h (-1 + sign i), h = sqrt(1/2) with split
    // precision in FP32 (mul_const): one rounded float h gives |w|^2 = 2 h^2 = 1 - 3.4e-8, the same shrink of these
    // outputs at every transform
    constexpr double h = 0.7071067811865476;
    const T sg = T(sign);
    o[1] = {mul_const(h, o[1].x - sg * o[1].y), mul_const(h, o[1].y + sg * o[1].x)};
    o[2] = times_i(o[2], sign);
    o[3] = {mul_const(h, -o[3].x - sg * o[3].y), mul_const(h, -o[3].y + sg * o[3].x)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = e[k] + o[k];
      v[k + 4] = e[k] - o[k];
    }
  } else if constexpr (R == 5) {
    odd_dft<5, T>(v, sign, kCos5, kSin5);
  } else if constexpr (R == 7) {
    odd_dft<7, T>(v, sign, kCos7, kSin7);
  } else if constexpr (R == 11) {
    odd_dft<11, T>(v, sign, kCos11, kSin11);
  }
}

constexpr int kMaxValues = 16;

template <int R, typename T>
__device__ void lds_stage(cx<T>* buf, const SeqLayout& lay, int n, int span, const cx<T>* table, const void* eps,
                          int sign) {
  constexpr int NB = kMaxValues / R;
  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const int m = n / R;
  const int tw_step = n / (span * R);
  int per_round = (nthreads * NB) / m;
  if (per_round < 1) per_round = 1;
  for (int s0 = 0; s0 < lay.count; s0 += per_round) {
    const int n_seq = min(per_round, lay.count - s0);
    const int total = n_seq * m;
    cx<T> v[NB][R];
    int out_base[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int b = tid + q * nthreads;
      out_base[q] = -1;
      if (b < total) {
        const int s = s0 + b / m;
        const int j = b % m;
        const int k = j % span;
        const int base = lay.base(s);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          cx<T> x = buf[base + (j + r * m) * lay.elem_stride];
          if (r > 0 && k > 0) {
            cx<T> w = table[r * k * tw_step];
            if (sign > 0) w = conj(w);
            if constexpr (sizeof(T) == 4) {
              const T e = static_cast<const T*>(eps)[r * k * tw_step];
              x = mul_hilo(w, {e * w.x, e * w.y}, x);
            } else {
              x = x * w;
            }
          }
          v[q][r] = x;
        }
        dft<R, T>(v[q], sign);
        out_base[q] = base + ((j / span) * span * R + k) * lay.elem_stride;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q)
      if (out_base[q] >= 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) buf[out_base[q] + r * span * lay.elem_stride] = v[q][r];
      }
    __syncthreads();
  }
}

// The whole transform (every stage). All threads of the workgroup must call it.
template <typename T>
__device__ void lds_fft(cx<T>* buf, const SeqLayout& lay, const FftPlan& plan, int sign) {
  const cx<T>* table = static_cast<const cx<T>*>(plan.table);
  int span = 1;
  for (int st = 0; st < plan.n_stages; ++st) {
    const int R = plan.radix[st];
    switch (R) {
      case 2: lds_stage<2, T>(buf, lay, plan.n, span, table, plan.eps, sign); break;
      case 3: lds_stage<3, T>(buf, lay, plan.n, span, table, plan.eps, sign); break;
      case 4: lds_stage<4, T>(buf, lay, plan.n, span, table, plan.eps, sign); break;
      case 5: lds_stage<5, T>(buf, lay, plan.n, span, table, plan.eps, sign); break;
      case 7: lds_stage<7, T>(buf, lay, plan.n, span, table, plan.eps, sign); break;
      case 8: lds_stage<8, T>(buf, lay, plan.n, span, table, plan.eps, sign); break;
      case 11: lds_stage<11, T>(buf, lay, plan.n, span, table, plan.eps, sign); break;
      default: break;
    }
    span *= R;
  }
}

}  // namespace spo_dev
