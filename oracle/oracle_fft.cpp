// TEST INFRASTRUCTURE — part of the CPU oracle (see oracle_fft.hpp).
//
// Stockham autosort over `batch` interleaved sequences. The work arrays hold real and imaginary parts apart
// (element i of sequence b at [i * batch + b]), so the innermost loop over b is unit-stride and vectorises.
#include <string>
#include "oracle_fft.hpp"

#include <cmath>
#include <stdexcept>

namespace oracle {

namespace {

constexpr long double kPiL = 3.14159265358979323846264338327950288L;

// In-register DFT of radix R with exponent sign `sign`; (cr[q], ci[q]) = exp(sign 2 pi i q / R).
template <int R>
inline void dft(double* vr, double* vi, int sign, const double* cr, const double* ci) {
  if constexpr (R == 2) {
    const double ar = vr[0], ai = vi[0], br = vr[1], bi = vi[1];
    vr[0] = ar + br;
    vi[0] = ai + bi;
    vr[1] = ar - br;
    vi[1] = ai - bi;
  } else if constexpr (R == 4) {
    const double a0r = vr[0] + vr[2], a0i = vi[0] + vi[2];
    const double a1r = vr[0] - vr[2], a1i = vi[0] - vi[2];
    const double b0r = vr[1] + vr[3], b0i = vi[1] + vi[3];
    const double dr = vr[1] - vr[3], di = vi[1] - vi[3];
    // (dr, di) * sign i
    const double b1r = sign < 0 ? di : -di;
    const double b1i = sign < 0 ? -dr : dr;
    vr[0] = a0r + b0r;
    vi[0] = a0i + b0i;
    vr[2] = a0r - b0r;
    vi[2] = a0i - b0i;
    vr[1] = a1r + b1r;
    vi[1] = a1i + b1i;
    vr[3] = a1r - b1r;
    vi[3] = a1i - b1i;
  } else {
    // Odd prime R: pairs (q, R - q) share cos and flip sin, so y_k and y_{R-k} come from the same sums.
    constexpr int H = (R - 1) / 2;
    double ar[H], ai[H], br[H], bi[H];
    for (int q = 1; q <= H; ++q) {
      ar[q - 1] = vr[q] + vr[R - q];
      ai[q - 1] = vi[q] + vi[R - q];
      br[q - 1] = vr[q] - vr[R - q];
      bi[q - 1] = vi[q] - vi[R - q];
    }
    double orr[R], oi[R];
    orr[0] = vr[0];
    oi[0] = vi[0];
    for (int q = 0; q < H; ++q) {
      orr[0] += ar[q];
      oi[0] += ai[q];
    }
    for (int k = 1; k <= H; ++k) {
      double pr = vr[0], pi = vi[0], qr = 0.0, qi = 0.0;
      for (int q = 1; q <= H; ++q) {
        const int idx = (k * q) % R;  // cr[idx] = cos(2 pi idx / R), ci[idx] = sign sin(2 pi idx / R)
        pr += ar[q - 1] * cr[idx];
        pi += ai[q - 1] * cr[idx];
        qr += br[q - 1] * ci[idx];
        qi += bi[q - 1] * ci[idx];
      }
      // y_k = p + i q, y_{R-k} = p - i q
      orr[k] = pr - qi;
      oi[k] = pi + qr;
      orr[R - k] = pr + qi;
      oi[R - k] = pi - qr;
    }
    for (int k = 0; k < R; ++k) {
      vr[k] = orr[k];
      vi[k] = oi[k];
    }
  }
}

template <int R>
void stage(const double* in_r, const double* in_i, double* out_r, double* out_i, int n, int span, long batch,
           int sign, const cplx* tw, const cplx* roots) {
  const int m = n / R;
  const int blocks = m / span;
  double cr[R], ci[R];
  for (int q = 0; q < R; ++q) {
    cr[q] = roots[q].real();
    ci[q] = sign < 0 ? roots[q].imag() : -roots[q].imag();
  }
  for (int p = 0; p < blocks; ++p)
    for (int k = 0; k < span; ++k) {
      const int j = p * span + k;
      const cplx* w = tw + static_cast<long>(k) * (R - 1);
      double tr[R], ti[R];
      tr[0] = 1.0;
      ti[0] = 0.0;
      for (int r = 1; r < R; ++r) {
        tr[r] = w[r - 1].real();
        ti[r] = sign < 0 ? w[r - 1].imag() : -w[r - 1].imag();
      }
      const long o = static_cast<long>(p) * span * R + k;
#pragma GCC ivdep
      for (long b = 0; b < batch; ++b) {
        double vr[R], vi[R];
        for (int r = 0; r < R; ++r) {
          const double xr = in_r[static_cast<long>(j + r * m) * batch + b];
          const double xi = in_i[static_cast<long>(j + r * m) * batch + b];
          vr[r] = xr * tr[r] - xi * ti[r];
          vi[r] = xr * ti[r] + xi * tr[r];
        }
        dft<R>(vr, vi, sign, cr, ci);
        for (int r = 0; r < R; ++r) {
          out_r[(o + static_cast<long>(r) * span) * batch + b] = vr[r];
          out_i[(o + static_cast<long>(r) * span) * batch + b] = vi[r];
        }
      }
    }
}

}  // namespace

Fft1d::Fft1d(int n) : n_(n) {
  if (n < 1) throw std::invalid_argument("Fft1d: length < 1");
  int m = n;
  while (m % 4 == 0) {
    radices_.push_back(4);
    m /= 4;
  }
  for (int p : {2, 3, 5, 7, 11, 13})
    while (m % p == 0) {
      radices_.push_back(p);
      m /= p;
    }
  if (m != 1) throw std::invalid_argument("Fft1d: length " + std::to_string(n) + " has a prime factor above 13");
  // Per stage: exp(-2 pi i r k / (span R)) for k < span, r = 1..R-1; then the roots of each radix.
  int span = 1;
  for (const int radix : radices_) {
    for (int k = 0; k < span; ++k)
      for (int r = 1; r < radix; ++r) {
        const long double a = -2.0L * kPiL * r * k / (static_cast<long double>(span) * radix);
        table_.emplace_back(static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a)));
      }
    for (int q = 0; q < radix; ++q) {
      const long double a = -2.0L * kPiL * q / radix;
      roots_.emplace_back(static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a)));
    }
    span *= radix;
  }
}

void Fft1d::run(cplx* x, int stride, int sign, cplx* work) const { run_batch(x, stride, 1, sign, work); }

void Fft1d::run_batch(cplx* x, long stride, int batch, int sign, cplx* work) const {
  const int n = n_;
  const long nb = static_cast<long>(n) * batch;
  double* ar = reinterpret_cast<double*>(work);
  double* ai = ar + nb;
  double* br = ai + nb;
  double* bi = br + nb;
  const double* xd = reinterpret_cast<const double*>(x);
  for (int i = 0; i < n; ++i)
    for (int e = 0; e < batch; ++e) {
      ar[static_cast<long>(i) * batch + e] = xd[2 * (i * stride + e)];
      ai[static_cast<long>(i) * batch + e] = xd[2 * (i * stride + e) + 1];
    }
  int span = 1;
  long offset = 0;
  long root_offset = 0;
  for (const int radix : radices_) {
    const cplx* tw = table_.data() + offset;
    const cplx* roots = roots_.data() + root_offset;
    switch (radix) {
      case 2: stage<2>(ar, ai, br, bi, n, span, batch, sign, tw, roots); break;
      case 3: stage<3>(ar, ai, br, bi, n, span, batch, sign, tw, roots); break;
      case 4: stage<4>(ar, ai, br, bi, n, span, batch, sign, tw, roots); break;
      case 5: stage<5>(ar, ai, br, bi, n, span, batch, sign, tw, roots); break;
      case 7: stage<7>(ar, ai, br, bi, n, span, batch, sign, tw, roots); break;
      case 11: stage<11>(ar, ai, br, bi, n, span, batch, sign, tw, roots); break;
      default: stage<13>(ar, ai, br, bi, n, span, batch, sign, tw, roots); break;
    }
    offset += static_cast<long>(span) * (radix - 1);
    root_offset += radix;
    std::swap(ar, br);
    std::swap(ai, bi);
    span *= radix;
  }
  double* xo = reinterpret_cast<double*>(x);
  for (int i = 0; i < n; ++i)
    for (int e = 0; e < batch; ++e) {
      xo[2 * (i * stride + e)] = ar[static_cast<long>(i) * batch + e];
      xo[2 * (i * stride + e) + 1] = ai[static_cast<long>(i) * batch + e];
    }
}

}  // namespace oracle
