// TEST INFRASTRUCTURE — part of the CPU oracle (see oracle_fft.hpp).
#include "oracle_fft.hpp"

#include <cmath>
#include <stdexcept>

namespace oracle {

namespace {

constexpr long double kPiL = 3.14159265358979323846264338327950288L;

struct C {
  double re, im;
};
inline C add(C a, C b) { return {a.re + b.re, a.im + b.im}; }
inline C sub(C a, C b) { return {a.re - b.re, a.im - b.im}; }
inline C mul(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C scale(C a, double s) { return {a.re * s, a.im * s}; }
// a * (sign i)
inline C rot90(C a, int sign) { return sign < 0 ? C{a.im, -a.re} : C{-a.im, a.re}; }

// In-register DFT of radix R with exponent sign `sign`; roots[q] = exp(sign 2 pi i q / R).
template <int R>
inline void dft(C* v, int sign, const C* roots) {
  if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
  } else if constexpr (R == 3) {
    const double c = -0.5, s = 0.86602540378443864676 * sign;
    const C t1 = add(v[1], v[2]), t2 = sub(v[1], v[2]);
    const C m = {v[0].re + c * t1.re, v[0].im + c * t1.im};
    const C js = {-s * t2.im, s * t2.re};
    v[0] = add(v[0], t1);
    v[1] = add(m, js);
    v[2] = sub(m, js);
  } else if constexpr (R == 4) {
    const C a0 = add(v[0], v[2]), a1 = sub(v[0], v[2]);
    const C b0 = add(v[1], v[3]), b1 = rot90(sub(v[1], v[3]), sign);
    v[0] = add(a0, b0);
    v[2] = sub(a0, b0);
    v[1] = add(a1, b1);
    v[3] = sub(a1, b1);
  } else {
    C out[R];
    for (int k = 0; k < R; ++k) {
      C s = v[0];
      for (int q = 1; q < R; ++q) s = add(s, mul(v[q], roots[(k * q) % R]));
      out[k] = s;
    }
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
}

template <int R>
void stage(const C* in, C* out, int n, int span, int sign, const C* tw, const C* roots) {
  const int m = n / R;
  const int blocks = m / span;
  for (int p = 0; p < blocks; ++p)
    for (int k = 0; k < span; ++k) {
      const int j = p * span + k;
      C v[R];
      v[0] = in[j];
      const C* w = tw + static_cast<long>(k) * (R - 1);
      for (int r = 1; r < R; ++r) {
        const C x = in[j + r * m];
        const C t = sign < 0 ? w[r - 1] : C{w[r - 1].re, -w[r - 1].im};
        v[r] = k == 0 ? x : mul(x, t);
      }
      dft<R>(v, sign, roots);
      C* o = out + static_cast<long>(p) * span * R + k;
      for (int r = 0; r < R; ++r) o[r * span] = v[r];
    }
}

// The same stage for `batch` interleaved sequences: element i of sequence b at in[i * batch + b].
template <int R>
void stage_batch(const C* in, C* out, int n, int span, int batch, int sign, const C* tw, const C* roots) {
  const int m = n / R;
  const int blocks = m / span;
  for (int p = 0; p < blocks; ++p)
    for (int k = 0; k < span; ++k) {
      const int j = p * span + k;
      const C* w = tw + static_cast<long>(k) * (R - 1);
      C t[R];
      for (int r = 1; r < R; ++r) t[r] = sign < 0 ? w[r - 1] : C{w[r - 1].re, -w[r - 1].im};
      C* o = out + (static_cast<long>(p) * span * R + k) * batch;
      for (int b = 0; b < batch; ++b) {
        C v[R];
        v[0] = in[static_cast<long>(j) * batch + b];
        for (int r = 1; r < R; ++r) {
          const C x = in[static_cast<long>(j + r * m) * batch + b];
          v[r] = k == 0 ? x : mul(x, t[r]);
        }
        dft<R>(v, sign, roots);
        for (int r = 0; r < R; ++r) o[static_cast<long>(r) * span * batch + b] = v[r];
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
  if (m != 1) throw std::invalid_argument("Fft1d: length with a prime factor above 13");
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

void Fft1d::run(cplx* x, int stride, int sign, cplx* work) const {
  const int n = n_;
  C* a = reinterpret_cast<C*>(work);
  C* b = a + n;
  for (int i = 0; i < n; ++i) a[i] = reinterpret_cast<const C*>(x)[static_cast<long>(i) * stride];
  int span = 1;
  long offset = 0;
  long root_offset = 0;
  for (const int radix : radices_) {
    C roots[16];
    for (int q = 0; q < radix; ++q) {
      const C r = reinterpret_cast<const C*>(roots_.data())[root_offset + q];
      roots[q] = sign < 0 ? r : C{r.re, -r.im};
    }
    const C* tw = reinterpret_cast<const C*>(table_.data()) + offset;
    switch (radix) {
      case 2: stage<2>(a, b, n, span, sign, tw, roots); break;
      case 3: stage<3>(a, b, n, span, sign, tw, roots); break;
      case 4: stage<4>(a, b, n, span, sign, tw, roots); break;
      case 5: stage<5>(a, b, n, span, sign, tw, roots); break;
      case 7: stage<7>(a, b, n, span, sign, tw, roots); break;
      case 11: stage<11>(a, b, n, span, sign, tw, roots); break;
      default: stage<13>(a, b, n, span, sign, tw, roots); break;
    }
    offset += static_cast<long>(span) * (radix - 1);
    root_offset += radix;
    std::swap(a, b);
    span *= radix;
  }
  for (int i = 0; i < n; ++i) reinterpret_cast<C*>(x)[static_cast<long>(i) * stride] = a[i];
}

void Fft1d::run_batch(cplx* x, long stride, int batch, int sign, cplx* work) const {
  const int n = n_;
  C* a = reinterpret_cast<C*>(work);
  C* b = a + static_cast<long>(n) * batch;
  for (int i = 0; i < n; ++i)
    for (int e = 0; e < batch; ++e) a[static_cast<long>(i) * batch + e] = reinterpret_cast<const C*>(x)[i * stride + e];
  int span = 1;
  long offset = 0;
  long root_offset = 0;
  for (const int radix : radices_) {
    C roots[16];
    for (int q = 0; q < radix; ++q) {
      const C r = reinterpret_cast<const C*>(roots_.data())[root_offset + q];
      roots[q] = sign < 0 ? r : C{r.re, -r.im};
    }
    const C* tw = reinterpret_cast<const C*>(table_.data()) + offset;
    switch (radix) {
      case 2: stage_batch<2>(a, b, n, span, batch, sign, tw, roots); break;
      case 3: stage_batch<3>(a, b, n, span, batch, sign, tw, roots); break;
      case 4: stage_batch<4>(a, b, n, span, batch, sign, tw, roots); break;
      case 5: stage_batch<5>(a, b, n, span, batch, sign, tw, roots); break;
      case 7: stage_batch<7>(a, b, n, span, batch, sign, tw, roots); break;
      case 11: stage_batch<11>(a, b, n, span, batch, sign, tw, roots); break;
      default: stage_batch<13>(a, b, n, span, batch, sign, tw, roots); break;
    }
    offset += static_cast<long>(span) * (radix - 1);
    root_offset += radix;
    std::swap(a, b);
    span *= radix;
  }
  for (int i = 0; i < n; ++i)
    for (int e = 0; e < batch; ++e) reinterpret_cast<C*>(x)[i * stride + e] = a[static_cast<long>(i) * batch + e];
}

}  // namespace oracle
