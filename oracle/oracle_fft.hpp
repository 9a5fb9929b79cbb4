// TEST INFRASTRUCTURE — part of the CPU oracle, not of the product. Only tests/, the smoke check and the
// CPU-baseline leg of bench.py load it, and only as the checker.
//
// FP64 complex FFT of any length whose prime factors are small (the reference's FFT sizes 2^h 3^i 5^j 7^k 11^m,
// src/fft/fft_sizes.cpp:10-34), in the sign convention of src/fft/fft.hpp:9-12: sign -1 computes
// y_k = sum_n x_n exp(-2 pi i k n / N), sign +1 the opposite exponent; no scaling. Stockham autosort, mixed radix,
// twiddles per stage evaluated in long double.
#pragma once

#include <complex>
#include <vector>

namespace oracle {

using cplx = std::complex<double>;

class Fft1d {
 public:
  explicit Fft1d(int n);
  int size() const { return n_; }
  // In place on x[0], x[stride], ...; work must hold 2 n values.
  void run(cplx* x, int stride, int sign, cplx* work) const;
  // In place on `batch` sequences at once: element i of sequence b at x[i * stride + b] (the sequences are
  // interleaved, batch elements contiguous); work must hold 2 n batch values.
  void run_batch(cplx* x, long stride, int batch, int sign, cplx* work) const;

 private:
  int n_;
  std::vector<int> radices_;
  std::vector<cplx> table_;  // per stage: exp(-2 pi i r k / (span R)), k < span, r = 1..R-1
  std::vector<cplx> roots_;  // per stage: exp(-2 pi i q / R), q < R
};

}  // namespace oracle
