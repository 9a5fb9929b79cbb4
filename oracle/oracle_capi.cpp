// TEST INFRASTRUCTURE — C entry points of the CPU oracle for the Python tests and bench.py's CPU-baseline leg.
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "oracle_fft.hpp"

#include "oracle.hpp"

extern "C" {

// A BackendFactory for spo_run_create (device SPO_DEVICE_EXTERNAL): input is a const spo::BackendInput*, the
// result a heap spo::PropagationBackend* owned by the caller.
void* oracle_backend_factory(const void* input) {
  auto backend = oracle::make_oracle_backend(*static_cast<const spo::BackendInput*>(input));
  return backend.release();
}

// FFT of one sequence in place (interleaved re/im), sign -1 or +1, no scaling: the oracle's FFT under test.
int oracle_fft(double* data, int n, int sign) {
  try {
    oracle::Fft1d f(n);
    std::vector<oracle::cplx> work(2 * static_cast<std::size_t>(n));
    f.run(reinterpret_cast<oracle::cplx*>(data), 1, sign, work.data());
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

}  // extern "C"
