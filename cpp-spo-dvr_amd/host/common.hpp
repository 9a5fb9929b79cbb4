// Shared numeric helpers of the host side: complex type, constants, the unit set the propagation uses,
// tolerant real comparisons and the error types thrown across the boundary.
#pragma once

#include <cmath>
#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace spo {

using cplx = std::complex<double>;

inline constexpr double kPi = 3.1415926535897932384626433832795028841971;

// Energy/mass/length factors of the propagation (src/core/unit_conversion.hpp:55, "legacy" set): the
// hartree<->eV factor is stored as hartree->eV and divided for the opposite direction.
struct Units {
  static constexpr double hartree_in_ev = 27.2114;
  static constexpr double amu_in_me = 1822.89;
  static constexpr double hartree_in_wavenumber = 219475.0;
  static constexpr double ev_to_hartree(double x) { return x / hartree_in_ev; }
  static constexpr double hartree_to_ev(double x) { return x * hartree_in_ev; }
  static constexpr double amu_to_me(double x) { return x * amu_in_me; }
};

// Absolute-tolerance comparisons (1e-12) used for every grid-placement decision
// (src/core/real_compare.hpp:14-33).
namespace tol {
inline constexpr double eps = 1.0e-12;
inline bool eq(double a, double b) { return std::abs(a - b) < eps; }
inline bool gt(double a, double b) { return (a - b) > eps; }
inline bool ge(double a, double b) { return eq(a, b) || (a - b) > eps; }
inline bool lt(double a, double b) { return !ge(a, b); }
inline bool le(double a, double b) { return !gt(a, b); }
}  // namespace tol

// x**n by binary powering, as libgcc's __powidf2 does (src/core/fortran_math.hpp:8-17).
inline double ipow(double x, int n) {
  unsigned m = n < 0 ? 0u - static_cast<unsigned>(n) : static_cast<unsigned>(n);
  double y = (m & 1u) ? x : 1.0;
  while ((m >>= 1) != 0) {
    x = x * x;
    if (m & 1u) y = y * x;
  }
  return n < 0 ? 1.0 / y : y;
}

// Machine epsilon by halving until 1 - eps rounds to 1 (src/core/machine_epsilon.cpp:8-20).
double machine_epsilon();

// Errors. The C ABI maps them to status codes: input 2, not supported 3, device 1.
struct InputError : std::runtime_error {
  InputError(const std::string& key, const std::string& what) : std::runtime_error(key + ": " + what) {}
};
struct NotSupportedError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Inclusive 0-based index range.
struct Range {
  int first = 0;
  int last = -1;
  int size() const { return last - first + 1; }
};

}  // namespace spo
