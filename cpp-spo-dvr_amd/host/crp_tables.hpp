// Constants of the CRP library shared by the host evaluation (crp.cpp) and the device one (csrc/kernels_crp.hpp):
// pi and sqrt(3) as the Fortran literals, cos/sin(pi/3) as gfortran folds them, and the Fourier coefficient tables
// of the interpolation over the surface sites (src/crplib/constants.hpp, src/crplib/fourier_tables.cpp; the DATA
// statements of FOUR3D and FOUR6D with their scale factors). Element (site, j) of a table is at site + n j.
#pragma once

#include <array>
#include <cstddef>

namespace spo::crp_tables {

constexpr double kPi = 3.1415926535897932384626433832795028841971;
constexpr double kSq3 = 1.732050807568877293500;
constexpr double kUsq3 = 1.0 / kSq3;
// bits 0x3FE0000000000001 and 0x3FEBB67AE8584CAA (src/crplib/constants.hpp:16-22)
constexpr double kCosPi3 = 0x1.0000000000001p-1;
constexpr double kSinPi3 = 0x1.bb67ae8584caap-1;

// Fourier tables of the interpolation over the sites, element (site, j) at site + n j (src/crplib/fourier_tables.cpp;
// Fortran DATA statements of FOUR3D and FOUR6D, with their scale factors). Integer weights written out.
constexpr std::array<double, 49> kAtomC6vWeights = {
    1.0, 3.0,  2.0,  6.0,  12.0, 6.0,  6.0,  2.0, -2.0, -2.0, 2.0,  -8.0, 8.0,  0.0, 2.0, -2.0, 4.0,
    -4.0, 4.0, 2.0,  -6.0, 2.0,  6.0,  -2.0, -6.0, 0.0, 0.0, 0.0, 2.0,  -2.0, -2.0, 2.0, 4.0, -4.0,
    0.0, 1.0,  -1.0, 2.0,  -2.0, -4.0, -2.0, 6.0, 0.666666666666667, 2.0, 1.333333333333, 4.0, -4.0, -2.0, -2.0};

constexpr double kA = 5.1961524227066318805600;
constexpr double kB = 20.7846096908265275222400;
constexpr double kC = 15.5884572681198956416800;
constexpr double kD = 10.3923048454132637611200;
constexpr std::array<double, 100> kAtomC3vWeights = {
    3.00, 9.00,  3.00,  9.00,  18.00,  18.00,  18.00,  9.00,  18.00,  3.00,    //
    6.00, -6.00, -3.00, 3.00,  -12.00, 24.00,  0.00,   3.00,  -12.00, -3.00,   //
    0.00, 0.00,  -kA,   -kA,   -kB,    0.00,   0.00,   kA,    kB,     kA,      //
    6.00, -6.00, 6.00,  -6.00, 6.00,   6.00,   -18.00, -6.00, 6.00,   6.00,    //
    6.00, 18.00, -3.00, -9.00, 0.00,   0.00,   0.00,   -9.00, 0.00,   -3.00,   //
    0.00, 0.00,  kA,    -kC,   0.00,   0.00,   0.00,   kC,    0.00,   -kA,     //
    6.00, -6.00, -3.00, 3.00,  6.00,   -12.00, 0.00,   3.00,  6.00,   -3.00,   //
    0.00, 0.00,  -kA,   -kA,   kD,     0.00,   0.00,   kA,    -kD,    kA,      //
    3.00, -3.00, 3.00,  -3.00, -6.00,  -6.00,  18.00,  -3.00, -6.00,  3.00,    //
    2.00, 6.00,  2.00,  6.00,  -6.00,  -6.00,  -6.00,  6.00,  -6.00,  2.00};

constexpr std::array<double, 144> kMoleculeWeights = {
    1, 1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  2, -2, -1, 1,  -2, 2,  -1, 1,  -1, 1,  -1, 1,
    2, 2,  -1, 1,  -2, -2, 1,  -1, -1, 1,  1,  -1, 2, -2, -1, -1, 2,  -2, 1,  1,  -1, -1, 1,  1,
    0, 0,  1,  1,  0,  0,  -1, 1,  -1, -1, 1,  -1, 0, 0,  1,  1,  0,  0,  1,  -1, -1, -1, -1, 1,
    0, 0,  -1, 1,  0,  0,  -1, -1, 1,  -1, 1,  1,  1, 1,  1,  -1, -1, -1, -1, 1,  1,  -1, -1, 1,
    1, -1, 1,  -1, -1, 1,  1,  -1, 1,  -1, 1,  -1, 1, -1, 1,  1,  1,  -1, -1, -1, 1,  1,  -1, -1,
    2, 2,  -1, -1, 2,  2,  -1, -1, -1, -1, -1, -1, 0, 0,  1,  -1, 0,  0,  -1, -1, -1, 1,  1,  1};

template <std::size_t N>
constexpr std::array<double, N> scaled(const std::array<double, N>& w, double divisor) {
  std::array<double, N> t{};
  for (std::size_t i = 0; i < N; ++i) t[i] = w[i] / divisor;
  return t;
}

constexpr std::array<double, 144> molecule_table() {
  constexpr double r1 = kUsq3 / 4.00;
  constexpr double r2 = 1.00 / 36.00;
  std::array<double, 144> t{};
  for (int j = 0; j < 12; ++j) {
    double factor = 3.0 * r2;
    if (j >= 4 && j <= 6) factor = r1;
    if (j == 10) factor = r2;
    if (j == 11) factor = r1 / 3.0;
    for (int i = 0; i < 12; ++i) t[i + 12 * j] = factor * kMoleculeWeights[i + 12 * j];
  }
  return t;
}

constexpr std::array<double, 49> kAtomC6vTable = scaled(kAtomC6vWeights, 36.00);
constexpr std::array<double, 100> kAtomC3vTable = scaled(kAtomC3vWeights, 108.00);
constexpr std::array<double, 144> kMoleculeTable = molecule_table();


}  // namespace spo::crp_tables
