// Native checkpoint and BK coefficient files, byte-compatible with the reference program's (src/io/checkpoint.hpp,
// format version 2), so that a run of either program continues from the other's files:
//   header:     16-byte magic, uint32 version, the SystemID (uint64 length + bytes), uint64 count and that many
//               (name: uint64 length + bytes, value: double) restart parameters;
//   checkpoint: double time, int32 n_rot n_xy n_zr nR nZSpecular, int32 surface index and previous index, uint8
//               fixed, double norm tolerance, the specular wave (ir, iz) and the scattering wave (irot, ixy, izr)
//               as complex<double>, uint64 end marker; written to <path>.tmp, then renamed;
//   BK file:    int32 n_states n_v, five int32 labels (J, mJ, n, m, order) per state, the channel energies (is, v),
//               then one record (double time, n_states n_v complex (is, v)) per analysis step.
// All little-endian (host order on x86-64). A file written for another input is refused with the parameter named.
#pragma once

#include <functional>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "analysis.hpp"
#include "config.hpp"

namespace spo {

// A missing, corrupt or mismatched checkpoint / BK file (the reference's CheckpointError, exit status 4).
struct CheckpointError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// n_phi: the run's number of phi points (a restart parameter).
void write_checkpoint(const std::string& path, const Config& config, int n_phi, double time,
                      const std::vector<cplx>& scattering, int n_rot, int n_xy, int n_zr,
                      const std::vector<cplx>& specular, const AnalysisSurfaceState& surface);

struct CheckpointInfo {
  double time = 0.0;
  AnalysisSurfaceState surface;
};

// Reads a checkpoint written for the same input into the two waves (sized by the caller).
CheckpointInfo read_checkpoint(const std::string& path, const Config& config, int n_phi, std::vector<cplx>& scattering,
                               int n_rot, int n_xy, int n_zr, std::vector<cplx>& specular);

// The BK coefficient file of a run: header, basis, then the (time, coefficients) records.
void write_bk_file(const std::string& path, const Config& config, int n_phi, const AnalysisBasis& basis,
                   const std::vector<std::pair<double, std::vector<cplx>>>& records);

// Calls on_record(time, coefficients) per record of a BK file written for the same input and basis; returns the
// time of the last record.
std::optional<double> read_bk_file(const std::string& path, const Config& config, int n_phi, const AnalysisBasis& basis,
                                   const std::function<void(double, const std::vector<cplx>&)>& on_record);

// Restart files in the Fortran program's layout (gfortran sequential unformatted: every record framed by int32
// byte counts; INTEGER 4 bytes, REAL 8, COMPLEX 16), so that runs continue across the two programs
// (src/io/fortran_files.hpp):
//   wave function (written as fort.9, read as fort.8): nproj = 2; the specular wave; (fixed, surface index + 1);
//               then one record per izr with the scattering block (irot, ixy);
//   BK file (written as fort.21, read as fort.22): (n_records, nstatw, ndftow, nrotw); (dtAna, kz, Z0, alpha,
//               Zana, E(vin, jin)); (E_trans, theta_in, phi_in, JIn, MJIn, initial channel + 1); six label arrays
//               of the analysed states (rotation + 1, diffraction + 1, n, m, J, mJ); the channel energies (is, v);
//               then one record of coefficients (is, v) per analysis step.
struct FortranBkHeader {
  int n_records = 0;  // the records the file will hold
  int n_rotations = 0, initial_channel = 0;
  double initial_energy = 0.0;  // hartree
};

void write_fortran_wave_function(const std::string& path, const std::vector<cplx>& scattering, int n_rot, int n_xy,
                                 int n_zr, const std::vector<cplx>& specular, const AnalysisSurfaceState& surface);

// The Fortran file keeps the surface position and whether it is fixed; the rest of the state comes from `fresh`
// (a new surface for the same input), with the previous index equal to the index.
AnalysisSurfaceState read_fortran_wave_function(const std::string& path, std::vector<cplx>& scattering, int n_rot,
                                                int n_xy, int n_zr, std::vector<cplx>& specular, int nr, int n_z_specular,
                                                const AnalysisSurfaceState& fresh);

void write_fortran_bk_file(const std::string& path, const Config& config, const AnalysisBasis& basis,
                           const FortranBkHeader& header, const std::vector<std::vector<cplx>>& records);

// Calls on_record(k dtAna, coefficients) for record k = 1, 2, ...; returns the number of records.
int read_fortran_bk_file(const std::string& path, const Config& config, const AnalysisBasis& basis, int n_rotations,
                         const std::function<void(double, const std::vector<cplx>&)>& on_record);

}  // namespace spo
