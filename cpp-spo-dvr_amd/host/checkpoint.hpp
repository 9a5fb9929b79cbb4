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

}  // namespace spo
