// Potential energy surfaces on the propagation path and the potential on the DVR grid of the interaction
// region.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/spo_hip.h"
#include "config.hpp"
#include "crp.hpp"
#include "scm.hpp"
#include "grids.hpp"
#include "symmetry.hpp"

namespace spo {

// V(x, y, Z, r, theta, phi) in eV; (x, y) skewed coordinates along the lattice vectors
// (src/potential/potential.hpp:25-70).
class Pes {
 public:
  virtual ~Pes() = default;
  virtual double value(double x, double y, double z, double r, double theta, double phi) const = 0;
};

// H2 Rydberg V(r) plus a 1/(Z^4 + 0.001) wall switched off by a cosine between Z = 0.1778 and 7
// (src/potential/hard_wall_potential.cpp:19-51).
class HardWallPes final : public Pes {
 public:
  double value(double x, double y, double z, double r, double theta, double phi) const override;
  static double gas_phase(double r);
};

// Dai-Light LEPS surface for H2/Cu(111) (src/potential/dai_light_cu111_potential.cpp:71-118).
class DaiLightPes final : public Pes {
 public:
  double value(double x, double y, double z, double r, double theta, double phi) const override;
};

// The CRP surface of a data set (SystemID CRPlibrary, src/potential/crp_potential.cpp): the QD wrapper of the CRP
// library, V in eV at skewed (x, y).
class CrpPes final : public Pes {
 public:
  explicit CrpPes(std::shared_ptr<const crp::QdSurface> surface) : surface_(std::move(surface)) {}
  double value(double x, double y, double z, double r, double theta, double phi) const override;

 private:
  std::shared_ptr<const crp::QdSurface> surface_;
};

// The surface of a CRP data set directory, read once per process and shared; InputError naming CRPlibDataSet for
// a missing or unusable data set.
std::shared_ptr<const crp::QdSurface> load_crp_surface(const std::string& directory);

// The CRP surface plus the SCM correction of a distorted slab (SystemID CRPDistrtSurface,
// src/potential/scm/scm_crp_potential.cpp).
class ScmPes final : public Pes {
 public:
  explicit ScmPes(std::shared_ptr<const scm::ScmSurface> surface) : surface_(std::move(surface)) {}
  double value(double x, double y, double z, double r, double theta, double phi) const override;

 private:
  std::shared_ptr<const scm::ScmSurface> surface_;
};

// The SCM surface of an input (its CRP data set, parameter and position files), read once per process.
std::shared_ptr<const scm::ScmSurface> load_scm_surface(const InputParameters& input);

std::unique_ptr<Pes> make_pes(const std::string& system_id, const std::string& crp_data_set = "");
std::unique_ptr<Pes> make_pes(const InputParameters& input);

// The device library's view (spo_hip_crp_surface) of a packed CRP surface; points into `packed`.
spo_hip_crp_surface crp_surface_view(const crp::Packed& packed);
// The device library's view of an SCM correction; `atoms` holds its atom table.
spo_hip_scm_correction scm_correction_view(const scm::ScmSurface& surface, std::vector<double>& atoms);

// V in hartree on (itp, irep, izr), itp = iphi + nphi * itheta, izr = ir + nr * iz < nz_potential * nr, and
// the gas-phase V(r) (src/potential/potential_grid.hpp:14-31).
struct PotentialGrid {
  int n_tp = 0, n_rep = 0, n_zr = 0;
  std::vector<double> values;
  std::vector<double> asymptotic;
  GridSymmetry symmetry;
};

// src/potential/potential_grid.cpp:65-159 (evaluation at the representatives, switching, limits).
PotentialGrid compute_potential_grid(const Pes& pes, const Config& config, const SpatialGrids& grids,
                                     const AngularGrid& angular);

// The same grid with the surface evaluated on the GPU (spo_hip_potential_grid); the gas-phase rows of the switch
// and the asymptote come from `pes` on the host. Throws NotSupportedError for a surface the device does not have.
PotentialGrid compute_potential_grid_on_device(const Pes& pes, const Config& config, const SpatialGrids& grids,
                                               const AngularGrid& angular);

}  // namespace spo
