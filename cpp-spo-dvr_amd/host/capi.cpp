// C ABI of the host driver (include/spo_host.h).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "../../include/spo_host.h"
#include "checkpoint.hpp"
#include "counting_backend.hpp"
#include "driver.hpp"
#include "json.hpp"
#include "potential.hpp"
#include "../../include/spo_hip.h"

struct spo_crp {
  std::shared_ptr<const spo::crp::QdSurface> qd;
};

struct spo_scm {
  std::shared_ptr<const spo::scm::ScmSurface> surface;
};

struct spo_run {
  std::shared_ptr<spo::RunSetup> setup;
  spo::SeamCallCounts counts;  // before propagation: the backend's decorator refers to it
  std::unique_ptr<spo::Propagation> propagation;
};

namespace {

thread_local std::string g_error;

int fail(int code, const char* what) {
  g_error = what;
  return code;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return SPO_OK;
  } catch (const spo::InputError& e) {
    return fail(SPO_E_INPUT, e.what());
  } catch (const spo::NotSupportedError& e) {
    return fail(SPO_E_NOT_SUPPORTED, e.what());
  } catch (const spo::CheckpointError& e) {
    return fail(SPO_E_CHECKPOINT, e.what());
  } catch (const spo::crp::DataError& e) {
    return fail(SPO_E_INPUT, e.what());
  } catch (const std::exception& e) {
    return fail(SPO_E_OTHER, e.what());
  }
}

void copy_complex(const std::vector<spo::cplx>& v, double* out) { std::memcpy(out, v.data(), v.size() * sizeof(spo::cplx)); }

// spo_run_create, with the set-up's potential grid replaced by `potential` (n values) when it is not NULL.
int create_run(const char* input_path, const char* overrides_json, int device, spo_backend_factory_fn factory,
               const double* potential, size_t n_potential, spo_run** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("spo_run_create: out is NULL");
    *out = nullptr;
    spo::json::Value doc = spo::json::parse_file(input_path);
    if (overrides_json && *overrides_json) {
      const spo::json::Value extra = spo::json::parse(overrides_json);
      for (const auto& [key, value] : extra.members) {
        bool replaced = false;
        for (auto& [k, v] : doc.members)
          if (k == key) {
            v = value;
            replaced = true;
          }
        if (!replaced) doc.members.emplace_back(key, value);
      }
    }
    auto run = std::make_unique<spo_run>();
    // On the GPU the potential grid is evaluated there too, unless SPO_HIP_POTENTIAL=host asks for the host build.
    const char* where = std::getenv("SPO_HIP_POTENTIAL");
    if (where && std::string(where) != "host" && std::string(where) != "device")
      throw std::invalid_argument("SPO_HIP_POTENTIAL must be host or device");
    const bool device_potential = device == SPO_DEVICE_HIP && !(where && std::string(where) == "host");
    run->setup = spo::make_run_setup(spo::read_input(doc), device_potential);
    if (potential) {
      std::vector<double>& v = run->setup->potential.values;
      if (n_potential != v.size())
        throw std::invalid_argument("spo_run_create_with_potential: " + std::to_string(n_potential) +
                                    " values for a potential grid of " + std::to_string(v.size()));
      for (size_t k = 0; k < n_potential; ++k)
        if (!std::isfinite(potential[k]))
          throw std::invalid_argument("spo_run_create_with_potential: value " + std::to_string(k) + " is not finite");
      v.assign(potential, potential + n_potential);
    }
    spo::BackendFactory make;
    if (device == SPO_DEVICE_HIP) {
      make = spo::make_hip_backend;
    } else if (device == SPO_DEVICE_EXTERNAL) {
      if (!factory) throw std::invalid_argument("spo_run_create: no factory for SPO_DEVICE_EXTERNAL");
      make = [factory](const spo::BackendInput& in) {
        void* b = factory(&in);
        if (!b) throw std::runtime_error("spo_run_create: the backend factory failed");
        return std::unique_ptr<spo::PropagationBackend>(static_cast<spo::PropagationBackend*>(b));
      };
    } else {
      throw spo::NotSupportedError("unknown device " + std::to_string(device));
    }
    // Every backend is made behind the counting decorator (spo_run_call_counts).
    spo::SeamCallCounts& counts = run->counts;
    spo::BackendFactory counted = [make, &counts](const spo::BackendInput& in) {
      return std::unique_ptr<spo::PropagationBackend>(std::make_unique<spo::CountingBackend>(make(in), counts));
    };
    run->propagation = std::make_unique<spo::Propagation>(run->setup, counted);
    *out = run.release();
  });
}

}  // namespace

extern "C" {

const char* spo_last_error(void) { return g_error.c_str(); }

int spo_run_create(const char* input_path, const char* overrides_json, int device, spo_backend_factory_fn factory,
                   spo_run** out) {
  return create_run(input_path, overrides_json, device, factory, nullptr, 0, out);
}

int spo_run_create_with_potential(const char* input_path, const char* overrides_json, int device,
                                  spo_backend_factory_fn factory, const double* potential, size_t n_potential,
                                  spo_run** out) {
  if (!potential) return fail(SPO_E_OTHER, "spo_run_create_with_potential: potential is NULL");
  return create_run(input_path, overrides_json, device, factory, potential, n_potential, out);
}

int spo_run_destroy(spo_run* run) {
  delete run;
  return SPO_OK;
}

int spo_run_get_info(const spo_run* run, spo_run_info* info) {
  return guarded([&] {
    const spo::RunSetup& s = *run->setup;
    const spo::InputParameters& p = s.config.input;
    info->n_rot = s.basis.size();
    info->nx = p.nx;
    info->ny = p.ny;
    info->nr = p.nr;
    info->nz = p.nz;
    info->nz_specular = p.nz_specular;
    info->nz_potential = s.grids.nz_potential;
    info->n_theta = s.angular.n_theta();
    info->n_phi = s.angular.n_phi();
    info->n_representatives = s.potential.n_rep;
    info->initial_channel = s.basis.initial;
    info->n_states = s.analysis.n_states();
    info->n_v = s.analysis.n_v;
    info->n_energies = s.energies.size();
    info->n_diffraction = s.analysis.n_diffraction;
    info->steps_per_analysis = s.config.derived.steps_per_analysis;
    info->analysis_steps_total = run->propagation->analysis_steps_total();
    info->analyses_per_average = s.config.derived.analyses_per_average;
    info->iz_analysis = s.grids.iz_analysis;
    info->time_step = p.time_step;
  });
}

void* spo_run_backend(spo_run* run) { return &run->propagation->backend(); }

int spo_run_call_counts(const spo_run* run, long* counts, int capacity) {
  return guarded([&] {
    if (!run || !counts || capacity < spo::SeamCallCounts::kCount)
      throw std::invalid_argument("spo_run_call_counts: NULL run or fewer than 12 slots");
    run->counts.to_array(counts);
  });
}

int spo_run_steps(spo_run* run, int n) {
  return guarded([&] { run->propagation->run(n); });
}

int spo_run_steps_batch(spo_run* const* runs, int n, int n_analysis_steps) {
  return guarded([&] {
    if (n < 0 || (n > 0 && !runs)) throw spo::InputError("runs", "bad arguments");
    std::vector<spo::Propagation*> p;
    for (int i = 0; i < n; ++i) {
      if (!runs[i]) throw spo::InputError("runs", "NULL run");
      p.push_back(runs[i]->propagation.get());
    }
    spo::Propagation::run_batch(p, n_analysis_steps);
  });
}

int spo_run_propagate(spo_run* run) {
  return guarded([&] { run->propagation->backend().propagate_analysis_step(); });
}

int spo_run_kinetic_and_potential(spo_run* run, int half_step, int kinetic_only) {
  return guarded([&] { run->propagation->backend().kinetic_and_potential(half_step != 0, kinetic_only != 0); });
}

int spo_run_analyse(spo_run* run, double t, double* bk) {
  return guarded([&] {
    std::vector<spo::cplx> c;
    auto& surface = const_cast<spo::AnalysisSurface&>(run->propagation->surface());
    surface.update(run->propagation->backend().specular());
    run->propagation->backend().analyse(surface, t, bk ? &c : nullptr);
    if (bk) copy_complex(c, bk);
  });
}

int spo_run_get_wave(spo_run* run, double* scattering, double* specular) {
  return guarded([&] {
    const spo::WavePacket& w = run->propagation->backend().wave_packet_on_host();
    if (scattering) copy_complex(w.scattering, scattering);
    if (specular) copy_complex(w.specular, specular);
  });
}

int spo_run_set_wave(spo_run* run, const double* scattering, const double* specular) {
  return guarded([&] {
    run->propagation->backend().read_wave_packet([&](spo::WavePacket& w) {
      if (scattering) std::memcpy(w.scattering.data(), scattering, w.scattering.size() * sizeof(spo::cplx));
      if (specular) std::memcpy(w.specular.data(), specular, w.specular.size() * sizeof(spo::cplx));
    });
  });
}

int spo_run_get_specular(spo_run* run, double* specular) {
  return guarded([&] { copy_complex(run->propagation->backend().specular(), specular); });
}

int spo_run_block_densities(spo_run* run, double* out) {
  return guarded([&] {
    const std::vector<double> d = run->propagation->backend().block_densities();
    std::memcpy(out, d.data(), d.size() * sizeof(double));
  });
}

int spo_run_z_plane(spo_run* run, int iz, double* out) {
  return guarded([&] { copy_complex(run->propagation->backend().z_plane(iz), out); });
}

int spo_run_site_column(spo_run* run, int ixy, double* out) {
  return guarded([&] { run->propagation->backend().site_column(ixy, reinterpret_cast<spo::cplx*>(out)); });
}

int spo_run_bk_record(const spo_run* run, int k, double* out) {
  return guarded([&] {
    const auto& r = run->propagation->bk_records();
    if (k < 0 || k >= static_cast<int>(r.size())) throw std::out_of_range("spo_run_bk_record: no such record");
    copy_complex(r[k], out);
  });
}

int spo_run_write_checkpoint(spo_run* run, const char* directory) {
  return guarded([&] {
    if (!directory || !*directory) throw std::invalid_argument("spo_run_write_checkpoint: no directory");
    run->propagation->write_checkpoint(directory);
  });
}

int spo_run_bk_count(const spo_run* run, int* count) {
  return guarded([&] { *count = static_cast<int>(run->propagation->bk_records().size()); });
}

int spo_run_averages(const spo_run* run, double* rows, int capacity, int* count) {
  return guarded([&] {
    const auto& a = run->propagation->averages();
    *count = static_cast<int>(a.size());
    for (int i = 0; i < *count && i < capacity; ++i) {
      rows[5 * i + 0] = a[i].time;
      rows[5 * i + 1] = a[i].norm;
      rows[5 * i + 2] = a[i].z_mean;
      rows[5 * i + 3] = a[i].r_mean;
      rows[5 * i + 4] = a[i].rest_norm;
    }
  });
}

int spo_run_surface(const spo_run* run, int* fixed, int* index, double* z_line) {
  return guarded([&] {
    const auto& s = run->propagation->surface();
    *fixed = s.fixed() ? 1 : 0;
    *index = s.index();
    *z_line = s.fixed_line().value_or(0.0);
  });
}

int spo_run_results(const spo_run* run, int zana_convention, double* s, double* p, double* p_reaction,
                    double* total_energy) {
  return guarded([&] {
    const auto results = zana_convention ? run->propagation->results(std::nullopt) : run->propagation->results();
    for (std::size_t ien = 0; ien < results.size(); ++ien) {
      const spo::EnergyResult& e = results[ien];
      const std::size_t n = e.s_matrix.size();
      if (s) std::memcpy(s + 2 * n * ien, e.s_matrix.data(), n * sizeof(spo::cplx));
      if (p) std::memcpy(p + n * ien, e.probability.data(), n * sizeof(double));
      if (p_reaction) p_reaction[ien] = e.p_reaction;
      if (total_energy) total_energy[ien] = e.total_energy;
    }
  });
}

int spo_run_s_sums(const spo_run* run, double* out) {
  return guarded([&] { copy_complex(run->propagation->backend().s_matrix_sums(), out); });
}

int spo_run_energy_resolved_shape(const spo_run* run, int* n_sites, int* n_energies) {
  return guarded([&] {
    const spo::EnergyResolvedSetup& e = run->propagation->energy_resolved_setup();
    *n_sites = static_cast<int>(e.site_ixy.size());
    *n_energies = static_cast<int>(e.total_hartree.size());
  });
}

int spo_run_energy_resolved(spo_run* run, double* sums, double* specular_sums) {
  return guarded([&] {
    std::vector<spo::cplx> a, b;
    run->propagation->energy_resolved(&a, &b);
    copy_complex(a, sums);
    copy_complex(b, specular_sums);
  });
}

int spo_run_array(const spo_run* run, const char* name, double* out, size_t capacity, size_t* count) {
  return guarded([&] {
    const spo::RunSetup& s = *run->setup;
    const std::string n(name);
    std::vector<double> v;
    if (n == "gas.energies") v = s.gas.energies;
    else if (n == "gas.wave_functions") v = s.gas.wave_functions;
    else if (n == "potential.values") v = s.potential.values;
    else if (n == "potential.asymptotic") v = s.potential.asymptotic;
    else if (n == "grids.x") v = s.grids.x.points;
    else if (n == "grids.y") v = s.grids.y.points;
    else if (n == "units.hartree_in_ev") v = {spo::Units::hartree_in_ev};
    else if (n == "switch.z") {
      if (s.config.input.use_switching) v = {s.config.input.z_min_switch, s.config.input.z_max_switch};
    }
    else if (n == "angular.thetas") v = s.angular.thetas;
    else if (n == "angular.weights") v = s.angular.theta_weights;
    else if (n == "angular.phis") v = s.angular.phis;
    else if (n == "basis.j") v.assign(s.basis.j.begin(), s.basis.j.end());
    else if (n == "basis.mj") v.assign(s.basis.mj.begin(), s.basis.mj.end());
    else if (n == "state.j") v.assign(s.analysis.state_j.begin(), s.analysis.state_j.end());
    else if (n == "state.mj") v.assign(s.analysis.state_mj.begin(), s.analysis.state_mj.end());
    else if (n == "state.n" || n == "state.m" || n == "state.order") {
      for (int is = 0; is < s.analysis.n_states(); ++is) {
        const spo::DiffractionChannel& c = s.analysis.channels[s.analysis.state_diffraction(is)];
        v.push_back(n == "state.n" ? c.n : n == "state.m" ? c.m : c.order);
      }
    } else if (n == "level.energies") {
      for (double e : s.gas.energies) v.push_back(e * spo::Units::hartree_in_ev);
    } else if (n == "symmetry.representatives") {
      v.assign(s.potential.symmetry.representatives.begin(), s.potential.symmetry.representatives.end());
    } else if (n == "symmetry.representative_of") {
      v.assign(s.potential.symmetry.representative_of.begin(), s.potential.symmetry.representative_of.end());
    } else if (n == "symmetry.operation_of") {
      v.assign(s.potential.symmetry.operation_of.begin(), s.potential.symmetry.operation_of.end());
    } else if (n == "symmetry.phi_index") {
      for (const auto& row : s.potential.symmetry.phi_index) v.insert(v.end(), row.begin(), row.end());
    } else if (n == "analysis.limits") {
      // jana, nviba, nana, nstata, inchna (1-based), as the reference's stdout prints them
      v = {double(s.gas.j_max_analysis), double(s.analysis.n_v), double(s.analysis.n_rotation()),
           double(s.analysis.n_states()), double(s.analysis.initial + 1)};
    } else if (n == "energies.total") {
      v = s.energies.total;
    } else if (n == "tables.z_optical") v = s.tables.z_optical;
    else if (n == "tables.r_optical") v = s.tables.r_optical;
    else if (n == "tables.zr_full" || n == "tables.zr_half" || n == "tables.xy_full" || n == "tables.xy_half" ||
             n == "tables.rotation_half") {
      // complex factors as interleaved (re, im)
      const std::vector<spo::cplx>& c = n == "tables.zr_full"   ? s.tables.zr_kinetic.full
                                        : n == "tables.zr_half" ? s.tables.zr_kinetic.half
                                        : n == "tables.xy_full" ? s.tables.xy_kinetic.full
                                        : n == "tables.xy_half" ? s.tables.xy_kinetic.half
                                                                : s.tables.rotation.half;
      for (const spo::cplx& z : c) {
        v.push_back(z.real());
        v.push_back(z.imag());
      }
    }
    else if (n == "angular.legendre") v = spo::legendre_matrix(s.angular, s.basis);
    else if (n == "grids.z") v = s.grids.z.points;
    else if (n == "grids.r") v = s.grids.r.points;
    else if (n == "derived.cell") {
      // lattice_x, lattice_y, skew (rad), total mass (m_e), dx, dy, dt, the Cartesian parallel momentum (x, y)
      const spo::DerivedParameters& d = s.config.derived;
      const double k_par = d.k_total * std::sin(d.theta_incidence);
      v = {d.system.lattice_x, d.system.lattice_y, d.system.skew, d.total_mass, d.dx, d.dy, s.config.input.time_step,
           k_par * std::cos(d.phi_incidence), k_par * std::sin(d.phi_incidence)};
    }
    else if (n == "initial.specular") {
      const double* d = reinterpret_cast<const double*>(s.initial_specular.data());
      v.assign(d, d + 2 * s.initial_specular.size());
    } else {
      throw std::invalid_argument("spo_run_array: unknown array " + n);
    }
    const bool complex = n == "initial.specular";
    *count = complex ? v.size() / 2 : v.size();
    if (out) {
      if (capacity < v.size()) throw std::invalid_argument("spo_run_array: capacity too small");
      std::memcpy(out, v.data(), v.size() * sizeof(double));
    }
  });
}

int spo_run_kernel_timing(spo_run* run, int on) {
  return guarded([&] { run->propagation->backend().set_kernel_timing(on != 0); });
}

int spo_run_kernel_times(spo_run* run, char* names, long* launches, double* total_ms, int capacity, int* count) {
  return guarded([&] {
    const auto times = run->propagation->backend().kernel_times();
    *count = static_cast<int>(times.size());
    for (int k = 0; k < *count && k < capacity; ++k) {
      std::strncpy(names + 48 * k, times[k].name.c_str(), 47);
      names[48 * k + 47] = '\0';
      launches[k] = times[k].launches;
      total_ms[k] = times[k].total_ms;
    }
  });
}

int spo_pes_value(const char* system_id, double x, double y, double z, double r, double theta, double phi,
                  double* out) {
  return guarded([&] { *out = spo::make_pes(system_id)->value(x, y, z, r, theta, phi); });
}

}  // extern "C"

namespace {

spo::crp::Units crp_units(int length, int energy) {
  if (length != 0 && length != 1) throw spo::InputError("length_unit", "0 (bohr) or 1 (angstrom)");
  if (energy < 2 || energy > 6) throw spo::InputError("energy_unit", "2 (hartree) .. 6 (J)");
  return {static_cast<spo::crp::LengthUnit>(length), static_cast<spo::crp::EnergyUnit>(energy)};
}

// Runs f(k) for k < n in parallel; the first exception is rethrown after the loop.
template <typename F>
void for_points(long n, F&& f) {
  std::exception_ptr error;
#pragma omp parallel for schedule(dynamic, 64)
  for (long k = 0; k < n; ++k) {
    try {
      f(k);
    } catch (...) {
#pragma omp critical
      if (!error) error = std::current_exception();
    }
  }
  if (error) std::rethrow_exception(error);
}

}  // namespace

extern "C" {

int spo_crp_open(const char* directory, spo_crp** out) {
  return guarded([&] { *out = new spo_crp{spo::load_crp_surface(directory)}; });
}

int spo_crp_close(spo_crp* crp) {
  delete crp;
  return SPO_OK;
}

int spo_crp_molecule(const spo_crp* crp, long n, const double* atoms, int length_unit, int energy_unit, double* out) {
  return guarded([&] {
    using namespace spo::crp;
    const Units u = crp_units(length_unit, energy_unit);
    const MoleculeSurface& s = crp->qd->surface();
    for_points(n, [&](long k) {
      const double* p = atoms + 6 * k;
      const Vec3 a{to_length(p[0], u.length, LengthUnit::bohr), to_length(p[1], u.length, LengthUnit::bohr),
                   to_length(p[2], u.length, LengthUnit::bohr)};
      const Vec3 b{to_length(p[3], u.length, LengthUnit::bohr), to_length(p[4], u.length, LengthUnit::bohr),
                   to_length(p[5], u.length, LengthUnit::bohr)};
      const MoleculeValue v = s.evaluate(a, b).value_or(MoleculeValue{0.00, {0.00, 0.00, 0.00}, {0.00, 0.00, 0.00}});
      double* o = out + 7 * k;
      o[0] = to_energy(v.v, EnergyUnit::hartree, u.energy);
      const double g[6] = {v.gradient_a.x, v.gradient_a.y, v.gradient_a.z, v.gradient_b.x, v.gradient_b.y,
                           v.gradient_b.z};
      for (int i = 0; i < 6; ++i) o[1 + i] = to_force(g[i], kInternalUnits, u);
    });
  });
}

int spo_crp_atom(const spo_crp* crp, long n, const double* positions, int length_unit, int energy_unit,
                 double* out) {
  return guarded([&] {
    using namespace spo::crp;
    const Units u = crp_units(length_unit, energy_unit);
    const AtomSurface& s = crp->qd->surface().atoms();
    for_points(n, [&](long k) {
      const double* p = positions + 3 * k;
      const AtomValue v = s.evaluate({to_length(p[0], u.length, LengthUnit::bohr),
                                      to_length(p[1], u.length, LengthUnit::bohr),
                                      to_length(p[2], u.length, LengthUnit::bohr)});
      double* o = out + 4 * k;
      o[0] = to_energy(v.v, EnergyUnit::hartree, u.energy);
      o[1] = to_force(v.gradient.x, kInternalUnits, u);
      o[2] = to_force(v.gradient.y, kInternalUnits, u);
      o[3] = to_force(v.gradient.z, kInternalUnits, u);
    });
  });
}

int spo_crp_qd(const spo_crp* crp, long n, const double* points, double* out) {
  return guarded([&] {
    for_points(n, [&](long k) {
      const double* p = points + 6 * k;
      out[k] = crp->qd->potential_ev(p[0], p[1], p[2], p[3], p[4], p[5]);
    });
  });
}

int spo_crp_gas_phase(const spo_crp* crp, long n, const double* zr, double* out) {
  return guarded([&] {
    const spo::crp::MoleculeSurface& s = crp->qd->surface();
    for_points(n, [&](long k) {
      const auto sw = s.switch_function(zr[2 * k]);
      const auto gp = s.gas_phase(zr[2 * k], zr[2 * k + 1]);
      double* o = out + 5 * k;
      o[0] = sw.f;
      o[1] = sw.df_dz;
      o[2] = gp.v;
      o[3] = gp.d_z;
      o[4] = gp.d_r;
    });
  });
}

int spo_crp_geometry(long n, const double* atoms, double* out) {
  return guarded([&] {
    for (long k = 0; k < n; ++k) {
      const double* p = atoms + 6 * k;
      const auto g = spo::crp::molecule_geometry({p[0], p[1], p[2]}, {p[3], p[4], p[5]});
      double* o = out + 6 * k;
      o[0] = g.x;
      o[1] = g.y;
      o[2] = g.z;
      o[3] = g.r;
      o[4] = g.theta;
      o[5] = g.phi;
    }
  });
}

int spo_crp_device_values(const spo_crp* crp, long n, const double* points, double* out) {
  return guarded([&] {
    const spo::crp::Packed packed = spo::crp::pack(*crp->qd);
    const spo_hip_crp_surface view = spo::crp_surface_view(packed);
    if (spo_hip_crp_values(&view, n, points, out) != SPO_HIP_OK)
      throw std::runtime_error(std::string("device CRP values: ") + spo_hip_last_error());
  });
}

int spo_scm_open(const char* crp_data_set, const char* parameter_file, const char* position_file, int configuration,
                 double cutoff, int rescale, spo_scm** out) {
  return guarded([&] {
    spo::InputParameters p;
    p.system_id = "CRPDistrtSurface";
    p.crp_data_set = crp_data_set;
    p.scm.parameter_file = parameter_file;
    p.scm.position_file = position_file;
    p.scm.configuration = configuration;
    p.scm.cutoff = cutoff;
    p.scm.rescale = rescale != 0;
    *out = new spo_scm{spo::load_scm_surface(p)};
  });
}

int spo_scm_close(spo_scm* scm) {
  delete scm;
  return SPO_OK;
}

int spo_scm_values(const spo_scm* scm, long n, const double* points, double* out) {
  return guarded([&] {
    const spo::scm::ScmSurface& s = *scm->surface;
    for_points(n, [&](long k) {
      const double* p = points + 6 * k;
      const auto xy = s.correction().to_cartesian(p[0], p[1]);
      double* o = out + 5 * k;
      o[0] = xy[0];
      o[1] = xy[1];
      o[2] = s.correction().value_hartree(xy[0], xy[1], p[2], p[3], p[4], p[5]);
      o[3] = s.crp().potential_ev(p[0] / s.crp_scale(), p[1] / s.crp_scale(), p[2], p[3], p[4], p[5]);
      o[4] = s.value_ev(p[0], p[1], p[2], p[3], p[4], p[5]);
    });
  });
}

int spo_scm_device_values(const spo_scm* scm, long n, const double* points, double* out) {
  return guarded([&] {
    const spo::crp::Packed packed = spo::crp::pack(scm->surface->crp());
    const spo_hip_crp_surface crp_view = spo::crp_surface_view(packed);
    std::vector<double> atoms;
    const spo_hip_scm_correction scm_view = spo::scm_correction_view(*scm->surface, atoms);
    if (spo_hip_scm_values(&crp_view, &scm_view, n, points, out) != SPO_HIP_OK)
      throw std::runtime_error(std::string("device SCM values: ") + spo_hip_last_error());
  });
}

int spo_scm_parameters(const spo_scm* scm, double* out) {
  return guarded([&] {
    const spo::scm::ScmSurface& s = *scm->surface;
    const auto one = s.correction().to_cartesian(0.0, 1.0);
    const double v[5] = {s.crp_scale(), s.lattice_constant(), one[0], one[1],
                         static_cast<double>(s.correction().atoms().size())};
    for (int i = 0; i < 5; ++i) out[i] = v[i];
  });
}

int spo_crp_parameters(const spo_crp* crp, double* out) {
  return guarded([&] {
    const spo::crp::MoleculeSurface& s = crp->qd->surface();
    const double v[9] = {s.lattice_constant(), s.z_asymptotic(), crp->qd->skewing_angle(), s.zso_min(), s.zso_max(),
                         s.r_min(), s.r_max(), s.z_min(),
                         s.symmetry() == spo::crp::Symmetry::c3v ? 3.0 : (s.symmetry() == spo::crp::Symmetry::c4v ? 4.0 : 6.0)};
    for (int i = 0; i < 9; ++i) out[i] = v[i];
  });
}

}  // extern "C"
