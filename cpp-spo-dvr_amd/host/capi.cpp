// C ABI of the host driver (include/spo_host.h).
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "../../include/spo_host.h"
#include "checkpoint.hpp"
#include "driver.hpp"
#include "json.hpp"

struct spo_run {
  std::shared_ptr<spo::RunSetup> setup;
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
  } catch (const std::exception& e) {
    return fail(SPO_E_OTHER, e.what());
  }
}

void copy_complex(const std::vector<spo::cplx>& v, double* out) { std::memcpy(out, v.data(), v.size() * sizeof(spo::cplx)); }

}  // namespace

extern "C" {

const char* spo_last_error(void) { return g_error.c_str(); }

int spo_run_create(const char* input_path, const char* overrides_json, int device, spo_backend_factory_fn factory,
                   spo_run** out) {
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
    run->propagation = std::make_unique<spo::Propagation>(run->setup, make);
    *out = run.release();
  });
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

int spo_run_steps(spo_run* run, int n) {
  return guarded([&] { run->propagation->run(n); });
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
