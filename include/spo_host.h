/* C ABI of the host driver (libspo_host.so): the set-up of a run from its JSON input, the time loop over a
 * propagation backend and the analysis results. Complex arrays are interleaved (re, im) doubles; layouts are
 * those of the reference's host side (src/propagation/wave_function.hpp:12-59).
 *
 * Every function returns 0 on success or a status (SPO_E_*), with the message in spo_last_error(). The status
 * codes are the reference program's exit codes for the same failures (src/app/run.cpp:596-607). */
#ifndef SPO_HOST_H
#define SPO_HOST_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPO_OK = 0,
  SPO_E_OTHER = 1,         /* device or other runtime error */
  SPO_E_INPUT = 2,         /* invalid input */
  SPO_E_NOT_SUPPORTED = 3, /* input or device not supported by the library */
  SPO_E_CHECKPOINT = 4     /* a checkpoint or BK file missing, corrupt or written for another input */
};

enum {
  SPO_DEVICE_HIP = 1,      /* the HIP backend (libspo_hip.so) */
  SPO_DEVICE_EXTERNAL = 2  /* a backend made by the caller's factory */
};

typedef struct spo_run spo_run;

/* A backend factory: input is a const spo::BackendInput*, the result a new spo::PropagationBackend* (owned by
 * the run afterwards), or NULL on failure. Replaces the reference's BackendFactory
 * (src/propagation/propagation_backend.hpp:73-74) at the C boundary. */
typedef void* (*spo_backend_factory_fn)(const void* backend_input);

/* Shapes and counts of a run. */
typedef struct spo_run_info {
  int n_rot, nx, ny, nr, nz, nz_specular, nz_potential;
  int n_theta, n_phi, n_representatives, initial_channel;
  int n_states, n_v, n_energies, n_diffraction;
  int steps_per_analysis, analysis_steps_total, analyses_per_average;
  int iz_analysis;
  double time_step;
} spo_run_info;

const char* spo_last_error(void);

/* Reads input_path (JSON, the reference's keys), merges overrides_json (a JSON object, may be NULL), builds the
 * set-up and the backend. factory is used with SPO_DEVICE_EXTERNAL only. */
int spo_run_create(const char* input_path, const char* overrides_json, int device, spo_backend_factory_fn factory,
                   spo_run** out);
/* spo_run_create with the potential grid of the set-up replaced by `potential` before the backend is built: n values
 * in hartree in the grid's own order (spo_run_array "potential.values": (itheta phi, representative, izr) with
 * (theta, phi) fastest, izr = iz NR + ir over the Z points of the interaction region), after the switch to the gas
 * phase and the VMin / VMax clamp. No counterpart in the reference: it propagates a grid built elsewhere (a potential
 * surrogate). A size that differs from the grid's, or a value that is not finite, is an error. */
int spo_run_create_with_potential(const char* input_path, const char* overrides_json, int device,
                                  spo_backend_factory_fn factory, const double* potential, size_t n_potential,
                                  spo_run** out);
int spo_run_destroy(spo_run* run);
int spo_run_get_info(const spo_run* run, spo_run_info* info);
/* The backend object (a spo::PropagationBackend*), for stage-level checks. */
void* spo_run_backend(spo_run* run);
/* The calls the run has made across the backend seam so far, after the reference's CountingBackend
 * (tests/counting_backend.hpp:35-113): counts[0..11] = propagate_analysis_step (batched steps included),
 * kinetic_and_potential, z_plane, block_densities, specular, wave_packet_on_host, read_wave_packet, site_column,
 * analyse, s_matrix_sums, add_earlier_analysis_step, energy_resolved_add. capacity >= 12. */
int spo_run_call_counts(const spo_run* run, long* counts, int capacity);

/* Analysis steps of the time loop: propagate, move the analysis surface, project and accumulate. */
int spo_run_steps(spo_run* run, int n_analysis_steps);
/* Up to n_analysis_steps analysis steps of n runs in lockstep: the propagations of one step together (the HIP
 * runs' launches interleaved on their own streams, spo_hip_propagate_batch), then each run's analysis; each run's
 * results equal spo_run_steps'. No counterpart in the reference: a batch of independent runs on one device. */
int spo_run_steps_batch(spo_run* const* runs, int n, int n_analysis_steps);
/* The backend's propagate_analysis_step() alone (no analysis). */
int spo_run_propagate(spo_run* run);
/* One T block of the schedule (specular grid, T_Zr, T_xy, then V unless kinetic_only) on the backend. */
int spo_run_kinetic_and_potential(spo_run* run, int half_step, int kinetic_only);
/* Analysis at time t of the current wave packet (surface update, projection, accumulation); bk may be NULL,
 * else n_states * n_v complex values (is, v). */
int spo_run_analyse(spo_run* run, double t, double* bk);

/* Wave packet in host layout: scattering (irot, ixy, izr) n_rot*nx*ny*nr*nz complex, specular (ir, iz). */
int spo_run_get_wave(spo_run* run, double* scattering, double* specular);
int spo_run_set_wave(spo_run* run, const double* scattering, const double* specular);
int spo_run_get_specular(spo_run* run, double* specular);
int spo_run_block_densities(spo_run* run, double* out);               /* n_zr values */
int spo_run_z_plane(spo_run* run, int iz, double* out);                /* n_rot*nx*ny*nr complex, (irot, ixy, ir) */
int spo_run_site_column(spo_run* run, int ixy, double* out);           /* n_zr*n_rot complex, (izr, irot) */

/* Recorded BK coefficients of analysis step k (0-based), n_states*n_v complex. */
int spo_run_bk_record(const spo_run* run, int k, double* out);
int spo_run_bk_count(const spo_run* run, int* count);
/* Writes the native restart files of the run so far into directory (created if missing): checkpoint.bin (wave
 * packet, time, analysis surface) and bk_coefficients.bin (every BK record, a restart's earlier ones first), in
 * the reference's format version 2 (src/io/checkpoint.cpp:216-309; files named as src/app/run.cpp:67-68).
 * A run made with NewCalc = false, RestartDir = directory and tStart = that time continues from them. */
int spo_run_write_checkpoint(spo_run* run, const char* directory);
/* Averages rows so far: count, then rows of (time, norm, z, r, rest). */
int spo_run_averages(const spo_run* run, double* rows, int capacity, int* count);
/* The analysis surface: fixed flag, index and Z of the fixed line (or 0). */
int spo_run_surface(const spo_run* run, int* fixed, int* index, double* z_line);

/* S-matrix results at every energy from the sums so far. s: n_energies*n_states*n_v complex (ien, v, is) with is
 * fastest; p likewise (doubles); p_reaction: n_energies; total_energy: n_energies (eV). zana_convention != 0
 * refers the initial state's phase to Zana instead of the fixed specular line. Any pointer may be NULL. */
int spo_run_results(const spo_run* run, int zana_convention, double* s, double* p, double* p_reaction,
                    double* total_energy);
/* The raw FP64 sums (ien, v, is), n_energies*n_states*n_v complex. */
int spo_run_s_sums(const spo_run* run, double* out);

/* Energy-resolved sums (EnergyResolvedSites / EnergyResolvedEnergies, src/analysis/energy_resolved.hpp:16-43): the
 * number of sites and energies (0 when the keys are not set), and the sums so far in the reference's layouts,
 * [site, energy, Z, r, rot] (rot fastest) and [energy, Z_specular, r] (r fastest). */
int spo_run_energy_resolved_shape(const spo_run* run, int* n_sites, int* n_energies);
int spo_run_energy_resolved(spo_run* run, double* sums, double* specular_sums);

/* Set-up arrays by name, for checks of the set-up: "gas.energies" (j + (jmax+1) v), "gas.wave_functions"
 * (ir, j, v), "potential.values" (itp, irep, izr), "potential.asymptotic", "angular.thetas", "angular.weights",
 * "angular.phis", "basis.j", "basis.mj", "state.j", "state.mj", "state.n", "state.m", "state.order",
 * "level.energies" (E(v, j) in eV, j + (jmax+1) v), "symmetry.representatives", "tables.z_optical",
 * "tables.r_optical", "initial.specular" (complex), "angular.legendre" (sqrt(w_t / nphi) P_J^m(theta_t), ordered
 * (mJ block, theta, J)), "grids.x", "grids.y" (skewed coordinates), "grids.z", "grids.r", "derived.cell" (lattice_x,
 * lattice_y, skew angle, total mass, dx, dy, dt, Cartesian parallel momentum x, y), "units.hartree_in_ev" (the
 * set-up's factor), "switch.z" (ZminSwitch, ZmaxSwitch; empty without UseSwitching). count receives the element count (doubles, or complex
 * values for complex arrays); out may be NULL to query the count. */
int spo_run_array(const spo_run* run, const char* name, double* out, size_t capacity, size_t* count);

/* Per-kernel device timing of the backend (HIP events around every launch); enabling clears the record.
 * spo_run_kernel_times fills up to capacity entries: names (48 chars each, NUL-terminated), launches and summed
 * milliseconds; count receives the number of kernels. A backend without device kernels reports none. */
int spo_run_kernel_timing(spo_run* run, int on);
int spo_run_kernel_times(spo_run* run, char* names, long* launches, double* total_ms, int capacity, int* count);

/* V(x, y, Z, r, theta, phi) in eV of a potential by SystemID. */
int spo_pes_value(const char* system_id, double x, double y, double z, double r, double theta, double phi,
                  double* out);

/* The CRP potential of a data set directory (SystemID CRPlibrary; the reference's CRP library, src/crplib/, as
 * src/potential/crp_potential.cpp uses it). C3v, C4v and C6v data sets; SPO_E_INPUT for a missing or unusable one.
 * Points are evaluated in parallel; every function takes n points and writes n records. */
typedef struct spo_crp spo_crp;
int spo_crp_open(const char* directory, spo_crp** out);
int spo_crp_close(spo_crp* crp);
/* length_unit 0 bohr / 1 angstrom, energy_unit 2 hartree / 3 eV / 4 kcal/mol / 5 kJ/mol / 6 J
 * (src/crplib/units.hpp:11-17). atoms[6k..] = xa ya za xb yb zb (Cartesian); out[7k..] = V, dV/dxa dya dza dxb dyb
 * dzb, all 0 where the surface gives up silently (src/crplib/wrapper.cpp, Calculate_CRP_MoleculeSurface_PES). */
int spo_crp_molecule(const spo_crp* crp, long n, const double* atoms, int length_unit, int energy_unit, double* out);
/* positions[3k..] = x y z; out[4k..] = V, dV/dx dy dz (Calculate_CRP_AtomSurface_PES). */
int spo_crp_atom(const spo_crp* crp, long n, const double* positions, int length_unit, int energy_unit, double* out);
/* points[6k..] = u v Z r theta phi (skewed bohr, rad); out[k] = V in eV (src/crplib/qd_wrapper.cpp:46-66). */
int spo_crp_qd(const spo_crp* crp, long n, const double* points, double* out);
/* The same values evaluated on the GPU (spo_hip_crp_values of libspo_hip.so on this surface's packed tables). */
int spo_crp_device_values(const spo_crp* crp, long n, const double* points, double* out);
/* zr[2k..] = Z r (bohr); out[5k..] = switch F, dF/dZ, V2D, dV2D/dZ, dV2D/dr (hartree). */
int spo_crp_gas_phase(const spo_crp* crp, long n, const double* zr, double* out);
/* atoms as for spo_crp_molecule (bohr); out[6k..] = centre x y Z, r, theta, phi (CAMBCORD). */
int spo_crp_geometry(long n, const double* atoms, double* out);
/* out[0..8] = lattice constant, Zasy, skewing angle, ZsoMin, ZsoMax, Rmin, Rmax, Zmin (bohr, rad), and 3, 4 or 6
 * for C3v, C4v or C6v. */
int spo_crp_parameters(const spo_crp* crp, double* out);

/* The CRP surface plus the SCM correction of a distorted slab snapshot (SystemID CRPDistrtSurface; the reference's
 * src/potential/scm/): params.dat, the slab position file, the 1-based snapshot, the cut-off (bohr) and whether the
 * slab is rescaled to the CRP lattice (EnableSurfExpan). */
typedef struct spo_scm spo_scm;
int spo_scm_open(const char* crp_data_set, const char* parameter_file, const char* position_file, int configuration,
                 double cutoff, int rescale, spo_scm** out);
int spo_scm_close(spo_scm* scm);
/* points[6k..] = u v Z r theta phi; out[5k..] = Cartesian x, y of (u, v), dV (hartree), the CRP at (u/s, v/s) (eV)
 * and their sum (eV), as the reference's SCM driver prints them. */
int spo_scm_values(const spo_scm* scm, long n, const double* points, double* out);
/* The total (eV) evaluated on the GPU (spo_hip_scm_values). */
int spo_scm_device_values(const spo_scm* scm, long n, const double* points, double* out);
/* out[0..4] = rescaling factor s (1 without rescaling), the grid's lattice constant, cos and sin of the skewing
 * angle, the number of slab atoms in use. */
int spo_scm_parameters(const spo_scm* scm, double* out);

#ifdef __cplusplus
}
#endif

#endif
