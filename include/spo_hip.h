/* C ABI of the MI355X split-operator propagation backend (libspo_hip.so).
 *
 * It replaces the reference's device backend below the PropagationBackend seam
 * (src/propagation/propagation_backend.hpp:34-59; the CUDA implementation it stands in for is
 * src/cuda/cuda_backend.cpp:51-199). Plain pointers and sizes only: complex arrays are interleaved (re, im)
 * doubles, integer tables are int32. Host layouts are those of src/propagation/wave_function.hpp:12-59:
 * scattering wave (irot, ixy, izr) with irot fastest, ixy = iy + ny * ix, izr = ir + nr * iz; specular wave
 * (ir, iz). The device keeps its own layout (izr fastest) and converts at upload / download.
 *
 * Every function returns 0 on success or a status (SPO_HIP_E_*) with the message in spo_hip_last_error().
 * Calls are synchronous: when one returns, the device work it started has finished. One backend is driven
 * from one host thread at a time. */
#ifndef SPO_HIP_H
#define SPO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPO_HIP_OK = 0,
  SPO_HIP_E_DEVICE = 1,        /* HIP runtime error (reference: CudaError -> exit 1, src/cuda/cuda_error.hpp:10-13) */
  SPO_HIP_E_ARGUMENT = 2,      /* inconsistent shapes or tables */
  SPO_HIP_E_NOT_SUPPORTED = 3  /* no usable device, too little memory, or a shape the library cannot run
                                  (reference: NotSupportedError -> exit 3, src/app/backend_factory.cpp:27, :52) */
};

/* Storage precision of the scattering wave and of the factors it is multiplied with. SPO_HIP_MIXED stores them
 * as complex64 (factors built in FP64 on the host and rounded once); the specular wave, the projection onto it,
 * the Balint-Kurti coefficients and the S-matrix sums stay FP64. */
enum { SPO_HIP_FP64 = 0, SPO_HIP_MIXED = 1 };

typedef struct spo_hip_backend spo_hip_backend;

/* Grid shape of one propagation (src/propagation/kinetic_operators.hpp:18-24 plus the angular and symmetry
 * extents). */
typedef struct spo_hip_shape {
  int n_rot, nx, ny, nr, nz, nz_specular, nz_potential;
  int n_theta, n_phi, n_representatives, n_operations;
  int initial_channel; /* index of (JIn, MJIn) in the rotational basis */
  int mj_max;          /* the mJ blocks are mJ = -mj_max .. mj_max */
  int steps_per_analysis;
  int r_absorber_first, r_absorber_last; /* r points inside the r absorber (specular grid) */
  int needs_shift;     /* the momentum shift of T_xy is not the identity */
  int precision;       /* SPO_HIP_FP64 */
} spo_hip_shape;

/* The factors of the split-operator step (src/propagation/propagation_tables.hpp:20-59), the potential
 * step's merged factors (src/propagation/potential_operator.hpp:23-26), the angular transform
 * (src/propagation/angular_transform.hpp:67-75) and exp(i A) with its symmetry maps
 * (src/symmetry/grid_symmetry.hpp:60-70). */
typedef struct spo_hip_tables {
  const double* zr_full; const double* zr_half;             /* (ir, iz), nr*nz complex */
  const double* xy_full; const double* xy_half;             /* ixy, complex */
  const double* specular_full; const double* specular_half; /* (ir, iz), nr*nz_specular complex */
  double fft_scale_z, fft_scale_r, fft_scale_x, fft_scale_y;
  const double* rotation_half;      /* (irot, ir) complex: exp(-i E_rot dt/2) */
  const double* before_interaction; /* (irot, ir) complex: exp(-V_opt,r dt) rot(dt/2) */
  const double* gas_phase_step;     /* (irot, ir) complex: exp(-V_opt,r dt) rot exp(-i vasy dt) rot */
  const double* specular_gas_phase; /* ir complex: rot exp(-i vasy dt) rot of the initial channel */
  const double* z_optical;          /* nz real */
  const double* r_optical;          /* nr real */
  const double* z_specular_optical; /* nz_specular real */
  const double* shift_forward; const double* shift_backward;                  /* ixy complex */
  const double* specular_to_scattering; const double* scattering_to_specular; /* ixy complex */
  const double* legendre;       /* sqrt(w_t / nphi) P_J^m(theta_t), ordered (mJ block, theta, J) */
  const int* block_first;       /* per mJ block: first basis index */
  const int* block_size;        /* per mJ block: number of J */
  const double* exp_potential;  /* exp(i A), A = -V dt, (itp, irep, izr) complex, izr < nz_potential * nr */
  const int* representative_of; /* per ixy */
  const int* operation_of;      /* per ixy */
  const int* phi_index;         /* (iphi, operation) */
} spo_hip_tables;

/* The Balint-Kurti analysis (src/analysis/balint_kurti.cpp:42-84) and the time-to-energy sums
 * (src/analysis/s_matrix.cpp:49-56). Coefficients are (is, v) with is = irotana * n_diffraction + idif. */
typedef struct spo_hip_analysis {
  int n_states, n_v, n_rotation, n_diffraction, n_energies;
  int iz_analysis, initial_state;
  const int* rotation_index;    /* per analysed rotational state: index in the basis */
  const int* diffraction_slot;  /* per analysed channel: iy + ny * ix of the (x, y) FFT output */
  const double* radial_weight;  /* (ir, irotana, v): phi_{v, j(irotana)}(r) * dR */
  const double* initial_radial; /* (ir, v): phi_{v, JIn}(r) */
  double dr, xy_scale, analysis_time_step;
  const double* total_energy_hartree; /* per energy */
} spo_hip_analysis;

/* The message of the last failed call on this thread (the reference throws CudaError / NotSupportedError with
 * the message, src/cuda/cuda_error.hpp:10-13). */
const char* spo_hip_last_error(void);

/* Number of HIP devices visible (0 on a machine without a GPU; never an error). Replaces the device probe of
 * the reference's backend factory (src/app/backend_factory.cpp:22-29, :52). */
int spo_hip_device_count(int* count);

/* Uploads the tables and the initial specular wave (ir, iz); the scattering wave starts at zero. Replaces the
 * BackendFactory call for a GPU device (src/propagation/propagation_backend.hpp:73-74; the CUDA backend's
 * constructor, src/cuda/cuda_backend.cpp:93). tables->exp_potential is exp(i A) with A = -V dt, built by the
 * caller in FP64 from the potential grid in hartree (the reference premultiplies by -dt inside its backends,
 * src/propagation/cpu_backend.cpp:17-28). */
int spo_hip_create(const spo_hip_shape* shape, const spo_hip_tables* tables, const spo_hip_analysis* analysis,
                   const double* initial_specular, spo_hip_backend** out);
/* Frees the backend (the reference backend's destructor; RAII buffers, src/cuda/device_buffer.hpp:25-53). */
int spo_hip_destroy(spo_hip_backend* backend);

/* Full wave packet, host layouts: upload replaces read_wave_packet (src/propagation/propagation_backend.hpp:39-42),
 * download replaces wave_packet_on_host (:43-45). */
int spo_hip_upload_wave(spo_hip_backend* backend, const double* scattering, const double* specular);
int spo_hip_download_wave(spo_hip_backend* backend, double* scattering, double* specular);

/* T(dt/2) V, (T V) x (n-1), T(dt/2) (src/propagation/split_operator_schedule.cpp:7-13). Replaces
 * propagate_analysis_step (src/propagation/propagation_backend.hpp:47-48). */
int spo_hip_propagate_analysis_step(spo_hip_backend* backend);
/* One analysis step of each of n distinct backends (independent runs on this device): their time-step blocks are
 * enqueued round robin, each run's in its own order on its own streams, so that runs too small to fill the GPU
 * overlap; returns when all are done. Each run's results are those of spo_hip_propagate_analysis_step. No
 * counterpart in the reference: the batch of initial states or distorted surfaces of one device. */
int spo_hip_propagate_batch(spo_hip_backend* const* backends, int n);
/* Graph mode on (1) or off (0): an analysis step is captured into a hipGraph at its next call (the T/V blocks on the
 * backend's two streams; the step's launches and arguments are the same every step) and replayed from then on,
 * with the results of the uncaptured step. Not with kernel timing. No counterpart in the reference. */
int spo_hip_graph_enable(spo_hip_backend* backend, int on);
/* A batch of n distinct backends stepped together through one hipGraph: at the first spo_hip_batch_step the blocks
 * of every member's analysis step are enqueued round robin on the members' streams (as spo_hip_propagate_batch)
 * inside one capture that forks from the first member's stream to the others and joins them back; every step
 * launches that graph once and returns when all members are done. Each member's results are those of
 * spo_hip_propagate_analysis_step. Not with kernel timing. A member may not be destroyed before the batch
 * (spo_hip_destroy refuses it). No counterpart in the reference: one launch per step for the batch of initial
 * states or distorted surfaces of one device. */
typedef struct spo_hip_batch spo_hip_batch;
int spo_hip_batch_create(spo_hip_backend* const* backends, int n, spo_hip_batch** out);
int spo_hip_batch_step(spo_hip_batch* batch);
int spo_hip_batch_destroy(spo_hip_batch* batch);
/* One T block: specular grid, T_Zr, T_xy, then V unless kinetic_only (src/propagation/split_operator.cpp:17-21);
 * the unit the reference's CUDA tests check stage by stage (tests/cuda/test_cuda_backend.cpp:94-106). */
int spo_hip_kinetic_and_potential(spo_hip_backend* backend, int half_step, int kinetic_only);

/* The specular wave (ir, iz) after the last step; replaces specular() (src/propagation/propagation_backend.hpp:50). */
int spo_hip_get_specular(spo_hip_backend* backend, double* specular);
/* nr blocks of Z index iz, layout (irot, ixy, ir); replaces z_plane(iz) (src/propagation/propagation_backend.hpp:51-52). */
int spo_hip_get_z_plane(spo_hip_backend* backend, int iz, double* out);
/* sum |psi|^2 over (irot, ixy) per izr, n_zr values; replaces block_densities()
 * (src/propagation/propagation_backend.hpp:53-55). */
int spo_hip_block_densities(spo_hip_backend* backend, double* out);
/* One (x, y) point, n_zr * n_rot complex values, layout (izr, irot); replaces site_column(ixy, out)
 * (src/propagation/propagation_backend.hpp:56-58). */
int spo_hip_site_column(spo_hip_backend* backend, int ixy, double* out);

/* Balint-Kurti coefficients at the analysis line; with surface_fixed the initial state's from the specular wave
 * at iz_specular. Then sums(is, v, ien) += c(is, v) exp(i E t) dtAna in FP64. coefficients (n_states * n_v
 * complex) may be NULL. Replaces BalintKurtiAnalyzer::coefficients on the host copy of z_plane
 * (src/analysis/balint_kurti.cpp:39-84) and SMatrixAccumulator::add (src/analysis/s_matrix.cpp:49-56): the
 * plane stays on the device and only the coefficients cross PCIe. */
int spo_hip_analyse(spo_hip_backend* backend, int surface_fixed, int iz_specular, double t, double* coefficients);
/* The sums, (is, v, ien) with is fastest: the input of SMatrixAccumulator::result_at and combine_s_matrix
 * (src/analysis/s_matrix.cpp:70-176), which stay on the host in FP64. */
int spo_hip_get_s_sums(spo_hip_backend* backend, double* out);

/* Energy-resolved sums at chosen (x, y) sites, replacing EnergyResolvedAccumulator
 * (src/analysis/energy_resolved.hpp:52-68, energy_resolved.cpp:62-91) with sums kept on the device: start zeroes
 * the FP64 sums of n_sites sites (site_ixy = iy + ny * ix) and n_energies energies; add sums one analysis step,
 * s += psi(site) exp(i E_tot t) dtAna and sigma += chi exp(i E_tot t) dtAna, given exp(i E_tot t) per energy
 * (n_energies complex); read returns them in the reference's layouts: sums [site, energy, Z, r, rot] (rot
 * fastest, n_sites * n_energies * nz * nr * n_rot complex) and specular_sums [energy, Z_specular, r] (r fastest). */
int spo_hip_energy_resolved_start(spo_hip_backend* backend, int n_sites, const int* site_ixy, int n_energies,
                                  double analysis_time_step);
int spo_hip_energy_resolved_add(spo_hip_backend* backend, const double* phase_factors);
int spo_hip_energy_resolved_read(spo_hip_backend* backend, double* sums, double* specular_sums);

/* Adds BK coefficients (n_states * n_v complex, (is, v)) of an analysis step of an earlier run at its time to the
 * S-matrix sums, as spo_hip_analyse does for its own; for a restart. Replaces SMatrixAccumulator::add of
 * WavePacketPropagation::add_earlier_analysis_step (src/propagation/propagation.hpp:88). */
int spo_hip_accumulate(spo_hip_backend* backend, double time, const double* coefficients);

/* Kernel timing with HIP events recorded on the backend's stream around every launch (no counterpart in the
 * reference; for the roofline measurement). Enabling clears the record. */
typedef struct spo_hip_kernel_time {
  char name[48];
  long launches;
  double total_ms;
} spo_hip_kernel_time;
int spo_hip_profile_enable(spo_hip_backend* backend, int on);
/* Per kernel: launches and summed event time since profiling was enabled; count receives the number of kernels. */
int spo_hip_profile_read(spo_hip_backend* backend, spo_hip_kernel_time* out, int capacity, int* count);

/* Potential energy surfaces evaluated on the device (the SystemIDs "HardWallPES" and "H2Cu111DaiLight";
 * src/potential/hard_wall_potential.cpp, src/potential/dai_light_cu111_potential.cpp). */
enum { SPO_HIP_PES_HARDWALL = 1, SPO_HIP_PES_DAILIGHT = 2, SPO_HIP_PES_CRP = 3 };

/* The PES in eV at n points, points = (x, y, z, r, theta, phi) per point in bohr and radians; replaces
 * Pes::value (src/potential/pes.hpp) for these surfaces. Needs no backend. */
int spo_hip_pes_values(int pes, long n, const double* points, double* values);

/* spo_hip_pes_values without any device allocation, for callers that cannot allocate device memory (tight
 * serving loops, code that forbids hipMalloc). No counterpart in the reference. The points and values pass
 * through a fixed scratch array in device memory (3.5 MiB per device, held until the process ends), so n is at
 * most 65536 per call (larger n: SPO_HIP_E_ARGUMENT). All calls on one device share that array: calls that
 * overlap on one device (from several host threads) race. Same results as spo_hip_pes_values. */
int spo_hip_pes_values_static(int pes, long n, const double* points, double* values);

/* The potential grid the propagation stores (src/potential/potential_grid.cpp:65-127): V in hartree at the
 * symmetry representatives, values[iphi + n_phi (itheta + n_theta (irep + n_representatives izr))] for
 * izr = ir + nr iz < nz_potential nr, with the cosine switch into gas_phase between z_min_switch and
 * z_max_switch (:35-56) and the VMin / VMax limits (:139-159). gas_phase (nr values, hartree) is the surface at
 * the switch's asymptote, which the caller evaluates. Replaces compute_potential_grid for the surfaces above;
 * needs no backend. */
typedef struct spo_hip_pes_grid {
  int pes; /* SPO_HIP_PES_* */
  int nx, ny, nr, nz_potential, n_theta, n_phi, n_representatives;
  const double* x; const double* y; const double* z; const double* r; /* z: nz_potential values */
  const double* theta; const double* phi;
  const int* representatives; /* full-grid ixy = iy + ny ix of each representative */
  int use_switching;
  double z_min_switch, z_max_switch;
  const double* gas_phase;
  int impose_v_min, impose_v_max;
  double v_min, v_max; /* hartree */
  const struct spo_hip_crp_surface* crp; /* SPO_HIP_PES_CRP: the surface; NULL otherwise */
  const struct spo_hip_scm_correction* scm; /* SPO_HIP_PES_CRP with a distorted slab: the SCM correction, or NULL */
} spo_hip_pes_grid;
int spo_hip_potential_grid(const spo_hip_pes_grid* grid, double* values);

/* A CRP surface (SystemID CRPlibrary, C3v or C6v data set) as flat host arrays: the splines and the atom cluster
 * that the reference's CRP library builds from the data set (src/crplib/hexagonal_atom_surface.cpp,
 * molecule_surface.cpp), in internal units (bohr, hartree). The device evaluates V alone, the value statements of
 * the QD wrapper (src/crplib/qd_wrapper.cpp:46-66): clamps, atom positions, V = Va + Vb + I below ZsoMin, the switch
 * to the gas-phase potential up to ZsoMax, eV, minus the free molecule at its last r. libspo_host.so fills it from
 * a data set (host/crp.cpp pack). */
typedef struct spo_hip_crp_surface {
  int c3v; /* 1: C3v (10 atom sites, 29 molecular cuts); 0: C6v (7, 18) or C4v (square below) */
  double lattice_constant, zso_min, zso_max, r_min, r_max, z_min, h2_asymptote;
  int n_corrugation; const double* corrugation; /* x[n] y[n] F''[n] */
  int n_cluster; const double* cluster;         /* x y z layer per atom, in summation order */
  const int* atom_offsets;                      /* n_sites + 1: site s has knots [off[s], off[s+1]) */
  const double* atom_splines;                   /* site s at 3 off[s]: x[k] y[k] F''[k] */
  int n_cuts; const int* cut_shape;             /* per cut: nr, nz, first cut with the same (r, Z) grid */
  const long* cut_offsets;                      /* n_cuts + 1: r[nr] Z[nz] f fxx fyy fxxyy (nr nz each) */
  const double* cut_data;
  int n_h2; const double* h2;                   /* x y F'' */
  int n_z_dependence; const double* z_dependence;
  /* C4v (square = 1; 16 molecular cuts, the 32-atom cluster, no atom splines): the cell's periodic 3D spline,
   * x[5] y[5] z[cell_nz] then f, f_xx, f_yy, f_xxyy, f_zz, f_zzxx, f_zzyy, f_zzxxyy (25 cell_nz each, x fastest) */
  int square, cell_nz; const double* cell_spline;
} spo_hip_crp_surface;

/* The SCM correction of a distorted slab snapshot (SystemID CRPDistrtSurface, src/potential/scm/): with it the
 * surface is V = CRP(u / scale, v / scale) + dV / 0.036749308286 (eV), dV in hartree the sum over the slab atoms of
 * pair functions at the distorted minus at the ideal positions (scm_correction.cpp:105-125). */
typedef struct spo_hip_scm_correction {
  int rescale;                     /* the H atoms are taken to the ideal lattice's frame (centre x, y divided by scale) */
  double scale;                    /* a_T / a_CRP, 1 without rescaling */
  double cos_skew, sin_skew;       /* Cartesian (u + cos v, sin v) of skewed (u, v) */
  double layers[2][18];            /* params.dat P1-P18 of layers 1 and 2 */
  int n_atoms; const double* atoms; /* distorted x y z, ideal x y z, layer (1 or 2) per atom, in summation order */
} spo_hip_scm_correction;

/* The CRP surface in eV at n points (u, v, Z, r, theta, phi: skewed bohr, rad), as the QD wrapper. */
int spo_hip_crp_values(const spo_hip_crp_surface* surface, long n, const double* points, double* values);
/* The same with the SCM correction (scm may be NULL). */
int spo_hip_scm_values(const spo_hip_crp_surface* surface, const spo_hip_scm_correction* scm, long n,
                       const double* points, double* values);

/* The potential surrogate: an MLP of 6 features -> 128 -> 128 -> 128 -> 1 with SiLU whose hidden layers run on
 * bf16 MFMA with FP32 accumulation, trained with Adam on the mean squared error. Parameters (FP32, this order):
 * W1 (6 x 128, row per input), b1 (128), W2 (128 x 128, row per input unit), b2, W3, b3, W4 (128), b4 (1).
 * x: 6 floats per row; batch: the rows of one training step, a multiple of 16. No counterpart in the reference:
 * the surrogate of potential grids for batches of distorted surfaces. */
enum { SPO_HIP_MLP_PARAMS = 34049 };
typedef struct spo_hip_mlp spo_hip_mlp;
int spo_hip_mlp_create(int batch, spo_hip_mlp** out); /* parameters 0: set them before use */
int spo_hip_mlp_destroy(spo_hip_mlp* mlp);
int spo_hip_mlp_set_params(spo_hip_mlp* mlp, const float* params); /* also clears Adam's moments */
int spo_hip_mlp_get_params(const spo_hip_mlp* mlp, float* params);
/* loss and gradient of the mean squared error over one batch of rows */
int spo_hip_mlp_gradient(spo_hip_mlp* mlp, const float* x, const float* t, float* loss, float* grad);
/* `steps` Adam steps over consecutive batches of the n rows (n a multiple of the batch; wraps around); losses[s]:
 * the batch loss of step s before its update */
int spo_hip_mlp_train(spo_hip_mlp* mlp, long n, const float* x, const float* t, int steps, float lr, float* losses);
int spo_hip_mlp_predict(spo_hip_mlp* mlp, long n, const float* x, float* y);

#ifdef __cplusplus
}
#endif

#endif
