// The MI355X propagation backend behind the C ABI of include/spo_hip.h: device buffers, FFT plans and the launch
// sequence of one analysis step. Kernels: kernels_ct.hpp (compile-time shapes) and kernels.hpp (any shape); LDS
// FFT engines: fft_ct.hpp, fft_lds.hpp.
//
// The scattering wave and the tables it is multiplied with are stored in T (double, or float in the mixed mode;
// the factors are built in FP64 on the host and rounded once). The specular wave, the projection onto it, the
// Balint-Kurti coefficients and the S-matrix sums are FP64 in both modes.
#include <hip/hip_runtime.h>


#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spo_hip.h"
#include "kernels_ct.hpp"
#include "kernels_mfma.hpp"
#include "kernels_pes.hpp"
#include "kernels_mlp.hpp"

using spo_dev::cx;

namespace {

thread_local std::string g_error;

struct Status : std::runtime_error {
  int code;
  Status(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Status(SPO_HIP_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIP_CHECK(call) check((call), #call)

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return SPO_HIP_OK;
  } catch (const Status& s) {
    g_error = s.what();
    return s.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return SPO_HIP_E_DEVICE;
  }
}

// Device memory owned by the backend.
class Buffer {
 public:
  Buffer() = default;
  explicit Buffer(std::size_t bytes) : bytes_(bytes) {
    if (bytes) HIP_CHECK(hipMalloc(&ptr_, bytes));
  }
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  Buffer(Buffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
  }
  Buffer& operator=(Buffer&& o) noexcept {
    std::swap(ptr_, o.ptr_);
    std::swap(bytes_, o.bytes_);
    return *this;
  }
  ~Buffer() {
    if (ptr_) (void)hipFree(ptr_);
  }
  template <typename T>
  T* as() const { return static_cast<T*>(ptr_); }

 private:
  void* ptr_ = nullptr;
  std::size_t bytes_ = 0;
};

void require_device() {
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1)
    throw Status(SPO_HIP_E_NOT_SUPPORTED, "no HIP device is visible");
}

Buffer upload(const void* host, std::size_t bytes) {
  Buffer b(bytes);
  if (bytes) HIP_CHECK(hipMemcpy(b.as<void>(), host, bytes, hipMemcpyHostToDevice));
  return b;
}

// A complex factor in the storage type: FP64 as is; complex64 rounded to nearest (tables applied at every step
// also carry the remainder, factor_lo, or for twiddles a modulus correction, DevicePlan::eps).
template <typename T>
cx<T> to_storage(double x, double y) {
  if constexpr (sizeof(T) == 8) {
    return {x, y};
  } else {
    return {static_cast<float>(x), static_cast<float>(y)};
  }
}

// Complex array (interleaved doubles) -> device array of cx<T>, optionally times a real factor (in FP64 first).
// Large complex64 tables (exp(i A)) are rounded on the device.
template <typename T>
Buffer upload_complex(const double* p, std::size_t n, double factor = 1.0) {
  if (sizeof(T) == 4 && n > (std::size_t(1) << 18)) {
    Buffer full = upload(p, n * sizeof(cx<double>));
    Buffer out(n * sizeof(cx<T>));
    hipLaunchKernelGGL(spo_dev::round_to_complex64_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256),
                       0, nullptr, full.as<const double>(), out.as<float>(), static_cast<long>(n), factor);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(nullptr));
    return out;
  }
  std::vector<cx<T>> v(n);
  for (std::size_t i = 0; i < n; ++i) v[i] = to_storage<T>(p[2 * i] * factor, p[2 * i + 1] * factor);
  return upload(v.data(), n * sizeof(cx<T>));
}
// complex64: per entry of a factor table uploaded by upload_complex, the FP32 remainder z - stored z (z the FP64
// value times factor); the kernels multiply by stored z + remainder (mul_hilo). Empty in FP64.
template <typename T>
Buffer factor_lo(const double* p, std::size_t n, double factor = 1.0) {
  if constexpr (sizeof(T) == 8) {
    return Buffer();
  } else {
    std::vector<cx<float>> lo(n);
    for (std::size_t i = 0; i < n; ++i) {
      const double x = p[2 * i] * factor, y = p[2 * i + 1] * factor;
      const cx<float> w = to_storage<float>(x, y);
      lo[i] = {static_cast<float>(x - static_cast<double>(w.x)), static_cast<float>(y - static_cast<double>(w.y))};
    }
    return upload(lo.data(), n * sizeof(cx<float>));
  }
}

template <typename T>
Buffer upload_real(const double* p, std::size_t n) {
  std::vector<T> v(p, p + n);
  return upload(v.data(), n * sizeof(T));
}

// Radices of the run-time engine: 8 while it divides, then 4, 2, 3, 5, 7, 11.
std::vector<int> radices_of(int n) {
  std::vector<int> r;
  int m = n;
  for (int p : {8, 4, 2, 3, 5, 7, 11})
    while (m % p == 0) {
      r.push_back(p);
      m /= p;
    }
  if (m != 1) throw Status(SPO_HIP_E_NOT_SUPPORTED, "FFT length " + std::to_string(n) + " has a prime factor above 11");
  if (r.size() > 12) throw Status(SPO_HIP_E_NOT_SUPPORTED, "FFT length " + std::to_string(n) + " has too many factors");
  return r;
}

// An FFT length: twiddle table W[m] = exp(-2 pi i m / n) on the device (in T) and the run-time plan.
template <typename T>
struct DevicePlan {
  Buffer table, eps;
  spo_dev::FftPlan plan{};
  explicit DevicePlan(int n) {
    const std::vector<int> r = radices_of(n);
    std::vector<cx<T>> w(n);
    std::vector<T> e(n, T(0));
    for (int m = 0; m < n; ++m) {
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / n;
      w[m] = to_storage<T>(static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a)));
      // complex64: the correction that makes the product's modulus exact (twiddle(), kernels_ct.hpp)
      const double m2 = static_cast<double>(w[m].x) * w[m].x + static_cast<double>(w[m].y) * w[m].y;
      e[m] = static_cast<T>(1.0 / std::sqrt(m2) - 1.0);
    }
    table = upload(w.data(), w.size() * sizeof(cx<T>));
    eps = upload(e.data(), e.size() * sizeof(T));
    plan.n = n;
    plan.n_stages = static_cast<int>(r.size());
    for (std::size_t i = 0; i < r.size(); ++i) plan.radix[i] = r[i];
    plan.table = table.as<void>();
    plan.eps = eps.as<void>();
  }
  const cx<T>* tw() const { return table.as<cx<T>>(); }
  // 1 / |w_m| - 1 of the stored twiddles (complex64 kernels only)
  const T* tw_eps() const { return eps.as<T>(); }
};

unsigned blocks_for(long n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

constexpr int kThreads = 256;

void set_lds(const void* kernel, std::size_t bytes) {
  HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
}

// ---- instantiations of the compile-time-shaped kernels ---------------------------------------------------------
// Phi lengths of the interaction kernel with the contractions on the vector units.
#define SPO_PHI_LENGTHS(X) X(48) X(22) X(10) X(12) X(8) X(56)
// T_Zr planes with register-resident passes (NZ, NR, A1, A2, B1, B2, threads): NR = A1 A2, NZ = B1 B2. Entries of
// one shape that differ only in the thread count compute identically; the launch tuner picks among them.
// Other factorisations of a shape (the radices of each axis) are further entries of that shape: they compute the same
// transforms with their roundings in another order, and the launch tuner picks among them as among block sizes.
#define SPO_ZR_REG_SHAPES(X)                                                                                 \
  X(96, 48, 3, 16, 12, 8, 192) X(96, 48, 3, 16, 8, 12, 192) X(96, 48, 4, 12, 8, 12, 192)                      \
  X(96, 48, 3, 16, 8, 12, 384) X(96, 48, 3, 16, 8, 12, 512) X(96, 48, 3, 16, 12, 8, 384)                       \
  X(96, 48, 12, 4, 8, 12, 384)                                                                                 \
  X(128, 32, 4, 8, 16, 8, 256) X(128, 32, 4, 8, 16, 8, 512) X(64, 16, 4, 4, 8, 8, 128) X(80, 16, 4, 4, 8, 10, 128) \
  X(176, 56, 7, 8, 11, 16, 256) X(176, 56, 7, 8, 11, 16, 512) X(96, 32, 4, 8, 8, 12, 256)                      \
  X(96, 32, 4, 8, 8, 12, 512)
// Interaction with the theta contractions on MFMA (phi length, wavefronts, phi radices R1 x R2 of the fused phi
// step); the staged-FFT variant is instantiated with 8 wavefronts.
#define SPO_MFMA_SHAPES(X)                                                                                   \
  X(48, 8, 12, 4) X(48, 4, 12, 4) X(48, 16, 12, 4) X(22, 8, 11, 2) X(12, 8, 4, 3) X(10, 8, 5, 2)             \
  X(8, 8, 4, 2) X(56, 8, 8, 7) X(56, 4, 8, 7)
// T_xy tiles with whole-line DFTs in registers (NX = NY a single radix; C izr per tile, C * NX threads).
#define SPO_XY_REG_SHAPES(X) X(16, 16, 32) X(16, 16, 16) X(16, 16, 8) X(8, 8, 32) X(8, 8, 16) X(4, 4, 64) X(4, 4, 32)
// The buffer-addressed interaction kernel (phi length, wavefronts, phi radices): the MFMA shapes, the phi steps of
// the reduced-JMax batch members (28 and 40 points) and two more factorisations of the 48-point step (16 x 3,
// 8 x 6) for the launch tuner.
#define SPO_BUF_SHAPES(X) SPO_MFMA_SHAPES(X) X(28, 8, 7, 4) X(40, 8, 10, 4) X(48, 8, 16, 3) X(48, 8, 8, 6)

// A launch variant of a kernel family: the kernel, its block size, its tile parameter (T_xy: izr per tile;
// interaction: wavefronts) and its LDS bytes.
struct Variant {
  const void* fn;
  int threads;
  int param;
  std::size_t lds;
};

template <typename T>
std::vector<Variant> zr_reg_variants(int nz, int nr) {
  std::vector<Variant> out;
#define X(A, B, A1, A2, B1, B2, TH)                                                                        \
  if (nz == A && nr == B)                                                                                  \
    out.push_back({reinterpret_cast<const void*>(spo_dev::zr_reg_kernel<T, A, B, A1, A2, B1, B2, TH>), TH, \
                   TH, spo_dev::zr_reg_lds<T>(A, B)});
  SPO_ZR_REG_SHAPES(X)
#undef X
  // Default (first) variant, as measured on MI355X at NZ 96 x NR 48: FP64 planes run best with the largest thread
  // count (512 threads: 3.42 ms per time step; 384: 3.44; 192: 4.2-4.5), complex64 planes with the smallest one, and
  // among those with the Z radices 12 x 8 (1.88 ms; 8 x 12: 1.92; r radices 4 x 12: 1.91; 6 x 8, 12 x 4 and the Z
  // radices 16 x 6 or 6 x 16: 1.98-2.15). Ties keep the order of the shape list.
  std::stable_sort(out.begin(), out.end(), [](const Variant& a, const Variant& b) {
    return sizeof(T) == 8 ? a.threads > b.threads : a.threads < b.threads;
  });
  return out;
}

template <typename T>
std::vector<Variant> xy_reg_variants(int nx, int ny) {
  std::vector<Variant> out;
  // FP64 tiles of 16 izr at 16 x 16 go through LDS in halves (xy_reg_kernel SPLIT): 35 KB instead of 70 KB.
#define X(A, B, CC)                                                                                          \
  if (nx == A && ny == B) {                                                                                  \
    constexpr bool split = sizeof(T) == 8 && A == 16 && CC == 16;                                            \
    out.push_back({reinterpret_cast<const void*>(spo_dev::xy_reg_kernel<T, A, B, CC, split>), CC * A, CC,    \
                   spo_dev::xy_reg_lds<T>(A, B, CC, split)});                                               \
  }
  SPO_XY_REG_SHAPES(X)
#undef X
  // Default (first) variant, as measured on MI355X: the widest tile in both precisions. FP64 at 16 x 16: tiles of
  // 32 izr 3.04 ms per time step, of 16 (through LDS in halves) 3.12, of 8 3.43; at 8 x 8: 32 izr 0.112, 16 izr
  // 0.115. Complex64 at 16 x 16: 32 izr 1.76, 16 izr 1.90, 8 izr 2.58.
  std::stable_sort(out.begin(), out.end(), [](const Variant& a, const Variant& b) { return a.param > b.param; });
  return out;
}

// fused: the phi step in three register passes (phi_step_fused) rather than the staged LDS FFTs (8 wavefronts).
template <typename T>
std::vector<Variant> interaction_mfma_variants(int nphi, int n_theta, bool fused) {
  constexpr int NT = 1;  // 8 columns: one 128-byte line of an irot row in FP64 (two complex64 tiles ran slower)
  std::vector<Variant> out;
  const std::size_t lds = spo_dev::interaction_mfma_lds<T>(8 * NT, n_theta, nphi, fused);
#define X(A, W, R1, R2)                                                                                      \
  if (nphi == A && (fused || W == 8))                                                                        \
    out.push_back({fused ? reinterpret_cast<const void*>(spo_dev::interaction_mfma_kernel<T, A, NT, W, R1, R2>) \
                         : reinterpret_cast<const void*>(spo_dev::interaction_mfma_kernel<T, A, NT, W>),     \
                   64 * W, W, lds});
  SPO_MFMA_SHAPES(X)
#undef X
  return out;
}

// The buffer-addressed interaction kernel: every MFMA shape with at most 8 mJ blocks per wavefront; bf16x3: the
// complex64 split-BF16x3 contraction variant (instantiated for the 48-point phi step).
template <typename T>
std::vector<Variant> interaction_buf_variants(int nphi, int n_theta, int n_blocks, bool bf16x3) {
  std::vector<Variant> out;
  const std::size_t lds = spo_dev::interaction_buf_lds<T>(n_theta, nphi);
  if (bf16x3) {
    if constexpr (sizeof(T) == 4) {
      if (nphi == 48 && n_blocks <= 64)
        out.push_back({reinterpret_cast<const void*>(spo_dev::interaction_buf_kernel<T, 48, 8, 12, 4, true>), 512,
                       8, lds});
    }
    return out;
  }
#define X(A, W, R1, R2)                                                                                         \
  if (nphi == A && n_blocks <= 8 * W)                                                                           \
    out.push_back({reinterpret_cast<const void*>(spo_dev::interaction_buf_kernel<T, A, W, R1, R2>), 64 * W, W, lds});
  SPO_BUF_SHAPES(X)
#undef X
  return out;
}

template <typename T>
const void* interaction_ct_symbol(int nphi) {
#define X(A) \
  if (nphi == A) return reinterpret_cast<const void*>(spo_dev::interaction_ct_kernel<T, A, kThreads>);
  SPO_PHI_LENGTHS(X)
#undef X
  return nullptr;
}

// Launch through the symbol with a packed argument list (hipLaunchKernel), so one call covers every instantiation.
void launch(const void* fn, unsigned grid, unsigned threads, std::size_t lds, hipStream_t stream, void** args) {
  HIP_CHECK(hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, stream));
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// Backend state, common part.
struct spo_hip_backend {
  spo_hip_shape shape{};
  spo_dev::Shape s{};
  hipStream_t stream = nullptr;
  // The specular grid's step runs on `side` (default priority, like `stream`), overlapping T_Zr and T_xy of the
  // scattering grid; `stream` waits for spec_ready before anything that touches the specular wave again.
  hipStream_t side = nullptr;
  hipEvent_t main_ready = nullptr, spec_ready = nullptr;
  // Graph mode (spo_hip_graph_enable): one analysis step captured from `stream` (with `side` joined through the
  // events) at its first use, then replayed. The step's launches and arguments do not change between steps.
  bool use_graph = false;
  hipGraphExec_t step_graph = nullptr;
  int batches = 0;  // spo_hip_batch objects this backend is a member of (it may not be destroyed before them)
  int max_lds = 65536;
  spo_hip_analysis analysis{};

  // FP64 in both modes: specular wave and its factors, the analysis.
  Buffer spec, spec_full, spec_half, spec_gas, z_spec_opt, r_opt_d, s2sc_d, sc2s_d;
  Buffer densities, partial, plane, staging, bk, sums;
  Buffer rot_index, dif_slot, radial_weight, initial_radial, energies, shift_f_d;
  std::unique_ptr<DevicePlan<double>> plan_r_d, plan_zsp, plan_x_d, plan_y_d;
  std::size_t staging_values = 0;

  // Kernel timing (spo_hip_profile_enable): one event pair per launch on `stream`.
  struct Timed {
    int kernel;
    hipEvent_t start, stop;
  };
  bool profiling = false;
  std::vector<std::string> kernel_names;
  std::vector<Timed> timed_launches;
  std::vector<hipEvent_t> event_pool;

  hipEvent_t take_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
  }
  // Runs `launch_fn` (one kernel launch on `stream`), bracketed by events when profiling.
  // `on`: the stream launch_fn launches on (default `stream`).
  template <typename F>
  void timed(const char* name, F&& launch_fn, hipStream_t on = nullptr) {
    if (!profiling) {
      launch_fn();
      HIP_CHECK(hipGetLastError());
      return;
    }
    const hipStream_t st = on ? on : stream;
    int id = 0;
    while (id < static_cast<int>(kernel_names.size()) && kernel_names[id] != name) ++id;
    if (id == static_cast<int>(kernel_names.size())) kernel_names.emplace_back(name);
    Timed t{id, take_event(), take_event()};
    HIP_CHECK(hipEventRecord(t.start, st));
    launch_fn();
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(t.stop, st));
    timed_launches.push_back(t);
  }

  virtual ~spo_hip_backend() {
    for (const Timed& t : timed_launches) {
      (void)hipEventDestroy(t.start);
      (void)hipEventDestroy(t.stop);
    }
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
    if (step_graph) (void)hipGraphExecDestroy(step_graph);
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipStreamDestroy(side);
    }
    if (main_ready) (void)hipEventDestroy(main_ready);
    if (spec_ready) (void)hipEventDestroy(spec_ready);
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }

  // One block of the schedule: the specular grid's kinetic step (half or full), the scattering grid's T_Zr and
  // T_xy (scattering: 0 none, 1 half, 2 full) and, unless kinetic_only, V.
  virtual void step_block(bool specular_half, int scattering, bool kinetic_only) = 0;
  void kinetic_and_potential(bool half, bool kinetic_only) { step_block(half, half ? 1 : 2, kinetic_only); }
  // Merged half steps (SPO_HIP_MERGE_HALF_STEPS=1, eager single-run steps): an analysis step leaves the scattering
  // wave's trailing T(dt/2) pending, and the next step's leading T(dt/2) applies both as one T(dt); the analysis
  // reads its plane from a T(dt/2) of the plane alone, and every other read of the wave applies the pending half
  // step first (flush_pending). Same results up to rounding; off by default, since a run split by a restart file
  // then no longer repeats the unsplit run bit for bit.
  bool merge_halves = false;
  bool pending_half = false;
  virtual void flush_pending() = 0;
  virtual void upload_wave(const double* scattering) = 0;
  virtual void download_wave(double* scattering) = 0;
  virtual void z_plane(int iz) = 0;
  virtual void block_densities() = 0;
  virtual void site_column(int ixy, double* out) = 0;
  virtual void bk_scattering() = 0;
  // One analysis step of the energy-resolved sums of the scattering wave (er_phase holds exp(i E t)).
  virtual void energy_resolved_scattering() = 0;

  // Energy-resolved sums: device layout [site][energy][irot][izr] and [energy][iz][ir].
  Buffer er_sums, er_specular, er_sites, er_phase;
  int er_n_sites = 0, er_n_energies = 0;
  double er_dt = 0.0;

  void specular_step(bool half, bool kinetic_only);
  void lines_d(cx<double>* data, spo_dev::Lines L, const DevicePlan<double>& p, int sign, int mode,
               const cx<double>* factor, int f_line, int f_elem);
  void analyse(int surface_fixed, int iz_specular, double t, double* coefficients);
};

void spo_hip_backend::lines_d(cx<double>* data, spo_dev::Lines L, const DevicePlan<double>& p, int sign, int mode,
                              const cx<double>* factor, int f_line, int f_elem) {
  L.lines_per_block = std::max(1, std::min(64, 65536 / (L.n * static_cast<int>(sizeof(cx<double>)))));
  const std::size_t lds = std::size_t(L.lines_per_block) * L.n * sizeof(cx<double>);
  const unsigned grid = blocks_for(L.count, L.lines_per_block);
  timed("specular_lines_fft", [&] {
    hipLaunchKernelGGL(spo_dev::lines_fft_kernel<double>, dim3(grid), dim3(kThreads), lds, side, data, L, p.plan,
                       sign, mode, factor, static_cast<const cx<double>*>(nullptr), f_line, f_elem);
  }, side);
}

// Specular grid: r lines, Z lines with the kinetic factor (FFT scales folded in), r lines back, then the gas phase
// beyond zasymp and the absorbers (src/propagation/specular_operator.cpp:40-71). FP64.
void spo_hip_backend::specular_step(bool half, bool kinetic_only) {
  // On `side`, after everything already on `stream` (the last projection wrote the specular wave).
  HIP_CHECK(hipEventRecord(main_ready, stream));
  HIP_CHECK(hipStreamWaitEvent(side, main_ready, 0));
  cx<double>* sp = spec.as<cx<double>>();
  spo_dev::Lines r_lines{s.nr, s.nzsp, s.nzsp, s.nr, 0, 1, 1};
  spo_dev::Lines z_lines{s.nzsp, s.nr, s.nr, 1, 0, s.nr, 1};
  lines_d(sp, r_lines, *plan_r_d, -1, 0, nullptr, 0, 0);
  lines_d(sp, z_lines, *plan_zsp, -1, 1, (half ? spec_half : spec_full).as<cx<double>>(), 1, s.nr);
  lines_d(sp, r_lines, *plan_r_d, +1, 0, nullptr, 0, 0);
  if (!kinetic_only) {
    timed("specular_potential", [&] {
      hipLaunchKernelGGL(spo_dev::specular_potential_kernel<double>, dim3(blocks_for(long(s.nr) * s.nzsp, kThreads)),
                         dim3(kThreads), 0, side, sp, s, spec_gas.as<cx<double>>(), z_spec_opt.as<double>(),
                         r_opt_d.as<double>());
    }, side);
  }
  HIP_CHECK(hipEventRecord(spec_ready, side));
}

void spo_hip_backend::analyse(int surface_fixed, int iz_specular, double t, double* coefficients) {
  const spo_hip_analysis& a = analysis;
  const std::size_t n = std::size_t(a.n_states) * a.n_v;
  timed("bk_zero", [&] {
    hipLaunchKernelGGL(spo_dev::zero_kernel, dim3(blocks_for(long(n), kThreads)), dim3(kThreads), 0, stream,
                       bk.as<cx<double>>(), long(n));
  });
  bk_scattering();
  if (surface_fixed) {
    if (iz_specular < 0 || iz_specular >= s.nzsp) throw Status(SPO_HIP_E_ARGUMENT, "specular line out of range");
    timed("bk_specular", [&] {
      hipLaunchKernelGGL(spo_dev::bk_specular_kernel, dim3(1), dim3(64), 0, stream, spec.as<cx<double>>(), s.nr,
                         iz_specular, a.n_v, a.n_states, a.initial_state, initial_radial.as<double>(), a.dr,
                         bk.as<cx<double>>());
    });
  }
  const long total = long(n) * a.n_energies;
  timed("s_accumulate", [&] {
    hipLaunchKernelGGL(spo_dev::s_accumulate_kernel, dim3(blocks_for(total, kThreads)), dim3(kThreads), 0, stream,
                       sums.as<cx<double>>(), bk.as<cx<double>>(), static_cast<int>(n), a.n_energies,
                       energies.as<double>(), t, a.analysis_time_step);
  });
  if (coefficients)
    HIP_CHECK(hipMemcpyAsync(coefficients, bk.as<void>(), n * sizeof(cx<double>), hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------------------------------------------
// The part that depends on the storage type of the scattering wave.
namespace {

template <typename T>
struct State final : spo_hip_backend {
  Buffer psi, zr_full, zr_half, xy_full, xy_half, rot_half, rot_half_t, before, gas_step, z_opt;
  // complex64: the FP32 remainders of the factor tables (factor_lo); empty in FP64
  Buffer zr_full_lo, zr_half_lo, xy_full_lo, xy_half_lo, before_lo, gas_step_lo, shift_f_lo, shift_b_lo, rot_lo,
      rot_lo_t;
  Buffer shift_f, shift_b, s2sc, legendre, block_first, block_size, block_offset, rot_block, rep_of, op_of, phi_index;
  Buffer exp_table, legendre_t;
  Buffer frag_fwd, frag_bwd, fwd_offset, bwd_offset, site_order, rot_pm, lo_fwd, lo_bwd, hi_fwd, hi_bwd;
  int max_jt = 1;
  int frag_fwd_count = 0, frag_bwd_count = 0;
  bool v_buf = false;  // v_variants are interaction_buf_kernel instantiations
  int legendre_size = 0;
  std::size_t zr_lds = 0, xy_lds = 0, v_lds = 0;
  std::unique_ptr<DevicePlan<T>> plan_z, plan_r, plan_x, plan_y, plan_phi;
  int xy_c = 1, v_c = 1;
  // Launch variants of the register-pass and MFMA kernels (identical arithmetic) and the one in use.
  std::vector<Variant> zr_variants, xy_variants, v_variants;
  int zr_pick = 0, xy_pick = 0, v_pick = 0;
  const void* v_fn = nullptr;
  bool zr_fused = false;

  Buffer ana_row;  // the analysis row of the scattering wave at the analysis time (merged half steps)

  void run_zr(const Variant& v, bool half, cx<T>* row_out = nullptr, int row = -1);
  void run_xy(const Variant& v, bool half, bool kinetic_only, cx<T>* data = nullptr,
              const spo_dev::Shape* on = nullptr);
  void run_interaction_mfma(const Variant& v);
  void tune();
  void scattering_kinetic(bool half, bool kinetic_only);

  void step_block(bool specular_half, int scattering, bool kinetic_only) override;
  void flush_pending() override;
  void upload_wave(const double* scattering) override;
  void download_wave(double* scattering) override;
  void z_plane(int iz) override;
  void block_densities() override;
  void site_column(int ixy, double* out) override;
  void bk_scattering() override;
  void energy_resolved_scattering() override;
  void lines(cx<T>* data, spo_dev::Lines L, const DevicePlan<T>& p, int sign, int mode, const cx<T>* factor,
             const cx<T>* factor_lo, int f_line, int f_elem);
};

template <typename T>
void State<T>::lines(cx<T>* data, spo_dev::Lines L, const DevicePlan<T>& p, int sign, int mode, const cx<T>* factor,
                     const cx<T>* factor_lo, int f_line, int f_elem) {
  L.lines_per_block = std::max(1, std::min(64, 65536 / (L.n * static_cast<int>(sizeof(cx<T>)))));
  const std::size_t lds = std::size_t(L.lines_per_block) * L.n * sizeof(cx<T>);
  const unsigned grid = blocks_for(L.count, L.lines_per_block);
  timed("zr_lines_fft", [&] {
    hipLaunchKernelGGL(spo_dev::lines_fft_kernel<T>, dim3(grid), dim3(kThreads), lds, stream, data, L, p.plan, sign,
                       mode, factor, factor_lo, f_line, f_elem);
  });
}

template <typename T>
void State<T>::run_zr(const Variant& v, bool half, cx<T>* row_out, int row) {
  cx<T>* ps = psi.as<cx<T>>();
  const cx<T>* tz = plan_z->tw();
  const cx<T>* tr = plan_r->tw();
  const cx<T>* zr = (half ? zr_half : zr_full).as<cx<T>>();
  const T* ez = plan_z->tw_eps();
  const T* er = plan_r->tw_eps();
  const cx<T>* fl = (half ? zr_half_lo : zr_full_lo).as<cx<T>>();
  void* args[] = {&ps, const_cast<cx<T>**>(&tz), const_cast<cx<T>**>(&tr), const_cast<cx<T>**>(&zr),
                  const_cast<T**>(&ez), const_cast<T**>(&er), const_cast<cx<T>**>(&fl), &row_out, &row};
  launch(v.fn, s.n_rot * s.n_xy, v.threads, v.lds, stream, args);
}

template <typename T>
void State<T>::run_xy(const Variant& v, bool half, bool kinetic_only, cx<T>* data, const spo_dev::Shape* on) {
  cx<T>* ps = data ? data : psi.as<cx<T>>();
  spo_dev::Shape sh = on ? *on : s;
  const int tiles = (sh.n_zr + v.param - 1) / v.param;
  const cx<T>* factor = (half ? xy_half : xy_full).as<cx<T>>();
  const cx<T>* sf = shift_f.as<cx<T>>();
  const cx<T>* sb = shift_b.as<cx<T>>();
  int needs_shift = shape.needs_shift;
  int apply_potential = kinetic_only ? 0 : 1;
  const T* zo = z_opt.as<T>();
  const cx<T>* bf = before.as<cx<T>>();
  const cx<T>* gs = gas_step.as<cx<T>>();
  spo_dev::XyLo<T> lo{(half ? xy_half_lo : xy_full_lo).as<cx<T>>(), shift_f_lo.as<cx<T>>(),
                      shift_b_lo.as<cx<T>>(), before_lo.as<cx<T>>(), gas_step_lo.as<cx<T>>()};
  void* args[] = {&ps, &sh, const_cast<cx<T>**>(&factor), const_cast<cx<T>**>(&sf), const_cast<cx<T>**>(&sb),
                  &needs_shift, &apply_potential, const_cast<T**>(&zo), const_cast<cx<T>**>(&bf),
                  const_cast<cx<T>**>(&gs), &lo};
  launch(v.fn, tiles * sh.n_rot, v.threads, v.lds, stream, args);
}

template <typename T>
void State<T>::run_interaction_mfma(const Variant& v) {
  cx<T>* ps = psi.as<cx<T>>();
  spo_dev::AngularTables at{legendre.as<double>(),  block_first.as<int>(), block_size.as<int>(),
                            block_offset.as<int>(), rot_block.as<int>(),   rep_of.as<int>(),
                            op_of.as<int>(),        phi_index.as<int>()};
  const int tiles = (s.n_zr_pot + 7) / 8;
  const int pair = s.n_xy % 8 == 0 ? 1 : 0;
  const cx<double>* sp = spec.as<cx<double>>();
  const cx<T>* et = exp_table.as<cx<T>>();
  const cx<T>* rh = rot_half_t.as<cx<T>>();
  const cx<T>* rl = rot_lo_t.as<cx<T>>();
  const cx<T>* ssc = s2sc.as<cx<T>>();
  const cx<T>* tp = plan_phi->tw();
  const T* ep = plan_phi->tw_eps();
  spo_dev::MfmaTables mt{fwd_offset.as<int>(), bwd_offset.as<int>(), site_order.as<int>(), pair};
  const T* ff = frag_fwd.as<T>();
  const T* fb = frag_bwd.as<T>();
  int mj = max_jt;
  const int grid = pair ? 2 * ((tiles + 1) / 2) * s.n_xy : tiles * s.n_xy;
  if (v_buf) {
    const cx<T>* rp = rot_pm.as<cx<T>>();
    const unsigned* lf = lo_fwd.as<unsigned>();
    const unsigned* lb = lo_bwd.as<unsigned>();
    const unsigned* hf = hi_fwd.as<unsigned>();
    const unsigned* hb = hi_bwd.as<unsigned>();
    int nf = frag_fwd_count, nbw = frag_bwd_count;
    void* args[] = {&ps, &s, const_cast<cx<T>**>(&tp), &at, &mt, const_cast<T**>(&ff), const_cast<T**>(&fb),
                    &nf, &nbw, const_cast<cx<T>**>(&et), const_cast<cx<T>**>(&rp), const_cast<unsigned**>(&lf),
                    const_cast<unsigned**>(&lb), const_cast<T**>(&ep), const_cast<unsigned**>(&hf),
                    const_cast<unsigned**>(&hb)};
    launch(v.fn, grid, v.threads, v.lds, stream, args);
    return;
  }
  void* args[] = {&ps, const_cast<cx<double>**>(&sp), &s, const_cast<cx<T>**>(&tp), &at, &mt,
                  const_cast<T**>(&ff), const_cast<T**>(&fb), &mj, const_cast<cx<T>**>(&et),
                  const_cast<cx<T>**>(&rh), const_cast<cx<T>**>(&ssc), const_cast<T**>(&ep),
                  const_cast<cx<T>**>(&rl)};
  launch(v.fn, grid, v.threads, v.lds, stream, args);
}

// The launch tuner. The variants of one family differ only in block sizes and tiles; the compiler may contract
// their multiply-adds differently, so their results can differ in the last bits, and the default is therefore
// fixed: the first variant of each family (ordered by what measured fastest on MI355X for the shape), which keeps
// repeated runs bitwise identical. SPO_HIP_TUNE=auto makes the first backend of a process with a given shape and
// precision time every variant on the (still zero) scattering wave, one warm-up and two timed launches each, and
// keep the fastest for the process; SPO_HIP_TUNE="zr=384,xy=8,interaction=16" forces variants by thread count
// (zr), izr per tile (xy) or wavefronts (interaction), and "interaction=#2" the variant at that place in the
// family's list (variants of one size that differ in their FFT radices).
std::map<std::string, int>& tune_cache() {
  static std::map<std::string, int> cache;
  return cache;
}

// The variant SPO_HIP_TUNE selects for a family, or -1 to time them (SPO_HIP_TUNE=auto).
int forced_variant(const std::vector<Variant>& v, const char* family, bool by_threads) {
  const char* env = std::getenv("SPO_HIP_TUNE");
  if (!env) return 0;
  const std::string spec = env;
  if (spec == "auto") return -1;
  const std::string key = std::string(family) + "=";
  const std::size_t at = spec.find(key);
  if (at == std::string::npos) return 0;
  const char* value = spec.c_str() + at + key.size();
  if (*value == '#') {  // "family=#k": the k-th variant of the family's list (0 = the default)
    const int k = std::atoi(value + 1);
    if (k < 0 || k >= static_cast<int>(v.size()))
      throw Status(SPO_HIP_E_ARGUMENT, std::string("SPO_HIP_TUNE: ") + family + " has " + std::to_string(v.size()) +
                                           " variants, no #" + std::to_string(k));
    return k;
  }
  const int want = std::atoi(value);
  for (std::size_t i = 0; i < v.size(); ++i)
    if ((by_threads ? v[i].threads : v[i].param) == want) return static_cast<int>(i);
  throw Status(SPO_HIP_E_ARGUMENT, std::string("SPO_HIP_TUNE: no ") + family + " variant " + std::to_string(want));
}

template <typename T>
void State<T>::tune() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  auto pick = [&](const std::vector<Variant>& v, const char* family, bool by_threads, auto&& run) -> int {
    if (v.size() < 2) return 0;
    const int forced = forced_variant(v, family, by_threads);
    if (forced >= 0) return forced;
    const std::string key = std::string(family) + ":" + std::to_string(s.nz) + "x" + std::to_string(s.nr) + "x" +
                            std::to_string(s.nx) + "x" + std::to_string(s.ny) + "x" + std::to_string(s.n_rot) +
                            "x" + std::to_string(s.n_phi) + ":" + std::to_string(sizeof(T)) + ":" +
                            std::to_string(dev);
    const auto hit = tune_cache().find(key);
    if (hit != tune_cache().end()) return hit->second;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    int best = 0;
    float best_ms = 0.0f;
    for (std::size_t i = 0; i < v.size(); ++i) {
      run(v[i]);
      HIP_CHECK(hipEventRecord(e0, stream));
      run(v[i]);
      run(v[i]);
      HIP_CHECK(hipEventRecord(e1, stream));
      HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (i == 0 || ms < best_ms) {
        best = static_cast<int>(i);
        best_ms = ms;
      }
    }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    tune_cache()[key] = best;
    if (std::getenv("SPO_HIP_TUNE_LOG"))
      std::fprintf(stderr, "spo_hip tune %s: %s=%d (#%d of %zu, %.3f ms for two launches)\n", key.c_str(), family,
                   by_threads ? v[best].threads : v[best].param, best, v.size(), best_ms);
    return best;
  };
  // The scattering wave is zero here, and every kernel maps zero to zero.
  zr_pick = pick(zr_variants, "zr", true, [&](const Variant& v) { run_zr(v, false); });
  xy_pick = pick(xy_variants, "xy", false, [&](const Variant& v) { run_xy(v, false, false); });
  v_pick = pick(v_variants, "interaction", false, [&](const Variant& v) { run_interaction_mfma(v); });
  HIP_CHECK(hipStreamSynchronize(stream));
}

template <typename T>
void State<T>::scattering_kinetic(bool half, bool kinetic_only) {
  cx<T>* ps = psi.as<cx<T>>();
  // T_Zr.
  const cx<T>* zr = (half ? zr_half : zr_full).as<cx<T>>();
  const cx<T>* zr_lo = (half ? zr_half_lo : zr_full_lo).as<cx<T>>();
  const int n_planes = s.n_rot * s.n_xy;
  if (!zr_variants.empty()) {
    timed("zr_plane", [&] { run_zr(zr_variants[zr_pick], half); });
  } else if (zr_fused) {
    timed("zr_plane_generic", [&] {
      hipLaunchKernelGGL(spo_dev::zr_plane_kernel<T>, dim3(n_planes), dim3(kThreads), std::size_t(s.n_zr) * sizeof(cx<T>),
                         stream, ps, s, plan_z->plan, plan_r->plan, zr, zr_lo);
    });
  } else {
    spo_dev::Lines r_lines{s.nr, s.nz * n_planes, s.nz, s.nr, s.n_zr, 1, 1};
    spo_dev::Lines z_lines{s.nz, s.nr * n_planes, s.nr, 1, s.n_zr, s.nr, 1};
    lines(ps, r_lines, *plan_r, -1, 0, nullptr, nullptr, 0, 0);
    lines(ps, z_lines, *plan_z, -1, 1, zr, zr_lo, 1, s.nr);
    lines(ps, r_lines, *plan_r, +1, 0, nullptr, nullptr, 0, 0);
  }
  // T_xy and the diagonal part of V.
  if (!xy_variants.empty()) {
    timed("xy_tile", [&] { run_xy(xy_variants[xy_pick], half, kinetic_only); });
  } else {
    const int tiles = (s.n_zr + xy_c - 1) / xy_c;
    const cx<T>* factor = (half ? xy_half : xy_full).as<cx<T>>();
    const T* zo = z_opt.as<T>();
    const spo_dev::XyLo<T> lo{(half ? xy_half_lo : xy_full_lo).as<cx<T>>(), shift_f_lo.as<cx<T>>(),
                              shift_b_lo.as<cx<T>>(), before_lo.as<cx<T>>(), gas_step_lo.as<cx<T>>()};
    timed("xy_tile_generic", [&] {
      hipLaunchKernelGGL(spo_dev::xy_kernel<T>, dim3(tiles * s.n_rot), dim3(kThreads), xy_lds, stream, ps, s, xy_c,
                         plan_x->plan, plan_y->plan, factor, shift_f.as<cx<T>>(), shift_b.as<cx<T>>(),
                         shape.needs_shift, kinetic_only ? 0 : 1, zo, before.as<cx<T>>(), gas_step.as<cx<T>>(), lo);
    });
  }
}

template <typename T>
void State<T>::flush_pending() {
  if (!pending_half) return;
  pending_half = false;
  scattering_kinetic(true, true);
}

template <typename T>
void State<T>::step_block(bool specular_half, int scattering, bool kinetic_only) {
  if (scattering == 0 && !kinetic_only) throw Status(SPO_HIP_E_ARGUMENT, "a potential step without its kinetic step");
  specular_step(specular_half, kinetic_only);
  if (scattering != 0) scattering_kinetic(scattering == 1, kinetic_only);
  cx<T>* ps = psi.as<cx<T>>();
  // The specular step (on `side`) is complete before the specular wave is read or written on `stream` again.
  HIP_CHECK(hipStreamWaitEvent(stream, spec_ready, 0));
  if (kinetic_only || s.n_zr_pot == 0) return;
  // V in the interaction region, then the projection onto the specular grid.
  spo_dev::AngularTables at{legendre.as<double>(),  block_first.as<int>(), block_size.as<int>(),
                            block_offset.as<int>(), rot_block.as<int>(),   rep_of.as<int>(),
                            op_of.as<int>(),        phi_index.as<int>()};
  const int tiles = (s.n_zr_pot + v_c - 1) / v_c;
  const std::size_t lds = v_lds;
  const cx<double>* sp = spec.as<cx<double>>();
  const cx<T>* et = exp_table.as<cx<T>>();
  const cx<T>* rh = rot_half.as<cx<T>>();
  const cx<T>* ssc = s2sc.as<cx<T>>();
  if (!v_variants.empty()) {
    timed("specular_coupling", [&] {
      hipLaunchKernelGGL(spo_dev::specular_coupling_kernel<T>, dim3(blocks_for(long(s.n_zr_pot) * s.n_xy, 256)),
                         dim3(256), 0, stream, ps, sp, s, rh, ssc);
    });
    timed("interaction", [&] { run_interaction_mfma(v_variants[v_pick]); });
  } else if (v_fn) {
    const cx<T>* tp = plan_phi->tw();
    const T* lt = legendre_t.as<T>();
    int c = v_c;
    int ls = legendre_size;
    const cx<T>* rl = rot_lo.as<cx<T>>();
    const T* ep = plan_phi->tw_eps();
    void* args[] = {&ps, const_cast<cx<double>**>(&sp), &s, &c, const_cast<cx<T>**>(&tp), &at,
                    const_cast<cx<T>**>(&et), const_cast<cx<T>**>(&rh), const_cast<cx<T>**>(&ssc),
                    const_cast<T**>(&lt), &ls, const_cast<cx<T>**>(&rl), const_cast<T**>(&ep)};
    timed("interaction", [&] { launch(v_fn, tiles * s.n_xy, kThreads, lds, stream, args); });
  } else {
    timed("interaction_generic", [&] {
      hipLaunchKernelGGL(spo_dev::interaction_kernel<T>, dim3(tiles * s.n_xy), dim3(kThreads), lds, stream, ps, sp, s,
                         v_c, plan_phi->plan, at, et, rh, ssc, rot_lo.as<cx<T>>());
    });
  }
  timed("projection", [&] {
    hipLaunchKernelGGL(spo_dev::projection_kernel<T>, dim3(blocks_for(s.n_zr_pot, 64)),
                       dim3(64 * spo_dev::kProjectionWaves), 0,
                       stream, ps, spec.as<cx<double>>(), s, sc2s_d.as<cx<double>>(), s2sc_d.as<cx<double>>());
  });
}

template <typename T>
void State<T>::upload_wave(const double* scattering) {
  pending_half = false;  // a new wave at the analysis time
  const std::size_t per_izr = std::size_t(s.n_rot) * s.n_xy;
  const int chunk = static_cast<int>(std::max<std::size_t>(1, staging_values / per_izr));
  for (int izr0 = 0; izr0 < s.n_zr; izr0 += chunk) {
    const int count = std::min(chunk, s.n_zr - izr0);
    HIP_CHECK(hipMemcpyAsync(staging.as<void>(), scattering + 2 * per_izr * izr0, per_izr * count * sizeof(cx<double>),
                             hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(spo_dev::host_to_device_kernel<T>, dim3(blocks_for(long(per_izr) * count, kThreads)),
                       dim3(kThreads), 0, stream, psi.as<cx<T>>(), s, staging.as<cx<double>>(), izr0, count);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(stream));
  }
}

template <typename T>
void State<T>::download_wave(double* scattering) {
  flush_pending();
  const std::size_t per_izr = std::size_t(s.n_rot) * s.n_xy;
  const int chunk = static_cast<int>(std::max<std::size_t>(1, staging_values / per_izr));
  for (int izr0 = 0; izr0 < s.n_zr; izr0 += chunk) {
    const int count = std::min(chunk, s.n_zr - izr0);
    hipLaunchKernelGGL(spo_dev::device_to_host_kernel<T>, dim3(blocks_for(long(per_izr) * count, kThreads)),
                       dim3(kThreads), 0, stream, psi.as<cx<T>>(), s, staging.as<cx<double>>(), izr0, count);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(scattering + 2 * per_izr * izr0, staging.as<void>(), per_izr * count * sizeof(cx<double>),
                             hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }
}

template <typename T>
void State<T>::z_plane(int iz) {
  flush_pending();
  const long n = long(s.n_rot) * s.n_xy * s.nr;
  hipLaunchKernelGGL(spo_dev::z_plane_kernel<T>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, stream,
                     psi.as<cx<T>>(), s, iz, plane.as<cx<double>>());
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void State<T>::block_densities() {
  flush_pending();
  // One part per site (n_zr x n_xy partial sums over irot), then the parts added in site order: enough workgroups
  // to cover the device, where a few long per-thread loops over sites left it latency-bound.
  const int chunk = 1;
  const int parts = s.n_xy;
  timed("densities", [&] {
    hipLaunchKernelGGL(spo_dev::densities_partial_kernel<T>, dim3(blocks_for(s.n_zr, kThreads), parts), dim3(kThreads),
                       0, stream, psi.as<cx<T>>(), s, chunk, partial.as<double>());
  });
  timed("densities_sum", [&] {
    hipLaunchKernelGGL(spo_dev::densities_sum_kernel, dim3(blocks_for(s.n_zr, kThreads)), dim3(kThreads), 0, stream,
                       partial.as<double>(), s.n_zr, parts, densities.as<double>());
  });
}

template <typename T>
void State<T>::site_column(int ixy, double* out) {
  flush_pending();
  const std::size_t n = std::size_t(s.n_zr) * s.n_rot;
  if constexpr (sizeof(T) == sizeof(double)) {
    HIP_CHECK(hipMemcpyAsync(out, psi.as<cx<T>>() + n * ixy, n * sizeof(cx<T>), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  } else {
    std::vector<cx<T>> tmp(n);
    HIP_CHECK(hipMemcpyAsync(tmp.data(), psi.as<cx<T>>() + n * ixy, n * sizeof(cx<T>), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    for (std::size_t i = 0; i < n; ++i) {
      out[2 * i] = tmp[i].x;
      out[2 * i + 1] = tmp[i].y;
    }
  }
}

template <typename T>
void State<T>::bk_scattering() {
  const spo_hip_analysis& a = analysis;
  const cx<T>* src = psi.as<cx<T>>();
  spo_dev::Shape sh = s;
  int iz = a.iz_analysis;
  if (pending_half && (zr_variants.empty() || xy_variants.empty())) flush_pending();
  if (pending_half) {
    // The analysis plane at the analysis time without advancing the wave: T_Zr(dt/2) of every (irot, ixy) plane
    // with only row iz_analysis written out, then T_xy(dt/2) of those rows (layout of psi with one Z point).
    if (!ana_row.as<void>()) ana_row = Buffer(std::size_t(s.nr) * s.n_rot * s.n_xy * sizeof(cx<T>));
    cx<T>* row = ana_row.as<cx<T>>();
    timed("zr_row", [&] { run_zr(zr_variants[zr_pick], true, row, a.iz_analysis); });
    sh.nz = 1;
    sh.n_zr = s.nr;
    sh.nzpot = 0;
    sh.n_zr_pot = 0;
    timed("xy_row", [&] { run_xy(xy_variants[xy_pick], true, true, row, &sh); });
    src = row;
    iz = 0;
  }
  timed("bk_scattering", [&] {
    hipLaunchKernelGGL(spo_dev::bk_scattering_kernel<T>, dim3(a.n_v * a.n_rotation), dim3(kThreads),
                       std::size_t(s.n_xy) * sizeof(cx<double>), stream, src, sh, iz,
                       a.n_rotation, a.n_diffraction, a.n_states, rot_index.as<int>(), dif_slot.as<int>(),
                       radial_weight.as<double>(), shift_f_d.as<cx<double>>(), a.xy_scale, plan_x_d->plan,
                       plan_y_d->plan, bk.as<cx<double>>());
  });
}

template <typename T>
void State<T>::energy_resolved_scattering() {
  flush_pending();
  const long block = long(s.n_zr) * s.n_rot;
  timed("energy_resolved", [&] {
    hipLaunchKernelGGL(spo_dev::energy_resolved_kernel<T>, dim3(blocks_for(block, 256), er_n_sites), dim3(256), 0,
                       stream, psi.as<cx<T>>(), block, er_sites.as<int>(), er_n_energies,
                       er_phase.as<cx<double>>(), er_dt, er_sums.as<cx<double>>());
  });
}

spo_dev::Shape make_shape(const spo_hip_shape* shape) {
  spo_dev::Shape s{};
  s.n_rot = shape->n_rot;
  s.nx = shape->nx;
  s.ny = shape->ny;
  s.n_xy = shape->nx * shape->ny;
  s.nr = shape->nr;
  s.nz = shape->nz;
  s.n_zr = shape->nr * shape->nz;
  s.nzsp = shape->nz_specular;
  s.nzpot = shape->nz_potential;
  s.n_zr_pot = shape->nz_potential * shape->nr;
  s.n_theta = shape->n_theta;
  s.n_phi = shape->n_phi;
  s.n_tp = shape->n_theta * shape->n_phi;
  s.n_rep = shape->n_representatives;
  s.ichn = shape->initial_channel;
  s.mj_max = shape->mj_max;
  s.n_blocks = 2 * shape->mj_max + 1;
  s.r_abs_first = shape->r_absorber_first;
  s.r_abs_last = shape->r_absorber_last;
  return s;
}

template <typename T>
std::unique_ptr<spo_hip_backend> create_state(const spo_hip_shape* shape, const spo_hip_tables* t,
                                              const spo_hip_analysis* a, const double* initial_specular) {
  auto b = std::make_unique<State<T>>();
  b->shape = *shape;
  b->s = make_shape(shape);
  const spo_dev::Shape& s = b->s;
  if (s.n_rot < 1 || s.nx < 2 || s.ny < 2 || s.nr < 2 || s.nz < 2 || s.nzsp < s.nz || s.nzpot < 0 || s.nzpot > s.nz ||
      s.n_theta < 1 || s.n_phi < 1 || s.n_rep < 1 || s.ichn < 0 || s.ichn >= s.n_rot || shape->n_operations < 1 ||
      shape->steps_per_analysis < 1 || s.r_abs_first < 0 || s.r_abs_last >= s.nr)
    throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_create: inconsistent shape");

  int device = 0;
  HIP_CHECK(hipGetDevice(&device));
  HIP_CHECK(hipDeviceGetAttribute(&b->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  const std::size_t psi_values = std::size_t(s.n_zr) * s.n_rot * s.n_xy;
  const std::size_t exp_values = std::size_t(s.n_tp) * s.n_rep * s.n_zr_pot;
  std::size_t free_bytes = 0, total_bytes = 0;
  HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
  const std::size_t need = (psi_values + exp_values) * sizeof(cx<T>) + (std::size_t(1) << 28);
  if (need > free_bytes)
    throw Status(SPO_HIP_E_NOT_SUPPORTED, "not enough device memory: need " + std::to_string(need >> 20) +
                                              " MiB, free " + std::to_string(free_bytes >> 20) + " MiB");
  HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreateWithFlags(&b->main_ready, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&b->spec_ready, hipEventDisableTiming));

  const std::size_t cT = sizeof(cx<T>), c16 = sizeof(cx<double>);
  // Zero slack behind the last site: the interaction kernels' phase-1 loads of a partial k step or of a column past
  // the interaction region read values past a site column without a branch (their products vanish: the matching
  // Legendre fragments are zero, or the column is never stored), at most 15 n_zr + 7 values past the end.
  const std::size_t psi_slack = 16 * std::size_t(s.n_zr) + 64;
  b->psi = Buffer((psi_values + psi_slack) * cT);
  HIP_CHECK(hipMemsetAsync(b->psi.template as<void>(), 0, (psi_values + psi_slack) * cT, b->stream));
  b->spec = upload(initial_specular, std::size_t(s.nr) * s.nzsp * c16);
  b->densities = Buffer(std::size_t(s.n_zr) * sizeof(double));
  b->partial = Buffer(std::size_t(s.n_zr) * s.n_xy * sizeof(double));
  b->plane = Buffer(std::size_t(s.n_rot) * s.n_xy * s.nr * c16);
  b->staging_values = std::max<std::size_t>(std::size_t(s.n_rot) * s.n_xy, std::size_t(4) << 20);
  b->staging = Buffer(b->staging_values * c16);

  // Kinetic factors with the forward and backward FFT scales folded in (multiplied in FP64, then stored).
  const double zr_f = t->fft_scale_z * t->fft_scale_r;
  const double zr_b = (1.0 / (s.nz * t->fft_scale_z)) * (1.0 / (s.nr * t->fft_scale_r));
  const double sp_b = (1.0 / (s.nzsp * t->fft_scale_z)) * (1.0 / (s.nr * t->fft_scale_r));
  const double xy_f = t->fft_scale_x * t->fft_scale_y;
  const double xy_b = (1.0 / (s.nx * t->fft_scale_x)) * (1.0 / (s.ny * t->fft_scale_y));
  auto fold = [](const double* p, std::size_t n, double f1, double f2) {
    std::vector<double> v(p, p + 2 * n);
    for (double& x : v) x = x * f1 * f2;
    return v;
  };
  {
    const auto zf = fold(t->zr_full, s.n_zr, zr_f, zr_b), zh = fold(t->zr_half, s.n_zr, zr_f, zr_b);
    b->zr_full = upload_complex<T>(zf.data(), s.n_zr);
    b->zr_full_lo = factor_lo<T>(zf.data(), s.n_zr);
    b->zr_half = upload_complex<T>(zh.data(), s.n_zr);
    b->zr_half_lo = factor_lo<T>(zh.data(), s.n_zr);
    const auto sf = fold(t->specular_full, std::size_t(s.nr) * s.nzsp, zr_f, sp_b);
    const auto sh = fold(t->specular_half, std::size_t(s.nr) * s.nzsp, zr_f, sp_b);
    b->spec_full = upload(sf.data(), sf.size() * sizeof(double));
    b->spec_half = upload(sh.data(), sh.size() * sizeof(double));
    const auto xf = fold(t->xy_full, s.n_xy, xy_f, xy_b), xh = fold(t->xy_half, s.n_xy, xy_f, xy_b);
    b->xy_full = upload_complex<T>(xf.data(), s.n_xy);
    b->xy_full_lo = factor_lo<T>(xf.data(), s.n_xy);
    b->xy_half = upload_complex<T>(xh.data(), s.n_xy);
    b->xy_half_lo = factor_lo<T>(xh.data(), s.n_xy);
  }
  b->rot_half = upload_complex<T>(t->rotation_half, std::size_t(s.n_rot) * s.nr);
  b->rot_lo = factor_lo<T>(t->rotation_half, std::size_t(s.n_rot) * s.nr);
  {
    // (irot, ir) -> (ir, irot), ir fastest, for the MFMA interaction kernel
    std::vector<double> rt(2 * std::size_t(s.n_rot) * s.nr);
    for (int ir = 0; ir < s.nr; ++ir)
      for (int irot = 0; irot < s.n_rot; ++irot)
        for (int c = 0; c < 2; ++c)
          rt[2 * (ir + std::size_t(s.nr) * irot) + c] = t->rotation_half[2 * (irot + std::size_t(s.n_rot) * ir) + c];
    b->rot_half_t = upload_complex<T>(rt.data(), rt.size() / 2);
    b->rot_lo_t = factor_lo<T>(rt.data(), rt.size() / 2);
  }
  b->before = upload_complex<T>(t->before_interaction, std::size_t(s.n_rot) * s.nr);
  b->gas_step = upload_complex<T>(t->gas_phase_step, std::size_t(s.n_rot) * s.nr);
  b->before_lo = factor_lo<T>(t->before_interaction, std::size_t(s.n_rot) * s.nr);
  b->gas_step_lo = factor_lo<T>(t->gas_phase_step, std::size_t(s.n_rot) * s.nr);
  b->spec_gas = upload(t->specular_gas_phase, std::size_t(s.nr) * c16);
  b->z_opt = upload_real<T>(t->z_optical, s.nz);
  b->r_opt_d = upload(t->r_optical, std::size_t(s.nr) * sizeof(double));
  b->z_spec_opt = upload(t->z_specular_optical, std::size_t(s.nzsp) * sizeof(double));
  b->shift_f = upload_complex<T>(t->shift_forward, s.n_xy);
  b->shift_b = upload_complex<T>(t->shift_backward, s.n_xy);
  b->shift_f_lo = factor_lo<T>(t->shift_forward, s.n_xy);
  b->shift_b_lo = factor_lo<T>(t->shift_backward, s.n_xy);
  b->shift_f_d = upload(t->shift_forward, std::size_t(s.n_xy) * c16);
  b->s2sc = upload_complex<T>(t->specular_to_scattering, s.n_xy);
  b->s2sc_d = upload(t->specular_to_scattering, std::size_t(s.n_xy) * c16);
  b->sc2s_d = upload(t->scattering_to_specular, std::size_t(s.n_xy) * c16);

  // Angular transform tables (the Legendre matrix stays FP64 and is rounded at use).
  std::vector<int> offsets(s.n_blocks), rot_block(s.n_rot, -1);
  int total = 0, covered = 0;
  for (int blk = 0; blk < s.n_blocks; ++blk) {
    offsets[blk] = total;
    const int first = t->block_first[blk], size = t->block_size[blk];
    if (first < 0 || size < 0 || first + size > s.n_rot) throw Status(SPO_HIP_E_ARGUMENT, "bad mJ block");
    for (int j = 0; j < size; ++j) rot_block[first + j] = blk;
    total += s.n_theta * size;
    covered += size;
  }
  if (covered != s.n_rot) throw Status(SPO_HIP_E_ARGUMENT, "the mJ blocks do not cover the basis");
  if (2 * s.mj_max + 1 > s.n_phi) throw Status(SPO_HIP_E_ARGUMENT, "nphi < 2 MJMax + 1");
  b->legendre = upload(t->legendre, std::size_t(total) * sizeof(double));
  b->legendre_t = upload_real<T>(t->legendre, std::size_t(total));
  b->legendre_size = total;
  b->block_first = upload(t->block_first, std::size_t(s.n_blocks) * sizeof(int));
  b->block_size = upload(t->block_size, std::size_t(s.n_blocks) * sizeof(int));
  b->block_offset = upload(offsets.data(), offsets.size() * sizeof(int));
  b->rot_block = upload(rot_block.data(), rot_block.size() * sizeof(int));
  for (int ixy = 0; ixy < s.n_xy; ++ixy)
    if (t->representative_of[ixy] < 0 || t->representative_of[ixy] >= s.n_rep || t->operation_of[ixy] < 0 ||
        t->operation_of[ixy] >= shape->n_operations)
      throw Status(SPO_HIP_E_ARGUMENT, "bad symmetry map");
  for (int k = 0; k < s.n_phi * shape->n_operations; ++k)
    if (t->phi_index[k] < 0 || t->phi_index[k] >= s.n_phi) throw Status(SPO_HIP_E_ARGUMENT, "bad phi map");
  b->rep_of = upload(t->representative_of, std::size_t(s.n_xy) * sizeof(int));
  b->op_of = upload(t->operation_of, std::size_t(s.n_xy) * sizeof(int));
  b->phi_index = upload(t->phi_index, std::size_t(s.n_phi) * shape->n_operations * sizeof(int));
  b->exp_table = upload_complex<T>(t->exp_potential, exp_values);

  // Analysis (FP64).
  b->analysis = *a;
  if (a->n_states != a->n_rotation * a->n_diffraction || a->n_v < 1 || a->n_energies < 1 || a->iz_analysis < 0 ||
      a->iz_analysis >= s.nz || a->initial_state < 0 || a->initial_state >= a->n_states || a->n_diffraction > s.n_xy)
    throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_create: inconsistent analysis");
  for (int k = 0; k < a->n_rotation; ++k)
    if (a->rotation_index[k] < 0 || a->rotation_index[k] >= s.n_rot) throw Status(SPO_HIP_E_ARGUMENT, "bad rotation index");
  for (int k = 0; k < a->n_diffraction; ++k)
    if (a->diffraction_slot[k] < 0 || a->diffraction_slot[k] >= s.n_xy) throw Status(SPO_HIP_E_ARGUMENT, "bad slot");
  b->rot_index = upload(a->rotation_index, std::size_t(a->n_rotation) * sizeof(int));
  b->dif_slot = upload(a->diffraction_slot, std::size_t(a->n_diffraction) * sizeof(int));
  b->radial_weight = upload(a->radial_weight, std::size_t(s.nr) * a->n_rotation * a->n_v * sizeof(double));
  b->initial_radial = upload(a->initial_radial, std::size_t(s.nr) * a->n_v * sizeof(double));
  b->energies = upload(a->total_energy_hartree, std::size_t(a->n_energies) * sizeof(double));
  b->analysis.rotation_index = nullptr;
  b->analysis.diffraction_slot = nullptr;
  b->analysis.radial_weight = nullptr;
  b->analysis.initial_radial = nullptr;
  b->analysis.total_energy_hartree = nullptr;
  b->bk = Buffer(std::size_t(a->n_states) * a->n_v * c16);
  const std::size_t n_sums = std::size_t(a->n_states) * a->n_v * a->n_energies;
  b->sums = Buffer(n_sums * c16);
  HIP_CHECK(hipMemsetAsync(b->sums.template as<void>(), 0, n_sums * c16, b->stream));

  // Plans, tile widths and the kernels for this shape.
  b->plan_z = std::make_unique<DevicePlan<T>>(s.nz);
  b->plan_r = std::make_unique<DevicePlan<T>>(s.nr);
  b->plan_x = std::make_unique<DevicePlan<T>>(s.nx);
  b->plan_y = std::make_unique<DevicePlan<T>>(s.ny);
  b->plan_phi = std::make_unique<DevicePlan<T>>(s.n_phi);
  b->plan_r_d = std::make_unique<DevicePlan<double>>(s.nr);
  b->plan_zsp = std::make_unique<DevicePlan<double>>(s.nzsp);
  b->plan_x_d = std::make_unique<DevicePlan<double>>(s.nx);
  b->plan_y_d = std::make_unique<DevicePlan<double>>(s.ny);
  const std::size_t lds_limit = static_cast<std::size_t>(b->max_lds);
  // T_Zr: the register-pass kernel (padded rows) if instantiated and it fits, else the run-time one, else lines.
  if (spo_dev::zr_reg_lds<T>(s.nz, s.nr) <= lds_limit) b->zr_variants = zr_reg_variants<T>(s.nz, s.nr);
  if (!b->zr_variants.empty()) {
    for (const Variant& v : b->zr_variants) set_lds(v.fn, v.lds);
  } else if (std::size_t(s.n_zr) * cT <= lds_limit) {
    b->zr_fused = true;
    b->zr_lds = std::size_t(s.n_zr) * cT;
    set_lds(reinterpret_cast<const void*>(spo_dev::zr_plane_kernel<T>), b->zr_lds);
  }
  // T_xy.
  for (const Variant& v : xy_reg_variants<T>(s.nx, s.ny))
    if (v.lds <= lds_limit) b->xy_variants.push_back(v);
  if (!b->xy_variants.empty()) {
    for (const Variant& v : b->xy_variants) set_lds(v.fn, v.lds);
  } else {
    b->xy_c = std::max(1, std::min(64, 4096 / s.n_xy));
    b->xy_lds = std::size_t(b->xy_c) * s.n_xy * cT;
    if (b->xy_lds > lds_limit) throw Status(SPO_HIP_E_NOT_SUPPORTED, "the (x, y) tile does not fit in LDS");
    set_lds(reinterpret_cast<const void*>(spo_dev::xy_kernel<T>), b->xy_lds);
  }
  // Interaction tile: the largest power of two of columns (at most 32) whose LDS fits in 80 KiB (two per CU).
  b->v_fn = s.n_theta <= spo_dev::kMaxTheta ? interaction_ct_symbol<T>(s.n_phi) : nullptr;
  auto v_bytes = [&](int c) {
    return b->v_fn ? spo_dev::interaction_ct_lds<T>(c, s.n_theta, s.n_phi, b->legendre_size)
                   : std::size_t(c) * (s.n_rot + s.n_tp) * cT;
  };
  int p2 = 1;
  while (p2 < 32 && v_bytes(2 * p2) <= 81920) p2 *= 2;
  b->v_c = p2;
  b->v_lds = v_bytes(p2);
  if (b->v_lds > lds_limit) throw Status(SPO_HIP_E_NOT_SUPPORTED, "an angular column does not fit in LDS");
  set_lds(b->v_fn ? b->v_fn : reinterpret_cast<const void*>(spo_dev::interaction_kernel<T>), b->v_lds);
  // Interaction with the theta contractions on MFMA: Legendre fragments in lane order, XCD-grouped site order.
  // SPO_HIP_INTERACTION selects a variant for comparisons: "buf" (the buffer-addressed MFMA kernel where its shape
  // conditions hold, else "mfma"; the complex64 default), "mfma" (MFMA contractions, fused register phi step; the
  // FP64 default: 4.33 against 4.89 ms per launch at the DZ workload on MI355X), "mfma_staged" (MFMA contractions,
  // staged LDS phi FFTs), "lds" (contractions on the vector units).
  {
    const char* env = std::getenv("SPO_HIP_INTERACTION");
    const std::string choice = env ? env : (sizeof(T) == 8 ? "mfma" : "buf");
    if (choice != "buf" && choice != "mfma" && choice != "mfma_staged" && choice != "lds")
      throw Status(SPO_HIP_E_ARGUMENT, "SPO_HIP_INTERACTION must be buf, mfma, mfma_staged or lds");
    const bool fused = (choice == "mfma" || choice == "buf") && s.n_theta <= 16;
    // SPO_HIP_CONTRACTION=bf16x3: the complex64 theta contractions as split-BF16x3 products (buffer kernel, 48-point
    // phi step); the default "fp32" is FP32 MFMA with the bf16 low part of L. FP64 ignores it.
    const char* contraction = std::getenv("SPO_HIP_CONTRACTION");
    if (contraction && std::string(contraction) != "fp32" && std::string(contraction) != "bf16x3")
      throw Status(SPO_HIP_E_ARGUMENT, "SPO_HIP_CONTRACTION must be fp32 or bf16x3");
    const bool bf16x3 = sizeof(T) == 4 && contraction && std::string(contraction) == "bf16x3";
    std::vector<Variant> variants;
    for (const Variant& v : interaction_mfma_variants<T>(s.n_phi, s.n_theta, fused))
      if (v.lds <= lds_limit) variants.push_back(v);
    // phi lengths with a buffer-kernel instantiation only (the reduced-JMax batch members' 28 and 40 points)
    const bool buf_candidate =
        choice == "buf" && !interaction_buf_variants<T>(s.n_phi, s.n_theta, s.n_blocks, bf16x3).empty();
    if ((!variants.empty() || buf_candidate) && choice != "lds") {
      const int n_tt = (s.n_theta + 15) / 16, k_bwd = (s.n_theta + 3) / 4;
      std::vector<int> fo(s.n_blocks), bo(s.n_blocks);
      int nf = 0, nbw = 0, max_nb = 0;
      for (int blk = 0; blk < s.n_blocks; ++blk) {
        const int nb = t->block_size[blk];
        fo[blk] = nf;
        nf += n_tt * ((nb + 3) / 4);
        bo[blk] = nbw;
        nbw += ((nb + 15) / 16) * k_bwd;
        max_nb = std::max(max_nb, nb);
      }
      std::vector<double> ff(64 * std::size_t(std::max(nf, 1)), 0.0), fb(64 * std::size_t(std::max(nbw, 1)), 0.0);
      for (int blk = 0; blk < s.n_blocks; ++blk) {
        const int nb = t->block_size[blk], ks = (nb + 3) / 4;
        const double* L = t->legendre + offsets[blk];  // (theta, J)
        for (int lane = 0; lane < 64; ++lane) {
          const int i = lane % 16, kk = lane / 16;
          for (int tt = 0; tt < n_tt; ++tt)
            for (int k = 0; k < ks; ++k) {
              const int th = 16 * tt + i, j = 4 * k + kk;
              if (th < s.n_theta && j < nb) ff[64 * std::size_t(fo[blk] + tt * ks + k) + lane] = L[th * nb + j];
            }
          for (int jt = 0; jt < (nb + 15) / 16; ++jt)
            for (int k = 0; k < k_bwd; ++k) {
              const int th = 4 * k + kk, j = 16 * jt + i;
              if (th < s.n_theta && j < nb) fb[64 * std::size_t(bo[blk] + jt * k_bwd + k) + lane] = L[th * nb + j];
            }
        }
      }
      std::vector<int> order(s.n_xy);
      for (int i = 0; i < s.n_xy; ++i) order[i] = i;
      if (s.n_xy % 8 == 0) {
        std::vector<int> sorted(order);
        std::stable_sort(sorted.begin(), sorted.end(), [&](int x, int y) {
          return t->representative_of[x] < t->representative_of[y];
        });
        const int per = s.n_xy / 8;
        for (int x = 0; x < 8; ++x)
          for (int k = 0; k < per; ++k) order[8 * k + x] = sorted[per * x + k];
      }
      b->frag_fwd = upload_real<T>(ff.data(), ff.size());
      b->frag_bwd = upload_real<T>(fb.data(), fb.size());
      b->fwd_offset = upload(fo.data(), fo.size() * sizeof(int));
      b->bwd_offset = upload(bo.data(), bo.size() * sizeof(int));
      b->site_order = upload(order.data(), order.size() * sizeof(int));
      b->max_jt = (max_nb + 15) / 16;
      b->frag_fwd_count = nf;
      b->frag_bwd_count = nbw;
      // The buffer-addressed kernel where its shape conditions hold: one theta tile, blocks of at most 16 J, a
      // site column addressable with 32-bit offsets.
      if (choice == "buf" && s.n_theta <= 16 && max_nb <= 16 &&
          std::size_t(s.n_zr) * s.n_rot * cT < (std::size_t(1) << 31)) {
        std::vector<Variant> buf;
        for (const Variant& v : interaction_buf_variants<T>(s.n_phi, s.n_theta, s.n_blocks, bf16x3))
          if (v.lds <= lds_limit) buf.push_back(v);
        if (bf16x3 && buf.empty())
          throw Status(SPO_HIP_E_NOT_SUPPORTED, "SPO_HIP_CONTRACTION=bf16x3: instantiated for the 48-point phi step only");
        if (!buf.empty()) {
          // Sign-split rotation factors, 16 zero rows past the last irot (the loads of a block's rows past the
          // basis stay inside the table).
          std::vector<spo_dev::RotEntry<T>> pm(2 * (std::size_t(s.n_rot) + 16) * s.nr, spo_dev::RotEntry<T>{});
          for (int irot = 0; irot < s.n_rot; ++irot)
            for (int ir = 0; ir < s.nr; ++ir) {
              const double x = t->rotation_half[2 * (irot + std::size_t(s.n_rot) * ir)];
              const double y = t->rotation_half[2 * (irot + std::size_t(s.n_rot) * ir) + 1];
              const cx<T> w = to_storage<T>(x, y);
              const std::size_t e0 = spo_dev::rot_pm_index(irot, ir, s.nr, 0);
              pm[e0].x = w.x;
              pm[e0].y = -w.y;
              pm[e0 + 1].x = w.x;
              pm[e0 + 1].y = w.y;
              if constexpr (sizeof(T) == 4) {
                const float lo_x = static_cast<float>(x - static_cast<double>(w.x));
                const float lo_y = static_cast<float>(y - static_cast<double>(w.y));
                pm[e0].lo_x = lo_x;
                pm[e0].lo_y = -lo_y;
                pm[e0 + 1].lo_x = lo_x;
                pm[e0 + 1].lo_y = lo_y;
              }
            }
          b->rot_pm = upload(pm.data(), pm.size() * sizeof(spo_dev::RotEntry<T>));
          if (sizeof(T) == 4) {
            // bf16 fragments of a function of L, lane (i, q) element t = contraction row 4 t + q
            // (legendre_low_product): forward A = L (theta = i, J = row), backward A = L^T (J = i, theta = row).
            auto bf16 = [](double x) -> std::uint16_t {
              const float f = static_cast<float>(x);
              std::uint32_t u;
              std::memcpy(&u, &f, 4);
              u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (the values are finite and small)
              return static_cast<std::uint16_t>(u >> 16);
            };
            auto bf16_value = [&](double x) {
              const std::uint32_t u = std::uint32_t(bf16(x)) << 16;
              float f;
              std::memcpy(&f, &u, 4);
              return static_cast<double>(f);
            };
            auto fragments = [&](auto&& piece, Buffer& fwd, Buffer& bwd) {
              std::vector<std::uint16_t> lf(256 * std::size_t(s.n_blocks), 0), lb(256 * std::size_t(s.n_blocks), 0);
              for (int blk = 0; blk < s.n_blocks; ++blk) {
                const int nb = t->block_size[blk];
                const double* L = t->legendre + offsets[blk];  // (theta, J)
                for (int lane = 0; lane < 64; ++lane)
                  for (int e = 0; e < 4; ++e) {
                    const int i = lane % 16, q = lane / 16, row = 4 * e + q;
                    if (i < s.n_theta && row < nb) lf[256 * std::size_t(blk) + 4 * lane + e] = bf16(piece(L[i * nb + row]));
                    if (row < s.n_theta && i < nb) lb[256 * std::size_t(blk) + 4 * lane + e] = bf16(piece(L[row * nb + i]));
                  }
              }
              fwd = upload(lf.data(), lf.size() * sizeof(std::uint16_t));
              bwd = upload(lb.data(), lb.size() * sizeof(std::uint16_t));
            };
            if (bf16x3) {
              // split-BF16x3: A_hi = bf16(L), A_lo = bf16(L - A_hi)
              fragments([&](double l) { return l; }, b->hi_fwd, b->hi_bwd);
              fragments([&](double l) { return l - bf16_value(l); }, b->lo_fwd, b->lo_bwd);
            } else {
              // the low part L - float(L) of the FP32 contraction
              fragments([](double l) { return l - static_cast<double>(static_cast<float>(l)); }, b->lo_fwd, b->lo_bwd);
            }
          }
          variants = buf;
          b->v_buf = true;
        }
      }
      if (!variants.empty()) {  // else the kernels chosen above (v_fn or the run-time-shaped one) stay
        b->v_variants = variants;
        b->v_c = 8;
        for (const Variant& v : variants) set_lds(v.fn, v.lds);
      }
    }
  }
  set_lds(reinterpret_cast<const void*>(spo_dev::lines_fft_kernel<T>), 65536);
  set_lds(reinterpret_cast<const void*>(spo_dev::lines_fft_kernel<double>), 65536);
  set_lds(reinterpret_cast<const void*>(spo_dev::bk_scattering_kernel<T>), std::size_t(s.n_xy) * c16);
  // SPO_HIP_MERGE_HALF_STEPS=1: merged half steps between analysis steps (see spo_hip_backend::merge_halves).
  if (const char* m = std::getenv("SPO_HIP_MERGE_HALF_STEPS")) {
    const std::string v = m;
    if (v != "0" && v != "1") throw Status(SPO_HIP_E_ARGUMENT, "SPO_HIP_MERGE_HALF_STEPS must be 0 or 1");
    b->merge_halves = v == "1";
  }
  HIP_CHECK(hipStreamSynchronize(b->stream));
  b->tune();
  return b;
}

}  // namespace

extern "C" {

const char* spo_hip_last_error(void) { return g_error.c_str(); }

int spo_hip_device_count(int* count) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return SPO_HIP_OK;
}

int spo_hip_create(const spo_hip_shape* shape, const spo_hip_tables* t, const spo_hip_analysis* a,
                   const double* initial_specular, spo_hip_backend** out) {
  return guarded([&] {
    if (!shape || !t || !a || !initial_specular || !out) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_create: NULL argument");
    *out = nullptr;
    require_device();
    std::unique_ptr<spo_hip_backend> b;
    if (shape->precision == SPO_HIP_FP64)
      b = create_state<double>(shape, t, a, initial_specular);
    else if (shape->precision == SPO_HIP_MIXED)
      b = create_state<float>(shape, t, a, initial_specular);
    else
      throw Status(SPO_HIP_E_NOT_SUPPORTED, "unsupported precision");
    *out = b.release();
  });
}

int spo_hip_destroy(spo_hip_backend* b) {
  return guarded([&] {
    if (b && b->batches > 0)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_destroy: the backend is still a member of a batch (destroy the batch first)");
    delete b;
  });
}

int spo_hip_upload_wave(spo_hip_backend* b, const double* scattering, const double* specular) {
  return guarded([&] {
    if (scattering) b->upload_wave(scattering);
    if (specular)
      HIP_CHECK(hipMemcpyAsync(b->spec.as<void>(), specular, std::size_t(b->s.nr) * b->s.nzsp * sizeof(cx<double>),
                               hipMemcpyHostToDevice, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_download_wave(spo_hip_backend* b, double* scattering, double* specular) {
  return guarded([&] {
    if (scattering) b->download_wave(scattering);
    if (specular)
      HIP_CHECK(hipMemcpyAsync(specular, b->spec.as<void>(), std::size_t(b->s.nr) * b->s.nzsp * sizeof(cx<double>),
                               hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_kinetic_and_potential(spo_hip_backend* b, int half_step, int kinetic_only) {
  return guarded([&] {
    b->flush_pending();
    b->kinetic_and_potential(half_step != 0, kinetic_only != 0);
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

}  // extern "C"

namespace {

// The blocks of one analysis step, [T/2 V], (n - 1) [T V], [T/2], enqueued on the backend's streams.
void enqueue_blocks(spo_hip_backend* b) {
  // Merged half steps only on eager steps (a captured graph replays the plain schedule).
  const bool merge = b->merge_halves && !b->use_graph;
  if (!merge) b->flush_pending();
  b->step_block(true, merge && b->pending_half ? 2 : 1, false);
  for (int k = 1; k < b->shape.steps_per_analysis; ++k) b->kinetic_and_potential(false, false);
  if (merge) {
    b->step_block(true, 0, true);  // the specular grid's trailing half step; the scattering grid's waits
    b->pending_half = true;
  } else {
    b->kinetic_and_potential(true, true);
  }
}

// One analysis step on the backend's stream: the blocks, or in graph mode the captured step (captured here the
// first time).
void enqueue_analysis_step(spo_hip_backend* b) {
  if (!b->use_graph) {
    enqueue_blocks(b);
    return;
  }
  if (!b->step_graph) {
    hipGraph_t graph = nullptr;
    HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_blocks(b);
    HIP_CHECK(hipStreamEndCapture(b->stream, &graph));
    const hipError_t e = hipGraphInstantiate(&b->step_graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_CHECK(e);
  }
  HIP_CHECK(hipGraphLaunch(b->step_graph, b->stream));
}

}  // namespace

extern "C" {

int spo_hip_graph_enable(spo_hip_backend* b, int on) {
  return guarded([&] {
    if (!b) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_graph_enable: NULL backend");
    if (on && b->profiling) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_graph_enable: kernel timing is on");
    b->flush_pending();
    HIP_CHECK(hipStreamSynchronize(b->stream));
    if (b->step_graph) {
      HIP_CHECK(hipGraphExecDestroy(b->step_graph));
      b->step_graph = nullptr;
    }
    b->use_graph = on != 0;
  });
}

int spo_hip_propagate_analysis_step(spo_hip_backend* b) {
  return guarded([&] {
    enqueue_analysis_step(b);
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_propagate_batch(spo_hip_backend* const* backends, int n) {
  return guarded([&] {
    if (n < 0 || (n > 0 && !backends)) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_propagate_batch: bad arguments");
    for (int i = 0; i < n; ++i) {
      if (!backends[i]) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_propagate_batch: NULL backend");
      for (int j = 0; j < i; ++j)
        if (backends[j] == backends[i]) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_propagate_batch: a backend twice");
    }
    // Block r of every run's schedule [T/2 V], (n - 1) [T V], [T/2], round robin over the runs: each run's launches
    // stay in its own order on its own streams, and the runs' kernels overlap where one alone does not fill the GPU.
    // Runs in graph mode launch their captured step; the others' blocks go round robin. A batch step runs the plain
    // schedule: a pending half step is applied first.
    for (int i = 0; i < n; ++i) backends[i]->flush_pending();
    for (int i = 0; i < n; ++i)
      if (backends[i]->use_graph) enqueue_analysis_step(backends[i]);
    int longest = 0;
    for (int i = 0; i < n; ++i)
      if (!backends[i]->use_graph) longest = std::max(longest, backends[i]->shape.steps_per_analysis + 1);
    for (int r = 0; r < longest; ++r)
      for (int i = 0; i < n; ++i) {
        spo_hip_backend* b = backends[i];
        const int blocks = b->shape.steps_per_analysis + 1;
        if (b->use_graph || r >= blocks) continue;
        b->kinetic_and_potential(r == 0 || r == blocks - 1, r == blocks - 1);
      }
    for (int i = 0; i < n; ++i) HIP_CHECK(hipStreamSynchronize(backends[i]->stream));
  });
}

}  // extern "C"

// A batch of backends stepped together through one hipGraph: at its first step, block r of every member's schedule
// is enqueued round robin on the members' streams (as spo_hip_propagate_batch does) inside one capture that starts
// on the first member's stream, forks to the others through an event and joins them back; later steps replay the
// graph with one launch. Each member's launches keep their order on its own streams, so a member's results are
// bitwise those of the member stepped alone.
struct spo_hip_batch {
  std::vector<spo_hip_backend*> members;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraphExec_t graph = nullptr;
  ~spo_hip_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (fork) (void)hipEventDestroy(fork);
    for (hipEvent_t e : joins) (void)hipEventDestroy(e);
    for (spo_hip_backend* b : members) --b->batches;
  }
};

namespace {

void enqueue_round_robin(spo_hip_backend* const* backends, int n) {
  int longest = 0;
  for (int i = 0; i < n; ++i) longest = std::max(longest, backends[i]->shape.steps_per_analysis + 1);
  for (int r = 0; r < longest; ++r)
    for (int i = 0; i < n; ++i) {
      spo_hip_backend* b = backends[i];
      const int blocks = b->shape.steps_per_analysis + 1;
      if (r >= blocks) continue;
      b->kinetic_and_potential(r == 0 || r == blocks - 1, r == blocks - 1);
    }
}

}  // namespace

extern "C" {

int spo_hip_batch_create(spo_hip_backend* const* backends, int n, spo_hip_batch** out) {
  return guarded([&] {
    if (!out || n < 1 || !backends) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_batch_create: bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
      if (!backends[i]) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_batch_create: NULL backend");
      if (backends[i]->profiling) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_batch_create: kernel timing is on");
      for (int j = 0; j < i; ++j)
        if (backends[j] == backends[i]) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_batch_create: a backend twice");
    }
    for (int i = 0; i < n; ++i) backends[i]->flush_pending();
    auto batch = std::make_unique<spo_hip_batch>();
    HIP_CHECK(hipEventCreateWithFlags(&batch->fork, hipEventDisableTiming));
    for (int i = 0; i < n; ++i) {
      hipEvent_t e = nullptr;
      HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      batch->joins.push_back(e);
    }
    for (int i = 0; i < n; ++i) {
      batch->members.push_back(backends[i]);
      ++backends[i]->batches;
    }
    *out = batch.release();
  });
}

int spo_hip_batch_step(spo_hip_batch* batch) {
  return guarded([&] {
    if (!batch) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_batch_step: NULL batch");
    std::vector<spo_hip_backend*>& m = batch->members;
    hipStream_t origin = m[0]->stream;
    for (spo_hip_backend* b : m)
      if (b->profiling) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_batch_step: kernel timing is on");
    for (spo_hip_backend* b : m) {  // the graph replays the plain schedule: pending half steps are applied first
      if (!b->pending_half) continue;
      b->flush_pending();
      HIP_CHECK(hipStreamSynchronize(b->stream));
    }
    if (!batch->graph) {
      // Every member's earlier work is complete before the capture (the members' streams join the capture only
      // through the fork event).
      for (spo_hip_backend* b : m) HIP_CHECK(hipStreamSynchronize(b->stream));
      hipGraph_t graph = nullptr;
      HIP_CHECK(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal));
      HIP_CHECK(hipEventRecord(batch->fork, origin));
      for (std::size_t i = 1; i < m.size(); ++i) HIP_CHECK(hipStreamWaitEvent(m[i]->stream, batch->fork, 0));
      enqueue_round_robin(m.data(), static_cast<int>(m.size()));
      for (std::size_t i = 1; i < m.size(); ++i) {
        HIP_CHECK(hipEventRecord(batch->joins[i], m[i]->stream));
        HIP_CHECK(hipStreamWaitEvent(origin, batch->joins[i], 0));
      }
      HIP_CHECK(hipStreamEndCapture(origin, &graph));
      const hipError_t e = hipGraphInstantiate(&batch->graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP_CHECK(e);
    }
    HIP_CHECK(hipGraphLaunch(batch->graph, origin));
    HIP_CHECK(hipStreamSynchronize(origin));
  });
}

int spo_hip_batch_destroy(spo_hip_batch* batch) {
  return guarded([&] {
    if (!batch) return;
    HIP_CHECK(hipStreamSynchronize(batch->members[0]->stream));
    delete batch;
  });
}

int spo_hip_get_specular(spo_hip_backend* b, double* out) {
  return guarded([&] {
    HIP_CHECK(hipMemcpyAsync(out, b->spec.as<void>(), std::size_t(b->s.nr) * b->s.nzsp * sizeof(cx<double>),
                             hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_get_z_plane(spo_hip_backend* b, int iz, double* out) {
  return guarded([&] {
    const spo_dev::Shape& s = b->s;
    if (iz < 0 || iz >= s.nz) throw Status(SPO_HIP_E_ARGUMENT, "z plane index out of range");
    b->z_plane(iz);
    const long n = long(s.n_rot) * s.n_xy * s.nr;
    HIP_CHECK(hipMemcpyAsync(out, b->plane.as<void>(), n * sizeof(cx<double>), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_block_densities(spo_hip_backend* b, double* out) {
  return guarded([&] {
    b->block_densities();
    HIP_CHECK(hipMemcpyAsync(out, b->densities.as<void>(), std::size_t(b->s.n_zr) * sizeof(double),
                             hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_site_column(spo_hip_backend* b, int ixy, double* out) {
  return guarded([&] {
    if (ixy < 0 || ixy >= b->s.n_xy) throw Status(SPO_HIP_E_ARGUMENT, "site index out of range");
    b->site_column(ixy, out);
  });
}

int spo_hip_analyse(spo_hip_backend* b, int surface_fixed, int iz_specular, double t, double* coefficients) {
  return guarded([&] { b->analyse(surface_fixed, iz_specular, t, coefficients); });
}

int spo_hip_get_s_sums(spo_hip_backend* b, double* out) {
  return guarded([&] {
    const spo_hip_analysis& a = b->analysis;
    const std::size_t n = std::size_t(a.n_states) * a.n_v * a.n_energies;
    HIP_CHECK(hipMemcpyAsync(out, b->sums.as<void>(), n * sizeof(cx<double>), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_energy_resolved_start(spo_hip_backend* b, int n_sites, const int* site_ixy, int n_energies,
                                  double analysis_time_step) {
  return guarded([&] {
    if (n_sites < 1 || n_energies < 1 || !site_ixy)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_energy_resolved_start: no sites or no energies");
    for (int i = 0; i < n_sites; ++i)
      if (site_ixy[i] < 0 || site_ixy[i] >= b->s.n_xy) throw Status(SPO_HIP_E_ARGUMENT, "site outside the (x, y) grid");
    const std::size_t block = std::size_t(b->s.n_zr) * b->s.n_rot;
    const std::size_t spec_n = std::size_t(b->s.nzsp) * b->s.nr;
    b->er_sums = Buffer(std::size_t(n_sites) * n_energies * block * sizeof(cx<double>));
    b->er_specular = Buffer(std::size_t(n_energies) * spec_n * sizeof(cx<double>));
    b->er_sites = upload(site_ixy, std::size_t(n_sites) * sizeof(int));
    b->er_phase = Buffer(std::size_t(n_energies) * sizeof(cx<double>));
    b->er_n_sites = n_sites;
    b->er_n_energies = n_energies;
    b->er_dt = analysis_time_step;
    HIP_CHECK(hipMemsetAsync(b->er_sums.as<void>(), 0, std::size_t(n_sites) * n_energies * block * sizeof(cx<double>),
                             b->stream));
    HIP_CHECK(hipMemsetAsync(b->er_specular.as<void>(), 0, std::size_t(n_energies) * spec_n * sizeof(cx<double>),
                             b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_energy_resolved_add(spo_hip_backend* b, const double* phase_factors) {
  return guarded([&] {
    if (b->er_n_sites == 0) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_energy_resolved_add before _start");
    HIP_CHECK(hipMemcpyAsync(b->er_phase.as<void>(), phase_factors, std::size_t(b->er_n_energies) * sizeof(cx<double>),
                             hipMemcpyHostToDevice, b->stream));
    b->energy_resolved_scattering();
    const int spec_n = b->s.nzsp * b->s.nr;
    b->timed("energy_resolved_specular", [&] {
      hipLaunchKernelGGL(spo_dev::energy_resolved_specular_kernel, dim3(blocks_for(spec_n, 256)), dim3(256), 0,
                         b->stream, b->spec.as<cx<double>>(), spec_n, b->er_n_energies, b->er_phase.as<cx<double>>(),
                         b->er_dt, b->er_specular.as<cx<double>>());
    });
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_energy_resolved_read(spo_hip_backend* b, double* sums, double* specular_sums) {
  return guarded([&] {
    if (b->er_n_sites == 0) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_energy_resolved_read before _start");
    const std::size_t n_zr = b->s.n_zr, n_rot = b->s.n_rot, block = n_zr * n_rot;
    const std::size_t pairs = std::size_t(b->er_n_sites) * b->er_n_energies;
    std::vector<cx<double>> dev(pairs * block);
    HIP_CHECK(hipMemcpyAsync(dev.data(), b->er_sums.as<void>(), dev.size() * sizeof(cx<double>), hipMemcpyDeviceToHost,
                             b->stream));
    HIP_CHECK(hipMemcpyAsync(specular_sums, b->er_specular.as<void>(),
                             std::size_t(b->er_n_energies) * b->s.nzsp * b->s.nr * sizeof(cx<double>),
                             hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    // [pair][irot][izr] -> [pair][izr][irot]
    // (the caller's buffer is only guaranteed 8-byte aligned: copy components)
    for (std::size_t pr = 0; pr < pairs; ++pr)
      for (std::size_t irot = 0; irot < n_rot; ++irot)
        for (std::size_t izr = 0; izr < n_zr; ++izr) {
          const cx<double> v = dev[pr * block + izr + n_zr * irot];
          double* o = sums + 2 * (pr * block + irot + n_rot * izr);
          o[0] = v.x;
          o[1] = v.y;
        }
  });
}

int spo_hip_profile_enable(spo_hip_backend* b, int on) {
  return guarded([&] {
    if (on && b->use_graph) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_profile_enable: graph mode is on");
    HIP_CHECK(hipStreamSynchronize(b->stream));
    for (const auto& t : b->timed_launches) {
      b->event_pool.push_back(t.start);
      b->event_pool.push_back(t.stop);
    }
    b->timed_launches.clear();
    b->profiling = on != 0;
  });
}

int spo_hip_profile_read(spo_hip_backend* b, spo_hip_kernel_time* out, int capacity, int* count) {
  return guarded([&] {
    HIP_CHECK(hipStreamSynchronize(b->stream));
    std::vector<spo_hip_kernel_time> agg(b->kernel_names.size());
    for (std::size_t k = 0; k < agg.size(); ++k) {
      std::memset(&agg[k], 0, sizeof(agg[k]));
      std::strncpy(agg[k].name, b->kernel_names[k].c_str(), sizeof(agg[k].name) - 1);
    }
    for (const auto& t : b->timed_launches) {
      float ms = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&ms, t.start, t.stop));
      agg[t.kernel].launches += 1;
      agg[t.kernel].total_ms += ms;
    }
    *count = static_cast<int>(agg.size());
    for (int k = 0; k < *count && k < capacity; ++k) out[k] = agg[k];
  });
}

int spo_hip_accumulate(spo_hip_backend* b, double time, const double* coefficients) {
  return guarded([&] {
    if (!b || !coefficients) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_accumulate: NULL argument");
    const spo_hip_analysis& a = b->analysis;
    const std::size_t n = std::size_t(a.n_states) * a.n_v;
    HIP_CHECK(hipMemcpyAsync(b->bk.as<void>(), coefficients, n * sizeof(cx<double>), hipMemcpyHostToDevice,
                             b->stream));
    const long total = long(n) * a.n_energies;
    hipLaunchKernelGGL(spo_dev::s_accumulate_kernel, dim3(blocks_for(total, kThreads)), dim3(kThreads), 0, b->stream,
                       b->sums.as<cx<double>>(), b->bk.as<cx<double>>(), static_cast<int>(n), a.n_energies,
                       b->energies.as<double>(), time, a.analysis_time_step);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

}  // extern "C"

// The potential surrogate's device state: FP32 master parameters, gradient, Adam moments, and activation buffers
// for `capacity` rows.
struct spo_hip_mlp {
  int batch = 0;
  long capacity = 0;
  int adam_step = 0;
  Buffer params, grad, m, v;
  Buffer x, t, y, dy, z1, a1, z2, a2, z3, a3, d1, d2;
};

namespace {

using spo_dev::kMlpWidth;

void mlp_forward(spo_hip_mlp& n, long rows, const float* x) {
  const float* p = n.params.as<float>();
  const long cells = rows * kMlpWidth;
  hipLaunchKernelGGL(spo_dev::mlp_input_kernel, dim3(blocks_for(cells, 256)), dim3(256), 0, nullptr, x, p, rows,
                     n.z1.as<float>(), n.a1.as<float>());
  const int tiles_j = kMlpWidth / 16;
  const dim3 grid(static_cast<unsigned>(rows / 16 * tiles_j));
  hipLaunchKernelGGL(spo_dev::mlp_gemm_kernel, grid, dim3(64), 0, nullptr, n.a1.as<float>(), long(kMlpWidth), 1L,
                     p + spo_dev::kMlpW2, long(kMlpWidth), 1L, kMlpWidth, tiles_j, n.z2.as<float>(), long(kMlpWidth), 1,
                     p + spo_dev::kMlpB2, n.a2.as<float>());
  hipLaunchKernelGGL(spo_dev::mlp_gemm_kernel, grid, dim3(64), 0, nullptr, n.a2.as<float>(), long(kMlpWidth), 1L,
                     p + spo_dev::kMlpW3, long(kMlpWidth), 1L, kMlpWidth, tiles_j, n.z3.as<float>(), long(kMlpWidth), 1,
                     p + spo_dev::kMlpB3, n.a3.as<float>());
  hipLaunchKernelGGL(spo_dev::mlp_output_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, nullptr, n.a3.as<float>(),
                     p, rows, n.y.as<float>());
  HIP_CHECK(hipGetLastError());
}

// The gradient of the mean squared error over `rows` rows into n.grad (after mlp_forward on the same rows).
void mlp_backward(spo_hip_mlp& n, long rows, const float* x, const float* t, float* loss) {
  const float* p = n.params.as<float>();
  float* g = n.grad.as<float>();
  const long cells = rows * kMlpWidth;
  const int tiles = kMlpWidth / 16;
  const long H = kMlpWidth;
  hipLaunchKernelGGL(spo_dev::mlp_loss_kernel, dim3(1), dim3(256), 0, nullptr, n.y.as<float>(), t, rows,
                     n.dy.as<float>(), loss);
  hipLaunchKernelGGL(spo_dev::mlp_column_sum_kernel, dim3(kMlpWidth), dim3(256), 0, nullptr, n.a3.as<float>(), H,
                     n.dy.as<float>(), rows, g + spo_dev::kMlpW4);
  hipLaunchKernelGGL(spo_dev::mlp_column_sum_kernel, dim3(1), dim3(256), 0, nullptr, n.dy.as<float>(), 1L,
                     static_cast<const float*>(nullptr), rows, g + spo_dev::kMlpB4);
  float* dz3 = n.d2.as<float>();  // dZ3, then dA2 / dZ2 in d1, then dA1 / dZ1 in d2
  hipLaunchKernelGGL(spo_dev::mlp_output_back_kernel, dim3(blocks_for(cells, 256)), dim3(256), 0, nullptr,
                     n.dy.as<float>(), p, n.z3.as<float>(), rows, dz3);
  // layer 3: dW3 = A2^T dZ3, db3, dA2 = dZ3 W3^T, dZ2 = dA2 SiLU'(z2)
  hipLaunchKernelGGL(spo_dev::mlp_gemm_kernel, dim3(tiles * tiles), dim3(64), 0, nullptr, n.a2.as<float>(), 1L, H,
                     dz3, H, 1L, static_cast<int>(rows), tiles, g + spo_dev::kMlpW3, H, 0,
                     static_cast<const float*>(nullptr), static_cast<float*>(nullptr));
  hipLaunchKernelGGL(spo_dev::mlp_column_sum_kernel, dim3(kMlpWidth), dim3(256), 0, nullptr, dz3, H,
                     static_cast<const float*>(nullptr), rows, g + spo_dev::kMlpB3);
  float* dz2 = n.d1.as<float>();
  hipLaunchKernelGGL(spo_dev::mlp_gemm_kernel, dim3(static_cast<unsigned>(rows / 16 * tiles)), dim3(64), 0, nullptr,
                     dz3, H, 1L, p + spo_dev::kMlpW3, 1L, H, kMlpWidth, tiles, dz2, H, 0,
                     static_cast<const float*>(nullptr), static_cast<float*>(nullptr));
  hipLaunchKernelGGL(spo_dev::mlp_silu_back_kernel, dim3(blocks_for(cells, 256)), dim3(256), 0, nullptr, dz2,
                     n.z2.as<float>(), cells, dz2);
  // layer 2
  hipLaunchKernelGGL(spo_dev::mlp_gemm_kernel, dim3(tiles * tiles), dim3(64), 0, nullptr, n.a1.as<float>(), 1L, H,
                     dz2, H, 1L, static_cast<int>(rows), tiles, g + spo_dev::kMlpW2, H, 0,
                     static_cast<const float*>(nullptr), static_cast<float*>(nullptr));
  hipLaunchKernelGGL(spo_dev::mlp_column_sum_kernel, dim3(kMlpWidth), dim3(256), 0, nullptr, dz2, H,
                     static_cast<const float*>(nullptr), rows, g + spo_dev::kMlpB2);
  float* dz1 = n.d2.as<float>();
  hipLaunchKernelGGL(spo_dev::mlp_gemm_kernel, dim3(static_cast<unsigned>(rows / 16 * tiles)), dim3(64), 0, nullptr,
                     dz2, H, 1L, p + spo_dev::kMlpW2, 1L, H, kMlpWidth, tiles, dz1, H, 0,
                     static_cast<const float*>(nullptr), static_cast<float*>(nullptr));
  hipLaunchKernelGGL(spo_dev::mlp_silu_back_kernel, dim3(blocks_for(cells, 256)), dim3(256), 0, nullptr, dz1,
                     n.z1.as<float>(), cells, dz1);
  // layer 1
  hipLaunchKernelGGL(spo_dev::mlp_input_grad_kernel, dim3(spo_dev::kMlpIn * kMlpWidth), dim3(256), 0, nullptr, x, dz1,
                     rows, g);
  hipLaunchKernelGGL(spo_dev::mlp_column_sum_kernel, dim3(kMlpWidth), dim3(256), 0, nullptr, dz1, H,
                     static_cast<const float*>(nullptr), rows, g + spo_dev::kMlpB1);
  HIP_CHECK(hipGetLastError());
}

void mlp_adam(spo_hip_mlp& n, float lr) {
  constexpr float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  ++n.adam_step;
  const float c1 = 1.0f - std::pow(beta1, static_cast<float>(n.adam_step));
  const float c2 = 1.0f - std::pow(beta2, static_cast<float>(n.adam_step));
  hipLaunchKernelGGL(spo_dev::mlp_adam_kernel, dim3(blocks_for(spo_dev::kMlpParams, 256)), dim3(256), 0, nullptr,
                     n.params.as<float>(), n.grad.as<float>(), n.m.as<float>(), n.v.as<float>(), spo_dev::kMlpParams,
                     lr, beta1, beta2, eps, c1, c2);
  HIP_CHECK(hipGetLastError());
}

}  // namespace

extern "C" {

int spo_hip_mlp_create(int batch, spo_hip_mlp** out) {
  return guarded([&] {
    if (!out) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_mlp_create: NULL out");
    if (batch < 16 || batch % 16 != 0 || batch > (1 << 22))
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_mlp_create: batch must be a multiple of 16 (16 .. 4194304)");
    require_device();
    auto n = std::make_unique<spo_hip_mlp>();
    n->batch = batch;
    n->capacity = std::max<long>(batch, 1L << 16);
    const std::size_t pb = spo_dev::kMlpParams * sizeof(float);
    n->params = Buffer(pb);
    n->grad = Buffer(pb);
    n->m = Buffer(pb);
    n->v = Buffer(pb);
    HIP_CHECK(hipMemset(n->params.as<void>(), 0, pb));
    HIP_CHECK(hipMemset(n->m.as<void>(), 0, pb));
    HIP_CHECK(hipMemset(n->v.as<void>(), 0, pb));
    const std::size_t rows = static_cast<std::size_t>(n->capacity);
    n->x = Buffer(rows * spo_dev::kMlpIn * sizeof(float));
    n->t = Buffer(rows * sizeof(float));
    n->y = Buffer(rows * sizeof(float));
    n->dy = Buffer(rows * sizeof(float));
    for (Buffer* b : {&n->z1, &n->a1, &n->z2, &n->a2, &n->z3, &n->a3, &n->d1, &n->d2})
      *b = Buffer(rows * kMlpWidth * sizeof(float));
    *out = n.release();
  });
}

int spo_hip_mlp_destroy(spo_hip_mlp* mlp) {
  delete mlp;
  return SPO_HIP_OK;
}

int spo_hip_mlp_set_params(spo_hip_mlp* mlp, const float* params) {
  return guarded([&] {
    if (!mlp || !params) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_mlp_set_params: NULL argument");
    HIP_CHECK(hipMemcpy(mlp->params.as<void>(), params, spo_dev::kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    const std::size_t pb = spo_dev::kMlpParams * sizeof(float);
    HIP_CHECK(hipMemset(mlp->m.as<void>(), 0, pb));
    HIP_CHECK(hipMemset(mlp->v.as<void>(), 0, pb));
    mlp->adam_step = 0;
  });
}

int spo_hip_mlp_get_params(const spo_hip_mlp* mlp, float* params) {
  return guarded([&] {
    if (!mlp || !params) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_mlp_get_params: NULL argument");
    HIP_CHECK(hipMemcpy(params, mlp->params.as<void>(), spo_dev::kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int spo_hip_mlp_gradient(spo_hip_mlp* mlp, const float* x, const float* t, float* loss, float* grad) {
  return guarded([&] {
    if (!mlp || !x || !t || !loss || !grad) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_mlp_gradient: NULL argument");
    const long rows = mlp->batch;
    HIP_CHECK(hipMemcpy(mlp->x.as<void>(), x, rows * spo_dev::kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(mlp->t.as<void>(), t, rows * sizeof(float), hipMemcpyHostToDevice));
    Buffer l(sizeof(float));
    mlp_forward(*mlp, rows, mlp->x.as<float>());
    mlp_backward(*mlp, rows, mlp->x.as<float>(), mlp->t.as<float>(), l.as<float>());
    HIP_CHECK(hipMemcpy(loss, l.as<void>(), sizeof(float), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(grad, mlp->grad.as<void>(), spo_dev::kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int spo_hip_mlp_train(spo_hip_mlp* mlp, long n, const float* x, const float* t, int steps, float lr, float* losses) {
  return guarded([&] {
    if (!mlp || !x || !t || !losses) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_mlp_train: NULL argument");
    if (n < mlp->batch || n % mlp->batch != 0 || steps < 0)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_mlp_train: n must be a positive multiple of the batch");
    Buffer dx = upload(x, std::size_t(n) * spo_dev::kMlpIn * sizeof(float));
    Buffer dt = upload(t, std::size_t(n) * sizeof(float));
    Buffer dl(std::size_t(std::max(steps, 1)) * sizeof(float));
    const long rows = mlp->batch;
    for (int s = 0; s < steps; ++s) {
      const long off = (static_cast<long>(s) * rows) % n;
      const float* xs = dx.as<float>() + off * spo_dev::kMlpIn;
      mlp_forward(*mlp, rows, xs);
      mlp_backward(*mlp, rows, xs, dt.as<float>() + off, dl.as<float>() + s);
      mlp_adam(*mlp, lr);
    }
    if (steps > 0) HIP_CHECK(hipMemcpy(losses, dl.as<void>(), steps * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int spo_hip_mlp_predict(spo_hip_mlp* mlp, long n, const float* x, float* y) {
  return guarded([&] {
    if (!mlp || (n > 0 && (!x || !y)) || n < 0) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_mlp_predict: bad arguments");
    for (long done = 0; done < n;) {
      const long chunk = std::min(mlp->capacity, n - done);
      const long rows = (chunk + 15) / 16 * 16;  // the tail padded with zero rows
      HIP_CHECK(hipMemset(mlp->x.as<void>(), 0, rows * spo_dev::kMlpIn * sizeof(float)));
      HIP_CHECK(hipMemcpy(mlp->x.as<void>(), x + done * spo_dev::kMlpIn, chunk * spo_dev::kMlpIn * sizeof(float),
                          hipMemcpyHostToDevice));
      mlp_forward(*mlp, rows, mlp->x.as<float>());
      HIP_CHECK(hipMemcpy(y + done, mlp->y.as<void>(), chunk * sizeof(float), hipMemcpyDeviceToHost));
      done += chunk;
    }
  });
}

int spo_hip_pes_values(int pes, long n, const double* points, double* values) {
  return guarded([&] {
    if (pes != SPO_HIP_PES_HARDWALL && pes != SPO_HIP_PES_DAILIGHT)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_pes_values: unknown surface " + std::to_string(pes));
    if (n < 0 || (n > 0 && (!points || !values))) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_pes_values: bad arrays");
    if (n == 0) return;
    require_device();
    Buffer p = upload(points, std::size_t(n) * 6 * sizeof(double));
    Buffer out(std::size_t(n) * sizeof(double));
    hipLaunchKernelGGL(spo_dev::pes_points_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, nullptr, pes, n,
                       p.as<double>(), out.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(values, out.as<void>(), std::size_t(n) * sizeof(double), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"

// Scratch of spo_hip_pes_values_static, one copy per device for the life of the process: the points (6 doubles
// each) of up to kPesStaticMax points, then their values. 3.5 MiB.
constexpr long kPesStaticMax = 1L << 16;
static __device__ __align__(256) double scratch[7 * kPesStaticMax];

namespace {

// The device address of `scratch` on each device, looked up once per device with that device current. In host
// code the name `scratch` is only a stand-in for the symbol; kernels and copies get this address instead.
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_address[kMaxDevices] = {};

double* scratch_on_current_device() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Status(SPO_HIP_E_NOT_SUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_address[dev] = static_cast<double*>(p);
  }
  return scratch_address[dev];
}

}  // namespace

extern "C" {

int spo_hip_pes_values_static(int pes, long n, const double* points, double* values) {
  return guarded([&] {
    if (pes != SPO_HIP_PES_HARDWALL && pes != SPO_HIP_PES_DAILIGHT)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_pes_values_static: unknown surface " + std::to_string(pes));
    if (n < 0 || (n > 0 && (!points || !values)))
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_pes_values_static: bad arrays");
    if (n > kPesStaticMax)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_pes_values_static: " + std::to_string(n) + " points, at most " +
                                           std::to_string(kPesStaticMax) + " per call");
    if (n == 0) return;
    require_device();
    double* p = scratch_on_current_device();
    double* out = p + 6 * kPesStaticMax;
    // Every value of the call is written by the kernel, so nothing of an earlier call's contents is read.
    HIP_CHECK(hipMemcpy(p, points, std::size_t(n) * 6 * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(spo_dev::pes_points_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, nullptr, pes, n, p, out);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(values, out, std::size_t(n) * sizeof(double), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"

namespace {

// A packed CRP surface checked against its own offsets and uploaded; dev() gives the kernels' view.
struct CrpUpload {
  Buffer corrugation, cluster, atom_offsets, atom_splines, cut_shape, cut_offsets, cut_data, h2, zdep, cell;
  spo_dev::CrpDev view{};

  explicit CrpUpload(const spo_hip_crp_surface& s) {
    const int n_basis = s.square ? 0 : (s.c3v ? 10 : 7), n_cuts = s.square ? 16 : (s.c3v ? 29 : 18);
    auto bad = [](const char* what) { throw Status(SPO_HIP_E_ARGUMENT, std::string("CRP surface: ") + what); };
    if (s.n_cuts != n_cuts) bad("the number of molecular cuts does not match the symmetry");
    if (s.n_corrugation < 4 || s.n_cluster < 1 || s.n_h2 < 4 || s.n_z_dependence < 4) bad("too few points");
    if (!s.corrugation || !s.cluster || !s.atom_offsets || !s.cut_shape || !s.cut_offsets || !s.cut_data || !s.h2 ||
        !s.z_dependence || (!s.square && !s.atom_splines) || (s.square && !s.cell_spline))
      bad("NULL array");
    if (s.square && s.cell_nz < 4) bad("the C4v cell spline needs at least 4 z points");
    if (s.atom_offsets[0] != 0) bad("atom offsets");
    for (int i = 0; i < n_basis; ++i)
      if (s.atom_offsets[i + 1] - s.atom_offsets[i] < 4) bad("an atom cut with fewer than 4 knots");
    if (s.cut_offsets[0] != 0) bad("cut offsets");
    for (int k = 0; k < n_cuts; ++k) {
      const long nr = s.cut_shape[3 * k], nz = s.cut_shape[3 * k + 1], g = s.cut_shape[3 * k + 2];
      if (nr < 4 || nz < 4 || g < 0 || g > k) bad("cut shape");
      if (g != k && (s.cut_shape[3 * g] != nr || s.cut_shape[3 * g + 1] != nz)) bad("cut group of another shape");
      if (s.cut_offsets[k + 1] - s.cut_offsets[k] != nr + nz + 4 * nr * nz) bad("cut block size");
    }
    corrugation = upload(s.corrugation, 3 * std::size_t(s.n_corrugation) * sizeof(double));
    cluster = upload(s.cluster, 4 * std::size_t(s.n_cluster) * sizeof(double));
    atom_offsets = upload(s.atom_offsets, (n_basis + 1) * sizeof(int));
    atom_splines = s.square ? Buffer(sizeof(double))
                            : upload(s.atom_splines, 3 * std::size_t(s.atom_offsets[n_basis]) * sizeof(double));
    if (s.square) cell = upload(s.cell_spline, (10 + std::size_t(s.cell_nz) + 8 * 25 * std::size_t(s.cell_nz)) * sizeof(double));
    cut_shape = upload(s.cut_shape, 3 * std::size_t(n_cuts) * sizeof(int));
    cut_offsets = upload(s.cut_offsets, (n_cuts + 1) * sizeof(long));
    cut_data = upload(s.cut_data, std::size_t(s.cut_offsets[n_cuts]) * sizeof(double));
    h2 = upload(s.h2, 3 * std::size_t(s.n_h2) * sizeof(double));
    zdep = upload(s.z_dependence, 3 * std::size_t(s.n_z_dependence) * sizeof(double));
    view = spo_dev::CrpDev{s.c3v, n_basis, n_cuts, s.n_corrugation, s.n_cluster, s.n_h2, s.n_z_dependence,
                           s.lattice_constant, s.zso_min, s.zso_max, s.r_min, s.r_max, s.z_min, s.h2_asymptote,
                           corrugation.as<double>(), cluster.as<double>(), atom_offsets.as<int>(),
                           atom_splines.as<double>(), cut_shape.as<int>(), cut_offsets.as<long>(),
                           cut_data.as<double>(), h2.as<double>(), zdep.as<double>(), s.square, s.cell_nz,
                           s.square ? cell.as<double>() : nullptr};
  }
};

// An SCM correction checked and uploaded (the atoms' layers 1 or 2, a positive scale).
struct ScmUpload {
  Buffer atoms;
  spo_dev::ScmDev view{};

  explicit ScmUpload(const spo_hip_scm_correction* m) {
    if (!m) return;
    if (m->n_atoms < 1 || !m->atoms) throw Status(SPO_HIP_E_ARGUMENT, "SCM correction: no atoms");
    if (!(m->scale > 0.0)) throw Status(SPO_HIP_E_ARGUMENT, "SCM correction: the scale must be positive");
    for (int i = 0; i < m->n_atoms; ++i)
      if (m->atoms[7 * i + 6] != 1.0 && m->atoms[7 * i + 6] != 2.0)
        throw Status(SPO_HIP_E_ARGUMENT, "SCM correction: an atom outside layers 1 and 2");
    atoms = upload(m->atoms, 7 * std::size_t(m->n_atoms) * sizeof(double));
    view.on = 1;
    view.rescale = m->rescale;
    view.scale = m->scale;
    view.cos_skew = m->cos_skew;
    view.sin_skew = m->sin_skew;
    for (int l = 0; l < 2; ++l)
      for (int k = 0; k < 18; ++k) view.layers[l][k] = m->layers[l][k];
    view.n_atoms = m->n_atoms;
    view.atoms = atoms.as<double>();
  }
};

void check_finite(const double* v, long n, const char* where) {
  for (long i = 0; i < n; ++i)
    if (std::isnan(v[i]))
      throw Status(SPO_HIP_E_ARGUMENT, std::string(where) + ": the CRP surface cannot be evaluated at point " +
                                           std::to_string(i) + " (an atom on a surface atom or outside the data)");
}

}  // namespace

extern "C" {

int spo_hip_crp_values(const spo_hip_crp_surface* surface, long n, const double* points, double* values) {
  return spo_hip_scm_values(surface, nullptr, n, points, values);
}

int spo_hip_scm_values(const spo_hip_crp_surface* surface, const spo_hip_scm_correction* scm, long n,
                       const double* points, double* values) {
  return guarded([&] {
    if (!surface) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_crp_values: NULL surface");
    if (n < 0 || (n > 0 && (!points || !values))) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_crp_values: bad arrays");
    if (n == 0) return;
    require_device();
    CrpUpload crp(*surface);
    ScmUpload correction(scm);
    Buffer p = upload(points, std::size_t(n) * 6 * sizeof(double));
    Buffer out(std::size_t(n) * sizeof(double));
    hipLaunchKernelGGL(spo_dev::crp_points_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, nullptr, crp.view,
                       correction.view, n, p.as<double>(), out.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(values, out.as<void>(), std::size_t(n) * sizeof(double), hipMemcpyDeviceToHost));
    check_finite(values, n, "spo_hip_crp_values");
  });
}

int spo_hip_potential_grid(const spo_hip_pes_grid* grid, double* values) {
  return guarded([&] {
    if (!grid || !values) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_potential_grid: NULL argument");
    const spo_hip_pes_grid& g = *grid;
    if (g.pes != SPO_HIP_PES_HARDWALL && g.pes != SPO_HIP_PES_DAILIGHT && g.pes != SPO_HIP_PES_CRP)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_potential_grid: unknown surface " + std::to_string(g.pes));
    if (g.pes == SPO_HIP_PES_CRP && !g.crp) throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_potential_grid: no CRP surface");
    if (g.nx <= 0 || g.ny <= 0 || g.nr <= 0 || g.nz_potential < 0 || g.n_theta <= 0 || g.n_phi <= 0 ||
        g.n_representatives <= 0 || g.n_representatives > g.nx * g.ny)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_potential_grid: bad extents");
    if (!g.x || !g.y || !g.r || !g.theta || !g.phi || !g.representatives || (g.nz_potential > 0 && !g.z) ||
        (g.use_switching && !g.gas_phase))
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_potential_grid: NULL table");
    for (int k = 0; k < g.n_representatives; ++k)
      if (g.representatives[k] < 0 || g.representatives[k] >= g.nx * g.ny)
        throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_potential_grid: representative outside the (x, y) grid");
    const long n_zr = long(g.nz_potential) * g.nr;
    const long n = long(g.n_theta) * g.n_phi * g.n_representatives * n_zr;
    if (n == 0) return;
    require_device();
    auto up = [](const void* h, std::size_t bytes) { return upload(h, bytes); };
    Buffer x = up(g.x, g.nx * sizeof(double)), y = up(g.y, g.ny * sizeof(double));
    Buffer z = up(g.z, g.nz_potential * sizeof(double)), r = up(g.r, g.nr * sizeof(double));
    Buffer th = up(g.theta, g.n_theta * sizeof(double)), ph = up(g.phi, g.n_phi * sizeof(double));
    Buffer reps = up(g.representatives, g.n_representatives * sizeof(int));
    Buffer gas = g.use_switching ? up(g.gas_phase, g.nr * sizeof(double)) : Buffer();
    Buffer out(std::size_t(n) * sizeof(double));
    std::optional<CrpUpload> crp;
    if (g.pes == SPO_HIP_PES_CRP) crp.emplace(*g.crp);
    const ScmUpload correction(g.pes == SPO_HIP_PES_CRP ? g.scm : nullptr);
    spo_dev::PesGrid d{g.pes, g.ny, g.nr, g.n_theta, g.n_phi, g.n_representatives, static_cast<int>(n_zr),
                       x.as<double>(), y.as<double>(), z.as<double>(), r.as<double>(), th.as<double>(),
                       ph.as<double>(), reps.as<int>(), g.use_switching, g.z_min_switch, g.z_max_switch,
                       g.use_switching ? gas.as<double>() : nullptr, g.impose_v_min, g.impose_v_max, g.v_min,
                       g.v_max, crp ? crp->view : spo_dev::CrpDev{}, correction.view};
    hipLaunchKernelGGL(spo_dev::potential_grid_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, nullptr, d,
                       out.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(values, out.as<void>(), std::size_t(n) * sizeof(double), hipMemcpyDeviceToHost));
    if (g.pes == SPO_HIP_PES_CRP) check_finite(values, n, "spo_hip_potential_grid");
  });
}

}  // extern "C"
