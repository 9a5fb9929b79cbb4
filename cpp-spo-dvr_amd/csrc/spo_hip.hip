// The MI355X propagation backend behind the C ABI of include/spo_hip.h: device buffers, FFT plans and the launch
// sequence of one analysis step. Kernels: kernels.hpp; LDS FFT engine: fft_lds.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spo_hip.h"
#include "kernels.hpp"

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
  Buffer(Buffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr; o.bytes_ = 0; }
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
  std::size_t bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  std::size_t bytes_ = 0;
};

Buffer upload(const void* host, std::size_t bytes) {
  Buffer b(bytes);
  if (bytes) HIP_CHECK(hipMemcpy(b.as<void>(), host, bytes, hipMemcpyHostToDevice));
  return b;
}

// Radices of a length: 8 while it divides, then 4, 2, 3, 5, 7, 11.
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

// An FFT plan with its twiddle table on the device.
struct DevicePlan {
  Buffer table;
  spo_dev::FftPlan plan{};
  explicit DevicePlan(int n) {
    const std::vector<int> r = radices_of(n);
    std::vector<cx<double>> w(n);
    for (int m = 0; m < n; ++m) {
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / n;
      w[m] = {static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a))};
    }
    table = upload(w.data(), w.size() * sizeof(cx<double>));
    plan.n = n;
    plan.n_stages = static_cast<int>(r.size());
    for (std::size_t i = 0; i < r.size(); ++i) plan.radix[i] = r[i];
    plan.table = table.as<void>();
  }
};

std::vector<cx<double>> complex_from(const double* p, std::size_t n) {
  std::vector<cx<double>> v(n);
  std::memcpy(v.data(), p, n * sizeof(cx<double>));
  return v;
}

unsigned blocks_for(long n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

constexpr int kThreads = 256;

}  // namespace

struct spo_hip_backend {
  spo_hip_shape shape{};
  spo_dev::Shape s{};
  hipStream_t stream = nullptr;
  int max_lds = 65536;

  Buffer psi, spec, densities, plane, staging, bk, sums;
  Buffer zr_full, zr_half, xy_full, xy_half, spec_full, spec_half;
  Buffer rot_half, before, gas_step, spec_gas, z_opt, r_opt, z_spec_opt;
  Buffer shift_f, shift_b, s2sc, sc2s;
  Buffer legendre, block_first, block_size, block_offset, rot_block, rep_of, op_of, phi_index, exp_table;
  Buffer rot_index, dif_slot, radial_weight, initial_radial, energies;
  std::unique_ptr<DevicePlan> plan_z, plan_r, plan_x, plan_y, plan_phi, plan_zsp;
  spo_hip_analysis analysis{};
  int xy_c = 1, v_c = 1;
  bool zr_fused = false;
  std::size_t staging_values = 0;

  void kinetic_and_potential(bool half, bool kinetic_only);
  void lines(cx<double>* data, spo_dev::Lines L, const DevicePlan& p, int sign, int mode, const cx<double>* factor,
             int f_line, int f_elem);
};

namespace {

void set_lds(const void* kernel, std::size_t bytes) {
  HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
}

}  // namespace

void spo_hip_backend::lines(cx<double>* data, spo_dev::Lines L, const DevicePlan& p, int sign, int mode,
                            const cx<double>* factor, int f_line, int f_elem) {
  L.lines_per_block = std::max(1, std::min(64, 65536 / (L.n * static_cast<int>(sizeof(cx<double>)))));
  const std::size_t lds = std::size_t(L.lines_per_block) * L.n * sizeof(cx<double>);
  const unsigned grid = blocks_for(L.count, L.lines_per_block);
  hipLaunchKernelGGL(spo_dev::lines_fft_kernel<double>, dim3(grid), dim3(kThreads), lds, stream, data, L, p.plan,
                     sign, mode, factor, f_line, f_elem);
  HIP_CHECK(hipGetLastError());
}

void spo_hip_backend::kinetic_and_potential(bool half, bool kinetic_only) {
  cx<double>* ps = psi.as<cx<double>>();
  cx<double>* sp = spec.as<cx<double>>();
  // Specular grid: r lines, Z lines with the kinetic factor, r lines back, then gas phase and absorbers.
  {
    spo_dev::Lines r_lines{s.nr, s.nzsp, s.nzsp, s.nr, 0, 1, 1};
    spo_dev::Lines z_lines{s.nzsp, s.nr, s.nr, 1, 0, s.nr, 1};
    lines(sp, r_lines, *plan_r, -1, 0, nullptr, 0, 0);
    lines(sp, z_lines, *plan_zsp, -1, 1, (half ? spec_half : spec_full).as<cx<double>>(), 1, s.nr);
    lines(sp, r_lines, *plan_r, +1, 0, nullptr, 0, 0);
    if (!kinetic_only) {
      hipLaunchKernelGGL(spo_dev::specular_potential_kernel<double>, dim3(blocks_for(long(s.nr) * s.nzsp, kThreads)),
                         dim3(kThreads), 0, stream, sp, s, spec_gas.as<cx<double>>(), z_spec_opt.as<double>(),
                         r_opt.as<double>());
      HIP_CHECK(hipGetLastError());
    }
  }
  // T_Zr.
  const cx<double>* zr = (half ? zr_half : zr_full).as<cx<double>>();
  const int n_planes = s.n_rot * s.n_xy;
  if (zr_fused) {
    hipLaunchKernelGGL(spo_dev::zr_plane_kernel<double>, dim3(n_planes), dim3(kThreads),
                       std::size_t(s.n_zr) * sizeof(cx<double>), stream, ps, s, plan_z->plan, plan_r->plan, zr);
    HIP_CHECK(hipGetLastError());
  } else {
    spo_dev::Lines r_lines{s.nr, s.nz * n_planes, s.nz, s.nr, s.n_zr, 1, 1};
    spo_dev::Lines z_lines{s.nz, s.nr * n_planes, s.nr, 1, s.n_zr, s.nr, 1};
    lines(ps, r_lines, *plan_r, -1, 0, nullptr, 0, 0);
    lines(ps, z_lines, *plan_z, -1, 1, zr, 1, s.nr);
    lines(ps, r_lines, *plan_r, +1, 0, nullptr, 0, 0);
  }
  // T_xy and the diagonal part of V.
  {
    const int tiles = (s.n_zr + xy_c - 1) / xy_c;
    const std::size_t lds = std::size_t(xy_c) * s.n_xy * sizeof(cx<double>);
    hipLaunchKernelGGL(spo_dev::xy_kernel<double>, dim3(tiles * s.n_rot), dim3(kThreads), lds, stream, ps, s, xy_c,
                       plan_x->plan, plan_y->plan, (half ? xy_half : xy_full).as<cx<double>>(),
                       shift_f.as<cx<double>>(), shift_b.as<cx<double>>(), shape.needs_shift, kinetic_only ? 0 : 1,
                       z_opt.as<double>(), before.as<cx<double>>(), gas_step.as<cx<double>>());
    HIP_CHECK(hipGetLastError());
  }
  if (kinetic_only || s.n_zr_pot == 0) return;
  // V in the interaction region, then the projection onto the specular grid.
  {
    spo_dev::AngularTables at{legendre.as<double>(),  block_first.as<int>(), block_size.as<int>(),
                              block_offset.as<int>(), rot_block.as<int>(),   rep_of.as<int>(),
                              op_of.as<int>(),        phi_index.as<int>()};
    const int tiles = (s.n_zr_pot + v_c - 1) / v_c;
    const std::size_t lds = std::size_t(v_c) * (s.n_rot + s.n_tp) * sizeof(cx<double>);
    hipLaunchKernelGGL(spo_dev::interaction_kernel<double>, dim3(tiles * s.n_xy), dim3(kThreads), lds, stream, ps, sp,
                       s, v_c, plan_phi->plan, at, exp_table.as<cx<double>>(), rot_half.as<cx<double>>(),
                       s2sc.as<cx<double>>());
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(spo_dev::projection_kernel<double>, dim3(blocks_for(s.n_zr_pot, kThreads)), dim3(kThreads), 0,
                       stream, ps, sp, s, sc2s.as<cx<double>>(), s2sc.as<cx<double>>());
    HIP_CHECK(hipGetLastError());
  }
}

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
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1)
      throw Status(SPO_HIP_E_NOT_SUPPORTED, "no HIP device is visible");
    if (shape->precision != SPO_HIP_FP64) throw Status(SPO_HIP_E_NOT_SUPPORTED, "unsupported precision");
    auto b = std::make_unique<spo_hip_backend>();
    b->shape = *shape;
    spo_dev::Shape& s = b->s;
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
    if (s.n_rot < 1 || s.nx < 2 || s.ny < 2 || s.nr < 2 || s.nz < 2 || s.nzsp < s.nz || s.nzpot < 0 || s.nzpot > s.nz ||
        s.n_theta < 1 || s.n_phi < 1 || s.n_rep < 1 || s.ichn < 0 || s.ichn >= s.n_rot || shape->n_operations < 1 ||
        shape->steps_per_analysis < 1)
      throw Status(SPO_HIP_E_ARGUMENT, "spo_hip_create: inconsistent shape");

    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    HIP_CHECK(hipDeviceGetAttribute(&b->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    const std::size_t psi_values = std::size_t(s.n_zr) * s.n_rot * s.n_xy;
    const std::size_t exp_values = std::size_t(s.n_tp) * s.n_rep * s.n_zr_pot;
    std::size_t free_bytes = 0, total_bytes = 0;
    HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
    const std::size_t need = (psi_values + exp_values) * sizeof(cx<double>) + (std::size_t(1) << 28);
    if (need > free_bytes)
      throw Status(SPO_HIP_E_NOT_SUPPORTED, "not enough device memory: need " + std::to_string(need >> 20) +
                                                " MiB, free " + std::to_string(free_bytes >> 20) + " MiB");
    HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));

    const std::size_t c16 = sizeof(cx<double>);
    b->psi = Buffer(psi_values * c16);
    HIP_CHECK(hipMemsetAsync(b->psi.as<void>(), 0, psi_values * c16, b->stream));
    b->spec = upload(initial_specular, std::size_t(s.nr) * s.nzsp * c16);
    b->densities = Buffer(std::size_t(s.n_zr) * sizeof(double));
    b->plane = Buffer(std::size_t(s.n_rot) * s.n_xy * s.nr * c16);
    b->staging_values = std::max<std::size_t>(std::size_t(s.n_rot) * s.n_xy, std::size_t(4) << 20);
    b->staging = Buffer(b->staging_values * c16);

    // Kinetic factors with the forward and backward FFT scales folded in.
    auto folded = [&](const double* f, std::size_t n, double fwd, double bwd) {
      std::vector<cx<double>> v = complex_from(f, n);
      for (auto& x : v) {
        x.x = x.x * fwd * bwd;
        x.y = x.y * fwd * bwd;
      }
      return upload(v.data(), n * c16);
    };
    const double zr_f = t->fft_scale_z * t->fft_scale_r;
    const double zr_b = (1.0 / (s.nz * t->fft_scale_z)) * (1.0 / (s.nr * t->fft_scale_r));
    const double sp_b = (1.0 / (s.nzsp * t->fft_scale_z)) * (1.0 / (s.nr * t->fft_scale_r));
    const double xy_f = t->fft_scale_x * t->fft_scale_y;
    const double xy_b = (1.0 / (s.nx * t->fft_scale_x)) * (1.0 / (s.ny * t->fft_scale_y));
    b->zr_full = folded(t->zr_full, s.n_zr, zr_f, zr_b);
    b->zr_half = folded(t->zr_half, s.n_zr, zr_f, zr_b);
    b->spec_full = folded(t->specular_full, std::size_t(s.nr) * s.nzsp, zr_f, sp_b);
    b->spec_half = folded(t->specular_half, std::size_t(s.nr) * s.nzsp, zr_f, sp_b);
    b->xy_full = folded(t->xy_full, s.n_xy, xy_f, xy_b);
    b->xy_half = folded(t->xy_half, s.n_xy, xy_f, xy_b);

    b->rot_half = upload(t->rotation_half, std::size_t(s.n_rot) * s.nr * c16);
    b->before = upload(t->before_interaction, std::size_t(s.n_rot) * s.nr * c16);
    b->gas_step = upload(t->gas_phase_step, std::size_t(s.n_rot) * s.nr * c16);
    b->spec_gas = upload(t->specular_gas_phase, std::size_t(s.nr) * c16);
    b->z_opt = upload(t->z_optical, std::size_t(s.nz) * sizeof(double));
    b->r_opt = upload(t->r_optical, std::size_t(s.nr) * sizeof(double));
    b->z_spec_opt = upload(t->z_specular_optical, std::size_t(s.nzsp) * sizeof(double));
    b->shift_f = upload(t->shift_forward, std::size_t(s.n_xy) * c16);
    b->shift_b = upload(t->shift_backward, std::size_t(s.n_xy) * c16);
    b->s2sc = upload(t->specular_to_scattering, std::size_t(s.n_xy) * c16);
    b->sc2s = upload(t->scattering_to_specular, std::size_t(s.n_xy) * c16);

    // Angular transform tables.
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
    b->exp_table = upload(t->exp_potential, exp_values * c16);

    // Analysis.
    b->analysis = *a;
    if (a->n_states != a->n_rotation * a->n_diffraction || a->n_v < 1 || a->n_energies < 1 ||
        a->iz_analysis < 0 || a->iz_analysis >= s.nz || a->initial_state < 0 || a->initial_state >= a->n_states ||
        a->n_diffraction > s.n_xy)
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
    HIP_CHECK(hipMemsetAsync(b->sums.as<void>(), 0, n_sums * c16, b->stream));

    // Plans and tile widths.
    b->plan_z = std::make_unique<DevicePlan>(s.nz);
    b->plan_r = std::make_unique<DevicePlan>(s.nr);
    b->plan_x = std::make_unique<DevicePlan>(s.nx);
    b->plan_y = std::make_unique<DevicePlan>(s.ny);
    b->plan_phi = std::make_unique<DevicePlan>(s.n_phi);
    b->plan_zsp = std::make_unique<DevicePlan>(s.nzsp);
    const std::size_t lds_limit = static_cast<std::size_t>(b->max_lds);
    b->zr_fused = std::size_t(s.n_zr) * c16 <= lds_limit;
    b->xy_c = std::max(1, std::min(64, 4096 / s.n_xy));
    int vc = static_cast<int>(81920 / ((s.n_rot + s.n_tp) * c16));
    int p2 = 1;
    while (p2 * 2 <= std::min(vc, 32)) p2 *= 2;
    b->v_c = std::max(1, p2);
    if (std::size_t(s.n_xy) * c16 > lds_limit || std::size_t(b->v_c) * (s.n_rot + s.n_tp) * c16 > lds_limit)
      throw Status(SPO_HIP_E_NOT_SUPPORTED, "a tile does not fit in LDS");
    set_lds(reinterpret_cast<const void*>(spo_dev::lines_fft_kernel<double>), 65536);
    set_lds(reinterpret_cast<const void*>(spo_dev::xy_kernel<double>), std::size_t(b->xy_c) * s.n_xy * c16);
    set_lds(reinterpret_cast<const void*>(spo_dev::interaction_kernel<double>),
            std::size_t(b->v_c) * (s.n_rot + s.n_tp) * c16);
    if (b->zr_fused) set_lds(reinterpret_cast<const void*>(spo_dev::zr_plane_kernel<double>), std::size_t(s.n_zr) * c16);
    set_lds(reinterpret_cast<const void*>(spo_dev::bk_scattering_kernel), std::size_t(s.n_xy) * c16);
    HIP_CHECK(hipStreamSynchronize(b->stream));
    *out = b.release();
  });
}

int spo_hip_destroy(spo_hip_backend* b) {
  if (!b) return SPO_HIP_OK;
  if (b->stream) {
    (void)hipStreamSynchronize(b->stream);
    (void)hipStreamDestroy(b->stream);
  }
  delete b;
  return SPO_HIP_OK;
}

int spo_hip_upload_wave(spo_hip_backend* b, const double* scattering, const double* specular) {
  return guarded([&] {
    const spo_dev::Shape& s = b->s;
    const std::size_t c16 = sizeof(cx<double>);
    if (scattering) {
      const std::size_t per_izr = std::size_t(s.n_rot) * s.n_xy;
      const int chunk = static_cast<int>(std::max<std::size_t>(1, b->staging_values / per_izr));
      for (int izr0 = 0; izr0 < s.n_zr; izr0 += chunk) {
        const int count = std::min(chunk, s.n_zr - izr0);
        HIP_CHECK(hipMemcpyAsync(b->staging.as<void>(), scattering + 2 * per_izr * izr0, per_izr * count * c16,
                                 hipMemcpyHostToDevice, b->stream));
        hipLaunchKernelGGL(spo_dev::host_to_device_kernel<double>, dim3(blocks_for(long(per_izr) * count, kThreads)),
                           dim3(kThreads), 0, b->stream, b->psi.as<cx<double>>(), s,
                           b->staging.as<cx<double>>(), izr0, count);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(b->stream));
      }
    }
    if (specular)
      HIP_CHECK(hipMemcpyAsync(b->spec.as<void>(), specular, std::size_t(s.nr) * s.nzsp * c16, hipMemcpyHostToDevice,
                               b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_download_wave(spo_hip_backend* b, double* scattering, double* specular) {
  return guarded([&] {
    const spo_dev::Shape& s = b->s;
    const std::size_t c16 = sizeof(cx<double>);
    if (scattering) {
      const std::size_t per_izr = std::size_t(s.n_rot) * s.n_xy;
      const int chunk = static_cast<int>(std::max<std::size_t>(1, b->staging_values / per_izr));
      for (int izr0 = 0; izr0 < s.n_zr; izr0 += chunk) {
        const int count = std::min(chunk, s.n_zr - izr0);
        hipLaunchKernelGGL(spo_dev::device_to_host_kernel<double>, dim3(blocks_for(long(per_izr) * count, kThreads)),
                           dim3(kThreads), 0, b->stream, b->psi.as<cx<double>>(), s,
                           b->staging.as<cx<double>>(), izr0, count);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(scattering + 2 * per_izr * izr0, b->staging.as<void>(), per_izr * count * c16,
                                 hipMemcpyDeviceToHost, b->stream));
        HIP_CHECK(hipStreamSynchronize(b->stream));
      }
    }
    if (specular)
      HIP_CHECK(hipMemcpyAsync(specular, b->spec.as<void>(), std::size_t(s.nr) * s.nzsp * c16, hipMemcpyDeviceToHost,
                               b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_kinetic_and_potential(spo_hip_backend* b, int half_step, int kinetic_only) {
  return guarded([&] {
    b->kinetic_and_potential(half_step != 0, kinetic_only != 0);
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_propagate_analysis_step(spo_hip_backend* b) {
  return guarded([&] {
    b->kinetic_and_potential(true, false);
    for (int k = 1; k < b->shape.steps_per_analysis; ++k) b->kinetic_and_potential(false, false);
    b->kinetic_and_potential(true, true);
    HIP_CHECK(hipStreamSynchronize(b->stream));
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
    const long n = long(s.n_rot) * s.n_xy * s.nr;
    hipLaunchKernelGGL(spo_dev::z_plane_kernel<double>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, b->stream,
                       b->psi.as<cx<double>>(), s, iz, b->plane.as<cx<double>>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, b->plane.as<void>(), n * sizeof(cx<double>), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_block_densities(spo_hip_backend* b, double* out) {
  return guarded([&] {
    const spo_dev::Shape& s = b->s;
    hipLaunchKernelGGL(spo_dev::densities_kernel<double>, dim3(blocks_for(s.n_zr, kThreads)), dim3(kThreads), 0,
                       b->stream, b->psi.as<cx<double>>(), s, b->densities.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, b->densities.as<void>(), std::size_t(s.n_zr) * sizeof(double), hipMemcpyDeviceToHost,
                             b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_site_column(spo_hip_backend* b, int ixy, double* out) {
  return guarded([&] {
    const spo_dev::Shape& s = b->s;
    if (ixy < 0 || ixy >= s.n_xy) throw Status(SPO_HIP_E_ARGUMENT, "site index out of range");
    const std::size_t n = std::size_t(s.n_zr) * s.n_rot;
    HIP_CHECK(hipMemcpyAsync(out, b->psi.as<cx<double>>() + n * ixy, n * sizeof(cx<double>), hipMemcpyDeviceToHost,
                             b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_analyse(spo_hip_backend* b, int surface_fixed, int iz_specular, double t, double* coefficients) {
  return guarded([&] {
    const spo_dev::Shape& s = b->s;
    const spo_hip_analysis& a = b->analysis;
    const std::size_t n = std::size_t(a.n_states) * a.n_v;
    hipLaunchKernelGGL(spo_dev::zero_kernel, dim3(blocks_for(long(n), kThreads)), dim3(kThreads), 0, b->stream,
                       b->bk.as<cx<double>>(), long(n));
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(spo_dev::bk_scattering_kernel, dim3(a.n_v * a.n_rotation), dim3(kThreads),
                       std::size_t(s.n_xy) * sizeof(cx<double>), b->stream, b->psi.as<cx<double>>(), s, a.iz_analysis,
                       a.n_rotation, a.n_diffraction, a.n_states, b->rot_index.as<int>(), b->dif_slot.as<int>(),
                       b->radial_weight.as<double>(), b->shift_f.as<cx<double>>(), a.xy_scale, b->plan_x->plan,
                       b->plan_y->plan, b->bk.as<cx<double>>());
    HIP_CHECK(hipGetLastError());
    if (surface_fixed) {
      if (iz_specular < 0 || iz_specular >= s.nzsp) throw Status(SPO_HIP_E_ARGUMENT, "specular line out of range");
      hipLaunchKernelGGL(spo_dev::bk_specular_kernel, dim3(1), dim3(64), 0, b->stream, b->spec.as<cx<double>>(), s.nr,
                         iz_specular, a.n_v, a.n_states, a.initial_state, b->initial_radial.as<double>(), a.dr,
                         b->bk.as<cx<double>>());
      HIP_CHECK(hipGetLastError());
    }
    const long total = long(n) * a.n_energies;
    hipLaunchKernelGGL(spo_dev::s_accumulate_kernel, dim3(blocks_for(total, kThreads)), dim3(kThreads), 0, b->stream,
                       b->sums.as<cx<double>>(), b->bk.as<cx<double>>(), static_cast<int>(n), a.n_energies,
                       b->energies.as<double>(), t, a.analysis_time_step);
    HIP_CHECK(hipGetLastError());
    if (coefficients)
      HIP_CHECK(hipMemcpyAsync(coefficients, b->bk.as<void>(), n * sizeof(cx<double>), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int spo_hip_get_s_sums(spo_hip_backend* b, double* out) {
  return guarded([&] {
    const spo_hip_analysis& a = b->analysis;
    const std::size_t n = std::size_t(a.n_states) * a.n_v * a.n_energies;
    HIP_CHECK(hipMemcpyAsync(out, b->sums.as<void>(), n * sizeof(cx<double>), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

}  // extern "C"
