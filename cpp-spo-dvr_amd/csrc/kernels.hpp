// Device kernels of one split-operator time step (CDNA4, wave64) and of the analysis.
//
// Device layout of the scattering wave: psi[izr + n_zr * (irot + n_rot * ixy)] (izr = ir + nr * iz fastest), so
//   - a (Z, r) plane of one (irot, ixy) is contiguous (the T_Zr pass reads whole planes),
//   - a tile of consecutive izr of one (irot, ixy) is a contiguous run (T_xy and V read C-wide runs).
// Specular wave spec[ir + nr * iz]. Every kernel is written for wave64 and LDS tiles; the FFT engine is
// fft_lds.hpp.
#pragma once

#include "fft_lds.hpp"

namespace spo_dev {

struct Shape {
  int n_rot, nx, ny, n_xy, nr, nz, n_zr, nzsp, nzpot, n_zr_pot;
  int n_theta, n_phi, n_tp, n_rep, ichn, mj_max, n_blocks;
  int r_abs_first, r_abs_last;
};

// w v for a factor w of a complex64 table whose entry idx has the FP32 remainder lo[idx] of the FP64 factor
// (mul_hilo); FP64 passes no lo.
template <typename T>
__device__ __forceinline__ cx<T> mul_factor(cx<T> w, cx<T> v, const cx<T>* lo, int idx) {
  if constexpr (sizeof(T) == 4) {
    return mul_hilo(w, lo[idx], v);
  } else {
    return w * v;
  }
}

// The complex64 remainders (mul_hilo) of the tables xy_reg_kernel multiplies with (NULL in FP64).
template <typename T>
struct XyLo {
  const cx<T>* factor;
  const cx<T>* shift_f;
  const cx<T>* shift_b;
  const cx<T>* before;
  const cx<T>* gas_step;
};

// ---------------------------------------------------------------------------------------------------------------
// Batched 1D FFTs of lines in global memory (the specular grid, and the T_Zr pass when a whole (Z, r) plane does
// not fit in LDS). Line l of the batch: offset (l % inner) * inner_stride + (l / inner) * outer_stride, element i
// at offset + i * elem_stride. mode 0: transform with sign_a; mode 1: transform (sign_a), multiply by
// factor[(l % inner) * f_line + i * f_elem] (complex64: with its remainder factor_lo, mul_factor), transform back
// (-sign_a). Each workgroup takes `lines_per_block` consecutive lines.
struct Lines {
  int n;
  int count, inner;
  long inner_stride, outer_stride, elem_stride;
  int lines_per_block;
};

template <typename T>
__global__ __launch_bounds__(256) void lines_fft_kernel(cx<T>* data, Lines L, FftPlan plan, int sign_a, int mode,
                                                        const cx<T>* factor, const cx<T>* factor_lo, int f_line,
                                                        int f_elem) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  const int l0 = blockIdx.x * L.lines_per_block;
  const int nl = min(L.lines_per_block, L.count - l0);
  const int total = nl * L.n;
  // tile layout: [line][i]
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int li = e / L.n, i = e % L.n;
    const int l = l0 + li;
    const long off = (l % L.inner) * L.inner_stride + (long)(l / L.inner) * L.outer_stride + i * L.elem_stride;
    tile[e] = data[off];
  }
  __syncthreads();
  SeqLayout lay{nl, nl, L.n, 0, 1};
  lds_fft<T>(tile, lay, plan, sign_a);
  if (mode == 1) {
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
      const int li = e / L.n, i = e % L.n;
      const int l = l0 + li;
      const int f = (l % L.inner) * f_line + i * f_elem;
      tile[e] = mul_factor(factor[f], tile[e], factor_lo, f);
    }
    __syncthreads();
    lds_fft<T>(tile, lay, plan, -sign_a);
  }
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int li = e / L.n, i = e % L.n;
    const int l = l0 + li;
    const long off = (l % L.inner) * L.inner_stride + (long)(l / L.inner) * L.outer_stride + i * L.elem_stride;
    data[off] = tile[e];
  }
}

// Specular grid after its kinetic step: exp(-i vasy dt) between rot(dt/2) for iz >= nzpot, then the Z and r
// absorbers (src/propagation/specular_operator.cpp:61-71).
template <typename T>
__global__ void specular_potential_kernel(cx<T>* spec, Shape s, const cx<T>* gas, const T* z_opt_sp,
                                          const T* r_opt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= s.nr * s.nzsp) return;
  const int ir = k % s.nr, iz = k / s.nr;
  cx<T> v = spec[k];
  if (iz >= s.nzpot) v = gas[ir] * v;
  v = scale(v, z_opt_sp[iz]);
  if (ir >= s.r_abs_first && ir <= s.r_abs_last) v = scale(v, r_opt[ir]);
  spec[k] = v;
}

// ---------------------------------------------------------------------------------------------------------------
// T_Zr: one (irot, ixy) plane per workgroup, the whole (Z, r) plane in LDS: 2D FFT, x exp(-i (T_Z + T_r) dt) with
// the FFT scales folded in (complex64: with its remainder factor_lo), inverse 2D FFT. One read and one write of psi.
template <typename T>
__global__ __launch_bounds__(256) void zr_plane_kernel(cx<T>* psi, Shape s, FftPlan plan_z, FftPlan plan_r,
                                                       const cx<T>* factor, const cx<T>* factor_lo) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* plane = psi + (long)blockIdx.x * s.n_zr;
  for (int e = threadIdx.x; e < s.n_zr; e += blockDim.x) tile[e] = plane[e];
  __syncthreads();
  const SeqLayout rows{s.nz, s.nz, s.nr, 0, 1};   // r sequences: contiguous, one per iz
  const SeqLayout cols{s.nr, s.nr, 1, 0, s.nr};   // Z sequences: stride nr, one per ir
  lds_fft<T>(tile, rows, plan_r, -1);
  lds_fft<T>(tile, cols, plan_z, -1);
  for (int e = threadIdx.x; e < s.n_zr; e += blockDim.x) tile[e] = mul_factor(factor[e], tile[e], factor_lo, e);
  __syncthreads();
  lds_fft<T>(tile, cols, plan_z, +1);
  lds_fft<T>(tile, rows, plan_r, +1);
  for (int e = threadIdx.x; e < s.n_zr; e += blockDim.x) plane[e] = tile[e];
}

// ---------------------------------------------------------------------------------------------------------------
// T_xy (+ the diagonal part of V): a tile of C consecutive izr of one irot, all (x, y) points, in LDS as
// [ixy][c]. Momentum shift, 2D FFT, x exp(-i T_xy dt) (scales folded in), inverse, shift back; then, unless
// kinetic_only, optZ (optR rot(dt/2) [exp(-i vasy dt) rot(dt/2) beyond zasymp]) per (irot, izr)
// (src/propagation/kinetic_operators.cpp:134-172, src/propagation/potential_operator.cpp:100-120).
template <typename T>
__global__ __launch_bounds__(256) void xy_kernel(cx<T>* psi, Shape s, int C, FftPlan plan_x, FftPlan plan_y,
                                                 const cx<T>* factor,
                          const cx<T>* shift_f, const cx<T>* shift_b, int needs_shift, int apply_potential,
                          const T* z_opt, const cx<T>* before, const cx<T>* gas_step, XyLo<T> lo) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  const int tiles_per_rot = (s.n_zr + C - 1) / C;
  const int irot = blockIdx.x / tiles_per_rot;
  const int izr0 = (blockIdx.x % tiles_per_rot) * C;
  const int cw = min(C, s.n_zr - izr0);
  const int total = C * s.n_xy;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int c = e % C, ixy = e / C;
    cx<T> v = {T(0), T(0)};
    if (c < cw) v = psi[izr0 + c + (long)s.n_zr * (irot + (long)s.n_rot * ixy)];
    if (needs_shift) v = mul_factor(shift_f[ixy], v, lo.shift_f, ixy);
    tile[e] = v;
  }
  __syncthreads();
  const SeqLayout ys{C * s.nx, C, 1, C * s.ny, C};        // for each (ix, c): iy at stride C
  const SeqLayout xs{C * s.ny, C * s.ny, 1, 0, C * s.ny};  // for each (iy, c): ix at stride C ny
  lds_fft<T>(tile, ys, plan_y, -1);
  lds_fft<T>(tile, xs, plan_x, -1);
  for (int e = threadIdx.x; e < total; e += blockDim.x) tile[e] = mul_factor(factor[e / C], tile[e], lo.factor, e / C);
  __syncthreads();
  lds_fft<T>(tile, xs, plan_x, +1);
  lds_fft<T>(tile, ys, plan_y, +1);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int c = e % C, ixy = e / C;
    if (c >= cw) continue;
    cx<T> v = tile[e];
    if (needs_shift) v = mul_factor(shift_b[ixy], v, lo.shift_b, ixy);
    const int izr = izr0 + c;
    if (apply_potential) {
      const int iz = izr / s.nr, ir = izr % s.nr;
      const int k = irot + s.n_rot * ir;
      const cx<T> u = iz >= s.nzpot ? mul_factor(gas_step[k], v, lo.gas_step, k) : mul_factor(before[k], v, lo.before, k);
      v = scale(u, z_opt[iz]);
    }
    psi[izr + (long)s.n_zr * (irot + (long)s.n_rot * ixy)] = v;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// V in the interaction region (iz < nzpot): a tile of C consecutive izr of one (x, y) column. The specular wave is
// added to the initial channel, then per mJ block the Gauss-Legendre sums to the theta grid, the phi FFT (sign +),
// x exp(i A) through the symmetry maps, the phi FFT back (sign -), the adjoint sums, and rot(dt/2)
// (src/propagation/potential_operator.cpp:121-131, :158-201; src/propagation/angular_transform.cpp:54-114).
// LDS: fbr [irot][c], dvr [c][theta][phi].
struct AngularTables {
  const double* legendre;  // (mJ block, theta, J)
  const int* block_first;
  const int* block_size;
  const int* block_offset;  // start of each block's (theta, J) matrix in legendre
  const int* rot_block;     // per irot: its mJ block
  const int* rep_of;
  const int* op_of;
  const int* phi_index;     // (iphi, op)
};

template <typename T>
__global__ __launch_bounds__(256) void interaction_kernel(cx<T>* psi, const cx<double>* spec, Shape s, int C,
                                                          FftPlan plan_phi, AngularTables at, const cx<T>* exp_table,
                                                          const cx<T>* rot_half, const cx<T>* s2sc,
                                                          const cx<T>* rot_lo) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* fbr = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* dvr = fbr + C * s.n_rot;
  const int tiles_per_xy = (s.n_zr_pot + C - 1) / C;
  const int ixy = blockIdx.x / tiles_per_xy;
  const int izr0 = (blockIdx.x % tiles_per_xy) * C;
  const int cw = min(C, s.n_zr_pot - izr0);
  const long col = (long)s.n_zr * s.n_rot * ixy;

  for (int e = threadIdx.x; e < C * s.n_rot; e += blockDim.x) {
    const int c = e % C, irot = e / C;
    cx<T> v = {T(0), T(0)};
    if (c < cw) {
      v = psi[izr0 + c + col + (long)s.n_zr * irot];
      if (irot == s.ichn) {
        const int izr = izr0 + c;
        const int ir = izr % s.nr, iz = izr / s.nr;
        const cx<double> in = spec[ir + s.nr * iz];
        v = v + rot_half[s.ichn + s.n_rot * ir] * (s2sc[ixy] * cx<T>{T(in.x), T(in.y)});
      }
    }
    fbr[e] = v;
  }
  const int dvr_size = C * s.n_theta * s.n_phi;
  for (int e = threadIdx.x; e < dvr_size; e += blockDim.x) dvr[e] = cx<T>{T(0), T(0)};
  __syncthreads();

  // FBR -> DVR, theta part: one output (c, theta, block) per thread iteration.
  const int n_out = C * s.n_theta * s.n_blocks;
  for (int e = threadIdx.x; e < n_out; e += blockDim.x) {
    const int c = e % C;
    const int th = (e / C) % s.n_theta;
    const int b = e / (C * s.n_theta);
    const int first = at.block_first[b], nj = at.block_size[b];
    const double* L = at.legendre + at.block_offset[b] + th * nj;
    cx<T> acc = {T(0), T(0)};
    for (int j = 0; j < nj; ++j) acc = acc + scale(fbr[(first + j) * C + c], T(L[j]));
    const int m = b - s.mj_max;
    const int slot = m >= 0 ? m : s.n_phi + m;
    dvr[(c * s.n_theta + th) * s.n_phi + slot] = acc;
  }
  __syncthreads();
  const SeqLayout rows{C * s.n_theta, C * s.n_theta, s.n_phi, 0, 1};
  lds_fft<T>(dvr, rows, plan_phi, +1);

  const int n_tp = s.n_tp;
  for (int e = threadIdx.x; e < dvr_size; e += blockDim.x) {
    const int ip = e % s.n_phi;
    const int th = (e / s.n_phi) % s.n_theta;
    const int c = e / (s.n_phi * s.n_theta);
    if (c >= cw) continue;
    const int rep = at.rep_of[ixy];
    const int op = at.op_of[ixy];
    const int itp = at.phi_index[ip + s.n_phi * op] + s.n_phi * th;
    dvr[e] = exp_table[itp + (long)n_tp * (rep + (long)s.n_rep * (izr0 + c))] * dvr[e];
  }
  __syncthreads();
  lds_fft<T>(dvr, rows, plan_phi, -1);

  // DVR -> FBR: one output (c, irot) per thread iteration, then rot(dt/2), then the store.
  for (int e = threadIdx.x; e < C * s.n_rot; e += blockDim.x) {
    const int c = e % C, irot = e / C;
    if (c >= cw) continue;
    const int b = at.rot_block[irot];
    const int first = at.block_first[b], nj = at.block_size[b];
    const int j = irot - first;
    const int m = b - s.mj_max;
    const int slot = m >= 0 ? m : s.n_phi + m;
    const double* L = at.legendre + at.block_offset[b] + j;
    cx<T> acc = {T(0), T(0)};
    for (int th = 0; th < s.n_theta; ++th) acc = acc + scale(dvr[(c * s.n_theta + th) * s.n_phi + slot], T(L[th * nj]));
    const int izr = izr0 + c;
    const int ir = izr % s.nr;
    psi[izr + col + (long)s.n_zr * irot] = mul_factor(rot_half[irot + s.n_rot * ir], acc, rot_lo, irot + s.n_rot * ir);
  }
}

// Projection of the initial channel onto the specular grid and its removal from the scattering grid
// (src/propagation/potential_operator.cpp:133-140). One workgroup per 64 interaction-region (Z, r) points, lanes
// along izr (each load is 64 consecutive values of one site); wavefront w sums the sites w, w + W, w + 2 W, ... and
// the W partial sums are added in wavefront order, a fixed order, so repeated runs agree bitwise.
// The projection is summed in FP64 and the specular wave kept in FP64 whatever the storage type of psi.
constexpr int kProjectionWaves = 16;

template <typename T>
__global__ __launch_bounds__(64 * kProjectionWaves) void projection_kernel(cx<T>* psi, cx<double>* spec, Shape s,
                                                                          const cx<double>* sc2s,
                                                                          const cx<double>* s2sc) {
  constexpr int W = kProjectionWaves;
  __shared__ cx<double> part[W][64];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int izr = blockIdx.x * 64 + lane;
  const bool on = izr < s.n_zr_pot;
  const long stride = (long)s.n_zr * s.n_rot;
  cx<T>* p = psi + izr + (long)s.n_zr * s.ichn;
  cx<double> acc = {0.0, 0.0};
  if (on)
    for (int ixy = w; ixy < s.n_xy; ixy += W) {
      const cx<T> v = p[stride * ixy];
      acc = acc + sc2s[ixy] * cx<double>{double(v.x), double(v.y)};
    }
  part[w][lane] = acc;
  __syncthreads();
  cx<double> chi = {0.0, 0.0};
  for (int k = 0; k < W; ++k) chi = chi + part[k][lane];
  if (!on) return;
  if (w == 0) spec[izr % s.nr + s.nr * (izr / s.nr)] = chi;
  for (int ixy = w; ixy < s.n_xy; ixy += W) {
    const cx<T> v = p[stride * ixy];
    const cx<double> r = cx<double>{double(v.x), double(v.y)} - s2sc[ixy] * chi;
    p[stride * ixy] = cx<T>{T(r.x), T(r.y)};
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Analysis. One workgroup per (v, irotana): overlap over r with phi_vj(r) dR and the momentum shift, (x, y) FFT
// (sign -, scale 1), pick the analysed diffraction slots (src/analysis/balint_kurti.cpp:52-72).
template <typename T>
__global__ void bk_scattering_kernel(const cx<T>* psi, Shape s, int iz_ana, int n_rot_ana, int n_dif,
                                     int n_states, const int* rot_index, const int* slots, const double* weights,
                                     const cx<double>* shift_f, double xy_scale, FftPlan plan_x, FftPlan plan_y,
                                     cx<double>* bk) {
  extern __shared__ unsigned char smem_raw[];
  cx<double>* ov = reinterpret_cast<cx<double>*>(smem_raw);
  const int task = blockIdx.x;
  const int v = task / n_rot_ana, irotana = task % n_rot_ana;
  const int irot = rot_index[irotana];
  const double* w = weights + (long)s.nr * task;
  for (int ixy = threadIdx.x; ixy < s.n_xy; ixy += blockDim.x) {
    const cx<T>* p = psi + (long)s.n_zr * (irot + (long)s.n_rot * ixy) + s.nr * iz_ana;
    cx<double> acc = {0.0, 0.0};
    for (int ir = 0; ir < s.nr; ++ir) {
      const cx<double> x = {double(p[ir].x), double(p[ir].y)};
      acc = acc + scale(shift_f[ixy] * x, w[ir]);
    }
    ov[ixy] = acc;
  }
  __syncthreads();
  const SeqLayout ys{s.nx, s.nx, s.ny, 0, 1};
  const SeqLayout xs{s.ny, s.ny, 1, 0, s.ny};
  lds_fft<double>(ov, ys, plan_y, -1);
  lds_fft<double>(ov, xs, plan_x, -1);
  for (int idif = threadIdx.x; idif < n_dif; idif += blockDim.x)
    bk[irotana * n_dif + idif + (long)n_states * v] = scale(ov[slots[idif]], xy_scale);
}

// The initial state from the specular wave at the fixed analysis surface (src/analysis/balint_kurti.cpp:74-82).
__global__ void bk_specular_kernel(const cx<double>* spec, int nr, int iz_sp, int n_v, int n_states, int initial,
                                   const double* phi_in, double dr, cx<double>* bk) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_v) return;
  cx<double> acc = {0.0, 0.0};
  for (int ir = 0; ir < nr; ++ir) acc = acc + scale(spec[ir + nr * iz_sp], phi_in[ir + nr * v]);
  bk[initial + (long)n_states * v] = scale(acc, dr);
}

// sums(k, ien) += c(k) exp(i E_ien t) dtAna, FP64 (src/analysis/s_matrix.cpp:49-56).
__global__ void s_accumulate_kernel(cx<double>* sums, const cx<double>* bk, int n, int n_e, const double* energies,
                                    double t, double dt_ana) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n_e) return;
  const int k = idx % n, ien = idx / n;
  double sn, cs;
  sincos(energies[ien] * t, &sn, &cs);
  const cx<double> ph = {cs, sn};
  sums[idx] = sums[idx] + scale(bk[k] * ph, dt_ana);
}

// ---------------------------------------------------------------------------------------------------------------
// Wave-packet utilities.
__global__ void zero_kernel(cx<double>* p, long n) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = cx<double>{0.0, 0.0};
}

template <typename T>
__global__ void densities_kernel(const cx<T>* psi, Shape s, double* out) {
  const int izr = blockIdx.x * blockDim.x + threadIdx.x;
  if (izr >= s.n_zr) return;
  double acc = 0.0;
  for (int ixy = 0; ixy < s.n_xy; ++ixy)
    for (int irot = 0; irot < s.n_rot; ++irot) {
      const cx<T> v = psi[izr + (long)s.n_zr * (irot + (long)s.n_rot * ixy)];
      const double a = hypot(double(v.x), double(v.y));
      acc = acc + a * a;
    }
  out[izr] = acc;
}

// out (irot, ixy, ir) of Z index iz.
template <typename T>
__global__ void z_plane_kernel(const cx<T>* psi, Shape s, int iz, cx<double>* out) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)s.n_rot * s.n_xy * s.nr;
  if (k >= n) return;
  const int irot = k % s.n_rot;
  const int ixy = (k / s.n_rot) % s.n_xy;
  const int ir = k / ((long)s.n_rot * s.n_xy);
  const cx<T> v = psi[ir + s.nr * iz + (long)s.n_zr * (irot + (long)s.n_rot * ixy)];
  out[k] = cx<double>{double(v.x), double(v.y)};
}

// Host layout chunk (irot, ixy, izr - izr0) <-> device layout, izr in [izr0, izr0 + count).
template <typename T>
__global__ void host_to_device_kernel(cx<T>* psi, Shape s, const cx<double>* chunk, int izr0, int count) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)s.n_rot * s.n_xy * count;
  if (k >= n) return;
  const int irot = k % s.n_rot;
  const int ixy = (k / s.n_rot) % s.n_xy;
  const int izr = izr0 + k / ((long)s.n_rot * s.n_xy);
  const cx<double> v = chunk[k];
  psi[izr + (long)s.n_zr * (irot + (long)s.n_rot * ixy)] = cx<T>{T(v.x), T(v.y)};
}

template <typename T>
__global__ void device_to_host_kernel(const cx<T>* psi, Shape s, cx<double>* chunk, int izr0, int count) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)s.n_rot * s.n_xy * count;
  if (k >= n) return;
  const int irot = k % s.n_rot;
  const int ixy = (k / s.n_rot) % s.n_xy;
  const int izr = izr0 + k / ((long)s.n_rot * s.n_xy);
  const cx<T> v = psi[izr + (long)s.n_zr * (irot + (long)s.n_rot * ixy)];
  chunk[k] = cx<double>{double(v.x), double(v.y)};
}

// Energy-resolved sums (src/analysis/energy_resolved.cpp:62-91): sums[site][ien][e] += (psi(site)[e] phase[ien])
// dtAna over the site's column e = izr + n_zr irot (the device layout of one ixy), in FP64.
template <typename T>
__global__ __launch_bounds__(256) void energy_resolved_kernel(const cx<T>* psi, long block, const int* site_ixy,
                                                              int n_energies, const cx<double>* phase, double dt,
                                                              cx<double>* sums) {
  const long e = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= block) return;
  const int site = blockIdx.y;
  const cx<T> p = psi[e + block * site_ixy[site]];
  const cx<double> v = {double(p.x), double(p.y)};
  cx<double>* out = sums + static_cast<long>(site) * n_energies * block + e;
  for (int ien = 0; ien < n_energies; ++ien) {
    cx<double>& acc = out[static_cast<long>(ien) * block];
    acc = acc + scale(v * phase[ien], dt);
  }
}

// The same for the specular wave: sums[ien][k] += (spec[k] phase[ien]) dtAna over its n values (ir, iz).
__global__ __launch_bounds__(256) void energy_resolved_specular_kernel(const cx<double>* spec, int n, int n_energies,
                                                                       const cx<double>* phase, double dt,
                                                                       cx<double>* sums) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const cx<double> v = spec[k];
  for (int ien = 0; ien < n_energies; ++ien) {
    cx<double>& acc = sums[static_cast<long>(ien) * n + k];
    acc = acc + scale(v * phase[ien], dt);
  }
}

}  // namespace spo_dev
