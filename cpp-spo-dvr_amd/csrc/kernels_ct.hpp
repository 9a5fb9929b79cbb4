// Hot-path kernels with compile-time FFT shapes (fft_ct.hpp), instantiated for the grid shapes the backend is
// tuned for; any other shape runs the run-time-shaped kernels of kernels.hpp. Same device layout:
// psi[izr + n_zr * (irot + n_rot * ixy)], izr = ir + nr * iz.
//
// LDS tiles are padded so that the strides between the sequences a wave touches at once are odd multiples of one
// complex value: strided 16-byte LDS accesses then spread over the banks instead of piling onto one.
#pragma once

#include "fft_ct.hpp"
#include "kernels.hpp"

namespace spo_dev {

// T_Zr: one (irot, ixy) plane of NZ x NR points per workgroup, in LDS as rows of NR + 1: r FFTs (rows), Z FFTs
// (columns), x exp(-i (T_Z + T_r) dt) with the FFT scales folded in, back. One read and one write of psi.
template <int NZ, int NR>
constexpr int zr_lds_elements() {
  return NZ * (NR + 1);
}

template <typename T, int NZ, int NR, int THREADS>
__global__ __launch_bounds__(THREADS) void zr_ct_kernel(cx<T>* psi, const cx<T>* table_z, const cx<T>* table_r,
                                                        const cx<T>* factor) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  constexpr int P = NZ * NR;
  constexpr int LD = NR + 1;
  cx<T>* plane = psi + static_cast<long>(blockIdx.x) * P;
#pragma unroll 4
  for (int e = threadIdx.x; e < P; e += THREADS) tile[(e / NR) * LD + e % NR] = plane[e];
  __syncthreads();
  using Rows = CtLayout<0, LD, 0, 1>;  // one r sequence per iz
  using Cols = CtLayout<0, 1, 0, LD>;  // one Z sequence per ir, stride NR + 1
  ct_fft<T, NR, THREADS, Rows>(tile, NZ, table_r, -1);
  ct_fft<T, NZ, THREADS, Cols>(tile, NR, table_z, -1);
  for (int e = threadIdx.x; e < P; e += THREADS) {
    cx<T>& v = tile[(e / NR) * LD + e % NR];
    v = factor[e] * v;
  }
  __syncthreads();
  ct_fft<T, NZ, THREADS, Cols>(tile, NR, table_z, +1);
  ct_fft<T, NR, THREADS, Rows>(tile, NZ, table_r, +1);
#pragma unroll 4
  for (int e = threadIdx.x; e < P; e += THREADS) plane[e] = tile[(e / NR) * LD + e % NR];
}

// T_xy (+ the diagonal part of V): C consecutive izr of one irot, all NX x NY points. LDS element (c, iy, ix) at
// ix * XS + iy * (C + 1) + c with XS = NY (C + 1) + 1 (both strides odd for even C).
template <int NX, int NY, int C>
constexpr int xy_lds_elements() {
  return NX * (NY * (C + 1) + 1);
}

template <typename T, int NX, int NY, int C, int THREADS>
__global__ __launch_bounds__(THREADS) void xy_ct_kernel(cx<T>* psi, Shape s, const cx<T>* table_x,
                                                        const cx<T>* table_y, const cx<T>* factor,
                                                        const cx<T>* shift_f, const cx<T>* shift_b, int needs_shift,
                                                        int apply_potential, const T* z_opt, const cx<T>* before,
                                                        const cx<T>* gas_step) {
  extern __shared__ unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  constexpr int NXY = NX * NY;
  constexpr int TOTAL = C * NXY;
  constexpr int YS = C + 1;
  constexpr int XS = NY * YS + 1;
  const int tiles_per_rot = (s.n_zr + C - 1) / C;
  const int irot = blockIdx.x / tiles_per_rot;
  const int izr0 = (blockIdx.x % tiles_per_rot) * C;
  const int cw = min(C, s.n_zr - izr0);
  const long rot_off = static_cast<long>(s.n_zr) * irot;
  const long xy_stride = static_cast<long>(s.n_zr) * s.n_rot;
  for (int e = threadIdx.x; e < TOTAL; e += THREADS) {
    const int c = e % C, ixy = e / C;
    cx<T> v = {T(0), T(0)};
    if (c < cw) v = psi[izr0 + c + rot_off + xy_stride * ixy];
    if (needs_shift) v = v * shift_f[ixy];
    tile[(ixy / NY) * XS + (ixy % NY) * YS + c] = v;
  }
  __syncthreads();
  using Ys = CtLayout<C, 1, XS, YS>;  // for each (c, ix): iy at stride C + 1
  using Xs = CtLayout<C, 1, YS, XS>;  // for each (c, iy): ix at stride XS
  ct_fft<T, NY, THREADS, Ys>(tile, C * NX, table_y, -1);
  ct_fft<T, NX, THREADS, Xs>(tile, C * NY, table_x, -1);
  for (int e = threadIdx.x; e < TOTAL; e += THREADS) {
    const int c = e % C, ixy = e / C;
    cx<T>& v = tile[(ixy / NY) * XS + (ixy % NY) * YS + c];
    v = factor[ixy] * v;
  }
  __syncthreads();
  ct_fft<T, NX, THREADS, Xs>(tile, C * NY, table_x, +1);
  ct_fft<T, NY, THREADS, Ys>(tile, C * NX, table_y, +1);
  for (int e = threadIdx.x; e < TOTAL; e += THREADS) {
    const int c = e % C, ixy = e / C;
    if (c >= cw) continue;
    cx<T> v = tile[(ixy / NY) * XS + (ixy % NY) * YS + c];
    if (needs_shift) v = v * shift_b[ixy];
    const int izr = izr0 + c;
    if (apply_potential) {
      const int iz = izr / s.nr, ir = izr % s.nr;
      const cx<T> f = iz >= s.nzpot ? gas_step[irot + s.n_rot * ir] : before[irot + s.n_rot * ir];
      v = scale(f * v, z_opt[iz]);
    }
    psi[izr + rot_off + xy_stride * ixy] = v;
  }
}

// V in the interaction region for C consecutive izr of one (x, y) column, phi FFTs of length NPHI.
//   - Workgroups run the (x, y) columns of one izr tile back to back (blockIdx = tile * n_xy + ixy), so the
//     exp(i A) rows of the tile's representatives stay in L2 while every column that maps onto them is processed.
//   - One thread per (c, mJ block) for both Legendre sums: it loads its block's FBR values of column c straight
//     from HBM into registers (C lanes read C consecutive izr), forms all theta rows, and after the phi FFTs forms
//     its block's FBR values again and stores them (rot(dt/2) applied). Only the DVR tile and the Legendre matrix
//     live in LDS.
//   - DVR (c, theta, phi) at (c * n_theta + theta) * (NPHI + 1) + phi.
constexpr int kMaxTheta = 16;

template <typename T, int NPHI, int THREADS>
__global__ __launch_bounds__(THREADS) void interaction_ct_kernel(cx<T>* psi, const cx<double>* spec, Shape s, int C,
                                                                 const cx<T>* table_phi, AngularTables at,
                                                                 const cx<T>* exp_table, const cx<T>* rot_half,
                                                                 const cx<T>* s2sc, const T* legendre,
                                                                 int legendre_size) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int LD = NPHI + 1;
  const int n_theta = s.n_theta;
  T* lmat = reinterpret_cast<T*>(smem_raw);
  int* phi_map = reinterpret_cast<int*>(lmat + ((legendre_size + 3) / 4) * 4);
  cx<T>* dvr = reinterpret_cast<cx<T>*>(phi_map + ((NPHI + 3) / 4) * 4);
  const int tile = blockIdx.x / s.n_xy;
  const int ixy = blockIdx.x % s.n_xy;
  const int izr0 = tile * C;
  const int cw = min(C, s.n_zr_pot - izr0);
  const long col = static_cast<long>(s.n_zr) * s.n_rot * ixy;

  for (int e = threadIdx.x; e < legendre_size; e += THREADS) lmat[e] = legendre[e];
  const int op = at.op_of[ixy];
  for (int e = threadIdx.x; e < NPHI; e += THREADS) phi_map[e] = at.phi_index[e + NPHI * op];
  // Slots of |mJ| > MJMax carry no basis function.
  const int n_unused = NPHI - s.n_blocks;
  for (int e = threadIdx.x; e < C * n_theta * n_unused; e += THREADS) {
    const int row = e / n_unused;
    dvr[row * LD + s.mj_max + 1 + e % n_unused] = cx<T>{T(0), T(0)};
  }
  __syncthreads();

  // FBR -> DVR, theta part.
  const int n_tasks = C * s.n_blocks;
  for (int tt = threadIdx.x; tt < n_tasks; tt += THREADS) {
    const int cc = tt % C, bb = tt / C;
    const int f0 = at.block_first[bb], nb = at.block_size[bb];
    const T* L = lmat + at.block_offset[bb];
    const int m = bb - s.mj_max;
    const int sl = m >= 0 ? m : NPHI + m;
    cx<T> acc[kMaxTheta];
#pragma unroll
    for (int th = 0; th < kMaxTheta; ++th) acc[th] = cx<T>{T(0), T(0)};
    if (cc < cw) {
      const int izr = izr0 + cc;
      for (int j = 0; j < nb; ++j) {
        const int irot = f0 + j;
        cx<T> f = psi[izr + col + static_cast<long>(s.n_zr) * irot];
        if (irot == s.ichn) {
          const int ir = izr % s.nr, iz = izr / s.nr;
          const cx<double> in = spec[ir + s.nr * iz];
          f = f + rot_half[s.ichn + s.n_rot * ir] * (s2sc[ixy] * cx<T>{T(in.x), T(in.y)});
        }
#pragma unroll
        for (int th = 0; th < kMaxTheta; ++th)
          if (th < n_theta) acc[th] = acc[th] + scale(f, L[th * nb + j]);
      }
    }
#pragma unroll
    for (int th = 0; th < kMaxTheta; ++th)
      if (th < n_theta) dvr[(cc * n_theta + th) * LD + sl] = acc[th];
  }
  __syncthreads();
  using Rows = CtLayout<0, LD, 0, 1>;
  ct_fft<T, NPHI, THREADS, Rows>(dvr, C * n_theta, table_phi, +1);

  const int rep = at.rep_of[ixy];
  const int n_tp = n_theta * NPHI;
  const int dvr_count = C * n_theta * NPHI;
  for (int e = threadIdx.x; e < dvr_count; e += THREADS) {
    const int ip = e % NPHI;
    const int row = e / NPHI;  // c * n_theta + theta
    const int th = row % n_theta;
    const int cc = row / n_theta;
    if (cc >= cw) continue;
    const int itp = phi_map[ip] + NPHI * th;
    cx<T>& v = dvr[row * LD + ip];
    v = exp_table[itp + static_cast<long>(n_tp) * (rep + static_cast<long>(s.n_rep) * (izr0 + cc))] * v;
  }
  __syncthreads();
  ct_fft<T, NPHI, THREADS, Rows>(dvr, C * n_theta, table_phi, -1);

  // DVR -> FBR, then rot(dt/2), stored straight to psi.
  for (int tt = threadIdx.x; tt < n_tasks; tt += THREADS) {
    const int cc = tt % C, bb = tt / C;
    if (cc >= cw) continue;
    const int f0 = at.block_first[bb], nb = at.block_size[bb];
    const T* L = lmat + at.block_offset[bb];
    const int m = bb - s.mj_max;
    const int sl = m >= 0 ? m : NPHI + m;
    cx<T> d[kMaxTheta];
#pragma unroll
    for (int th = 0; th < kMaxTheta; ++th)
      d[th] = th < n_theta ? dvr[(cc * n_theta + th) * LD + sl] : cx<T>{T(0), T(0)};
    const int izr = izr0 + cc;
    const int ir = izr % s.nr;
    for (int j = 0; j < nb; ++j) {
      cx<T> acc = {T(0), T(0)};
#pragma unroll
      for (int th = 0; th < kMaxTheta; ++th)
        if (th < n_theta) acc = acc + scale(d[th], L[th * nb + j]);
      const int irot = f0 + j;
      psi[izr + col + static_cast<long>(s.n_zr) * irot] = rot_half[irot + s.n_rot * ir] * acc;
    }
  }
}

// LDS bytes of the interaction kernel for tile width C.
template <typename T>
inline std::size_t interaction_ct_lds(int C, int n_theta, int nphi, int legendre_size) {
  return ((legendre_size + 3) / 4) * 4 * sizeof(T) + ((nphi + 3) / 4) * 4 * sizeof(int) +
         std::size_t(C) * n_theta * (nphi + 1) * sizeof(cx<T>);
}

// Block densities in two deterministic passes: partial sums over a chunk of (x, y) points per izr, then the
// chunks in order (the reference sums a block in memory order, src/propagation/averages.cpp:19-30).
template <typename T>
__global__ void densities_partial_kernel(const cx<T>* psi, Shape s, int chunk, double* partial) {
  const int izr = blockIdx.x * blockDim.x + threadIdx.x;
  const int part = blockIdx.y;
  if (izr >= s.n_zr) return;
  const int x0 = part * chunk, x1 = min(s.n_xy, x0 + chunk);
  double acc = 0.0;
  for (int ixy = x0; ixy < x1; ++ixy)
    for (int irot = 0; irot < s.n_rot; ++irot) {
      const cx<T> v = psi[izr + static_cast<long>(s.n_zr) * (irot + static_cast<long>(s.n_rot) * ixy)];
      const double a = hypot(double(v.x), double(v.y));
      acc = acc + a * a;
    }
  partial[izr + static_cast<long>(s.n_zr) * part] = acc;
}

__global__ void densities_sum_kernel(const double* partial, int n_zr, int parts, double* out) {
  const int izr = blockIdx.x * blockDim.x + threadIdx.x;
  if (izr >= n_zr) return;
  double acc = 0.0;
  for (int p = 0; p < parts; ++p) acc = acc + partial[izr + static_cast<long>(n_zr) * p];
  out[izr] = acc;
}

}  // namespace spo_dev
