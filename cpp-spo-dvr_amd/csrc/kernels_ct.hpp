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

// T_Zr with register-resident Stockham passes: NR = A1 * A2 and NZ = B1 * B2, two radix passes per axis, each pass
// a set of whole butterflies per thread with one LDS exchange between passes. The first r pass reads its
// butterflies straight from HBM and the last inverse r pass writes straight to HBM; the last forward Z pass, the
// factor exp(-i (T_Z + T_r) dt) (FFT scales folded in) and the first inverse Z pass run on the same registers
// (the inverse runs its radices in reverse order, so its first pass takes exactly the elements the forward pass
// left in each thread). Seven passes, six LDS round trips, versus a tile in and out of LDS around every stage.
//
// Stockham positions for N = R1 * R2: pass 1 (radix R1) takes x[j + r R2] (j < R2) and writes j R1 + r; pass 2
// (radix R2) takes j + r R1 (j < R1), twiddles by W_N^{r j} and writes back to the same places, in natural order.
// Plane in LDS: row iz at iz * (NR + 1) (complex64: followed by the NR + NZ twiddles). The HBM passes (1 and 7) run lanes along j so that 16 lanes touch
// consecutive addresses; with A1 odd, pass 1's LDS writes (stride A1) also spread over the banks. Passes 2 and 6
// run lanes along iz (stride NR + 1, odd) and the Z passes along kr (stride 1).
template <int NBUT, int THREADS>
constexpr int per_thread() {
  return (NBUT + THREADS - 1) / THREADS;
}

// SKIP_ZERO: skip the (exact) multiplications by 1 of j == 0, worth a branch when j is uniform over a wavefront.
// Table entry k sits at table[k * TS]. In FP32 the table entries cannot all be of modulus 1 (no pair of floats near
// an eighth turn (h, h) is), so each comes with eps[k * ES] = 1 / |w_k| - 1 (a few 1e-8, from FP64) and the
// twiddle is applied as w + eps w in hi + lo form (mul_hilo): the twiddles then keep the norm to ~1e-15 instead of
// shrinking the modes they multiply by up to ~3e-8 at every transform. Their phase errors need no correction: the
// inverse transform applies the conjugates, which undo them. FP64 passes no eps.
template <typename T, int R, bool SKIP_ZERO = true, int TS = 1, int ES = 1>
__device__ __forceinline__ void twiddle(cx<T>* v, const cx<T>* table, int j, int sign, const T* eps = nullptr) {
  if (SKIP_ZERO && j == 0) return;
#pragma unroll
  for (int r = 1; r < R; ++r) {
    cx<T> w = table[r * j * TS];
    if (sign > 0) w = conj(w);
    if constexpr (sizeof(T) == 4) {
      const T e = eps[r * j * ES];
      v[r] = mul_hilo(w, {e * w.x, e * w.y}, v[r]);
    } else {
      v[r] = v[r] * w;
    }
  }
}

template <typename T>
constexpr bool zr_twiddles_in_lds() {
  return sizeof(T) == 4;
}

// LDS bytes of zr_reg_kernel (complex64: the twiddles and their remainders behind the plane).
template <typename T>
inline std::size_t zr_reg_lds(int nz, int nr) {
  return std::size_t(nz) * (nr + 1) * sizeof(cx<T>) +
         (zr_twiddles_in_lds<T>() ? std::size_t(nz + nr) * 2 * sizeof(cx<T>) : 0);
}

// twiddle() with the complex64 remainders already formed: lo[k * LS] = eps_k w_k (the part of the twiddle that
// restores its modulus), applied through mul_hilo. FP64 passes no lo.
template <typename T, int R, bool SKIP_ZERO = true, int TS = 1, int LS = 1>
__device__ __forceinline__ void twiddle_lo(cx<T>* v, const cx<T>* table, const cx<T>* lo, int j, int sign) {
  if (SKIP_ZERO && j == 0) return;
#pragma unroll
  for (int r = 1; r < R; ++r) {
    cx<T> w = table[r * j * TS];
    if constexpr (sizeof(T) == 4) {
      cx<T> l = lo[r * j * LS];
      if (sign > 0) {
        w = conj(w);
        l = conj(l);
      }
      v[r] = mul_hilo(w, l, v[r]);
    } else {
      if (sign > 0) w = conj(w);
      v[r] = v[r] * w;
    }
  }
}

// eps_z, eps_r: the complex64 twiddle tables' modulus corrections (twiddle()); factor_lo: the factor's FP32
// remainders (mul_hilo). Unused in FP64.
// row >= 0: psi is only read, and Z row `row` of each transformed plane (NR values) goes to row_out[plane * NR + ir]
// (the inverse r passes run for that row alone); row < 0: psi is transformed in place and row_out is unused.
template <typename T, int NZ, int NR, int A1, int A2, int B1, int B2, int THREADS>
__global__ __launch_bounds__(THREADS) void zr_reg_kernel(cx<T>* psi, const cx<T>* table_z, const cx<T>* table_r,
                                                         const cx<T>* factor, const T* eps_z, const T* eps_r,
                                                         const cx<T>* factor_lo, cx<T>* row_out, int row) {
  static_assert(A1 * A2 == NR && B1 * B2 == NZ, "zr_reg_kernel: radices must multiply to the lengths");
  extern __shared__ unsigned char smem_raw[];
  cx<T>* t = reinterpret_cast<cx<T>*>(smem_raw);
  constexpr int P = NZ * NR;
  constexpr int LD = NR + 1;
  cx<T>* plane = psi + static_cast<long>(blockIdx.x) * P;
  const int tid = threadIdx.x;
  // complex64: the twiddle tables live in LDS behind the plane (lower latency than the cache for the in-pass
  // twiddles); FP64 keeps them in the cache, where they do not compete with the plane for LDS bandwidth (measured
  // on MI355X: 1.89 -> 1.73 ms and 3.85 -> 4.15 ms per time step respectively).
  // complex64: each twiddle's remainder eps w is formed once here, so that the passes spend no products on it.
  const cx<T>* lo_r = nullptr;
  const cx<T>* lo_z = nullptr;
  if constexpr (zr_twiddles_in_lds<T>()) {
    cx<T>* tw_r = t + NZ * LD;
    cx<T>* tw_z = tw_r + NR;
    cx<T>* l_r = tw_z + NZ;
    cx<T>* l_z = l_r + NR;
    for (int e = tid; e < NR; e += THREADS) {
      const cx<T> w = table_r[e];
      const T k = eps_r[e];
      tw_r[e] = w;
      l_r[e] = {k * w.x, k * w.y};
    }
    for (int e = tid; e < NZ; e += THREADS) {
      const cx<T> w = table_z[e];
      const T k = eps_z[e];
      tw_z[e] = w;
      l_z[e] = {k * w.x, k * w.y};
    }
    table_r = tw_r;
    table_z = tw_z;
    lo_r = l_r;
    lo_z = l_z;
  }

  // Pass 1, r forward, radix A1: butterfly (iz, j < A2) from HBM.
  {
    constexpr int N = NZ * A2, Q = per_thread<N, THREADS>();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int j = b % A2, iz = b / A2;
      cx<T> v[A1];
#pragma unroll
      for (int r = 0; r < A1; ++r) v[r] = plane[iz * NR + j + r * A2];
      rdft<A1, T>(v, -1);
#pragma unroll
      for (int r = 0; r < A1; ++r) t[iz * LD + j * A1 + r] = v[r];
    }
  }
  __syncthreads();
  // Pass 2, r forward, radix A2, in place: butterfly (iz, j < A1), lanes along iz (row stride NR + 1).
  {
    constexpr int N = NZ * A1, Q = per_thread<N, THREADS>();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int iz = b % NZ, j = b / NZ;
      cx<T> v[A2];
#pragma unroll
      for (int r = 0; r < A2; ++r) v[r] = t[iz * LD + j + r * A1];
      twiddle_lo<T, A2>(v, table_r, lo_r, j, -1);
      rdft<A2, T>(v, -1);
#pragma unroll
      for (int r = 0; r < A2; ++r) t[iz * LD + j + r * A1] = v[r];
    }
  }
  __syncthreads();
  // Pass 3, Z forward, radix B1: butterfly (kr, j < B2); reads all, then writes all.
  {
    constexpr int N = NR * B2, Q = per_thread<N, THREADS>();
    cx<T> v[Q][B1];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int kr = b % NR, j = b / NR;
#pragma unroll
      for (int r = 0; r < B1; ++r) v[q][r] = t[(j + r * B2) * LD + kr];
      rdft<B1, T>(v[q], -1);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int kr = b % NR, j = b / NR;
#pragma unroll
      for (int r = 0; r < B1; ++r) t[(j * B1 + r) * LD + kr] = v[q][r];
    }
  }
  __syncthreads();
  // Pass 4, Z forward radix B2 (twiddled), x factor, Z inverse radix B2: butterfly (kr, j < B1).
  {
    constexpr int N = NR * B1, Q = per_thread<N, THREADS>();
    cx<T> v[Q][B2];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int kr = b % NR, j = b / NR;
#pragma unroll
      for (int r = 0; r < B2; ++r) v[q][r] = t[(j + r * B1) * LD + kr];
      twiddle_lo<T, B2>(v[q], table_z, lo_z, j, -1);
      rdft<B2, T>(v[q], -1);
#pragma unroll
      for (int r = 0; r < B2; ++r) {
        const int f = (j + r * B1) * NR + kr;
        v[q][r] = mul_factor(factor[f], v[q][r], factor_lo, f);
      }
      rdft<B2, T>(v[q], +1);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int kr = b % NR, j = b / NR;
#pragma unroll
      for (int r = 0; r < B2; ++r) t[(j * B2 + r) * LD + kr] = v[q][r];
    }
  }
  __syncthreads();
  // Pass 5, Z inverse, radix B1, in place: butterfly (kr, j < B2).
  {
    constexpr int N = NR * B2, Q = per_thread<N, THREADS>();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int kr = b % NR, j = b / NR;
      cx<T> v[B1];
#pragma unroll
      for (int r = 0; r < B1; ++r) v[r] = t[(j + r * B2) * LD + kr];
      twiddle_lo<T, B1>(v, table_z, lo_z, j, +1);
      rdft<B1, T>(v, +1);
#pragma unroll
      for (int r = 0; r < B1; ++r) t[(j + r * B2) * LD + kr] = v[r];
    }
  }
  __syncthreads();
  // Pass 6, r inverse, radix A2: butterfly (iz, j < A1), lanes along iz; reads all, then writes all.
  {
    constexpr int N = NZ * A1, Q = per_thread<N, THREADS>();
    cx<T> v[Q][A2];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int iz = b % NZ, j = b / NZ;
      if (row >= 0 && iz != row) continue;
#pragma unroll
      for (int r = 0; r < A2; ++r) v[q][r] = t[iz * LD + j + r * A1];
      rdft<A2, T>(v[q], +1);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int iz = b % NZ, j = b / NZ;
      if (row >= 0 && iz != row) continue;
#pragma unroll
      for (int r = 0; r < A2; ++r) t[iz * LD + j * A2 + r] = v[q][r];
    }
  }
  __syncthreads();
  // Pass 7, r inverse, radix A1 (twiddled): butterfly (iz, j < A2) to HBM.
  {
    constexpr int N = NZ * A2, Q = per_thread<N, THREADS>();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      if (N % THREADS != 0 && b >= N) break;
      const int j = b % A2, iz = b / A2;
      if (row >= 0 && iz != row) continue;
      cx<T> v[A1];
#pragma unroll
      for (int r = 0; r < A1; ++r) v[r] = t[iz * LD + j + r * A2];
      twiddle_lo<T, A1>(v, table_r, lo_r, j, +1);
      rdft<A1, T>(v, +1);
      if (row >= 0) {
        cx<T>* out = row_out + static_cast<long>(blockIdx.x) * NR;
#pragma unroll
        for (int r = 0; r < A1; ++r) out[j + r * A2] = v[r];
      } else {
#pragma unroll
        for (int r = 0; r < A1; ++r) plane[iz * NR + j + r * A2] = v[r];
      }
    }
  }
}

// T_xy (+ the diagonal part of V) for grids whose NX and NY are single radices (whole-line DFTs in registers):
// C consecutive izr of one irot, C * NX threads. Thread (c, ix) loads its y line of NY values straight from HBM,
// transforms it and parks it in LDS; thread (c, iy) then takes an x line, transforms it forward, multiplies the
// T_xy factor, transforms it back and parks it again; thread (c, ix) finishes its y line and stores it to HBM with
// the shift and the potential factors applied. Two LDS round trips, two barriers. LDS element (c, iy, ix) at
// ix * XS + iy * (C + 1) + c, XS = NY (C + 1) + 1.
// SPLIT: each round trip goes through LDS in two halves of the y lines (rows iy < NY / 2, then the rest; four
// barriers each), so that the tile takes half the LDS and twice as many workgroups fit on a CU. At least three
// wavefronts per SIMD are asked for (FP64 at 16 x 16 needs 169 VGPRs otherwise, two wavefronts per SIMD).
template <int N>
constexpr bool single_radix() {
  return is_radix(N);
}

// LDS bytes of xy_reg_kernel.
template <typename T>
constexpr std::size_t xy_reg_lds(int nx, int ny, int c, bool split) {
  return std::size_t(nx) * ((split ? ny / 2 : ny) * (c + 1) + 1) * sizeof(cx<T>);
}

template <typename T, int NX, int NY, int C, bool SPLIT = false>
__global__ __launch_bounds__(C * NX) __attribute__((amdgpu_waves_per_eu(3))) void xy_reg_kernel(cx<T>* psi, Shape s, const cx<T>* factor,
                                                         const cx<T>* shift_f, const cx<T>* shift_b, int needs_shift,
                                                         int apply_potential, const T* z_opt, const cx<T>* before,
                                                         const cx<T>* gas_step, XyLo<T> lo) {
  static_assert(single_radix<NX>() && single_radix<NY>() && NX == NY, "xy_reg_kernel: one radix per axis, NX = NY");
  static_assert(!SPLIT || NY % 2 == 0, "xy_reg_kernel: halves of the y lines");
  extern __shared__ unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  constexpr int YS = C + 1;
  constexpr int NH = SPLIT ? NY / 2 : NY;  // y rows held in LDS at once
  constexpr int XS = NH * YS + 1;
  const int tiles_per_rot = (s.n_zr + C - 1) / C;
  const int irot = blockIdx.x / tiles_per_rot;
  const int izr0 = (blockIdx.x % tiles_per_rot) * C;
  const int cw = min(C, s.n_zr - izr0);
  const long rot_off = static_cast<long>(s.n_zr) * irot;
  const long xy_stride = static_cast<long>(s.n_zr) * s.n_rot;
  const int c = threadIdx.x % C;
  const int a = threadIdx.x / C;  // ix in the y phases, iy in the x phase
  const bool live = c < cw;
  cx<T>* g = psi + izr0 + c + rot_off;

  cx<T> v[NY];
#pragma unroll
  for (int iy = 0; iy < NY; ++iy) {
    const int ixy = a * NY + iy;
    v[iy] = live ? g[xy_stride * ixy] : cx<T>{T(0), T(0)};
    if (needs_shift) v[iy] = mul_factor(shift_f[ixy], v[iy], lo.shift_f, ixy);
  }
  rdft<NY, T>(v, -1);
  cx<T> w[NX];
  if constexpr (!SPLIT) {
#pragma unroll
    for (int iy = 0; iy < NY; ++iy) tile[a * XS + iy * YS + c] = v[iy];
    __syncthreads();
#pragma unroll
    for (int ix = 0; ix < NX; ++ix) w[ix] = tile[ix * XS + a * YS + c];
  } else {
    // Half h holds the rows iy = h NH .. h NH + NH - 1; the thread of row a = iy reads in its half's round.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h) __syncthreads();  // the previous round's reads are done
#pragma unroll
      for (int iy = 0; iy < NH; ++iy) tile[a * XS + iy * YS + c] = v[h * NH + iy];
      __syncthreads();
      if (a / NH == h) {
#pragma unroll
        for (int ix = 0; ix < NX; ++ix) w[ix] = tile[ix * XS + (a - h * NH) * YS + c];
      }
    }
  }
  rdft<NX, T>(w, -1);
#pragma unroll
  for (int ix = 0; ix < NX; ++ix) w[ix] = mul_factor(factor[ix * NY + a], w[ix], lo.factor, ix * NY + a);
  rdft<NX, T>(w, +1);
  if constexpr (!SPLIT) {
#pragma unroll
    for (int ix = 0; ix < NX; ++ix) tile[ix * XS + a * YS + c] = w[ix];
    __syncthreads();
#pragma unroll
    for (int iy = 0; iy < NY; ++iy) v[iy] = tile[a * XS + iy * YS + c];
  } else {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      __syncthreads();  // every thread is done with the LDS of the previous round
      if (a / NH == h) {
#pragma unroll
        for (int ix = 0; ix < NX; ++ix) tile[ix * XS + (a - h * NH) * YS + c] = w[ix];
      }
      __syncthreads();
#pragma unroll
      for (int iy = 0; iy < NH; ++iy) v[h * NH + iy] = tile[a * XS + iy * YS + c];
    }
  }
  rdft<NY, T>(v, +1);
  if (!live) return;
  const int izr = izr0 + c;
  const int iz = izr / s.nr, ir = izr % s.nr;
  cx<T> f = {T(1), T(0)};
  cx<T> f_lo = {T(0), T(0)};  // the FP32 remainder of f (complex64)
  T zo = T(1);
  if (apply_potential) {
    const int k = irot + s.n_rot * ir;
    f = iz >= s.nzpot ? gas_step[k] : before[k];
    if constexpr (sizeof(T) == 4) f_lo = iz >= s.nzpot ? lo.gas_step[k] : lo.before[k];
    zo = z_opt[iz];
  }
#pragma unroll
  for (int iy = 0; iy < NY; ++iy) {
    const int ixy = a * NY + iy;
    cx<T> u = v[iy];
    if (needs_shift) u = mul_factor(shift_b[ixy], u, lo.shift_b, ixy);
    if (apply_potential) {
      u = mul_hilo(f, f_lo, u);
      u = scale(u, zo);
    }
    g[xy_stride * ixy] = u;
  }
}

template <int NX, int NY, int C>
constexpr int xy_reg_lds_elements() {
  return NX * (NY * (C + 1) + 1);
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
                                                                 int legendre_size, const cx<T>* rot_lo,
                                                                 const T* eps_phi) {
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
  ct_fft<T, NPHI, THREADS, Rows>(dvr, C * n_theta, table_phi, eps_phi, +1);

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
  ct_fft<T, NPHI, THREADS, Rows>(dvr, C * n_theta, table_phi, eps_phi, -1);

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
      psi[izr + col + static_cast<long>(s.n_zr) * irot] = mul_factor(rot_half[irot + s.n_rot * ir], acc, rot_lo,
                                                                      irot + s.n_rot * ir);
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
