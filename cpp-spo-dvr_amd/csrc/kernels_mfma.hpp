// The interaction-region potential step with the Gauss-Legendre theta contractions on the matrix cores.
//
// Per workgroup: C = 8 NT consecutive izr (< nz_potential * nr) of one (x, y) site, WAVES wavefronts.
//   1. FBR -> DVR, theta part: for every mJ block m, D[theta, col] = sum_J L_m[theta, J] c[J, col] as 16 x 16 x 4
//      MFMA products. The B operand (4 J x 16 real columns: re and im of 8 complex columns) is loaded straight from
//      HBM, 16 lanes reading 8 consecutive izr of one irot (128 B in FP64), all of a wavefront's B values before
//      its first product; A is a pre-arranged fragment of the Legendre matrix, one value per lane. D goes to the
//      DVR tile in LDS at phi slot m (mod n_phi).
//   2. phi FFTs (sign +1) over every (theta, col) row, x exp(i A[rep(ixy), theta, phi_map(op(ixy), phi)]), phi
//      FFTs back (sign -1): three register passes (phi_step_fused), or staged LDS FFTs.
//   3. DVR -> FBR: C'[J, col] = sum_theta L_m[theta, J] D'[theta, col] (A = L_m transposed, B read from the tile),
//      x rot(dt/2) (read from a copy of the table with ir fastest, so that the 8 columns of a lane group take
//      consecutive factors), stored straight to HBM.
// The initial channel's specular coupling term rot(dt/2) s2sc(ixy) chi(Z, r) is added to its FBR value as it is
// loaded. Workgroups run tile-major over a site order that groups the sites of one symmetry representative on one
// XCD (blockIdx mod 8), so the exp(i A) rows of a representative are fetched into one L2.
//
// DVR tile: row theta * C + col, phi along the row. In FP64 with the fused phi step and n_phi a multiple of 16 the
// rows are n_phi long and element phi of row w sits at column phi ^ (w mod 16): the 16 rows a wavefront touches at
// one phi (phases 1 and 3), or at one column (the first and last phi passes, lanes along rows), land on 16
// distinct bank groups without
// padding, which keeps the FP64 tile of 8 x 13 x 48 values at two workgroups per CU. Otherwise rows are padded to
// n_phi + 1 (an odd stride spreads the rows over the banks just as well); with the fused phi step the padding
// column, never part of a row's data, holds the phi twiddles (rows 0 .. n_phi - 1) and the phi map (rows n_phi ..
// 2 n_phi - 1, the slot's first int), the tile having at least 2 n_phi rows: the complex64 tile of 8 x 13 x 48
// values then takes 40768 bytes in all, four workgroups per CU.
//
// MFMA 16x16x4 operand layouts (lane l of 64): A[i][k] = lane i + 16 k, B[k][j] = lane j + 16 k; result
// element v of lane l is D[l / 16 + 4 v][l % 16] for f64 and D[4 (l / 16) + v][l % 16] for f32.
#pragma once

#include <type_traits>

#include "fft_ct.hpp"
#include "kernels.hpp"
#include "kernels_ct.hpp"

namespace spo_dev {

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t fma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return lane / 16 + 4 * v; }
};

template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t fma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane / 16) + v; }
};

// The value of the neighbouring lane (lane ^ 1) through a DPP quad permutation [1, 0, 3, 2] (a VALU operand
// modifier, no LDS traffic); the real and imaginary parts of a column sit in adjacent lanes.
__device__ __forceinline__ float swap_adjacent_lanes(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double swap_adjacent_lanes(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(bits), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(bits >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Pre-arranged Legendre fragments (built on the host), 64 values per fragment in lane order:
//   forward  [fwd_offset[b] + tt * k_fwd(b) + k]: lane (i, kk) holds L_m[theta = 16 tt + i][J = 4 k + kk]
//   backward [bwd_offset[b] + jt * k_bwd + k]:    lane (i, kk) holds L_m[theta = 4 k + kk][J = 16 jt + i]
// with i = lane % 16, kk = lane / 16, zero outside the block; k_fwd(b) = ceil(nb / 4), k_bwd = ceil(n_theta / 4).
struct MfmaTables {
  const int* fwd_offset;
  const int* bwd_offset;
  const int* site_order;  // position -> ixy, the sites of one symmetry representative at positions of one XCD
  int pair_tiles;         // see block_tile_site
};

// The (izr tile, site) of a workgroup. Without pairing: tile-major over the site order (blockIdx = tile * n_xy +
// position). With pairing (n_xy a multiple of 8): the workgroups of tiles 2k and 2k + 1 of a site are 8 apart in
// blockIdx, so they run on the same XCD at about the same time and their half-line (64-byte) complex64 rows
// combine in its L2 before they reach HBM.
__device__ __forceinline__ void block_tile_site(int b, int n_xy, const MfmaTables& mt, int& tile, int& ixy) {
  if (mt.pair_tiles) {
    const int x = b % 8, q = b / 8, per = n_xy / 8;
    const int rest = q / 2;
    tile = 2 * (rest / per) + q % 2;
    ixy = mt.site_order[8 * (rest % per) + x];
  } else {
    tile = b / n_xy;
    ixy = mt.site_order[b % n_xy];
  }
}

// Position of DVR element (row, phi) in the tile.
template <int NPHI, bool SWZ>
struct DvrTile {
  static constexpr int ld = SWZ ? NPHI : NPHI + 1;
  static __device__ __forceinline__ int at(int row, int phi) { return row * ld + (SWZ ? (phi ^ (row & 15)) : phi); }
};

// FP64 tiles are swizzled (unpadded, to fit two workgroups per CU); complex64 tiles are padded, so that the fused
// phi passes address their elements as a per-thread base plus compile-time offsets.
template <typename T, int NPHI, int R1>
constexpr bool dvr_swizzled() {
  return sizeof(T) == 8 && R1 > 0 && NPHI % 16 == 0;
}

// The phi step in three register passes over the DVR rows (NPHI = R1 R2; rows = C n_theta <= C 16): radix R1
// forward; radix R2 forward, x exp(i A) in natural phi order, radix R2 backward on the same registers; radix R1
// backward. Butterfly b: row b % (16 C), lanes along rows. Sign +1 towards the DVR, -1 back; W = exp(-2 pi i m/NPHI).
template <int NPHI, int R1, int C, int THREADS>
constexpr int phi_exp_per_thread() {
  return per_thread<16 * C * R1, THREADS>();
}

// The exp(i A) values a thread multiplies in the second pass of phi_step_fused (R2 per butterfly, natural phi
// order), from the exp rows of the tile's first izr (exp_rows, wavefront-uniform; row stride exp_stride per izr).
// Rows past the tile (row >= rows, or a column cc >= cw past the interaction region) read the last valid row
// instead: their products are never stored, and every load stays inside the table without a branch. The offsets
// are 32-bit byte offsets from the uniform base.
// phi map entry k sits at phi_map[k * PS].
template <typename T, int NPHI, int R1, int R2, int C, int THREADS, int PS, int Q>
__device__ __forceinline__ void phi_exp_load(cx<T> (&ev)[Q][R2], int rows, int cw, const int* phi_map,
                                             const cx<T>* exp_rows, int exp_stride) {
  static_assert(Q == phi_exp_per_thread<NPHI, R1, C, THREADS>(), "exp values per thread");
  const unsigned char* base = reinterpret_cast<const unsigned char*>(exp_rows);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const unsigned b = threadIdx.x + q * THREADS;
    const unsigned row = min(b / R1, static_cast<unsigned>(rows - 1)), j = b % R1;
    const int th = row / C, cc = min(static_cast<int>(row % C), cw - 1);
    const unsigned off = static_cast<unsigned>(exp_stride * cc + NPHI * th);
#pragma unroll
    for (int r = 0; r < R2; ++r)
      ev[q][r] = *reinterpret_cast<const cx<T>*>(base + (off + static_cast<unsigned>(phi_map[(j + r * R1) * PS])) *
                                                            static_cast<unsigned>(sizeof(cx<T>)));
  }
}

// ev holds the exp(i A) values of phi_exp_load once load_ev() has run; load_ev is called at the start of the second
// pass (a no-op when the caller loaded them earlier). complex64: eps[k * ES] is twiddle k's modulus correction
// (twiddle(), kernels_ct.hpp); FP64 passes none.
template <typename T, int NPHI, int R1, int R2, int C, int THREADS, class Tile, int TS, int ES, int QE, class LoadEv>
__device__ __forceinline__ void phi_step_fused(cx<T>* dvr, int rows, const cx<T>* table, const T* eps,
                                               cx<T> (&ev)[QE][R2], LoadEv load_ev) {
  static_assert(R1 * R2 == NPHI, "phi radices");
  static_assert(QE == phi_exp_per_thread<NPHI, R1, C, THREADS>(), "exp values per thread");
  constexpr int ROWS_MAX = 16 * C;
  const int tid = threadIdx.x;
  {
    constexpr int Q = per_thread<ROWS_MAX * R2, THREADS>();
    cx<T> v[Q][R1];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      const int row = b % ROWS_MAX, j = b / ROWS_MAX;
      if (row < rows && j < R2) {
#pragma unroll
        for (int r = 0; r < R1; ++r) v[q][r] = dvr[Tile::at(row, j + r * R2)];
        rdft<R1, T>(v[q], +1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      const int row = b % ROWS_MAX, j = b / ROWS_MAX;
      if (row < rows && j < R2) {
#pragma unroll
        for (int r = 0; r < R1; ++r) dvr[Tile::at(row, j * R1 + r)] = v[q][r];
      }
    }
    __syncthreads();
  }
  {
    constexpr int Q = per_thread<ROWS_MAX * R1, THREADS>();
    cx<T> v[Q][R2];
    load_ev();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const unsigned b = tid + q * THREADS;
      const int row = b / R1, j = b % R1;  // lanes along j: the exp(i A) loads of a row share cache lines
      if (row < rows) {
#pragma unroll
        for (int r = 0; r < R2; ++r) v[q][r] = dvr[Tile::at(row, j + r * R1)];
        twiddle<T, R2, false, TS, ES>(v[q], table, j, +1, eps);
        rdft<R2, T>(v[q], +1);
#pragma unroll
        for (int r = 0; r < R2; ++r) v[q][r] = ev[q][r] * v[q][r];
        rdft<R2, T>(v[q], -1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const unsigned b = tid + q * THREADS;
      const int row = b / R1, j = b % R1;
      if (row < rows) {
#pragma unroll
        for (int r = 0; r < R2; ++r) dvr[Tile::at(row, j * R2 + r)] = v[q][r];
      }
    }
    __syncthreads();
  }
  {
    constexpr int Q = per_thread<ROWS_MAX * R2, THREADS>();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b = tid + q * THREADS;
      const int row = b % ROWS_MAX, j = b / ROWS_MAX;
      if (row < rows && j < R2) {
        cx<T> v[R1];
#pragma unroll
        for (int r = 0; r < R1; ++r) v[r] = dvr[Tile::at(row, j + r * R2)];
        twiddle<T, R1, true, TS, ES>(v, table, j, -1, eps);
        rdft<R1, T>(v, -1);
#pragma unroll
        for (int r = 0; r < R1; ++r) dvr[Tile::at(row, j + r * R2)] = v[r];
      }
    }
    __syncthreads();
  }
}

template <typename T, int NPHI, int NT, int WAVES, int R1 = 0, int R2 = 0>
__global__ __launch_bounds__(64 * WAVES) void interaction_mfma_kernel(
    cx<T>* psi, const cx<double>* spec, Shape s, const cx<T>* table_phi, AngularTables at, MfmaTables mt,
    const T* frag_fwd, const T* frag_bwd, int max_jt, const cx<T>* exp_table, const cx<T>* rot_t,
    const cx<T>* s2sc, const T* eps_phi, const cx<T>* rot_lo_t) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  using Tile = DvrTile<NPHI, dvr_swizzled<T, NPHI, R1>()>;
  constexpr int C = 8 * NT;
  constexpr int THREADS = 64 * WAVES;
  extern __shared__ unsigned char smem_raw[];
  // LDS: phi map, phi twiddles, DVR tile; or the DVR tile alone with the two tables in its padding column.
  constexpr bool kPackTables = R1 > 0 && !dvr_swizzled<T, NPHI, R1>();
  constexpr int TS = kPackTables ? Tile::ld : 1;                                        // twiddle stride
  constexpr int PS = kPackTables ? Tile::ld * static_cast<int>(sizeof(cx<T>) / sizeof(int)) : 1;  // phi map stride
  constexpr int kTablesBytes = kPackTables ? 0 : ((NPHI * sizeof(int) + 15) / 16) * 16 + NPHI * sizeof(cx<T>);
  cx<T>* dvr = reinterpret_cast<cx<T>*>(smem_raw + kTablesBytes);
  int* phi_map = kPackTables ? reinterpret_cast<int*>(dvr + NPHI * Tile::ld + NPHI) : reinterpret_cast<int*>(smem_raw);
  cx<T>* tw_phi = kPackTables ? dvr + NPHI
                              : reinterpret_cast<cx<T>*>(smem_raw + ((NPHI * sizeof(int) + 15) / 16) * 16);
  T* dvr_t = reinterpret_cast<T*>(dvr);

  const int n_theta = s.n_theta;
  int tile, ixy;
  block_tile_site(blockIdx.x, s.n_xy, mt, tile, ixy);
  const int izr0 = tile * C;
  if (izr0 >= s.n_zr_pot) return;  // the padding workgroup of an odd tile count (whole workgroup, no barrier yet)
  const int cw = min(C, s.n_zr_pot - izr0);
  // This site's column of psi (n_zr * n_rot values, so 32-bit offsets within it) as components.
  T* site_t = reinterpret_cast<T*>(psi + static_cast<long>(s.n_zr) * s.n_rot * ixy);
  const int tid = threadIdx.x;
  const int lane = tid % 64;
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
  const int n = lane % 16;   // B / D column: complex column n / 2 of the N tile, component n % 2
  const int kk = lane / 16;  // B row within the k step
  const int ri = n % 2;
  const int rows = C * n_theta;

  // The phi map (and for complex64 the twiddles) are loaded here and written to LDS only after phase 1, so that
  // their latency does not hold up phase 1's first loads. (FP64 writes its twiddles at once: holding them through
  // phase 1 would spill.)
  static_assert(NPHI <= THREADS, "one phi entry per thread");
  constexpr bool kLateTwiddles = sizeof(T) == 4;
  const int op = at.op_of[ixy];
  int map_entry = 0;
  T eps_entry = T(0);
  cx<T> tw_entry = {};
  if (tid < NPHI) {
    map_entry = at.phi_index[tid + NPHI * op];
    if constexpr (sizeof(T) == 4) eps_entry = eps_phi[tid];
    if constexpr (kLateTwiddles)
      tw_entry = table_phi[tid];
    else
      tw_phi[tid * TS] = table_phi[tid];
  }
  // Slots of |mJ| > MJMax carry no basis function.
  const int n_unused = NPHI - s.n_blocks;
  for (int e = tid; e < rows * n_unused; e += THREADS) dvr[Tile::at(e / n_unused, s.mj_max + 1 + e % n_unused)] = {};

  // 1. FBR -> DVR (theta). Task = (mJ block, theta tile, N tile); tasks are wavefront-uniform. The initial
  // channel's specular coupling term has been added to psi beforehand (specular_coupling_kernel).
  const int n_tt = (n_theta + 15) / 16;
  const int n_tasks_f = s.n_blocks * n_tt * NT;
  constexpr int kPrefetchTasks = 8;
  // One theta tile, one N tile, at most 16 J per block: every lane's addresses are a per-lane constant plus a
  // wavefront-uniform offset, and all of a wavefront's B values are loaded before its first product.
  const bool fast = n_tt == 1 && max_jt == 1 && s.n_blocks <= kPrefetchTasks * WAVES;
  // Per N tile: this lane's complex column and whether it lies inside the tile.
  int cc_of[NT];
  bool live_of[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    cc_of[nt] = nt * 8 + n / 2;
    live_of[nt] = cc_of[nt] < cw;
  }
  if (fast) {
    // Task = mJ block (all N tiles); B element (J = kk) of block offset 0, and the phase-1 store positions of this
    // lane's 4 theta rows, per N tile.
    int lane_b[NT], dvr_off[NT][4], dvr_swz[NT][4];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      lane_b[nt] = 2 * (izr0 + cc_of[nt] + s.n_zr * kk) + ri;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = M::row(lane, v) * C + cc_of[nt];
        dvr_off[nt][v] = 2 * row * Tile::ld + ri;
        dvr_swz[nt][v] = Tile::at(row, 0) - row * Tile::ld;  // the row's column swizzle (0 without swizzling)
      }
    }
    // B values of every k step of a task, wavefront-uniform; the rows past the block (J >= nb) and the columns past
    // the interaction region read neighbouring finite values (or the zero slack behind psi): the former meet zero
    // fragment entries, the latter are never stored.
    T bv[kPrefetchTasks][NT][4];
#pragma unroll
    for (int i = 0; i < kPrefetchTasks; ++i) {
      const int bb = wave + i * WAVES;
      const int f0 = bb < s.n_blocks ? at.block_first[bb] : 0;
      const int nb = bb < s.n_blocks ? at.block_size[bb] : 0;
      const T* src = site_t + 2L * s.n_zr * f0;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int k = 0; k < 4; ++k)
          bv[i][nt][k] = k < (nb + 3) / 4 ? src[lane_b[nt] + 8 * s.n_zr * k] : T(0);
    }
    // The A fragments of task i + 1 are loaded while the products of task i run.
    auto load_frags = [&](int bb, T (&a)[4]) {
      const int k_steps = bb < s.n_blocks ? (at.block_size[bb] + 3) / 4 : 0;
      const T* frag = frag_fwd + 64L * (bb < s.n_blocks ? mt.fwd_offset[bb] : 0) + lane;
#pragma unroll
      for (int k = 0; k < 4; ++k) a[k] = k < k_steps ? frag[64 * k] : T(0);
    };
    T fa[4];
    load_frags(wave, fa);
#pragma unroll
    for (int i = 0; i < kPrefetchTasks; ++i) {
      const int bb = wave + i * WAVES;
      if (bb >= s.n_blocks) break;
      T next[4];
      load_frags(bb + WAVES, next);
      const int nb = at.block_size[bb];
      const int m = bb - s.mj_max;
      const int slot = m >= 0 ? m : NPHI + m;
      const int k_steps = (nb + 3) / 4;
      acc_t acc[NT] = {};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < k_steps) {
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) acc[nt] = M::fma(fa[k], bv[i][nt][k], acc[nt]);
        }
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (M::row(lane, v) < n_theta) dvr_t[dvr_off[nt][v] + 2 * (slot ^ dvr_swz[nt][v])] = acc[nt][v];
#pragma unroll
      for (int k = 0; k < 4; ++k) fa[k] = next[k];
    }
  } else {
    for (int task = wave; task < n_tasks_f; task += WAVES) {
      const int nt = task % NT;
      const int tt = (task / NT) % n_tt;
      const int bb = task / (NT * n_tt);
      const int f0 = at.block_first[bb], nb = at.block_size[bb];
      const int m = bb - s.mj_max;
      const int slot = m >= 0 ? m : NPHI + m;
      const int cc = nt * 8 + n / 2;
      const int k_steps = (nb + 3) / 4;
      const T* frag = frag_fwd + 64L * (mt.fwd_offset[bb] + tt * k_steps);
      acc_t acc = {};
      for (int k = 0; k < k_steps; ++k) {
        const int j = 4 * k + kk;
        const T b = (cc < cw && j < nb) ? site_t[2 * (izr0 + cc + s.n_zr * (f0 + j)) + ri] : T(0);
        acc = M::fma(frag[64 * k + lane], b, acc);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int th = 16 * tt + M::row(lane, v);
        if (th < n_theta) dvr_t[2 * Tile::at(th * C + cc, slot) + ri] = acc[v];
      }
    }
  }
  if (tid < NPHI) {
    phi_map[tid * PS] = map_entry;
    // complex64: twiddle tid's modulus correction in the second word of its phi map slot (packed tables only)
    if constexpr (sizeof(T) == 4 && kPackTables) reinterpret_cast<T*>(phi_map)[tid * PS + 1] = eps_entry;
    if constexpr (kLateTwiddles) tw_phi[tid * TS] = tw_entry;
  }
  __syncthreads();

  // 2. phi FFTs, the potential phase, phi FFTs back. The exp(i A) values are loaded at the start of the second
  // pass, so that their latency overlaps its LDS reads and twiddles.
  const int rep = at.rep_of[ixy];
  const int n_tp = n_theta * NPHI;
  if constexpr (R1 > 0) {
    cx<T> ev[phi_exp_per_thread<NPHI, R1, C, THREADS>()][R2];
    const cx<T>* exp_rows = exp_table + n_tp * (rep + static_cast<long>(s.n_rep) * izr0);
    phi_step_fused<T, NPHI, R1, R2, C, THREADS, Tile, TS, PS>(dvr, rows, tw_phi, reinterpret_cast<const T*>(phi_map) + 1,
                                                               ev, [&] {
      phi_exp_load<T, NPHI, R1, R2, C, THREADS, PS>(ev, rows, cw, phi_map, exp_rows, n_tp * s.n_rep);
    });
  } else {
    using Rows = CtLayout<0, Tile::ld, 0, 1>;
    ct_fft<T, NPHI, THREADS, Rows>(dvr, rows, tw_phi, eps_phi, +1);
    for (int e = tid; e < rows * NPHI; e += THREADS) {
      const int ip = e % NPHI;
      const int row = e / NPHI;  // theta * C + col
      const int th = row / C;
      const int cc = row % C;
      if (cc >= cw) continue;
      cx<T>& v = dvr[Tile::at(row, ip)];
      v = exp_table[phi_map[ip] + NPHI * th + static_cast<long>(n_tp) * (rep + static_cast<long>(s.n_rep) * (izr0 + cc))] * v;
    }
    __syncthreads();
    ct_fft<T, NPHI, THREADS, Rows>(dvr, rows, tw_phi, eps_phi, -1);
  }

  // 3. DVR -> FBR, x rot(dt/2), to HBM. Task = (mJ block, J tile, N tile); max_jt = ceil(largest block / 16).
  const int k_bwd = (n_theta + 3) / 4;
  if (fast) {
    // k_bwd <= 4 (n_theta <= 16), one J tile per block, all N tiles per task.
    int izr_of[NT], ir_of[NT], b_off[NT][4], b_swz[NT][4];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      izr_of[nt] = izr0 + cc_of[nt];
      ir_of[nt] = izr_of[nt] % s.nr;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int row = (4 * k + kk) * C + cc_of[nt];
        b_off[nt][k] = 2 * row * Tile::ld + ri;
        b_swz[nt][k] = Tile::at(row, 0) - row * Tile::ld;
      }
    }
    for (int bb = wave; bb < s.n_blocks; bb += WAVES) {
      const int f0 = at.block_first[bb], nb = at.block_size[bb];
      const int m = bb - s.mj_max;
      const int slot = m >= 0 ? m : NPHI + m;
      // Every HBM / L2 operand of the task first (A fragments, rotation factors), then the LDS operands.
      const T* frag = frag_bwd + 64L * mt.bwd_offset[bb] + lane;
      T a[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) a[k] = k < k_bwd ? frag[64 * k] : T(0);
      cx<T> rot[NT][4];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = M::row(lane, v);
          rot[nt][v] = j < nb ? rot_t[ir_of[nt] + s.nr * (f0 + j)] : cx<T>{T(0), T(0)};
        }
      T b[NT][4];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int k = 0; k < 4; ++k)
          b[nt][k] = (k < k_bwd && 4 * k + kk < n_theta) ? dvr_t[b_off[nt][k] + 2 * (slot ^ b_swz[nt][k])] : T(0);
      acc_t acc[NT] = {};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < k_bwd) {
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) acc[nt] = M::fma(a[k], b[nt][k], acc[nt]);
        }
      T* dst = site_t + 2L * s.n_zr * f0;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const T mine = acc[nt][v];
          const T other = swap_adjacent_lanes(mine);
          const int j = M::row(lane, v);
          const cx<T> c = ri ? cx<T>{other, mine} : cx<T>{mine, other};
          if (live_of[nt] && j < nb) {
            const cx<T> out = mul_factor(rot[nt][v], c, rot_lo_t, ir_of[nt] + s.nr * (f0 + j));
            dst[2 * (izr_of[nt] + s.n_zr * j) + ri] = ri ? out.y : out.x;
          }
        }
      }
    }
  } else {
    int ir_of[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) ir_of[nt] = (izr0 + nt * 8 + n / 2) % s.nr;
    for (int bt = wave; bt < s.n_blocks * NT; bt += WAVES)
      for (int jt = 0; jt < max_jt; ++jt) {
        const int nt = bt % NT;
        const int bb = bt / NT;
        const int f0 = at.block_first[bb], nb = at.block_size[bb];
        if (16 * jt >= nb) break;  // wave-uniform
        const int m = bb - s.mj_max;
        const int slot = m >= 0 ? m : NPHI + m;
        const int cc = nt * 8 + n / 2;
        const T* frag = frag_bwd + 64L * (mt.bwd_offset[bb] + jt * k_bwd);
        acc_t acc = {};
        for (int k = 0; k < k_bwd; ++k) {
          const int th = 4 * k + kk;
          const T b = th < n_theta ? dvr_t[2 * Tile::at(th * C + cc, slot) + ri] : T(0);
          acc = M::fma(frag[64 * k + lane], b, acc);
        }
        const int izr = izr0 + cc;
        const int ir = ir_of[nt];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const T mine = acc[v];
          const T other = __shfl_xor(mine, 1);
          const int j = 16 * jt + M::row(lane, v);
          const cx<T> c = ri ? cx<T>{other, mine} : cx<T>{mine, other};
          if (cc < cw && j < nb) {
            const int irot = f0 + j;
            const cx<T> out = mul_factor(rot_t[ir + s.nr * irot], c, rot_lo_t, ir + s.nr * irot);
            site_t[2 * (izr + s.n_zr * irot) + ri] = ri ? out.y : out.x;
          }
        }
      }
  }
}

// ---- buffer addressing ----------------------------------------------------------------------------------------
// A raw buffer resource (CDNA dword 3 0x00020000: 32-bit elements, no swizzle) over `bytes` bytes. Loads and stores
// take a 32-bit per-lane offset and a wavefront-uniform offset in an SGPR, so a per-lane offset computed once serves
// every task of a loop while the uniform part moves on the scalar unit; loads at or past `bytes` return 0 and
// stores there are dropped.
using rsrc_t = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ rsrc_t buffer_rsrc(const void* base, long bytes) {
  const int n = static_cast<int>(bytes < 0x7fffffffL ? bytes : 0x7fffffffL);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, n, 0x00020000);
}

template <typename V>
__device__ __forceinline__ V buffer_load(rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  } else if constexpr (sizeof(V) == 8) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  } else {
    static_assert(sizeof(V) == 16, "buffer_load: 4, 8 or 16 bytes");
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  }
}

template <typename V>
__device__ __forceinline__ void buffer_store(V v, rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(V) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
  } else if constexpr (sizeof(V) == 8) {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
  } else {
    static_assert(sizeof(V) == 16, "buffer_store: 4, 8 or 16 bytes");
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, voff, soff, 0);
  }
}

// The low part of the complex64 Legendre contraction. The FP32 fragments hold L rounded to float, L_hi; summing
// products of the same rounded matrix forward (L_hi) and back (L_hi^T) makes L_hi^T L_hi differ from the identity
// by ~1e-8, the same at every time step, which moves the norm of a rotational state linearly with the step count.
// The remainder L_lo = L - L_hi (|L_lo| <= 2^-24 |L|) is added as one 16 x 16 x 16 bf16 product: L_lo rounded to
// bf16 (host) times the same four B values of the lane truncated to bf16 (their top 16 bits), accumulated into the
// FP32 result; its error, ~2^-33 |L| |B|, leaves L_hi + L_lo orthogonal to ~1e-10.
// Operand layout of v_mfma_f32_16x16x16_bf16: lane l holds A[l % 16][4 (l / 16) + t] and B[4 (l / 16) + t][l % 16]
// for t = 0 .. 3 (bf16 element t in bits 16 t .. 16 t + 15). The contraction index is relabelled as
// kappa = 4 (l / 16) + t <-> row 4 t + l / 16 of the FP32 k steps, so B element t is the lane's value of k step t,
// and the host arranges the fragments of L_lo (512 bytes per mJ block, lane order) the same way.
typedef unsigned lo_frag_t __attribute__((ext_vector_type(2)));

// The lane's L_lo fragment of mJ block bb (four bf16 values).
__device__ __forceinline__ lo_frag_t legendre_low_fragment(rsrc_t lo, int bb) {
  return buffer_load<lo_frag_t>(lo, (threadIdx.x % 64) * 8u, static_cast<unsigned>(bb) * 512u);
}

// acc + L_lo B with the lane's fragment a of L_lo (legendre_low_fragment).
__device__ __forceinline__ Mfma<float>::acc_t legendre_low_product(lo_frag_t a, const float (&b)[4],
                                                                  Mfma<float>::acc_t acc) {
  typedef short s4 __attribute__((ext_vector_type(4)));
  // the top halves of (b0, b1) and (b2, b3): element t = b[t] truncated to bf16
  const lo_frag_t bb16 = {__builtin_amdgcn_perm(__float_as_uint(b[1]), __float_as_uint(b[0]), 0x07060302u),
                          __builtin_amdgcn_perm(__float_as_uint(b[3]), __float_as_uint(b[2]), 0x07060302u)};
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, a), __builtin_bit_cast(s4, bb16), acc, 0,
                                                   0, 0);
}


// The split-BF16x3 variant of the complex64 Legendre contraction (SPO_HIP_CONTRACTION=bf16x3): L and the four B
// values of the lane are each split into bf16 pieces, L = A_hi + A_lo (host) and b = b_hi + b_lo (here, both
// rounded to nearest), and A_hi b_hi + A_hi b_lo + A_lo b_hi is accumulated in FP32 by three 16 x 16 x 16 bf16
// products (smallest first; A_lo b_lo is dropped). The contraction index is relabelled as in legendre_low_product; the
// host arranges the A_hi and A_lo fragments (512 bytes each per mJ block) the same way. Each piece carries 8
// significant bits, so L takes part with about 16 bits: its relative error, ~2^-17, is ~100 times that of an FP32
// L, and the same at every step.
__device__ __forceinline__ Mfma<float>::acc_t legendre_bf16x3(rsrc_t hi, rsrc_t lo, int bb, const float (&b)[4]) {
  typedef short s4 __attribute__((ext_vector_type(4)));
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
  typedef float f4 __attribute__((ext_vector_type(4)));
  const unsigned off = (threadIdx.x % 64) * 8u, base = static_cast<unsigned>(bb) * 512u;
  const u2 a_hi = buffer_load<u2>(hi, off, base);
  const u2 a_lo = buffer_load<u2>(lo, off, base);
  const f4 x = {b[0], b[1], b[2], b[3]};
  const bf4 x_hi = __builtin_convertvector(x, bf4);
  const f4 rest = x - __builtin_convertvector(x_hi, f4);
  const bf4 x_lo = __builtin_convertvector(rest, bf4);
  Mfma<float>::acc_t acc = {};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, a_lo), __builtin_bit_cast(s4, x_hi), acc,
                                                  0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, a_hi), __builtin_bit_cast(s4, x_lo), acc,
                                                  0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, a_hi), __builtin_bit_cast(s4, x_hi), acc,
                                                   0, 0, 0);
}

// An entry of the sign-split rotation table rot_pm (16 bytes): (Re w, -Im w) for the lane holding Re c and
// (Re w, Im w) for the lane holding Im c; in complex64 followed by the FP32 remainders of the same two values
// (the FP64 factor minus its nearest complex64 value, mul_hilo in kernels_ct.hpp).
template <typename T>
struct alignas(16) RotEntry {
  T x, y;
};
template <>
struct alignas(16) RotEntry<float> {
  float x, y, lo_x, lo_y;
};

// The interaction step of one (8 izr, site) tile with buffer addressing, for shapes where every mJ block has at
// most 16 J, n_theta <= 16 and the blocks fit 8 tasks per wavefront. Same arithmetic as interaction_mfma_kernel's
// fast path (same MFMA products in the same order, same phi step), organised so that the vector units do little
// besides the phi FFTs:
//   - the B operands of phase 1 and the A fragments of both contractions come through buffer loads whose per-lane
//     offsets are computed once; the block's base moves in an SGPR;
//   - phase 3 multiplies by rot(dt/2) in two operations per value: the lane holding Re c takes (Re w, -Im w) and
//     the lane holding Im c takes (Re w, Im w) from a sign-split copy of the table (rot_pm, (irot, ir, lane parity)),
//     so Re and Im of w c are each one product and one fused multiply-add with the neighbour's component;
//   - results go straight to HBM through buffer stores at per-lane offsets computed once.
// The DVR tile has padded rows (n_phi + 1) whose padding column holds the phi twiddles and the phi map, as the
// complex64 tile of interaction_mfma_kernel; the tile has at least 2 n_phi rows.
// BF16X3 (complex64 only): the theta contractions as split-BF16x3 products (legendre_bf16x3) instead of FP32 MFMA
// with the bf16 low part; lo_fwd / lo_bwd then hold the A_lo fragments and hi_fwd / hi_bwd the A_hi ones.
// Complex64 asks for eight wavefronts per SIMD (at most 64 VGPRs): with the phase-3 fragment prefetch the compiler
// would otherwise take 65 and fit one workgroup fewer per CU (measured on MI355X at the DZ shape: 3.02 ms per
// launch against 3.11 without the prefetch and 3.24 with it at six wavefronts per SIMD).
template <typename T, int NPHI, int WAVES, int R1, int R2, bool BF16X3 = false>
__global__ __launch_bounds__(64 * WAVES, sizeof(T) == 8 ? 1 : 8) void interaction_buf_kernel(
    cx<T>* psi, Shape s, const cx<T>* table_phi, AngularTables at, MfmaTables mt, const T* frag_fwd,
    const T* frag_bwd, int frag_fwd_count, int frag_bwd_count, const cx<T>* exp_table, const cx<T>* rot_pm,
    const unsigned* lo_fwd, const unsigned* lo_bwd, const T* eps_phi, const unsigned* hi_fwd,
    const unsigned* hi_bwd) {
  static_assert(!BF16X3 || sizeof(T) == 4, "split-BF16x3 is a complex64 variant");
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  using Tile = DvrTile<NPHI, false>;
  constexpr int C = 8;
  constexpr int THREADS = 64 * WAVES;
  constexpr int TASKS = 8;  // mJ blocks per wavefront in phase 1 (all B operands loaded up front)
  constexpr int LD = Tile::ld;
  constexpr int TS = LD;
  constexpr int PS = LD * static_cast<int>(sizeof(cx<T>) / sizeof(int));
  constexpr unsigned CB = sizeof(cx<T>);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* dvr = reinterpret_cast<cx<T>*>(smem_raw);
  cx<T>* tw_phi = dvr + NPHI;                                          // row k, padding column
  int* phi_map = reinterpret_cast<int*>(dvr + NPHI * LD + NPHI);       // row n_phi + k, padding column

  const int n_theta = s.n_theta;
  int tile, ixy;
  block_tile_site(blockIdx.x, s.n_xy, mt, tile, ixy);
  const int izr0 = tile * C;
  if (izr0 >= s.n_zr_pot) return;  // the padding workgroup of an odd tile count (whole workgroup, no barrier yet)
  const int cw = min(C, s.n_zr_pot - izr0);
  const int tid = threadIdx.x;
  const int lane = tid % 64;
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
  const int n = lane % 16;   // B / D column: complex column n / 2, component n % 2
  const int kk = lane / 16;  // B row within a k step
  const int ri = n % 2;
  const int col = n / 2;
  const bool live = col < cw;
  const int rows = C * n_theta;
  const unsigned zr_bytes = static_cast<unsigned>(s.n_zr) * CB;

  const rsrc_t site = buffer_rsrc(psi + static_cast<long>(s.n_zr) * s.n_rot * ixy, static_cast<long>(s.n_zr) * s.n_rot * CB);
  const rsrc_t ffwd = buffer_rsrc(frag_fwd, 64L * frag_fwd_count * sizeof(T));
  const rsrc_t fbwd = buffer_rsrc(frag_bwd, 64L * frag_bwd_count * sizeof(T));
  // rot_pm entries are 16 bytes: (Re w, +-Im w) in FP64, (Re w, +-Im w) and their remainders in complex64
  // (RotEntry)
  const rsrc_t rots = buffer_rsrc(rot_pm, 2L * s.n_rot * s.nr * 16);
  const rsrc_t lo_f = buffer_rsrc(lo_fwd, sizeof(T) == 4 ? 512L * s.n_blocks : 0);
  const rsrc_t lo_b = buffer_rsrc(lo_bwd, sizeof(T) == 4 ? 512L * s.n_blocks : 0);
  const rsrc_t hi_f = buffer_rsrc(hi_fwd, BF16X3 ? 512L * s.n_blocks : 0);
  const rsrc_t hi_b = buffer_rsrc(hi_bwd, BF16X3 ? 512L * s.n_blocks : 0);

  // The phi map (and for complex64 the twiddles) go to LDS after phase 1, so that their loads do not hold up its
  // first loads; FP64 writes its twiddles at once (holding them through phase 1 would spill).
  constexpr bool kLateTwiddles = sizeof(T) == 4;
  const int op = at.op_of[ixy];
  int map_entry = 0;
  T eps_entry = T(0);
  cx<T> tw_entry = {};
  if (tid < NPHI) {
    map_entry = at.phi_index[tid + NPHI * op];
    if constexpr (sizeof(T) == 4) eps_entry = eps_phi[tid];
    if constexpr (kLateTwiddles)
      tw_entry = table_phi[tid];
    else
      tw_phi[tid * TS] = table_phi[tid];
  }
  const int n_unused = NPHI - s.n_blocks;
  for (int e = tid; e < rows * n_unused; e += THREADS) dvr[Tile::at(e / n_unused, s.mj_max + 1 + e % n_unused)] = {};

  // 1. FBR -> DVR (theta), task = mJ block bb = wave + i WAVES: B[k][n] = psi(J = f0 + 4 k + kk, izr0 + col).ri.
  // Rows past the block read the next block's values (their fragment entries are zero) or 0 past the column.
  const unsigned b_lane = static_cast<unsigned>(kk) * zr_bytes + static_cast<unsigned>(n) * sizeof(T);
  T bv[TASKS][4];
#pragma unroll
  for (int i = 0; i < TASKS; ++i) {
    const int bb = wave + i * WAVES;
    if (bb < s.n_blocks) {
      const unsigned base = (static_cast<unsigned>(at.block_first[bb]) * s.n_zr + izr0) * CB;
#pragma unroll
      for (int k = 0; k < 4; ++k) bv[i][k] = buffer_load<T>(site, b_lane, base + 4u * k * zr_bytes);
    }
  }
  // This lane's DVR tile position of row theta = M::row(lane, 0), column col, component ri (theta rows of the
  // other v follow at compile-time distances).
  const unsigned d_lane = (static_cast<unsigned>(M::row(lane, 0) * C + col) * LD) * CB + ri * sizeof(T);
  constexpr unsigned d_v = (sizeof(T) == 8 ? 4u : 1u) * C * LD * CB;  // tile bytes between v and v + 1
  unsigned char* dvr_b = smem_raw;
#pragma unroll
  for (int i = 0; i < TASKS; ++i) {
    const int bb = wave + i * WAVES;
    if (bb >= s.n_blocks) break;
    const int k_steps = (at.block_size[bb] + 3) / 4;
    const unsigned fbase = static_cast<unsigned>(mt.fwd_offset[bb]) * 64u * sizeof(T);
    const int m = bb - s.mj_max;
    const int slot = m >= 0 ? m : NPHI + m;
    acc_t acc = {};
    if constexpr (BF16X3) {
      acc = legendre_bf16x3(hi_f, lo_f, bb, bv[i]);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < k_steps)
          acc = M::fma(buffer_load<T>(ffwd, lane * sizeof(T), fbase + k * 64u * sizeof(T)), bv[i][k], acc);
      if constexpr (sizeof(T) == 4) acc = legendre_low_product(legendre_low_fragment(lo_f, bb), bv[i], acc);
    }
    unsigned char* d = dvr_b + d_lane + slot * CB;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (M::row(lane, v) < n_theta) *reinterpret_cast<T*>(d + v * d_v) = acc[v];
  }
  if (tid < NPHI) {
    phi_map[tid * PS] = map_entry;
    if constexpr (sizeof(T) == 4) reinterpret_cast<T*>(phi_map)[tid * PS + 1] = eps_entry;
    if constexpr (kLateTwiddles) tw_phi[tid * TS] = tw_entry;
  }
  __syncthreads();

  // 2. phi FFTs, the potential phase, phi FFTs back.
  {
    const int rep = at.rep_of[ixy];
    const int n_tp = n_theta * NPHI;
    cx<T> ev[phi_exp_per_thread<NPHI, R1, C, THREADS>()][R2];
    const cx<T>* exp_rows = exp_table + n_tp * (rep + static_cast<long>(s.n_rep) * izr0);
    phi_step_fused<T, NPHI, R1, R2, C, THREADS, Tile, TS, PS>(dvr, rows, tw_phi, reinterpret_cast<const T*>(phi_map) + 1,
                                                               ev, [&] {
      phi_exp_load<T, NPHI, R1, R2, C, THREADS, PS>(ev, rows, cw, phi_map, exp_rows, n_tp * s.n_rep);
    });
  }

  // 3. DVR -> FBR (theta), x rot(dt/2), to HBM. B[k][n] = DVR(theta = 4 k + kk, col).ri at the block's phi slot;
  // rows theta >= n_theta read the last theta row instead (written in phase 2, so finite; their fragment entries
  // are zero). The tile holds no row past it: a read there would meet whatever the LDS holds.
  const int k_bwd = (n_theta + 3) / 4;
  unsigned t_lane[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int th = 4 * k + kk < n_theta ? 4 * k + kk : n_theta - 1;
    t_lane[k] = (static_cast<unsigned>(th * C + col) * LD) * CB + ri * sizeof(T);
  }
  // Per v: the store offset of J = M::row(lane, v) (past the tile's columns: beyond every block, dropped by the
  // row test below) and the rotation factor's offset in rot_pm.
  unsigned st_lane[4], rot_lane[4];
  const int ir0 = izr0 % s.nr;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int j = M::row(lane, v);
    int ir = ir0 + col;
    if (ir >= s.nr) ir -= s.nr;
    st_lane[v] = (static_cast<unsigned>(j) * s.n_zr + col) * CB + ri * sizeof(T);
    rot_lane[v] = ((static_cast<unsigned>(j) * s.nr + ir) * 2u + ri) * 16u;
  }
  // The A fragments, rotation factors (and L_lo fragment) of the wavefront's next block are loaded before this
  // block's products, so that their latency overlaps them. Past the basis the last block's are loaded (never used).
  auto load_bwd = [&](int bb, T (&a)[4], lo_frag_t& lo) {
    const int bc = min(bb, s.n_blocks - 1);
    const unsigned fbase = static_cast<unsigned>(mt.bwd_offset[bc]) * 64u * sizeof(T);
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = k < k_bwd ? buffer_load<T>(fbwd, lane * sizeof(T), fbase + k * 64u * sizeof(T)) : T(0);
    if constexpr (sizeof(T) == 4 && !BF16X3) lo = legendre_low_fragment(lo_b, bc);
  };
  T a[4];
  lo_frag_t lo = {};
  if (wave < s.n_blocks) load_bwd(wave, a, lo);
  for (int bb = wave; bb < s.n_blocks; bb += WAVES) {
    const int f0 = at.block_first[bb], nb = at.block_size[bb];
    const int m = bb - s.mj_max;
    const int slot = m >= 0 ? m : NPHI + m;
    RotEntry<T> w[4];
    const unsigned rbase = static_cast<unsigned>(f0) * s.nr * 2u * 16u;
#pragma unroll
    for (int v = 0; v < 4; ++v) w[v] = buffer_load<RotEntry<T>>(rots, rot_lane[v], rbase);
    T na[4];
    lo_frag_t nlo = {};
    const bool more = bb + WAVES < s.n_blocks;  // wavefront-uniform
    if (more) load_bwd(bb + WAVES, na, nlo);
    T b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = k < k_bwd ? *reinterpret_cast<const T*>(dvr_b + t_lane[k] + slot * CB) : T(0);
    acc_t acc = {};
    if constexpr (BF16X3) {
      acc = legendre_bf16x3(hi_b, lo_b, bb, b);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < k_bwd) acc = M::fma(a[k], b[k], acc);
      if constexpr (sizeof(T) == 4) acc = legendre_low_product(lo, b, acc);
    }
    const unsigned sbase = (static_cast<unsigned>(f0) * s.n_zr + izr0) * CB;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const T mine = acc[v];
      const T other = swap_adjacent_lanes(mine);
      T out;
      if constexpr (sizeof(T) == 4)  // remainder products first, the nearest values fused on top (mul_hilo)
        out = fmaf(w[v].x, mine, fmaf(w[v].y, other, fmaf(w[v].lo_x, mine, w[v].lo_y * other)));
      else
        out = w[v].x * mine + w[v].y * other;
      if (live && M::row(lane, v) < nb) buffer_store<T>(out, site, st_lane[v], sbase);
    }
    if (more) {
#pragma unroll
      for (int k = 0; k < 4; ++k) a[k] = na[k];
      lo = nlo;
    }
  }
}

// Host-side twin of the rot_pm layout: entry ((irot * nr + ir) * 2 + parity) = (Re w, -Im w) for parity 0 and
// (Re w, Im w) for parity 1, w = rot(dt/2)[irot, ir].
inline std::size_t rot_pm_index(int irot, int ir, int nr, int parity) {
  return (static_cast<std::size_t>(irot) * nr + ir) * 2 + parity;
}

// LDS bytes of interaction_buf_kernel: the padded tile of max(8 n_theta, 2 n_phi) rows.
template <typename T>
inline std::size_t interaction_buf_lds(int n_theta, int nphi) {
  return std::size_t(std::max(8 * n_theta, 2 * nphi)) * (nphi + 1) * sizeof(cx<T>);
}

// The initial channel's specular coupling term of the interaction region, added to psi before the MFMA interaction
// kernel: psi[ichn](izr, ixy) += rot(dt/2)[ichn, ir] (s2sc[ixy] chi(ir, iz)) for izr < nz_potential * nr.
template <typename T>
__global__ __launch_bounds__(256) void specular_coupling_kernel(cx<T>* psi, const cx<double>* spec, Shape s,
                                                                 const cx<T>* rot_half, const cx<T>* s2sc) {
  const long e = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= static_cast<long>(s.n_zr_pot) * s.n_xy) return;
  const int izr = static_cast<int>(e % s.n_zr_pot);
  const int ixy = static_cast<int>(e / s.n_zr_pot);
  const int ir = izr % s.nr, iz = izr / s.nr;
  const cx<double> in = spec[ir + s.nr * iz];
  cx<T>& f = psi[izr + static_cast<long>(s.n_zr) * (s.ichn + static_cast<long>(s.n_rot) * ixy)];
  f = f + rot_half[s.ichn + s.n_rot * ir] * (s2sc[ixy] * cx<T>{T(in.x), T(in.y)});
}

// LDS bytes of the MFMA interaction kernel (phi map, phi twiddles, DVR tile of C n_theta rows).
template <typename T>
inline std::size_t interaction_mfma_lds(int C, int n_theta, int nphi, bool fused) {
  const bool swizzled = sizeof(T) == 8 && fused && nphi % 16 == 0;
  if (fused && !swizzled)  // the tables in the padding column of a tile of at least 2 n_phi rows
    return std::size_t(std::max(C * n_theta, 2 * nphi)) * (nphi + 1) * sizeof(cx<T>);
  return ((nphi * sizeof(int) + 15) / 16) * 16 + std::size_t(nphi) * sizeof(cx<T>) +
         std::size_t(C) * n_theta * (swizzled ? nphi : nphi + 1) * sizeof(cx<T>);
}

}  // namespace spo_dev
