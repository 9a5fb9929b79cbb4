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
// pass (a no-op when the caller loaded them earlier).
template <typename T, int NPHI, int R1, int R2, int C, int THREADS, class Tile, int TS, int QE, class LoadEv>
__device__ __forceinline__ void phi_step_fused(cx<T>* dvr, int rows, const cx<T>* table, cx<T> (&ev)[QE][R2],
                                               LoadEv load_ev) {
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
        twiddle<T, R2, false, TS>(v[q], table, j, +1);
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
        twiddle<T, R1, true, TS>(v, table, j, -1);
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
    const cx<T>* s2sc) {
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
  cx<T> tw_entry = {};
  if (tid < NPHI) {
    map_entry = at.phi_index[tid + NPHI * op];
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
    phi_step_fused<T, NPHI, R1, R2, C, THREADS, Tile, TS>(dvr, rows, tw_phi, ev, [&] {
      phi_exp_load<T, NPHI, R1, R2, C, THREADS, PS>(ev, rows, cw, phi_map, exp_rows, n_tp * s.n_rep);
    });
  } else {
    using Rows = CtLayout<0, Tile::ld, 0, 1>;
    ct_fft<T, NPHI, THREADS, Rows>(dvr, rows, tw_phi, +1);
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
    ct_fft<T, NPHI, THREADS, Rows>(dvr, rows, tw_phi, -1);
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
            const cx<T> out = rot[nt][v] * c;
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
            const cx<T> out = rot_t[ir + s.nr * irot] * c;
            site_t[2 * (izr + s.n_zr * irot) + ri] = ri ? out.y : out.x;
          }
        }
      }
  }
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
