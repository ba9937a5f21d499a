// firex_jacobi.hip — temporally blocked Jacobi pressure sweeps for gfx950 and
// their autotuner.
//
// Jacobi sweep (solvePressureGlobal, src/sim/firex/shaders.ts:375-403):
//   p'(c) = (wx (p[x-1] + p[x+1]) + wy (p[y-1] + p[y+1]) + wz (p[z-1] + p[z+1]) - div(c))
//           * (1 / (2 (wx + wy + wz)))
// with the weighted sum contracted to fma(wz, sz, fma(wy, sy, wx sx))
// (jacobi_sum)
// with clamped (Neumann) neighbours on the side walls and floor, zero
// (Dirichlet) above the open top, and p' = 0 on the top row y = ny - 1.
//
// One launch advances K sweeps.  A block owns a 64 x H column tile (x on the
// 64 lanes of a wave, H = W * R rows in y: R rows per thread, W waves stacked
// in y) plus one z chunk, and streams the chunk's planes bottom to top.
// Sweep level s works on plane (t - s) while level s - 1 works on plane
// (t - s + 1), so each level needs only the three most recent planes of the
// level below it.  Those stay in registers (per thread: planes z - 1, z, z + 1
// for its own columns); x neighbours come from the adjacent lanes, y
// neighbours from the thread's own rows or, at wave edges, from a small
// double-buffered LDS exchange (one barrier per streamed plane).  Each level
// loses one column/row of validity at the tile edges, so a tile outputs
// (64 - 2K) x (H - 2K) columns; halo columns are recomputed by the
// neighbouring tile.  HBM traffic per launch is one read of p and div and one
// write of p per output cell, instead of K of each.
//
// Every cell update uses the same f32 operations in the same order as one
// sweep of the reference kernel, so K fused sweeps are bit-identical to K
// separate sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "firex_jacobi.h"
#include "firex_kernels.h"

namespace fx {

namespace {

// Word 3 of a gfx9 buffer resource for raw 32-bit accesses.
constexpr int kBufferWord3 = 0x00020000;

template <int P> struct PIO;
template <> struct PIO<0> {
  using Word = float;
  static constexpr uint32_t kBytes = 4;
  __device__ static inline float load(const Word* p, uint32_t i) { return p[i]; }
  __device__ static inline void store(Word* p, uint32_t i, float v) { p[i] = v; }
  __device__ static inline float round(float v) { return v; }
  __device__ static inline float bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
  }
  __device__ static inline void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)voff, (int)soff, 0);
  }
};
template <> struct PIO<1> {
  using Word = _Float16;
  static constexpr uint32_t kBytes = 2;
  __device__ static inline float load(const Word* p, uint32_t i) { return (float)p[i]; }
  __device__ static inline void store(Word* p, uint32_t i, float v) { p[i] = (_Float16)fx_keep_f32(v); }
  __device__ static inline float round(float v) { return (float)(_Float16)fx_keep_f32(v); }
  __device__ static inline float bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return (float)__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0));
  }
  __device__ static inline void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float v) {
    const _Float16 h = (_Float16)fx_keep_f32(v);
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, h), r, (int)voff, (int)soff, 0);
  }
};

__device__ inline uint32_t xcd_remap(uint32_t hw, uint32_t total) {
  uint32_t q = total >> 3, r = total & 7u;
  uint32_t xcd = hw & 7u, slot = hw >> 3;
  uint32_t base = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return base + slot;
}

template <int P, int K, int R, int W>
__global__ __launch_bounds__(64 * W) void k_jacobi_tb(const typename PIO<P>::Word* __restrict__ pin,
                                                      typename PIO<P>::Word* __restrict__ pout,
                                                      const float* __restrict__ div, FxStep st, int chunk_z,
                                                      int tiles_x, int tiles_y, int zero_start) {
  constexpr int H = W * R;
  constexpr int TXI = 64 - 2 * K;
  constexpr int TYI = H - 2 * K;
  // [parity][level][wave][row 0 | row R-1][lane]
  __shared__ float edge[2][K][W][2][64];

  const int nx = st.nx, ny = st.ny, nz = st.nz;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint32_t chunks = (uint32_t)(nz + chunk_z - 1) / (uint32_t)chunk_z;
  const uint32_t total = (uint32_t)tiles_x * (uint32_t)tiles_y * chunks;
  const uint32_t b = xcd_remap(blockIdx.x, total);
  const int tile_x = (int)(b % (uint32_t)tiles_x);
  const uint32_t rest = b / (uint32_t)tiles_x;
  const int tile_y = (int)(rest % (uint32_t)tiles_y);
  const int chunk = (int)(rest / (uint32_t)tiles_y);

  const int gx = tile_x * TXI - K + lane;
  const int gy0 = tile_y * TYI - K + wave * R;
  const int z0 = chunk * chunk_z;
  const int z1 = min(z0 + chunk_z, nz);
  const uint32_t plane = (uint32_t)nx * (uint32_t)ny;
  const int gxc = min(max(gx, 0), nx - 1);
  uint32_t rowoff[R];
#pragma unroll
  for (int r = 0; r < R; ++r) rowoff[r] = (uint32_t)gxc + (uint32_t)nx * (uint32_t)min(max(gy0 + r, 0), ny - 1);

  const bool x_lo = gx == 0, x_hi = gx == nx - 1;
  const bool out_lane = lane >= K && lane < 64 - K && gx < nx;
  const float wx = st.wx, wy = st.wy, wz = st.wz, inv = st.jacobi_inv;

  float A[K][R], B[K][R], C[K][R], D[K + 1][R];
#pragma unroll
  for (int s = 0; s < K; ++s)
#pragma unroll
    for (int r = 0; r < R; ++r) A[s][r] = B[s][r] = C[s][r] = 0.0f;
#pragma unroll
  for (int s = 0; s <= K; ++s)
#pragma unroll
    for (int r = 0; r < R; ++r) D[s][r] = 0.0f;

  const int lo0 = max(z0 - K, 0);
  const int hi0 = min(z1 - 1 + K, nz - 1);
  const int t_end = z1 - 1 + K;
  for (int t = lo0; t <= t_end; ++t) {
    const int par = t & 1;
    const bool act0 = t <= hi0;
    if (act0) {
      const uint32_t base = (uint32_t)t * plane;
#pragma unroll
      for (int r = 0; r < R; ++r) C[0][r] = zero_start ? 0.0f : PIO<P>::load(pin, base + rowoff[r]);
    }
#pragma unroll
    for (int s = K; s >= 2; --s)
#pragma unroll
      for (int r = 0; r < R; ++r) D[s][r] = D[s - 1][r];
    if (t >= 1 && t - 1 < nz) {
      const uint32_t base = (uint32_t)(t - 1) * plane;
#pragma unroll
      for (int r = 0; r < R; ++r) D[1][r] = div[base + rowoff[r]];
    }

    bool active[K + 1];
    active[0] = act0;
#pragma unroll
    for (int s = 1; s <= K; ++s) {
      const int z = t - s;
      const int lo_s = max(z0 - (K - s), 0);
      const int hi_s = min(z1 - 1 + (K - s), nz - 1);
      active[s] = z >= lo_s && z <= hi_s;
      if (!active[s]) continue;
      const bool z_lo = z == 0, z_hi = z == nz - 1;
      float res[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int gy = gy0 + r;
        const float ctr = B[s - 1][r];
        float left = __shfl_up(ctr, 1, 64);
        float right = __shfl_down(ctr, 1, 64);
        if (x_lo) left = ctr;
        if (x_hi) right = ctr;
        float down, up;
        if (r > 0) down = B[s - 1][r - 1];
        else down = wave > 0 ? edge[par][s - 1][wave - 1][1][lane] : ctr;
        if (r < R - 1) up = B[s - 1][r + 1];
        else up = wave < W - 1 ? edge[par][s - 1][wave + 1][0][lane] : ctr;
        if (gy <= 0) down = ctr;
        if (gy + 1 >= ny) up = 0.0f;
        const float zl = z_lo ? ctr : A[s - 1][r];
        const float zh = z_hi ? ctr : C[s - 1][r];
        const float nb = jacobi_sum(wx, left + right, wy, down + up, wz, zl + zh);
        float v = (nb - D[s][r]) * inv;
        if (gy == ny - 1) v = 0.0f;
        res[r] = PIO<P>::round(v);
      }
      if (s < K) {
#pragma unroll
        for (int r = 0; r < R; ++r) C[s][r] = res[r];
      } else if (out_lane) {
        const uint32_t base = (uint32_t)z * plane;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = wave * R + r;
          const int gy = gy0 + r;
          if (row >= K && row < H - K && gy < ny) PIO<P>::store(pout, base + (uint32_t)gx + (uint32_t)nx * (uint32_t)gy, res[r]);
        }
      }
    }
#pragma unroll
    for (int s = 0; s < K; ++s) {
      if (active[s]) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          A[s][r] = B[s][r];
          B[s][r] = C[s][r];
        }
      }
      edge[par ^ 1][s][wave][0][lane] = B[s][0];
      edge[par ^ 1][s][wave][1][lane] = B[s][R - 1];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- ring kernel
// The same K-level pipeline, arranged to spend as few instructions as
// possible per cell update:
//   * the z loop is unrolled by U, so each level's planes live in register
//     rings addressed at compile time (level 0: 3 + PF planes, levels
//     1..K-1: 3 planes, div: D planes), and the unrolled blocks between the
//     pipeline's fill and its end have no branch at all;
//   * with PF = 1, level s works on plane t - s - 1 while plane t is being
//     loaded, so the loads of a plane have a whole iteration to arrive (the
//     barrier that ends an iteration waits for LDS only); with PF = 0 it
//     works on plane t - s, which needs fewer registers;
//   * x neighbours come from zero-filling DPP wave shifts, y neighbours at
//     wave edges from LDS, both without selects;
//   * the clamped (Neumann) walls -- x = 0, x = nx - 1, the floor y = 0 and
//     both z walls -- are mirror images: halo cells outside the grid load
//     the cell mirrored across the wall (-1 - i for cell -1 - i, 2n - 1 - i
//     for cell i >= n).  The plain update then keeps every such cell equal
//     to its mirror image level after level (at the wall both sums add the
//     same two values, and f32 addition commutes), which is exactly the
//     reference's clamped read, so wall tiles need no per-level fix-ups and
//     lose validity from the outside in like any other halo.  The top row
//     y = ny - 1 is set to zero after every level; the last y tile ends
//     there.
// Tiles overlap by 2K columns/rows (K at the floor); an output cell is
// only written by the tile whose valid range holds it.  With 16-wave
// blocks the columns left at the right wall may go to a 32-column tile
// whose blocks carry two y tiles, one per half wave (RingLayout).  Needs
// nx >= 64, ny >= W * R and nz > K.

// lane i <- lane i - 1 (lane 0 <- 0)
__device__ inline float shl_from_left(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}
// lane i <- lane i + 1 (lane 63 <- 0)
__device__ inline float shr_from_right(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
}
// Tiles of width T along an axis of n cells with halo K: tile i starts at
// -K + i (T - 2K); the last tile ends at n - 1 + K (top = 0) or at n - 1
// (top = 1: no halo above the last cell).
__host__ __device__ inline int ring_tiles(int n, int T, int K, int top) {
  const int span = n + K + (top ? 0 : K);
  return span <= T ? 1 : 1 + (span - T + (T - 2 * K) - 1) / (T - 2 * K);
}

// Origin and output range [lo, hi) of tile i of `count` (count >= 2).
__device__ inline void ring_tile(int i, int count, int n, int T, int K, int top, int& origin, int& lo, int& hi) {
  if (i == count - 1) {
    origin = n + (top ? 0 : K) - T;
    hi = n;
  } else {
    origin = -K + i * (T - 2 * K);
    hi = origin + T - K;
  }
  lo = i == 0 ? 0 : origin + K;
}

// Layout of a ring launch (RingLayout); lane_rows: the 16-wave shape, whose
// addressing is per lane and so allows the narrow tile.
__host__ __device__ inline RingLayout ring_layout(int nx, int ny, int K, int H, bool lane_rows) {
  RingLayout l;
  l.tiles_y = ring_tiles(ny, H, K, 1);
  const int step = 64 - 2 * K;
  const int full = nx / step, left = nx - full * step;
  if (lane_rows && full >= 1 && left > 0 && left <= 32 - 2 * K) {
    l.tiles_x = full;
    l.narrow = 1;
    l.pairs = (l.tiles_y + 1) / 2;
  } else {
    l.tiles_x = ring_tiles(nx, 64, K, 0);
    l.narrow = 0;
    l.pairs = 0;
  }
  return l;
}
__host__ __device__ inline uint32_t ring_columns(const RingLayout& l) {
  return (uint32_t)l.tiles_x * (uint32_t)l.tiles_y + (uint32_t)l.pairs;
}

// Mirror of cell v across the walls of an axis of n cells (-K <= v < n + K, n > K).
__device__ inline int mirror(int v, int n) { return v < 0 ? -1 - v : v >= n ? 2 * n - 1 - v : v; }

constexpr int kRingMaxSweeps = 5;  // K = 6 no longer unrolls fully (spills to scratch)

// Ring sizes: level 0 keeps M0 planes, div D planes (D >= K + 1 + PF); the
// z loop is unrolled by U, a multiple of 2 (LDS parity), 3, M0 and D.
template <int K, int PF>
struct RingDims {
  static constexpr int M0 = 3 + PF;
  static constexpr int D = PF ? (K + 2 <= 4 ? 4 : K + 2 <= 6 ? 6 : 12) : (K + 1 <= 3 ? 3 : 6);
  static constexpr int U = PF ? 12 : 6;
  static constexpr int NF = (2 * K + PF + U - 1) / U;  // unrolled blocks holding the fill iterations
  static_assert(K >= 1 && K + 1 + PF <= D, "sweeps per launch out of range");
  static_assert(U % 2 == 0 && U % 3 == 0 && U % M0 == 0 && U % D == 0, "ring sizes");
};

// The grid and stencil one ring launch works on.
struct RingGrid {
  int32_t nx, ny, nz;
  float wx, wy, wz, inv;
};

// One block of a ring launch: logical block b of ring_columns(lay) * chunks
// (chunk-major; within a chunk the full tiles x fastest, then the narrow
// tile's y pairs).
template <int P, int K, int R, int W, int PF>
__device__ __forceinline__ void ring_block(const typename PIO<P>::Word* __restrict__ pin,
                                           typename PIO<P>::Word* __restrict__ pout, const float* __restrict__ div,
                                           const RingGrid& st, int chunk_z, const RingLayout& lay, int zero_start,
                                           uint32_t b) {
  constexpr int H = W * R;
  // 16-wave blocks address every row per lane (VGPR offsets), which the
  // narrow tile's two half-wave y tiles need; 8-wave blocks keep the row
  // offsets wave-uniform (SGPRs), which costs fewer registers for R >= 4.
  constexpr bool kLaneRows = W == kRingWideWaves;
  constexpr int M0 = RingDims<K, PF>::M0;
  constexpr int D = RingDims<K, PF>::D;
  constexpr int U = RingDims<K, PF>::U;
  constexpr int LV = K > 1 ? K - 1 : 1;   // levels 1..K-1 kept in rings
  __shared__ float edge[2][K][W][2][64];  // [parity][level][wave][row 0 | row R-1][lane]

  const int nx = st.nx, ny = st.ny, nz = st.nz;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int tiles_x = lay.tiles_x, tiles_y = lay.tiles_y;
  const uint32_t cols = ring_columns(lay);
  const int chunk = (int)(b / cols);
  const uint32_t c = b % cols;
  const uint32_t n_full = (uint32_t)tiles_x * (uint32_t)tiles_y;
  const bool narrow_block = kLaneRows && c >= n_full;  // block-uniform
  int ox, xlo, xhi, gx, tile_y;
  bool half_used = true;
  if (narrow_block) {
    // columns nx + K - 32 .. nx + K - 1 on both half waves; lanes 0-31 run
    // y tile 2q, lanes 32-63 y tile 2q + 1 (absent for the last pair of an
    // odd tile count: that half repeats the last tile and stores nothing).
    // The wave-wide DPP shifts pass values between lanes 31 and 32, both
    // outermost halo columns, which lose validity after the first level
    // like the zero-filled lanes 0 and 63.
    ox = nx + K - 32;
    xlo = tiles_x * (64 - 2 * K);
    xhi = nx;
    gx = ox + (lane & 31);
    const int ty = 2 * (int)(c - n_full) + (lane >> 5);
    half_used = ty < tiles_y;
    tile_y = half_used ? ty : tiles_y - 1;
  } else {
    const int tile_x = (int)(c % (uint32_t)tiles_x);
    tile_y = (int)(c / (uint32_t)tiles_x);
    ring_tile(tile_x, tiles_x + lay.narrow, nx, 64, K, 0, ox, xlo, xhi);
    gx = ox + lane;
  }
  int oy, ylo, yhi;
  ring_tile(tile_y, tiles_y, ny, H, K, 1, oy, ylo, yhi);
  const bool fix_top = tile_y == tiles_y - 1 && wave == W - 1;  // row R - 1 is y = ny - 1
  const int gy0 = oy + wave * R;
  const int z0 = chunk * chunk_z;
  const int z1 = min(z0 + chunk_z, nz);
  // Odd chunks stream downwards: the pipeline runs on stream coordinates
  // (planes z0 - K .. z1 + K - 1 in order) and zr = z0 + z1 - 1 maps them to
  // grid planes back to front.  Neighbouring chunks then read their shared
  // halo planes at the same time (both at the start or both at the end of
  // their streams), which lets the second read hit in L2.  The update is
  // symmetric in z (below + above commutes), so the result is the same.
  const int zr = (chunk & 1) ? z0 + z1 - 1 : 0;
  const int zsign = (chunk & 1) ? -1 : 1;
  const uint32_t plane = (uint32_t)nx * (uint32_t)ny;
  const uint32_t cells = plane * (uint32_t)nz;
  // Buffer resources: per-row byte offsets in VGPRs, the plane offset in an
  // SGPR.  A zero start reads the pressure through an empty resource, which
  // returns 0 for every access.
  const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
      (void*)pin, (short)0, zero_start ? 0 : (int)(cells * PIO<P>::kBytes), kBufferWord3);
  const __amdgpu_buffer_rsrc_t rout =
      __builtin_amdgcn_make_buffer_rsrc((void*)pout, (short)0, (int)(cells * PIO<P>::kBytes), kBufferWord3);
  const __amdgpu_buffer_rsrc_t rdiv =
      __builtin_amdgcn_make_buffer_rsrc((void*)div, (short)0, (int)(cells * 4u), kBufferWord3);
  // Addressing: byte offset = voff (VGPR) + plane offset + srow (SGPR).
  // 8-wave blocks: voff is the (mirrored) column, srow the wave-uniform
  // (mirrored) row; 16-wave blocks: voff holds column and row, srow is 0.
  // Pressure and div offsets differ only with f16 pressure.
  const uint32_t cx = (uint32_t)mirror(gx, nx);
  uint32_t voff[R], voffd[R], srow[R], srowd[R];
  bool y_out[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int y = gy0 + r;
    const uint32_t row = (uint32_t)nx * (uint32_t)mirror(y, ny);
    voff[r] = (kLaneRows ? cx + row : cx) * PIO<P>::kBytes;
    voffd[r] = (kLaneRows ? cx + row : cx) * 4u;
    srow[r] = kLaneRows ? 0u : row * PIO<P>::kBytes;
    srowd[r] = kLaneRows ? 0u : row * 4u;
    y_out[r] = y >= ylo && y < yhi && half_used;
  }
  const bool x_out = gx >= xlo && gx < xhi;
  // Store offsets: voff where (x, y) is one of the tile's output cells, else
  // 2 GiB, past the end of every field the ring kernel runs on (< 2 GiB), so
  // the buffer store is discarded by the hardware range check (with or
  // without the plane and row offsets, which stay below 2 GiB, the sum
  // neither wraps nor lands in range).  Stores then need no branches.
  uint32_t vst[R];
#pragma unroll
  for (int r = 0; r < R; ++r) vst[r] = x_out && y_out[r] ? voff[r] : 0x80000000u;
  const float wx = st.wx, wy = st.wy, wz = st.wz, inv = st.inv;
  // Row R - 1 takes its last step as fma(nb - div, top_mul, top_add): on the
  // top row y = ny - 1 that is (nb - div) * 0 + (+0) = +0 whatever the sign
  // of nb - div, the forced zero of the open top; on every other row it is
  // (nb - div) * inv + (-0), which equals the product bit for bit (x + (-0)
  // = x for every x, zeros included).  So no per-level select is needed.
  const float top_mul = fix_top ? 0.0f : inv;
  const float top_add = fix_top ? 0.0f : -0.0f;
  // wave-edge neighbours; wave 0's "down" and wave W-1's "up" are halo rows
  // or the zeroed top row, so any in-bounds slot will do
  const int wdn = wave > 0 ? wave - 1 : W - 1;
  const int wup = wave < W - 1 ? wave + 1 : 0;

  float P0[M0][R], Pr[LV][3][R], Dv[D][R];
#pragma unroll
  for (int q = 0; q < M0; ++q)
#pragma unroll
    for (int r = 0; r < R; ++r) P0[q][r] = 0.0f;
#pragma unroll
  for (int s = 0; s < LV; ++s)
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int r = 0; r < R; ++r) Pr[s][q][r] = 0.0f;
#pragma unroll
  for (int q = 0; q < D; ++q)
#pragma unroll
    for (int r = 0; r < R; ++r) Dv[q][r] = 0.0f;

  // Iteration t loads plane t and level s computes plane t - s - PF, so the
  // chunk [z0, z1) takes planes z0 - K .. z1 + K - 1 and (z1 - z0) + 2K + PF
  // iterations.  Level s is needed on planes z0 - (K - s) .. z1 - 1 + (K - s)
  // only, which it reaches at iteration 2s + PF: the fill iterations before
  // that feed nothing and skip the level (their ring slots and edge rows hold
  // values no needed plane reads).  The fill iterations lie in the first NF
  // unrolled blocks, peeled so that the skips are resolved at compile time,
  // and level K then computes exactly the chunk's own planes z0 .. z1 - 1.
  // The blocks after them run branch-free (the DPP shifts of a level stay in
  // the basic block of the add that consumes them and fold into it); the
  // last, partial block checks the stream's end before each iteration.
  const int t0 = z0 - K;
  const int n_it = (z1 - z0) + 2 * K + PF;
  constexpr int NF = RingDims<K, PF>::NF;
  auto iteration = [&](const int tb, const int ph, const bool fill) __attribute__((always_inline)) {
    const int t = t0 + tb + ph;
    const int par = ph & 1;  // U is even: the parity of the iteration
    {
      // Planes outside the grid load their mirror image.  (With PF = 1
      // the last iteration's load is not used.  It stays unconditional: a
      // load that may or may not have been issued makes the compiler's
      // memory counter waits stall on the newest loads.)
      const uint32_t pbase = (uint32_t)min(max(mirror(zr + zsign * t, nz), 0), nz - 1) * plane;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        P0[ph % M0][r] = PIO<P>::bload(rin, voff[r], pbase * PIO<P>::kBytes + srow[r]);
        Dv[ph % D][r] = __builtin_bit_cast(
            float, __builtin_amdgcn_raw_buffer_load_b32(rdiv, (int)voffd[r], (int)(pbase * 4u + srowd[r]), 0));
      }
    }
#pragma unroll
    for (int s = 1; s <= K; ++s) {
      if (fill && 2 * s + PF > tb + ph) continue;
      const int z = t - s - PF;
      // ring slots (plane - t0 mod ring size) of planes z - 1, z, z + 1 of
      // level s - 1, and of z's div
      const int m0 = s == 1 ? M0 : 3;
      const int sa = ((ph - s - PF - 1) % m0 + m0) % m0;
      const int sc = ((ph - s - PF) % m0 + m0) % m0;
      const int sb = ((ph - s - PF + 1) % m0 + m0) % m0;
      const int sd = ((ph - s - PF) % D + D) % D;
      const int lv = s >= 2 ? s - 2 : 0;
      const float* below = s == 1 ? P0[sa] : Pr[lv][sa];
      const float* centre = s == 1 ? P0[sc] : Pr[lv][sc];
      const float* above = s == 1 ? P0[sb] : Pr[lv][sb];
      float res[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float ctr = centre[r];
        const float left = shl_from_left(ctr);
        const float right = shr_from_right(ctr);
        const float down = r > 0 ? centre[r - 1] : edge[par][s - 1][wdn][1][lane];
        const float up = r < R - 1 ? centre[r + 1] : edge[par][s - 1][wup][0][lane];
        const float nb = jacobi_sum(wx, left + right, wy, down + up, wz, below[r] + above[r]);
        res[r] = PIO<P>::round(r == R - 1 ? fxm_fma(nb - Dv[sd][r], top_mul, top_add) : (nb - Dv[sd][r]) * inv);
      }
      if (s < K) {
        // level s keeps planes in slots (plane - t0) % 3
        const int qc = ((ph - s - PF) % 3 + 3) % 3;
#pragma unroll
        for (int r = 0; r < R; ++r) Pr[s - 1][qc][r] = res[r];
      } else {
        // z0 <= z < z1 here (see above)
        const uint32_t pbase = (uint32_t)(zr + zsign * z) * plane * PIO<P>::kBytes;
#pragma unroll
        for (int r = 0; r < R; ++r) PIO<P>::bstore(rout, vst[r], pbase + srow[r], res[r]);
      }
    }
    // Publish rows 0 and R-1 of the plane each level's neighbours use next
    // iteration (level q: plane t - q - PF), then swap parity.
    {
      const int q0 = ((ph - PF) % M0 + M0) % M0;
      edge[par ^ 1][0][wave][0][lane] = P0[q0][0];
      edge[par ^ 1][0][wave][1][lane] = P0[q0][R - 1];
    }
#pragma unroll
    for (int s = 1; s < K; ++s) {
      const int q = ((ph - s - PF) % 3 + 3) % 3;
      edge[par ^ 1][s][wave][0][lane] = Pr[s - 1][q][0];
      edge[par ^ 1][s][wave][1][lane] = Pr[s - 1][q][R - 1];
    }
    __syncthreads();
  };
  // Two blocks share a CU, and the SIMDs issue the older block's waves
  // first: at UHD the first block on a CU finished in about 20 us, the
  // second in about 27.  So waves start at priority 3 and step down as their
  // block passes half and three quarters of its stream, which hands the
  // issue slots to the block that is behind (29.1 -> 28.7 us per launch).
  __builtin_amdgcn_s_setprio(3);
  // Every exit below is block-uniform: all waves leave together.
#pragma unroll
  for (int fb = 0; fb < NF; ++fb)
#pragma unroll
    for (int ph = 0; ph < U; ++ph) {
      if (fb * U + ph >= n_it) return;
      iteration(fb * U, ph, true);
    }
  int tb = NF * U;
  for (; tb + U <= n_it; tb += U) {
    if (4 * tb >= 3 * n_it) __builtin_amdgcn_s_setprio(0);
    else if (2 * tb >= n_it) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int ph = 0; ph < U; ++ph) iteration(tb, ph, false);
  }
#pragma unroll
  for (int ph = 0; ph < U; ++ph) {
    if (tb + ph >= n_it) return;
    iteration(tb, ph, false);
  }
}

template <int P, int K, int R, int W, int PF>
__global__ __launch_bounds__(64 * W) void k_jacobi_ring(const typename PIO<P>::Word* __restrict__ pin,
                                                        typename PIO<P>::Word* __restrict__ pout,
                                                        const float* __restrict__ div, RingGrid st, int chunk_z,
                                                        RingLayout lay, int zero_start) {
  const uint32_t chunks = (uint32_t)(st.nz + chunk_z - 1) / (uint32_t)chunk_z;
  const uint32_t b = xcd_remap(blockIdx.x, ring_columns(lay) * chunks);
  ring_block<P, K, R, W, PF>(pin, pout, div, st, chunk_z, lay, zero_start, b);
}

// Several plumes' ring launches as one launch: the blocks are numbered plume
// after plume (RingBatchEntry::block_begin) and each block runs the tile of
// its plume that the plume's own launch would give it, so every plume's
// result is the same as alone.
template <int P, int K, int R, int W, int PF>
__global__ __launch_bounds__(64 * W) void k_jacobi_ring_batch(RingBatch batch) {
  const uint32_t bg = xcd_remap(blockIdx.x, batch.total_blocks);
  int i = 0;
  while (i + 1 < batch.count && bg >= batch.entry[i + 1].block_begin) ++i;  // block-uniform
  const RingBatchEntry& q = batch.entry[i];
  const RingGrid g{q.nx, q.ny, q.nz, q.wx, q.wy, q.wz, q.inv};
  ring_block<P, K, R, W, PF>((const typename PIO<P>::Word*)q.pin, (typename PIO<P>::Word*)q.pout, q.div, g, q.chunk_z,
                             q.layout, q.zero_start, bg - q.block_begin);
}

template <int P, int K, int R, int W, int PF>
void launch_ring(const JacobiPlan& plan, const void* pin, void* pout, const float* div, const FxStep& st,
                 int zero_start, hipStream_t s) {
  const RingLayout lay = ring_layout(st.nx, st.ny, K, W * R, W == kRingWideWaves);
  const int chunks = (st.nz + plan.chunk_z - 1) / plan.chunk_z;
  const uint32_t blocks = ring_columns(lay) * (uint32_t)chunks;
  using Word = typename PIO<P>::Word;
  const RingGrid g{st.nx, st.ny, st.nz, st.wx, st.wy, st.wz, st.jacobi_inv};
  k_jacobi_ring<P, K, R, W, PF><<<blocks, 64 * W, 0, s>>>((const Word*)pin, (Word*)pout, div, g, plan.chunk_z, lay,
                                                       zero_start);
}

template <int P, int R, int W, int PF>
void launch_ring_rw(int k, const JacobiPlan& plan, const void* pin, void* pout, const float* div, const FxStep& st,
                    int zero_start, hipStream_t s) {
  switch (k) {
    case 1: launch_ring<P, 1, R, W, PF>(plan, pin, pout, div, st, zero_start, s); break;
    case 2: launch_ring<P, 2, R, W, PF>(plan, pin, pout, div, st, zero_start, s); break;
    case 3: launch_ring<P, 3, R, W, PF>(plan, pin, pout, div, st, zero_start, s); break;
    default:
      // the prefetch variant runs K <= 3 only (jacobi_plan_valid)
      if constexpr (PF == 0) {
        if (k == 4)
          launch_ring<P, 4, R, W, PF>(plan, pin, pout, div, st, zero_start, s);
        else
          launch_ring<P, 5, R, W, PF>(plan, pin, pout, div, st, zero_start, s);
      }
      break;
  }
}

// Ring variants: tile height (8 waves of R = 4, 6 or 8 rows: 64 x 32, 64 x
// 48 or 64 x 64 cells) and prefetch distance.  Taller tiles recompute fewer
// halo rows (at K = 3, 26 of 32 rows are output, 58 of 64) but give fewer
// blocks and hold more registers.
template <int P, int R>
void launch_ring_pf(int k, const JacobiPlan& plan, const void* pin, void* pout, const float* div, const FxStep& st,
                    int zero_start, hipStream_t s) {
  if (plan.prefetch)
    launch_ring_rw<P, R, 8, 1>(k, plan, pin, pout, div, st, zero_start, s);
  else
    launch_ring_rw<P, R, 8, 0>(k, plan, pin, pout, div, st, zero_start, s);
}

template <int P>
void launch_ring_k(int k, const JacobiPlan& plan, const void* pin, void* pout, const float* div, const FxStep& st,
                   int zero_start, hipStream_t s) {
  if (plan.waves == kRingWideWaves) {
    // 16 waves of 2 rows (K <= kRingWideMaxSweeps, jacobi_plan_valid)
    switch (k) {
      case 1: launch_ring<P, 1, 2, kRingWideWaves, 0>(plan, pin, pout, div, st, zero_start, s); break;
      case 2: launch_ring<P, 2, 2, kRingWideWaves, 0>(plan, pin, pout, div, st, zero_start, s); break;
      case 3: launch_ring<P, 3, 2, kRingWideWaves, 0>(plan, pin, pout, div, st, zero_start, s); break;
      default: launch_ring<P, 4, 2, kRingWideWaves, 0>(plan, pin, pout, div, st, zero_start, s); break;
    }
    return;
  }
  if (plan.rows == 8)
    launch_ring_pf<P, 8>(k, plan, pin, pout, div, st, zero_start, s);
  else if (plan.rows == 6)
    launch_ring_pf<P, 6>(k, plan, pin, pout, div, st, zero_start, s);
  else
    launch_ring_pf<P, 4>(k, plan, pin, pout, div, st, zero_start, s);
}

// Batched launches run the 16-wave shape only.
template <int P, int K>
void launch_ring_batch(RingBatch& b, hipStream_t s) {
  constexpr int R = 2, W = kRingWideWaves;
  uint32_t total = 0;
  for (int i = 0; i < b.count; ++i) {
    RingBatchEntry& e = b.entry[i];
    e.layout = ring_layout(e.nx, e.ny, K, W * R, true);
    const int chunks = (e.nz + e.chunk_z - 1) / e.chunk_z;
    e.block_begin = total;
    total += ring_columns(e.layout) * (uint32_t)chunks;
  }
  b.total_blocks = total;
  k_jacobi_ring_batch<P, K, R, W, 0><<<total, 64 * W, 0, s>>>(b);
}

template <int P>
void launch_ring_batch_k(int k, RingBatch& b, hipStream_t s) {
  switch (k) {
    case 1: launch_ring_batch<P, 1>(b, s); break;
    case 2: launch_ring_batch<P, 2>(b, s); break;
    case 3: launch_ring_batch<P, 3>(b, s); break;
    default: launch_ring_batch<P, 4>(b, s); break;
  }
}

template <int P, int K>
void launch_tb(const JacobiPlan& plan, const void* pin, void* pout, const float* div, const FxStep& st, int zero_start,
               hipStream_t s) {
  constexpr int R = 4, W = 8;
  const int TXI = 64 - 2 * K, TYI = W * R - 2 * K;
  const int tiles_x = (st.nx + TXI - 1) / TXI;
  const int tiles_y = (st.ny + TYI - 1) / TYI;
  const int chunks = (st.nz + plan.chunk_z - 1) / plan.chunk_z;
  const uint32_t blocks = (uint32_t)tiles_x * (uint32_t)tiles_y * (uint32_t)chunks;
  using Word = typename PIO<P>::Word;
  k_jacobi_tb<P, K, R, W><<<blocks, 64 * W, 0, s>>>((const Word*)pin, (Word*)pout, div, st, plan.chunk_z, tiles_x,
                                                    tiles_y, zero_start);
}

template <int P>
void launch_tb_k(int k, const JacobiPlan& plan, const void* pin, void* pout, const float* div, const FxStep& st,
                 int zero_start, hipStream_t s) {
  switch (k) {
    case 1: launch_tb<P, 1>(plan, pin, pout, div, st, zero_start, s); break;
    case 2: launch_tb<P, 2>(plan, pin, pout, div, st, zero_start, s); break;
    case 3: launch_tb<P, 3>(plan, pin, pout, div, st, zero_start, s); break;
    case 4: launch_tb<P, 4>(plan, pin, pout, div, st, zero_start, s); break;
    case 5: launch_tb<P, 5>(plan, pin, pout, div, st, zero_start, s); break;
    case 6: launch_tb<P, 6>(plan, pin, pout, div, st, zero_start, s); break;
    case 7: launch_tb<P, 7>(plan, pin, pout, div, st, zero_start, s); break;
    default: launch_tb<P, 8>(plan, pin, pout, div, st, zero_start, s); break;
  }
}

// ----------------------------------------------------- projection + divergence
}  // namespace

bool jacobi_plan_valid(const JacobiPlan& plan) {
  const int max_k = plan.kernel == JACOBI_KERNEL_RING ? kRingMaxSweeps : 8;
  const bool wide = plan.kernel == JACOBI_KERNEL_RING && plan.waves == kRingWideWaves && plan.rows == 2 &&
                    plan.prefetch == 0 && plan.sweeps_per_launch <= kRingWideMaxSweeps;
  const bool shape = plan.kernel == JACOBI_KERNEL_RING
      ? wide || ((plan.rows == 4 || plan.rows == 6 || plan.rows == 8) && plan.waves == 8 &&
                 (plan.prefetch == 0 || (plan.prefetch == 1 && plan.sweeps_per_launch <= 3)))
      : plan.rows == 4 && plan.waves == 8;
  return plan.sweeps_per_launch >= 1 && plan.sweeps_per_launch <= max_k && shape && plan.chunk_z >= 1 &&
         (plan.kernel == JACOBI_KERNEL_SHIFT || plan.kernel == JACOBI_KERNEL_RING);
}

bool jacobi_plan_batchable(const JacobiPlan& plan) {
  return jacobi_plan_valid(plan) && plan.kernel == JACOBI_KERNEL_RING && plan.waves == kRingWideWaves;
}

void launch_jacobi_ring_batch(const JacobiPlan& plan, int f16, int sweeps, RingBatch& batch, hipStream_t s) {
  const int k = sweeps < plan.sweeps_per_launch ? sweeps : plan.sweeps_per_launch;
  if (f16)
    launch_ring_batch_k<1>(k, batch, s);
  else
    launch_ring_batch_k<0>(k, batch, s);
}

int jacobi_plan_blocks(const JacobiPlan& p, const int32_t dims[3]) {
  const int K = p.sweeps_per_launch;
  const int tiles = p.kernel == JACOBI_KERNEL_RING
      ? (int)ring_columns(ring_layout(dims[0], dims[1], K, p.waves * p.rows, p.waves == kRingWideWaves))
      : ((dims[0] + 63 - 2 * K) / (64 - 2 * K)) * ((dims[1] + 31 - 2 * K) / (32 - 2 * K));
  return tiles * ((dims[2] + p.chunk_z - 1) / p.chunk_z);
}

JacobiPlan make_jacobi_plan(const int32_t dims[3], int32_t sweeps_per_launch, int32_t chunk_z, int32_t kernel,
                            int32_t tile_height, int32_t prefetch, int32_t ring_waves) {
  JacobiPlan p;
  // the ring kernel's tile height (8 waves of 4, 6 or 8 rows); the shift
  // kernel keeps 32
  const int rows = tile_height == 48 ? 6 : tile_height == 64 ? 8 : 4;
  if (dims[1] >= p.waves * rows) p.rows = rows;
  // The ring kernel addresses fields through buffer resources with 32-bit
  // byte offsets: fields up to 2 GiB.
  const uint64_t field_bytes = (uint64_t)dims[0] * (uint64_t)dims[1] * (uint64_t)dims[2] * 4u;
  const bool ring_ok = dims[0] >= 64 && dims[1] >= p.waves * p.rows && dims[2] >= 3 && field_bytes < (1ull << 31);
  p.kernel = (kernel == JACOBI_KERNEL_RING && ring_ok) ? JACOBI_KERNEL_RING
           : kernel == JACOBI_KERNEL_SHIFT ? JACOBI_KERNEL_SHIFT
           : ring_ok ? JACOBI_KERNEL_RING : JACOBI_KERNEL_SHIFT;
  if (p.kernel != JACOBI_KERNEL_RING) p.rows = 4;
  if (p.kernel == JACOBI_KERNEL_RING && ring_waves == kRingWideWaves && p.rows == 4) {
    p.rows = 2;  // the same 32-row tile as 16 waves of 2 rows
    p.waves = kRingWideWaves;
  }
  // the ring kernel's mirrored z halo needs nz >= K + 2
  const int max_k = p.kernel == JACOBI_KERNEL_RING ? (dims[2] - 2 < kRingMaxSweeps ? dims[2] - 2 : kRingMaxSweeps) : 8;
  const int dflt = p.kernel == JACOBI_KERNEL_RING ? 3 : 4;
  const int k = sweeps_per_launch > 0 ? sweeps_per_launch : dflt;
  p.sweeps_per_launch = k > max_k ? max_k : k;
  if (p.waves == kRingWideWaves && p.sweeps_per_launch > kRingWideMaxSweeps) p.sweeps_per_launch = kRingWideMaxSweeps;
  p.prefetch = p.kernel == JACOBI_KERNEL_RING && prefetch == 2 && p.sweeps_per_launch <= 3 && p.waves == 8 ? 1 : 0;
  if (chunk_z > 0) {
    p.chunk_z = chunk_z;
    return p;
  }
  // Enough z chunks to give every CU a couple of blocks.
  p.chunk_z = dims[2];
  while (p.chunk_z > 8 && jacobi_plan_blocks(p, dims) < 512) p.chunk_z = (p.chunk_z + 1) / 2;
  return p;
}

void launch_jacobi_tb(const JacobiPlan& plan, int f16, const JacobiLaunch& l, void* const buffers[2],
                      const float* div, const FxStep& st, hipStream_t s) {
  const void* pin = buffers[l.src];
  void* pout = buffers[1 - l.src];
  const int zs = l.zero_start ? 1 : 0;
  if (plan.kernel == JACOBI_KERNEL_RING) {
    if (f16)
      launch_ring_k<1>(l.sweeps, plan, pin, pout, div, st, zs, s);
    else
      launch_ring_k<0>(l.sweeps, plan, pin, pout, div, st, zs, s);
  } else if (f16) {
    launch_tb_k<1>(l.sweeps, plan, pin, pout, div, st, zs, s);
  } else {
    launch_tb_k<0>(l.sweeps, plan, pin, pout, div, st, zs, s);
  }
}

// ------------------------------------------------------------------ tuner

namespace {

// Tuning data: a divergence field of hashed values in [-0.3, 0.36) (f32
// patterns that switch like a real field; an all-zero field runs faster
// than real data and ranks plans differently).
__global__ __launch_bounds__(256) void k_tune_pattern(float* __restrict__ div, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    div[i] = (float)(h >> 16) * 1e-5f - 0.3f;
  }
}

// device, dims, f16, K, chunk, kernel, tile height, prefetch, ring waves
using TuneKey = std::tuple<int, int, int, int, int, int, int, int, int, int, int>;
struct TuneEntry {
  JacobiPlan plan;
  float us_per_sweep;
};
std::mutex g_tune_mutex;
std::map<TuneKey, TuneEntry> g_tune_cache;

// z chunk sizes worth timing: about 10% apart, giving 200..3000 blocks.  A
// grid too small for 200 blocks at chunks of 4 planes or more is searched
// from 16 blocks, down to 1-plane chunks: there more blocks pay for the halo
// planes a short chunk adds (64 x 96 x 64 at depth 4: 2.1 us per sweep on
// 2-plane chunks against 3.3 us on the 6-plane chunks of the 200-block rule).
std::vector<int> chunk_candidates(int nz, long tiles) {
  std::vector<int> out;
  const long lo = tiles * ((nz + 3) / 4) < 200 ? 16 : 200;
  for (int n = 1; n <= nz; ++n) {
    const int c = (nz + n - 1) / n;
    if (c < (lo < 200 ? 1 : 4)) break;
    const long blocks = tiles * ((nz + c - 1) / c);
    if (blocks < lo || blocks > 3000) continue;
    if (!out.empty() && c * 10 > out.back() * 9) continue;
    out.push_back(c);
  }
  if (out.empty()) out.push_back(std::max(1, std::min(nz, 16)));
  return out;
}

}  // namespace

int tune_jacobi_plan(const int32_t dims[3], int f16, int32_t sweeps_per_launch, int32_t chunk_z, int32_t kernel,
                     int32_t tile_height, int32_t prefetch, int32_t ring_waves, void* const buffers[2], float* div,
                     hipStream_t s, JacobiPlan* out, float* us_per_sweep) {
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return e;
  const TuneKey key{device, dims[0], dims[1], dims[2], f16, sweeps_per_launch, chunk_z, kernel, tile_height,
                    prefetch, ring_waves};
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      *out = it->second.plan;
      if (us_per_sweep) *us_per_sweep = it->second.us_per_sweep;
      return hipSuccess;
    }
  }
  // Candidates: the default plan first, then every kernel / depth / chunk /
  // prefetch combination the fixed parameters allow.
  std::vector<JacobiPlan> cands;
  cands.push_back(make_jacobi_plan(dims, sweeps_per_launch, chunk_z, kernel, tile_height, prefetch, ring_waves));
  const int kernels[2] = {JACOBI_KERNEL_RING, JACOBI_KERNEL_SHIFT};
  for (int kern : kernels) {
    if (kernel != JACOBI_KERNEL_AUTO && kernel != kern) continue;
    const JacobiPlan probe = make_jacobi_plan(dims, 1, 1, kern);
    if (probe.kernel != kern) continue;  // this grid does not allow it
    // the ring kernel, where allowed, is always the faster one
    if (kern == JACOBI_KERNEL_SHIFT && kernel == JACOBI_KERNEL_AUTO && cands[0].kernel == JACOBI_KERNEL_RING) continue;
    const int kmax = kern == JACOBI_KERNEL_RING ? kRingMaxSweeps : 4;
    for (int K = sweeps_per_launch > 0 ? sweeps_per_launch : 2; K <= (sweeps_per_launch > 0 ? sweeps_per_launch : kmax);
         ++K) {
      // taller ring tiles are searched only when asked for: at UHD (192 x
      // 288 x 192) they recompute fewer halo rows yet run slower (9.9, 11.1
      // and 12.3 us per sweep for 32, 48 and 64 rows), as they hold more
      // registers and give fewer blocks
      for (int height : {32, 48, 64}) {
        if (tile_height > 0 ? height != tile_height : height != 32) continue;
        // block shapes of a 32-row ring tile: 8 waves of 4 rows or 16 of 2
        for (int waves : {8, kRingWideWaves}) {
          if (waves == kRingWideWaves && (kern != JACOBI_KERNEL_RING || height != 32)) continue;
          if (ring_waves > 0 && waves != ring_waves) continue;
          JacobiPlan base = make_jacobi_plan(dims, K, 1, kern, height, 0, waves);
          if (base.sweeps_per_launch != K || base.rows * base.waves != height || base.waves != waves) continue;
          base.chunk_z = 1;
          const long tiles = jacobi_plan_blocks(base, dims) / dims[2];
          std::vector<int> chunks = chunk_z > 0 ? std::vector<int>{chunk_z} : chunk_candidates(dims[2], tiles);
          for (int c : chunks) {
            for (int pf = 0; pf <= (kern == JACOBI_KERNEL_RING && K <= 3 && waves == 8 ? 1 : 0); ++pf) {
              if (prefetch > 0 && pf != prefetch - 1) continue;
              JacobiPlan p = base;
              p.chunk_z = c;
              p.prefetch = pf;
              if (jacobi_plan_valid(p)) cands.push_back(p);
            }
          }
        }
      }
    }
  }

  FxStep st;
  memset(&st, 0, sizeof st);
  st.nx = dims[0];
  st.ny = dims[1];
  st.nz = dims[2];
  st.wx = st.wy = st.wz = 1.0f;
  st.jacobi_inv = 1.0f / 6.0f;
  const size_t cells = (size_t)dims[0] * dims[1] * dims[2];
  const size_t pbytes = (cells * (f16 ? 2 : 4) + 3) / 4 * 4;
  if ((e = hipMemsetAsync(buffers[0], 0, pbytes, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(buffers[1], 0, pbytes, s)) != hipSuccess) return e;
  k_tune_pattern<<<2048, 256, 0, s>>>(div, (uint64_t)cells);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipEvent_t ev[2];
  if ((e = hipEventCreate(&ev[0])) != hipSuccess) return e;
  if ((e = hipEventCreate(&ev[1])) != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return e;
  }
  auto time_plan = [&](const JacobiPlan& p, int reps, float* us) -> hipError_t {
    const JacobiLaunch l0{p.sweeps_per_launch, 0, false}, l1{p.sweeps_per_launch, 1, false};
    launch_jacobi_tb(p, f16, l0, buffers, div, st, s);
    hipError_t err = hipEventRecord(ev[0], s);
    for (int r = 0; r < reps && err == hipSuccess; ++r) launch_jacobi_tb(p, f16, (r & 1) ? l0 : l1, buffers, div, st, s);
    if (err == hipSuccess) err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(ev[1], s);
    if (err == hipSuccess) err = hipEventSynchronize(ev[1]);
    float ms = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, ev[0], ev[1]);
    *us = ms * 1000.0f / (float)(reps * p.sweeps_per_launch);
    return err;
  };
  // Warm the clocks, then a quick pass over all candidates, chunk sizes
  // within 15 % of the two best (time is not smooth in the chunk size: the
  // block count steps), and a longer pass over the best few.
  std::vector<std::pair<float, int>> score;
  float us = 0;
  e = time_plan(cands[0], 12, &us);
  // Repetitions scaled so that a measurement spans at least about 100 us
  // (quick pass) and 400 us (final pass): on small grids a launch takes a
  // few microseconds and three of them do not rank the candidates.
  const float launch_us = us * (float)cands[0].sweeps_per_launch;
  const int quick = std::max(3, std::min(32, (int)(100.0f / std::max(launch_us, 1.0f)) + 1));
  const int final_reps = std::max(8, std::min(128, (int)(400.0f / std::max(launch_us, 1.0f)) + 1));
  for (size_t i = 0; i < cands.size() && e == hipSuccess; ++i) {
    e = time_plan(cands[i], quick, &us);
    score.emplace_back(us, (int)i);
  }
  if (e == hipSuccess && chunk_z <= 0) {
    std::sort(score.begin(), score.end());
    const size_t first = cands.size();
    for (size_t j = 0; j < std::min<size_t>(score.size(), 2); ++j) {
      const JacobiPlan b = cands[score[j].second];
      const int step = std::max(1, b.chunk_z / 16);
      for (int c = std::max(1, b.chunk_z * 85 / 100); c <= std::min(dims[2], b.chunk_z * 115 / 100); c += step) {
        bool seen = false;
        for (const JacobiPlan& q : cands)
          seen = seen || (q.kernel == b.kernel && q.sweeps_per_launch == b.sweeps_per_launch &&
                          q.prefetch == b.prefetch && q.rows == b.rows && q.waves == b.waves && q.chunk_z == c);
        if (seen) continue;
        JacobiPlan p = b;
        p.chunk_z = c;
        cands.push_back(p);
      }
    }
    for (size_t i = first; i < cands.size() && e == hipSuccess; ++i) {
      e = time_plan(cands[i], quick, &us);
      score.emplace_back(us, (int)i);
    }
  }
  if (e == hipSuccess) {
    std::sort(score.begin(), score.end());
    const size_t top = std::min<size_t>(score.size(), 6);
    for (size_t j = 0; j < top && e == hipSuccess; ++j) {
      float a = 0, b = 0;
      e = time_plan(cands[score[j].second], final_reps, &a);
      if (e == hipSuccess) e = time_plan(cands[score[j].second], final_reps, &b);
      score[j].first = std::min(a, b);
    }
    std::sort(score.begin(), score.begin() + top);
  }
  (void)hipEventDestroy(ev[0]);  // teardown: a failure here changes nothing
  (void)hipEventDestroy(ev[1]);
  if (e != hipSuccess) return e;
  const TuneEntry best{cands[score[0].second], score[0].first};
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune_cache[key] = best;
  }
  *out = best.plan;
  if (us_per_sweep) *us_per_sweep = best.us_per_sweep;
  return hipSuccess;
}


}  // namespace fx
