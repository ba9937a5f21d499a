// firex_jacobi.h — temporally blocked Jacobi sweeps and their autotuner.
#pragma once

#include <hip/hip_runtime.h>

#include "fx_common.h"

namespace fx {

// Launch geometry of the temporally blocked sweep kernel.
enum { JACOBI_KERNEL_AUTO = 0, JACOBI_KERNEL_SHIFT = 1, JACOBI_KERNEL_RING = 2 };

struct JacobiPlan {
  int32_t sweeps_per_launch = 4;  // K: sweeps advanced per launch
  int32_t rows = 4;               // R: y rows per thread (RING: 4, 6 or 8)
  int32_t waves = 8;              // W: waves per block (stacked in y)
  int32_t chunk_z = 16;           // output planes per block
  int32_t kernel = JACOBI_KERNEL_SHIFT;  // SHIFT: any grid, K <= 8; RING: nx >= 64, ny >= W * R, K <= 5
  int32_t prefetch = 0;           // RING: load plane t + 1 one iteration before its first use
};
// RING block shapes: 8 waves of 4, 6 or 8 rows (32-, 48- or 64-row tiles,
// K <= 5), or 16 waves of 2 rows (32-row tiles, K <= 4, no prefetch: fewer
// registers per thread, twice the waves per CU, per-lane row addressing and
// with it the narrow wall tile of RingLayout).
constexpr int kRingWideWaves = 16;
constexpr int kRingWideMaxSweeps = 4;

// kernel: JACOBI_KERNEL_* (AUTO picks RING where the grid allows it);
// sweeps_per_launch, chunk_z and tile_height (RING: 32, 48 or 64 rows per
// tile): 0 = default; prefetch (RING, K <= 3): 0 = default (off), 1 = off,
// 2 = on.
// ring_waves (RING): 0 or 8 = 8 waves per block, 16 = the 16-wave shape.
JacobiPlan make_jacobi_plan(const int32_t dims[3], int32_t sweeps_per_launch, int32_t chunk_z = 0,
                            int32_t kernel = JACOBI_KERNEL_AUTO, int32_t tile_height = 0, int32_t prefetch = 0,
                            int32_t ring_waves = 0);

// One launch of the temporally blocked kernel: `sweeps` sweeps from
// buffers[src] (or from the zero field, without reading it) into
// buffers[1 - src].
struct JacobiLaunch {
  int32_t sweeps;
  int32_t src;
  bool zero_start;
};

// Blocks one launch of the plan runs on a grid of dims.
int jacobi_plan_blocks(const JacobiPlan& plan, const int32_t dims[3]);

// f16 selects f16 pressure storage: every sweep's result is rounded to f16,
// as pressureShaders.ts:116 stores it.
void launch_jacobi_tb(const JacobiPlan& plan, int f16, const JacobiLaunch& l, void* const buffers[2],
                      const float* div, const FxStep& st, hipStream_t s);

// Host-side check of a plan against the kernel's assumptions.
bool jacobi_plan_valid(const JacobiPlan& plan);

// Tiles of a RING launch.  Full tiles are 64 columns wide and output
// 64 - 2K columns each, placed from the left wall (the last at the right
// wall).  With the 16-wave shape, where the columns left over at the right
// wall number at most 32 - 2K, one 32-column tile covers them instead
// (narrow = 1): a block of that tile runs two y tiles side by side, lanes
// 0-31 on y tile 2q and lanes 32-63 on y tile 2q + 1 (pair q), so that no
// half wave computes columns that feed no output.  At 192 columns and
// K = 4: three full tiles and a narrow one, 224 columns per row computed
// instead of 256.  Blocks: (tiles_x * tiles_y + pairs) per z chunk.
struct RingLayout {
  int32_t tiles_x;  // full 64-column tiles
  int32_t tiles_y;
  int32_t narrow;   // 1: a narrow tile at the right wall
  int32_t pairs;    // y-tile pairs of the narrow tile ((tiles_y + 1) / 2)
};

// One launch of the RING kernel over several grids: each entry is one
// grid's launch (its buffers, dims, stencil and chunk height), and the
// launch's blocks are numbered entry after entry, so every grid gets exactly
// the blocks, and the result, of its own launch.  The table travels in the
// kernel arguments (no device allocation).
constexpr int kRingBatchMax = 16;
struct RingBatchEntry {
  const void* pin;  // source pressure (not read when zero_start)
  void* pout;
  const float* div;
  int32_t nx, ny, nz;
  float wx, wy, wz, inv;
  int32_t chunk_z;
  int32_t zero_start;
  RingLayout layout;     // set by launch_jacobi_ring_batch
  uint32_t block_begin;  // set by launch_jacobi_ring_batch
};
struct RingBatch {
  int32_t count;
  uint32_t total_blocks;
  RingBatchEntry entry[kRingBatchMax];
};

// Whether a plan runs in batched launches: the 16-wave RING shape.
bool jacobi_plan_batchable(const JacobiPlan& plan);

// One batched launch of `sweeps` (<= plan.sweeps_per_launch) sweeps; every
// entry's grid must satisfy the plan (RING: nx >= 64, ny >= 32, nz >= K + 2,
// fields under 2 GiB).  Fills the entries' tiles and block offsets.
void launch_jacobi_ring_batch(const JacobiPlan& plan, int f16, int sweeps, RingBatch& batch, hipStream_t s);

// Autotuner: times candidate plans for a grid of dims on the current device
// and returns the fastest in *out (and its time per sweep in microseconds).
// sweeps_per_launch / chunk_z / tile_height / prefetch > 0 and kernel !=
// AUTO are kept fixed; the rest is searched.  buffers (two pressure fields in the given storage) and
// div are scratch: their contents are overwritten.  Results are cached per
// device, shape, storage and fixed parameters, so a second solver of the
// same shape does not time again.  Returns a hipError_t (0 on success).
int tune_jacobi_plan(const int32_t dims[3], int f16, int32_t sweeps_per_launch, int32_t chunk_z, int32_t kernel,
                     int32_t tile_height, int32_t prefetch, int32_t ring_waves, void* const buffers[2], float* div,
                     hipStream_t s, JacobiPlan* out, float* us_per_sweep);

}  // namespace fx
