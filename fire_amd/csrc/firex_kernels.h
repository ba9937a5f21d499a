// firex_kernels.h — host-side launchers of the gfx950 Fire-X kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fx_common.h"

enum { FX_STORAGE_F32 = 0, FX_STORAGE_F16 = 1, FX_STORAGE_BF16 = 2 };

namespace fx {

struct SimulateArgs {
  const float4* flow_in;
  const void* species_in;
  float4* flow_out;
  void* species_out;
  float* reaction;
  const uint32_t* liquid;  // may be null when st.has_liquid == 0
  float* tables;           // sim_table_bytes(dims) of scratch: per-step coordinate tables
  FxStep st;
};

// Bytes of the per-step coordinate tables simulate needs for a grid.
size_t sim_table_bytes(const int32_t dims[3]);

void launch_initialize(int storage, float4* flow, void* species, uint32_t cells, float ambient, float oxygen,
                       hipStream_t s);
void launch_fill_u32(void* p, uint64_t words, uint32_t value, hipStream_t s);
void launch_copy_u128(void* dst, const void* src, uint64_t n16, hipStream_t s);
void launch_simulate(int storage, const SimulateArgs& a, hipStream_t s);
void launch_divergence(const float4* flow, float* div, const FxStep& st, hipStream_t s);
// packFields into out: three consecutive volumes of 4 x f16 per cell
// (flow, chemistry, media); liquid may be null (zero field).
void launch_pack_fields(int storage, const float4* flow, const void* species, const float* reaction,
                        const uint32_t* liquid, const float* div, void* out, const FxStep& st, int divergence_view,
                        hipStream_t s);
void launch_jacobi(int f16, const void* pin, void* pout, const float* div, const FxStep& st, hipStream_t s);
void launch_project(int f16, const float4* fin, const void* p, float4* fout, const FxStep& st, hipStream_t s);
void launch_reduce_field(int storage, int kind, int component, const void* data, uint32_t cells, double* partial,
                         double* out, hipStream_t s);
void launch_reduce_residual(int f16, const void* p, const float* div, const FxStep& st, double* partial, double* out,
                            hipStream_t s);
void launch_reduce_flow_divergence(const float4* flow, const FxStep& st, double* partial, double* out, hipStream_t s);

// SPH droplet passes (one wave per particle for the neighbour sums).
void launch_sph_density(const float4* particles, float4* density, const FxDroplets& d, hipStream_t s);
void launch_sph_integrate(const float4* particles_in, float4* particles_out, const float4* density,
                          const float4* flow, const FxDroplets& d, hipStream_t s);
void launch_liquid_deposit(const float4* current, const float4* previous, uint32_t* liquid, const FxDroplets& d,
                           hipStream_t s);

// Field checksum (fx_checksum_term summed over the words of `bytes` stored
// bytes; data may be null for an all-zero field) into *out (device memory),
// through up to 1024 block partials in `partial`.
void launch_checksum(const void* data, uint64_t bytes, uint64_t* partial, uint64_t* out, hipStream_t s);

// The volume ray marcher: packed = the three render volumes (4 x f16 per
// cell, flow / chemistry / media), detail = 64^3 rgba8 texels, out = width x
// height rgba16f pixels, row 0 at the top.
void launch_render_volume(const void* packed, const uint32_t* detail, const float* bricks, void* out,
                          const FxRender& r, hipStream_t s);
// Occupancy bricks of the media volume (bnx x bny x bnz floats): the largest
// media component over each brick's texels and one texel beyond, read by
// the beauty view's empty-space skipping.
void launch_render_bricks(const void* packed, float* bricks, const FxRender& r, hipStream_t s);
// The presentation after the ray march: with cinematic, the quarter-resolution
// bloom (bloom: ceil(w/4) x ceil(h/4) rgba16f) and the tone-map composite;
// without, the scene image stored as 8 bits.  frame: w x h rgba8 pixels.
void launch_render_post(const void* scene, void* bloom, int width, int height, float exposure, int cinematic,
                        uint32_t* frame, hipStream_t s);

// Droplet streaks (dropletVertex / dropletFragment) over an image: particles
// = the live droplet buffer (count records of 8 f32), target = width x height
// rgba16f (the scene image) or, with unorm8, rgba8 (the standard canvas).
// Nothing is drawn outside the beauty and liquid views.  work:
// sprite_work_bytes(count) bytes.
uint64_t sprite_work_bytes(int count);
void launch_droplet_streaks(const void* particles, int count, const FxRender& r, void* target, int unorm8,
                            void* work, hipStream_t s);

constexpr uint32_t kReducePartials = 1024;
// Blocks (= partials) an FP64 reduction over `cells` cells launches: at most
// kReducePartials for any cells < 2^32.
uint32_t reduce_block_count(uint32_t cells);
// In-place sqrt of one double (the L2 norms of an on-device reduction).
void launch_sqrt_f64(double* value, hipStream_t s);

}  // namespace fx
