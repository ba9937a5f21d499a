// fx_common.h — data shared by the host side and the gfx950 kernels of the
// Fire-X solver: the per-step constant block and the storage formats.
#pragma once

#include <stdint.h>

#include "fx_math.h"

// The 44-float uniform block of the reference (FireXEngine.ts:2047-2099;
// WGSL struct Parameters, src/sim/firex/shaders.ts:2-14).  Slot indices:
enum FxUniformSlot {
  FXU_GRID_X = 0, FXU_GRID_Y = 1, FXU_GRID_Z = 2, FXU_CELLS = 3,
  FXU_DT = 4, FXU_SIM_TIME = 5, FXU_FRAME = 6, FXU_TIME_STEP = 7,
  FXU_FUEL_RATE = 8, FXU_OXYGEN_RATE = 9, FXU_HEAT_EFFICIENCY = 10, FXU_VORTICITY = 11,
  FXU_WATER_FLOW = 12, FXU_SPRAY_ANGLE = 13, FXU_WATER_ENABLED = 14, FXU_AMBIENT = 15,
  FXU_CAMERA_YAW = 16, FXU_CAMERA_PITCH = 17, FXU_CAMERA_DISTANCE = 18, FXU_ASPECT = 19,
  FXU_CANVAS_W = 20, FXU_CANVAS_H = 21, FXU_RAY_STEPS = 22, FXU_PARTICLES = 23,
  FXU_VIEW_MODE = 24, FXU_AIM_HEIGHT = 25, FXU_NOZZLE = 26, FXU_EXPOSURE = 27,
  FXU_BURNER_SIZE = 28, FXU_BURNER_DEPTH = 29, FXU_SOURCE_THICKNESS = 30, FXU_FIRE_POWER = 31,
  FXU_DOMAIN_X = 32, FXU_DOMAIN_Y = 33, FXU_DOMAIN_Z = 34,
  FXU_SOURCE_LIFT = 36, FXU_BUOYANCY = 37, FXU_PERSISTENCE = 38, FXU_VIEW_ZOOM = 39,
  FXU_ENTRAINMENT = 40, FXU_FRONT_DEFINITION = 41, FXU_CINEMATIC = 42, FXU_OPTICAL_DETAIL = 43,
  FXU_COUNT = 44
};

#define FX_MAX_EMITTERS 8

// Per-step constants handed to the kernels by value.  Every member is either
// a uniform slot or a subexpression of the reference kernels that depends on
// uniforms only; the host computes the latter with the same f32 operations,
// in the same order, as the per-cell code of shaders.ts would.
struct FxStep {
  int32_t nx, ny, nz;
  int32_t has_liquid;
  float gx, gy, gz;            // params.grid.xyz
  float dt, time;              // params.time.x, params.time.y
  float mx, my, mz;            // cellMetric() = grid / domainExtent()
  float ex, ey, ez;            // domainExtent() = max(domain, 0.001)
  float ambient;               // params.water.w (300 K)
  float chem_x, chem_y, chem_z, chem_w;
  float mix_velocity;          // clamp(dt * 0.32, 0, 0.12)
  float mix_temperature;       // clamp(dt * 0.18, 0, 0.08)
  float mix_species;           // clamp(dt * 0.12, 0, 0.06)
  float fire_power, power_root, thermal_power;
  float swirl_tz, swirl_tx;    // time * 2.3, time * 1.9
  float lift_clamp;            // clamp(fire.x, 0, 4)
  float entrainment, ambient_oxygen;
  float inv_ambient_oxygen;    // 1 / ambient_oxygen (f32)
  float heat_release;          // 2150 + 1450 * chemistry.z
  float max_dt;                // max(dt, 0.0001)
  float inv_max_dt;            // 1 / max_dt (f32)
  float vortex_k;              // 0.00135 * chemistry.w * dt
  float turbulence_base;       // 0.075 + 0.16 * chemistry.w
  float tt31, tt53, tt41, tt27, tt47;  // time * {3.1, 5.3, 4.1, 2.7, 4.7}
  float buoyancy_clamp;        // clamp(fire.y, 0, 4)
  float spray_x, spray_y, spray_z;     // normalize(sprayTarget - nozzleOrigin)
  float persistence;           // clamp(fire.z, 0.4, 3)
  float ambient4;              // pow(water.w, 4)
  float radiation_denominator; // 1005 * persistence
  float inv_radiation_denominator;  // 1 / radiation_denominator (f32)
  float decay_products, decay_vapor;   // exp(-dt * 0.075 / q), exp(-dt * 0.14 / q)
  // reference emitter (shaders.ts:101-133)
  float em_cx, em_cy, em_cz;
  float em_rx2, em_rz2, em_vdenom, em_rx_safe, em_rz_safe, em_cinematic;
  float et52, et44, et17, et13, et37;  // time * {5.2, 4.4, 1.7, 1.3, 3.7}
  // seeded emitters (addition)
  int32_t seeded_count;
  float seeded[FX_MAX_EMITTERS][5];    // cx cy cz radius amplitude (cells)
  float jitter, velocity_noise;
  uint64_t seed, frame;
  // pressure stencil
  float wx, wy, wz, jacobi_inv;        // metric^2, 1 / (2 (wx + wy + wz))
};

// Per-step constants of the droplet passes (uniform slots of
// FIREX_PARTICLE_WGSL, shaders.ts:514-530); every derived value is computed
// per thread, in WGSL order.
struct FxDroplets {
  uint32_t count;              // u32(params.render.w)
  int32_t nx, ny, nz;
  float dt, frame;             // params.time.x, params.time.z
  float gx, gy, gz;            // params.grid.xyz
  float ex, ey, ez;            // domainExtent() = max(domain, 0.001)
  float water_flow, spray_angle, water_enabled;  // params.water.xyz
  float aim, nozzle;           // params.controls.yz
};

// Per-frame constants of the volume ray marcher (uniform slots read by
// volumeFragment, qualityShaders.ts:407-575, and the output size).
struct FxRender {
  int32_t width, height;       // output pixels (render.xy)
  int32_t steps;               // min(u32(render.z), 384)
  int32_t nx, ny, nz;
  float gx, gy, gz;            // params.grid.xyz
  float ex, ey, ez;            // domainExtent()
  float yaw, pitch, cam_distance, aspect;  // params.camera
  float render_z;              // params.render.z (ray steps as f32)
  float view_mode;             // params.controls.x
  float exposure;              // params.controls.w
  float front_definition;      // params.mixing.y
  float cinematic;             // params.mixing.z
  float detail_target;         // params.mixing.w
  float time;                  // params.time.y
  float chem_y, chem_z;        // params.chemistry.yz
  float zoom;                  // params.fire.w
  float canvas_height;         // params.render.y
  float nozzle;                // params.controls.z
  int32_t bnx, bny, bnz;       // occupancy bricks per axis (kFxBrick texels each)
};

// Edge of an occupancy brick of the media volume, in texels.
#define kFxBrick 8

// Droplet streak of one droplet in NDC (dropletVertex): quad corner (a, b)
// sits at (cx, cy) + normal * a * radius + tangent * b * half, normal =
// (-ty, tx).
struct FxSprite {
  float cx, cy, tx, ty, radius, half, mass, pad;
};
// Streak rasterizer bins: a bin is kFxSpriteBin x kFxSpriteBin pixels, a
// tile (one block) kFxSpriteTile x kFxSpriteTile.
#define kFxSpriteBin 128
#define kFxSpriteTile 16
#define kFxSpriteMaxBins ((16384 / kFxSpriteBin) * (16384 / kFxSpriteBin))

// Edge of the periodic detail volume (FireXEngine.ts:1622) and the WGSL
// constant 1.0 / 2.2 as an f32 (a const-expression, evaluated exactly and
// rounded once).
#define kFxDetailSize 64
#define kFxInverseGamma ((float)(1.0 / 2.2))

#if defined(__HIP__)
// Keeps an f32 value as computed.  Without it the compiler may fold
// "f32 multiply, then convert to f16" into one mixed-precision FMA that
// rounds once, straight to f16, where the reference rounds to f32 first and
// then to f16 (pressureShaders.ts:115-116).  The empty asm only pins the
// register; it emits no instruction.
__device__ static inline float fx_keep_f32(float v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Weighted neighbour sum of a Jacobi update, wx sx + wy sy + wz sz with each
// multiply-add contracted into one fused multiply-add (v_fma_f32, rounded
// once): fma(wz, sz, fma(wy, sy, wx sx)).  Every Jacobi kernel and the f32
// oracle use this form.
__device__ static inline float jacobi_sum(float wx, float sx, float wy, float sy, float wz, float sz) {
  return __builtin_fmaf(wz, sz, __builtin_fmaf(wy, sy, wx * sx));
}
#endif

// Floats of simulate's per-step coordinate tables for an nx x ny x nz grid
// (layout: SimTab in firex_kernels.hip).
static inline uint64_t fx_sim_table_floats(const int32_t d[3]) {
  const uint64_t nx = (uint64_t)d[0], ny = (uint64_t)d[1], nz = (uint64_t)d[2];
  return 5 * nx * nz + 5 * ny + nx + nz + ny * nz + nx * ny;
}

// Counter-based seeded noise: splitmix64 finaliser on (seed, frame, cell, stream).
#if defined(__HIP__)
__host__ __device__
#endif
static inline uint64_t fx_mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Field checksum (an addition, for auditable captures): the stored bytes
// read as little-endian u32 words w_i (a final partial word zero-padded),
// checksum = sum over i of fx_mix64((i << 32) | w_i), mod 2^64.  Integer
// addition commutes, so any summation order gives the same value; a change
// of any word changes its term (fx_mix64 is a bijection).  Fields hold at
// most 2^32 words.
#if defined(__HIP__)
__host__ __device__
#endif
static inline uint64_t fx_checksum_term(uint64_t i, uint32_t w) { return fx_mix64((i << 32) | (uint64_t)w); }

#if defined(__HIP__)
__host__ __device__
#endif
static inline float fx_seeded_uniform(uint64_t seed, uint64_t frame, uint64_t cell, uint32_t stream) {
  uint64_t k = fx_mix64(seed ^ fx_mix64(frame));
  k = fx_mix64(k ^ (cell * 4ull + stream));
  return (float)(uint32_t)(k >> 40) * 5.9604644775390625e-08f;
}

// bf16 round-to-nearest-even (NaN becomes the canonical quiet NaN).
#if defined(__HIP__)
__host__ __device__
#endif
static inline uint16_t fx_float_to_bf16_bits(float v) {
  uint32_t u = fxm_float_to_bits(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

#if defined(__HIP__)
__host__ __device__
#endif
static inline float fx_bf16_bits_to_float(uint16_t b) { return fxm_bits_to_float((uint32_t)b << 16); }
