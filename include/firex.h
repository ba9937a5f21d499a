/*
 * firex.h — C ABI of the MI355X (gfx950) Fire-X solver library, libfirex_hip.so.
 *
 * The library is a drop-in for the compute half of the reference engine
 * (class FireXEngine, src/sim/firex/FireXEngine.ts:370): it owns every device
 * buffer of one plume, advances it with the Fire-X fixed step and hands fields
 * back to the caller, including the SPH droplet passes that feed its liquid
 * field.  Rendering and video export are not part of it.
 *
 * Conventions
 *   - Every function returning int returns FIREX_OK (0) or a FIREX_ERR_* code;
 *     firex_last_error() then holds a human-readable message (thread-local).
 *   - Fields are laid out x fastest: index = x + nx * (y + ny * z)
 *     (cellIndex, src/sim/firex/shaders.ts:47-51).  Vector fields hold 4
 *     components per cell.
 *   - Plain pointers and sizes only; the caller owns every host buffer.
 *   - One solver handle must not be used from two threads at once
 *     (the reference is single-threaded, FireXEngine.ts:1882-1889).
 */
#ifndef FIREX_H
#define FIREX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FIREX_ABI_VERSION 1

typedef struct firex_solver firex_solver;

/* ---- status codes ------------------------------------------------------ */
enum {
  FIREX_OK = 0,
  FIREX_ERR_INVALID = 1,        /* bad argument (thrown Error in the reference) */
  FIREX_ERR_UNSUPPORTED = 2,    /* grid fails preflight (gridConfiguration.ts:297-350) */
  FIREX_ERR_BUDGET = 3,         /* transactional peak above the memory guard (FireXEngine.ts:2823-2853) */
  FIREX_ERR_OUT_OF_MEMORY = 4,  /* device allocation failed; no partial state (FireXEngine.ts:1891-1915) */
  FIREX_ERR_DEVICE = 5,         /* no usable HIP device / HIP call failed */
  FIREX_ERR_ABORTED = 6,        /* caller's abort flag was set (FireXEngine.ts:2571-2574) */
  FIREX_ERR_DEVICE_LOST = 7     /* sticky: a device fault was seen earlier (FireXEngine.ts:538-542) */
};

/* ---- fields ------------------------------------------------------------ */
enum {
  FIREX_FIELD_FLOW = 0,            /* vec4 f32 (u, v, w, T): flow[0] after a step        */
  FIREX_FIELD_SPECIES = 1,         /* vec4 (fuel, O2, products, vapor): species[current]  */
  FIREX_FIELD_PRESSURE = 2,        /* f32 primary pressure (warm-started)                 */
  FIREX_FIELD_CORRECTION = 3,      /* f32 last residual-pressure solution                 */
  FIREX_FIELD_DIVERGENCE = 4,      /* f32 divergence buffer (residual RHS after a step)   */
  FIREX_FIELD_REACTION = 5,        /* f32 heat-release diagnostic                         */
  FIREX_FIELD_LIQUID = 6,          /* u32 fixed-point liquid (x 4096), input of simulate  */
  FIREX_FIELD_FLOW_PREDICTED = 7,  /* vec4 f32 flow[1]                                    */
  /* render volumes written by firex_pack_fields (4 x f16 per cell; read as f32/f64) */
  FIREX_FIELD_PACKED_FLOW = 8,     /* (u, v, w, clamp((T - 300) / 2600, 0, 1))            */
  FIREX_FIELD_PACKED_CHEMISTRY = 9,/* max(species, 0)                                     */
  FIREX_FIELD_PACKED_MEDIA = 10,   /* (reaction, soot, liquid, display value)             */
  /* per-particle fields: particle_count records (not cells) */
  FIREX_FIELD_PARTICLES = 11,      /* 8 f32: positionAge (domain fractions, age s),
                                      velocityMass (m/s, mass); the current buffer       */
  FIREX_FIELD_PARTICLE_DENSITY = 12,/* 4 f32: density, near density, pressure, near p.   */
  /* image of the last firex_render_volume: width * height pixels of 4 x f16
     (rgba16float), row 0 at the top */
  FIREX_FIELD_SCENE = 13,
  /* the last firex_render_frame: width * height rgba8 pixels, row 0 at the top */
  FIREX_FIELD_FRAME = 14
};

/* Host dtypes of firex_read_field / firex_write_field.  FIREX_DTYPE_RAW
 * (read only) copies the field's stored bytes verbatim (firex_get_field_layout
 * gives their format and size). */
enum { FIREX_DTYPE_F32 = 0, FIREX_DTYPE_F64 = 1, FIREX_DTYPE_U32 = 2, FIREX_DTYPE_RAW = 3 };

/* Storage formats of a field on the device. */
enum { FIREX_KIND_F32 = 0, FIREX_KIND_F16 = 1, FIREX_KIND_BF16 = 2, FIREX_KIND_U32 = 3, FIREX_KIND_U8 = 4 };

typedef struct firex_field_layout {
  int32_t kind;          /* FIREX_KIND_* of each component                    */
  int32_t components;    /* per record                                         */
  int64_t records;       /* cells, or particles for the per-particle fields    */
  int64_t stored_bytes;  /* records * components * element size               */
} firex_field_layout;

/* Storage of the advected species field. Velocity, temperature, pressure and
 * all arithmetic stay f32. */
enum { FIREX_STORAGE_F32 = 0, FIREX_STORAGE_F16 = 1, FIREX_STORAGE_BF16 = 2 };

/* Pressure path on heavy grids (>= 64 Mi cells, gridConfiguration.ts:105). */
enum {
  FIREX_HEAVY_AUTO = 0,  /* reference selection: f16 pressure storage at >= 64 Mi cells */
  FIREX_HEAVY_OFF = 1,   /* always the warm-started f32 path                           */
  FIREX_HEAVY_F16 = 2,   /* force f16 pressure storage (pressureShaders.ts)            */
  FIREX_HEAVY_F32 = 3    /* force the compact f32 path: shared buffers, cold start     */
};

/* Kernel plan.  Both plans give bit-identical fields. */
enum {
  FIREX_PLAN_REFERENCE = 0,  /* one launch per reference pass, 2N + 6 per step         */
  FIREX_PLAN_FUSED = 1       /* temporally blocked Jacobi, fused projection/divergence */
};

typedef struct firex_config {
  int32_t dims[3];               /* nx, ny, nz                                       */
  int32_t pressure_iterations;   /* primary Jacobi sweeps N                          */
  int32_t correction_iterations; /* residual Jacobi sweeps                           */
  double time_step;              /* configured fixed step (QualityConfiguration)     */
  int32_t maximum_substeps;      /* FixedStepClock steps per live frame              */
  int32_t species_storage;       /* FIREX_STORAGE_*                                  */
  int32_t heavy_mode;            /* FIREX_HEAVY_*                                    */
  int32_t plan;                  /* FIREX_PLAN_*                                     */
  int32_t cinematic;             /* uniform mixing.z: 1 for offline tiers / cinematic */
  int32_t device;                /* HIP device ordinal                               */
  uint64_t memory_budget_bytes;  /* guard on current + replacement bytes; 0 = none   */
  int32_t jacobi_sweeps_per_launch; /* FUSED plan temporal depth; 0 = tuned default  */
  int32_t capture_graphs;        /* 1: everything after simulate runs as a captured
                                    hipGraph (one launch per step for that part)     */
  int32_t jacobi_chunk_z;        /* FUSED plan: z planes per block; 0 = tuned default */
  int32_t jacobi_kernel;         /* FUSED plan: 0 auto, 1 register-shift kernel (any
                                    grid), 2 register-ring kernel (nx >= 64, ny >= 32) */
  int32_t jacobi_autotune;       /* FUSED plan: 1 = time candidate Jacobi plans on the
                                    device at creation and keep the fastest; fields
                                    above left at 0 are searched (cached per shape)  */
  int32_t particle_count;        /* SPH droplets of the resource set (384-4096 by
                                    tier, FireXEngine.ts:187-277); 0 = none         */
  int32_t ray_steps;             /* ray-march samples per pixel (raySteps by tier,
                                    80-384; firex_render_volume caps it at 384)    */
  int32_t jacobi_tile_height;    /* FUSED plan, ring kernel: y rows per tile, 32, 48
                                    or 64 (needs ny >= it); 0 = tuned default      */
  int32_t jacobi_prefetch;       /* FUSED plan, ring kernel: 0 = tuned default, 1 =
                                    off, 2 = load each plane one iteration early
                                    (K <= 3)                                       */
  int32_t jacobi_ring_waves;     /* FUSED plan, ring kernel: waves per block, 8 or
                                    16 (16 waves of 2 rows: 32-row tiles, K <= 4);
                                    0 = tuned default                              */
} firex_config;

/* Seeded Gaussian emitter (an addition: replaces the reference emitter()). */
typedef struct firex_emitter {
  float center[3];   /* cell units; cell i spans [i, i+1)                  */
  float radius;      /* cells; source = amplitude * exp(-|p - c|^2 / r^2)  */
  float amplitude;
} firex_emitter;

typedef struct firex_info {
  int32_t dims[3];
  int64_t cell_count;
  double simulation_time;       /* seconds of simulated time so far          */
  int64_t frame_index;          /* steps taken since the last reset           */
  int32_t species_index;        /* which species buffer is current            */
  int32_t pressure_index;       /* which pressure buffer is current           */
  int32_t heavy;                /* 1 if the >= 64 Mi cell path is active      */
  int32_t f16_pressure;         /* 1 if pressure is stored as f16             */
  int32_t species_storage;
  int32_t plan;
  uint64_t device_bytes;        /* bytes this solver holds on the device      */
  int32_t jacobi_kernel;        /* FUSED plan in use: kernel, sweeps per      */
  int32_t jacobi_sweeps_per_launch; /* launch, z chunk and prefetch distance   */
  int32_t jacobi_chunk_z;
  int32_t jacobi_prefetch;
  float jacobi_tuned_us_per_sweep; /* autotuner's time per sweep; 0 = untuned */
  int32_t particle_count;       /* SPH droplets (firex_config.particle_count) */
  int32_t jacobi_tile_height;   /* FUSED plan in use: y rows per tile          */
  int32_t jacobi_ring_waves;    /* FUSED plan in use: waves per ring block     */
} firex_info;

typedef struct firex_memory_estimate {
  int64_t cell_count;
  int64_t grid_bytes;
  int64_t particle_bytes;
  int64_t fixed_bytes;
  int64_t total_bytes;
  int64_t largest_storage_buffer_bytes;
} firex_memory_estimate;

typedef struct firex_device_limits {
  int64_t max_texture_dimension_3d;
  int64_t max_storage_buffer_binding_size;
  int64_t max_buffer_size;
} firex_device_limits;

/* Preflight issue bits (codes of preflightFireXGrid, gridConfiguration.ts:297-350). */
enum {
  FIREX_ISSUE_INVALID_DIMENSION = 1u << 0,
  FIREX_ISSUE_UNALIGNED_DIMENSION = 1u << 1,
  FIREX_ISSUE_TEXTURE_DIMENSION_LIMIT = 1u << 2,
  FIREX_ISSUE_STORAGE_BINDING_LIMIT = 1u << 3,
  FIREX_ISSUE_BUFFER_SIZE_LIMIT = 1u << 4,
  FIREX_ISSUE_MEMORY_BUDGET = 1u << 5
};

/* FP64 reductions (additions; the reference computes none). */
enum {
  FIREX_REDUCE_SUM_FUEL = 0,
  FIREX_REDUCE_SUM_OXYGEN = 1,
  FIREX_REDUCE_SUM_PRODUCTS = 2,
  FIREX_REDUCE_SUM_VAPOR = 3,
  FIREX_REDUCE_SUM_TEMPERATURE = 4,
  FIREX_REDUCE_SUM_REACTION = 5,
  FIREX_REDUCE_RESIDUAL_L2 = 6,       /* ||div - L(corr)||_2 over pressure unknowns  */
  FIREX_REDUCE_FLOW_DIVERGENCE_L2 = 7 /* ||D(flow)||_2 of the current flow field      */
};

/* ---- library ------------------------------------------------------------ */
const char* firex_last_error(void);
int firex_abi_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int firex_device_count(void);

/* ---- configuration (host only) ---------------------------------------- */
/* Defaults: balanced live tier (FireXEngine.ts:199-208), f32 storage,
 * reference plan, heavy auto, cinematic 0, device 0, no guard. */
void firex_config_init(firex_config* config);
/* Fill dims, iterations, time_step, maximum_substeps and cinematic from a tier
 * name: live "low" "balanced" "high" "maximum" "cinematic"
 * (FIREX_QUALITY, FireXEngine.ts:187-238) or offline "hd" "qhd" "uhd"
 * (FIREX_OFFLINE_QUALITY, FireXEngine.ts:240-277). */
int firex_config_from_tier(firex_config* config, const char* tier);

/* estimateFireXGridMemory, gridConfiguration.ts:272-291. */
int firex_estimate_memory(const int32_t dims[3], int64_t particle_count, firex_memory_estimate* out);
/* preflightFireXGrid, gridConfiguration.ts:297-350.  memory_budget_bytes < 0
 * disables the aggregate guard (the reference's null); *issues receives the
 * FIREX_ISSUE_* bits; returns FIREX_OK when no bit is set. */
int firex_preflight(const int32_t dims[3], const firex_device_limits* limits, int64_t particle_count,
                    int64_t memory_budget_bytes, uint32_t* issues);
/* splitOfflineSimulationDelta, src/sim/offlineTimeline.ts:5-20.  Writes at most
 * cap steps; *count receives the full step count. */
int firex_split_offline_delta(double delta_seconds, double maximum_step_seconds, double* steps, int32_t cap,
                              int32_t* count);
/* FixedStepClock.advance, src/sim/FixedStepClock.ts:16-33: adds the bounded
 * delta to *accumulator and returns the number of steps to run; -1 (and
 * FIREX_ERR_INVALID in firex_last_error) for an invalid clock. */
int32_t firex_fixed_step_clock_advance(double step_seconds, int32_t maximum_steps, double* accumulator,
                                       double delta_seconds);
/* The clamp setParameter applies to one numeric key (FireXEngine.ts:937-1032):
 * *out receives the stored value given the previous one; FIREX_ERR_INVALID
 * for a key the solver does not know. */
int firex_clamp_parameter(const char* name, double value, double previous, double* out);
/* resolveFireXOfflineDimensions, FireXEngine.ts:279-290: per-axis maximum of
 * the tier preset and the override (override_dims may be NULL). */
int firex_resolve_offline_dims(const char* tier, const int32_t* override_dims, int32_t out[3]);

/* ---- solver lifetime --------------------------------------------------- */
/* Allocate and initialise a plume (createResourceSet + initializeFields,
 * FireXEngine.ts:1536-1974).  Fails without partial state. */
int firex_create(const firex_config* config, firex_solver** out);
void firex_destroy(firex_solver* solver);
/* Transactional reconfiguration of a live solver: the resource half of
 * setGridDimensions, setOfflineRenderTier and setQuality
 * (FireXEngine.ts:799-831, 863-927) through rebuildResourcesSafely
 * (:1891-1915).  If the new configuration keeps the buffer set (same grid,
 * particle count, species storage, pressure path, plan and Jacobi settings),
 * only the per-step settings change (iterations, time step and substeps,
 * cinematic flag, ray steps, budget), the fields are kept and *rebuilt = 0.
 * Otherwise the estimated bytes of the current and the replacement set
 * (estimateFireXGridMemory of each) must fit memory_guard_bytes (0 = no
 * guard; FIREX_ERR_BUDGET with the reference's message, assertGridSupported
 * :2823-2853), the replacement is built while the current set stays live,
 * installed with installResourceSet's reset (:1917-1931: parameters kept;
 * clock, time, indices and fields started again) and only then is the old set
 * freed; *rebuilt = 1.  On any failure the solver is left exactly as it was.
 * The device cannot change; plumes on a serving stream cannot be rebuilt. */
int firex_reconfigure(firex_solver* solver, const firex_config* config, uint64_t memory_guard_bytes,
                      int32_t* rebuilt);
/* reset(), FireXEngine.ts:843-852. */
int firex_reset(firex_solver* solver);

/* setParameter(name, number), FireXEngine.ts:937-1032: same keys and clamps;
 * non-finite values keep the previous value; unknown keys are ignored. */
int firex_set_parameter(firex_solver* solver, const char* name, double value);
/* setParameter(name, boolean) for "waterEnabled". */
int firex_set_parameter_bool(firex_solver* solver, const char* name, int32_t value);
int firex_get_parameter(firex_solver* solver, const char* name, double* value);
/* Seeded emitters (count 0 restores the reference emitter; at most 8). */
int firex_set_emitters(firex_solver* solver, const firex_emitter* emitters, int32_t count, float jitter,
                       float velocity_noise, uint64_t seed);

/* ---- stepping ----------------------------------------------------------- */
/* count solver steps of dt seconds each (writeUniforms(dt) +
 * encodeSimulationStep, FireXEngine.ts:2043-2276); enqueued, not waited for. */
int firex_step(firex_solver* solver, double dt, int32_t count);
/* renderOfflineFrame(delta, present = false), FireXEngine.ts:643-697: splits
 * delta into exact steps no longer than time_step, runs them and waits.
 * abort_flag may be NULL; when *abort_flag != 0 before a step, returns
 * FIREX_ERR_ABORTED. */
int firex_render_offline_frame(firex_solver* solver, double delta_seconds, const volatile int32_t* abort_flag);
/* frame(t, dt) simulation half, FireXEngine.ts:1064-1101 with FixedStepClock
 * (src/sim/FixedStepClock.ts:16-33); *steps_taken may be NULL. */
int firex_frame(firex_solver* solver, double delta_seconds, int32_t* steps_taken);
/* Wait for all enqueued work. */
int firex_sync(firex_solver* solver);

/* ---- fields -------------------------------------------------------------- */
/* Copy a field to host memory (waits for enqueued steps).  bytes must equal
 * cells * components * sizeof(dtype), or particle_count * components *
 * sizeof(dtype) for the per-particle fields. */
int firex_read_field(firex_solver* solver, int32_t field, void* dst, size_t bytes, int32_t dtype);
/* Overwrite a field from host memory (seeding and tests).  Writing
 * FIREX_FIELD_PARTICLES fills both particle buffers, as
 * writeParticleInitialState does (FireXEngine.ts:2030-2031). */
int firex_write_field(firex_solver* solver, int32_t field, const void* src, size_t bytes, int32_t dtype);
/* Stored format and size of a field (for FIREX_DTYPE_RAW reads and dumps). */
int firex_get_field_layout(firex_solver* solver, int32_t field, firex_field_layout* out);
/* 64-bit checksum of a field's stored bytes, computed on the device (an
 * addition for auditable captures; the capture manifest of record-fire.mjs,
 * scripts/record-fire.mjs:425-440, has no per-field record).  Definition:
 * the bytes read as little-endian u32 words w_i (a final partial word
 * zero-padded), checksum = sum over i of splitmix64((i << 32) | w_i) mod
 * 2^64, where splitmix64(z) adds 0x9e3779b97f4a7c15 and applies the
 * finaliser (z ^ z >> 30) * 0xbf58476d1ce4e5b9, (z ^ z >> 27) *
 * 0x94d049bb133111eb, z ^ z >> 31.  Order-independent, so deterministic. */
int firex_field_checksum(firex_solver* solver, int32_t field, uint64_t* out);
/* The same checksum of host bytes (verifying dumped fields; no device). */
int firex_checksum_bytes(const void* data, size_t bytes, uint64_t* out);

/* packFields (src/sim/firex/qualityShaders.ts:37-73) as encodeRender runs it
 * (FireXEngine.ts:2585-2605): fills the three render volumes (flow,
 * chemistry, media; rgba16float, 24 B/cell, allocated on the first call) from
 * the current state.  With viewMode 7 (the divergence view) it first
 * recomputes the divergence of the final flow into the divergence field, as
 * the presentation divergence pass does.  Read the volumes with
 * firex_read_field(FIREX_FIELD_PACKED_*). */
int firex_pack_fields(firex_solver* solver);

/* The volume ray marcher of the cinematic scene pass (volumeFragment,
 * src/sim/firex/qualityShaders.ts:374-575, drawn by encodeRender,
 * FireXEngine.ts:2607-2640).  Runs firex_pack_fields, then marches one ray
 * per pixel through the three render volumes into an rgba16float HDR image of
 * width x height (the reference's cinematic scene texture), readable as
 * FIREX_FIELD_SCENE.  Camera, view mode, exposure, zoom and optical detail
 * come from the solver's parameters (cameraYaw, cameraPitch, cameraDistance,
 * viewMode, exposure, viewZoom, frontDefinition, opticalDetailTarget).
 * Texture filtering is defined in f32 arithmetic (linear, clamp-to-edge for
 * the volumes, repeat for the 64^3 detail volume).  With the cinematic
 * presentation and a live particle field (water on), the droplet streaks
 * (dropletVertex / dropletFragment, qualityShaders.ts:156-232) are drawn over
 * the image in instance order, as the reference's scene pass does
 * (FireXEngine.ts:2636-2639); rasterization and blending are defined in f32
 * arithmetic (a pixel is covered when its centre's quad coordinates lie
 * strictly inside (-1, 1)^2; src * a + dst * (1 - a), stored as f16). */
typedef struct firex_render_config {
  int32_t width, height;  /* output pixels, each in [1, 16384]                 */
  int32_t ray_steps;      /* 0 = firex_config.ray_steps                         */
  int32_t reserved[5];
} firex_render_config;
int firex_render_volume(firex_solver* solver, const firex_render_config* config);
/* A presented frame (encodeRender, FireXEngine.ts:2607-2680): the ray march
 * above, then, with the cinematic presentation (firex_config.cinematic), the
 * quarter-resolution bloom and the filmic tone-map / dither composite
 * (qualityShaders.ts:580-671); without it, the scene image itself.  Stored as
 * rgba8 (round(clamp(c, 0, 1) * 255), ties to even), readable as
 * FIREX_FIELD_FRAME.  The standard presentation then draws the droplet
 * streaks into the frame (FireXEngine.ts:2686-2689), blending in f32 with
 * the destination read as byte / 255.  The standard path stores the scene
 * image rounded to f16 first, where the reference's standard pipeline writes
 * its f32 fragment colour to the canvas: a byte can differ where that
 * rounding moves c * 255 across a half step. */
int firex_render_frame(firex_solver* solver, const firex_render_config* config);

/* FP64 reduction on the device (FIREX_REDUCE_*).  FIREX_REDUCE_RESIDUAL_L2
 * fails with FIREX_ERR_INVALID after firex_pack_fields ran the divergence view
 * (viewMode 7): that view replaces the residual solve's right-hand side in the
 * divergence field with the final flow's divergence, as the reference's
 * presentation divergence pass does (FireXEngine.ts:2588-2596); the next step
 * restores it. */
int firex_reduce(firex_solver* solver, int32_t which, double* out);
/* firex_reduce without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference (it computes no
 * reductions).  The result (the same FP64 value firex_reduce returns) is
 * written to device_out, device-accessible memory the caller owns, on the
 * solver's stream and without waiting: read it after firex_sync or order
 * later work on that stream.  The block partials live in one 8 KiB array per
 * GPU inside the library (grids up to 2^32 - 1 cells; larger are rejected).
 * Every call on a GPU shares that array, so calls that overlap on one GPU
 * (two solvers on different streams) race; calls on one solver are ordered
 * by its stream. */
int firex_reduce_noalloc(firex_solver* solver, int32_t which, double* device_out);
int firex_get_info(firex_solver* solver, firex_info* info);

/* ---- instrumentation (additions: the reference has no GPU timestamps) -- */
enum {
  FIREX_PASS_SIMULATE = 0,            /* advection + sources                      */
  FIREX_PASS_DIVERGENCE = 1,
  FIREX_PASS_JACOBI = 2,              /* one entry per Jacobi launch              */
  FIREX_PASS_PROJECT = 3,
  FIREX_PASS_PROJECT_DIVERGENCE = 4,  /* no longer launched (kept for the layout) */
  FIREX_PASS_FILL_COPY = 5,           /* clears and copyFlow                      */
  FIREX_PASS_PACK = 6,                /* firex_pack_fields                        */
  FIREX_PASS_DROPLETS = 7,            /* SPH density, integration, deposition     */
  FIREX_PASS_RENDER = 8,              /* firex_render_volume's ray march          */
  FIREX_PASS_PRESENT = 9,             /* firex_render_frame's bloom + composite   */
  FIREX_PASS_STREAKS = 10,            /* droplet streaks over the scene or frame  */
  FIREX_PASS_COUNT = 12
};
typedef struct firex_pass_timing {
  double milliseconds[FIREX_PASS_COUNT];  /* summed device time per pass kind     */
  int64_t launches[FIREX_PASS_COUNT];
  int64_t jacobi_sweeps;                  /* sweeps covered by the Jacobi launches */
  int64_t steps;
} firex_pass_timing;
/* Enable (1) or disable (0) HIP-event timing of every launch; enabling clears
 * the totals.  Timing adds an event pair per launch on the solver's stream. */
int firex_set_pass_timing(firex_solver* solver, int32_t enabled);
/* Wait for the device and return the totals since timing was enabled. */
int firex_get_pass_timing(firex_solver* solver, firex_pass_timing* out);

#ifdef __cplusplus
}
#endif

#endif /* FIREX_H */
