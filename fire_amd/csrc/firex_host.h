// firex_host.h — host-side logic of the Fire-X solver library (no device code).
#pragma once

#include <stdint.h>

#include <string>

#include "../../include/firex.h"
#include "fx_common.h"

namespace fx {

// Thread-local error message behind firex_last_error().
int set_error(int code, const std::string& message);
void clear_error();

// Engine parameters with the reference defaults (Inferno, FireXEngine.ts:422-451).
struct Params {
  double burnerSize = 3.1, burnerDepth = 2.7, sourceThickness = 1.35, firePower = 2.5;
  double sourceLift = 1.6, buoyancyScale = 2.15, flamePersistence = 1.45, airEntrainment = 1.65;
  double fuelRate = 0.30, oxygenRate = 1, heatEfficiency = 0.86, vorticity = 3.6;
  double waterFlow = 0, sprayAngle = 10;
  bool waterEnabled = false;
  double cameraYaw = -0.34, cameraPitch = 0.03, cameraDistance = 4.9;
  double viewMode = 0, aimHeight = 0.42, nozzleType = 0, exposure = 1.25, frontDefinition = 1.65;
  double domainWidth = 1.15, domainHeight = 1.35, domainDepth = 1.15, viewZoom = 2.1;
  double opticalDetailTarget = 1024;
};

// setParameter, FireXEngine.ts:937-1032.  Returns false for unknown keys.
bool params_set(Params& p, const char* name, double value);
bool params_set_bool(Params& p, const char* name, bool value);
bool params_get(const Params& p, const char* name, double* value);
double* params_field(Params& p, const char* name);  // numeric keys only

// writeUniforms, FireXEngine.ts:2043-2113.
void write_uniforms(const Params& p, const int32_t dims[3], double dt, double simulation_time, double frame_index,
                    double time_step, bool cinematic, float out[FXU_COUNT], int32_t particle_count = 0,
                    int32_t canvas_width = 0, int32_t canvas_height = 0, int32_t ray_steps = 0);

// createPeriodicDetailVolume, FireXEngine.ts:311-355: size^3 rgba8 texels.
void periodic_detail_volume(int size, uint8_t* out);
// The ray marcher's uniform slots and output size.
FxRender make_render(const float u[FXU_COUNT], const int32_t dims[3], int32_t width, int32_t height);

// waterSimulationActive, FireXEngine.ts:2537-2539.
inline bool water_active(const Params& p) { return p.waterEnabled && p.waterFlow > 0.001; }
// writeParticleInitialState, FireXEngine.ts:1976-2034: count records of 8
// floats (positionAge, velocityMass) into out.
void particle_initial_state(const Params& p, int32_t count, bool active, float* out);
// The droplet passes' uniform slots.
FxDroplets make_droplets(const float u[FXU_COUNT], const int32_t dims[3]);

struct EmitterConfig {
  int32_t count = 0;
  float data[FX_MAX_EMITTERS][5] = {};
  float jitter = 0, velocity_noise = 0;
  uint64_t seed = 0;
};

// Per-step constants for the kernels: uniform slots plus the uniform-only
// subexpressions of shaders.ts:101-342, evaluated in the same f32 order.
FxStep make_step(const float u[FXU_COUNT], const int32_t dims[3], const EmitterConfig& em, uint64_t frame,
                 bool has_liquid);

// FixedStepClock, src/sim/FixedStepClock.ts:1-38.
struct FixedStepClock {
  double step_seconds = 1.0 / 30.0;
  int32_t maximum_steps = 1;
  double accumulator = 0;
  int32_t advance(double delta_seconds);  // returns the number of steps to run
  void reset() { accumulator = 0; }
};

// Bytes this library allocates on the device for one plume.
uint64_t device_bytes(const int32_t dims[3], int32_t species_storage, bool f16_pressure, bool compact_f32,
                      bool with_liquid);
// Particle buffers (two of 32 B records, one of 16 B density records).
uint64_t particle_device_bytes(int32_t count);
bool heavy_grid(const int32_t dims[3]);
constexpr uint64_t kHeavyCellThreshold = 64ull * 1024 * 1024;  // gridConfiguration.ts:105
// FP64 reduction scratch: 1024 block partials plus 8 results.
constexpr uint64_t kReduceScratchBytes = (1024 + 8) * sizeof(double);

}  // namespace fx
