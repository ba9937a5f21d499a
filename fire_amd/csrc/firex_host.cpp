// firex_host.cpp — host-side logic of the Fire-X solver library: tier tables,
// memory model and preflight, offline timeline, fixed-step clock, parameter
// clamps, uniform packing and the per-step constants handed to the kernels.
#include "firex_host.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>

namespace fx {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& message) {
  g_last_error = message;
  return code;
}

void clear_error() { g_last_error.clear(); }

// ------------------------------------------------------------------ params
static double finite_or(double v, double fallback) { return std::isfinite(v) ? v : fallback; }
static double clampd(double v, double lo, double hi) { return std::max(lo, std::min(hi, v)); }
static double detail_target(double t) { return t >= 768 ? 1024 : t >= 384 ? 512 : t >= 192 ? 256 : 0; }

bool params_set(Params& p, const char* name, double v) {
  const std::string n(name ? name : "");
  if (n == "burnerSize") p.burnerSize = clampd(finite_or(v, p.burnerSize), 0.35, 3.2);
  else if (n == "burnerDepth") p.burnerDepth = clampd(finite_or(v, p.burnerDepth), 0.35, 3.2);
  else if (n == "sourceThickness") p.sourceThickness = clampd(finite_or(v, p.sourceThickness), 0.35, 2.5);
  else if (n == "firePower") p.firePower = clampd(finite_or(v, p.firePower), 0.1, 4);
  else if (n == "sourceLift") p.sourceLift = clampd(finite_or(v, p.sourceLift), 0, 4);
  else if (n == "buoyancyScale") p.buoyancyScale = clampd(finite_or(v, p.buoyancyScale), 0, 4);
  else if (n == "flamePersistence") p.flamePersistence = clampd(finite_or(v, p.flamePersistence), 0.4, 3);
  else if (n == "airEntrainment") p.airEntrainment = clampd(finite_or(v, p.airEntrainment), 0, 3);
  else if (n == "fuelRate") p.fuelRate = finite_or(v, p.fuelRate);
  else if (n == "oxygenRate") p.oxygenRate = finite_or(v, p.oxygenRate);
  else if (n == "heatEfficiency") p.heatEfficiency = finite_or(v, p.heatEfficiency);
  else if (n == "vorticity") p.vorticity = finite_or(v, p.vorticity);
  else if (n == "waterFlow") p.waterFlow = finite_or(v, p.waterFlow);
  else if (n == "sprayAngle") p.sprayAngle = finite_or(v, p.sprayAngle);
  else if (n == "cameraYaw") p.cameraYaw = finite_or(v, p.cameraYaw);
  else if (n == "cameraPitch") p.cameraPitch = finite_or(v, p.cameraPitch);
  else if (n == "cameraDistance") p.cameraDistance = finite_or(v, p.cameraDistance);
  else if (n == "viewMode") p.viewMode = clampd(std::round(finite_or(v, p.viewMode)), 0, 7);
  else if (n == "aimHeight") p.aimHeight = clampd(finite_or(v, p.aimHeight), 0.08, 0.82);
  else if (n == "nozzleType") p.nozzleType = finite_or(v, p.nozzleType) >= 0.5 ? 1 : 0;
  else if (n == "exposure") p.exposure = clampd(finite_or(v, p.exposure), 0.25, 3);
  else if (n == "frontDefinition") p.frontDefinition = clampd(finite_or(v, p.frontDefinition), 0.5, 2.5);
  else if (n == "opticalDetailTarget") p.opticalDetailTarget = detail_target(finite_or(v, p.opticalDetailTarget));
  else if (n == "domainWidth") p.domainWidth = clampd(finite_or(v, p.domainWidth), 0.5, 4);
  else if (n == "domainHeight") p.domainHeight = clampd(finite_or(v, p.domainHeight), 0.5, 6);
  else if (n == "domainDepth") p.domainDepth = clampd(finite_or(v, p.domainDepth), 0.5, 4);
  else if (n == "viewZoom") p.viewZoom = clampd(finite_or(v, p.viewZoom), 0.5, 5);
  else return false;
  return true;
}

bool params_set_bool(Params& p, const char* name, bool value) {
  const std::string n(name ? name : "");
  if (n == "waterEnabled") {
    p.waterEnabled = value;
    return true;
  }
  return false;
}

double* params_field(Params& p, const char* name) {
  const std::string n(name ? name : "");
  struct Entry {
    const char* key;
    double* v;
  };
  const Entry table[] = {
      {"burnerSize", &p.burnerSize}, {"burnerDepth", &p.burnerDepth}, {"sourceThickness", &p.sourceThickness},
      {"firePower", &p.firePower}, {"sourceLift", &p.sourceLift}, {"buoyancyScale", &p.buoyancyScale},
      {"flamePersistence", &p.flamePersistence}, {"airEntrainment", &p.airEntrainment}, {"fuelRate", &p.fuelRate},
      {"oxygenRate", &p.oxygenRate}, {"heatEfficiency", &p.heatEfficiency}, {"vorticity", &p.vorticity},
      {"waterFlow", &p.waterFlow}, {"sprayAngle", &p.sprayAngle}, {"cameraYaw", &p.cameraYaw},
      {"cameraPitch", &p.cameraPitch}, {"cameraDistance", &p.cameraDistance}, {"viewMode", &p.viewMode},
      {"aimHeight", &p.aimHeight}, {"nozzleType", &p.nozzleType}, {"exposure", &p.exposure},
      {"frontDefinition", &p.frontDefinition}, {"opticalDetailTarget", &p.opticalDetailTarget},
      {"domainWidth", &p.domainWidth}, {"domainHeight", &p.domainHeight}, {"domainDepth", &p.domainDepth},
      {"viewZoom", &p.viewZoom},
  };
  for (const auto& e : table)
    if (n == e.key) return e.v;
  return nullptr;
}

bool params_get(const Params& p, const char* name, double* value) {
  if (name && std::string(name) == "waterEnabled") {
    *value = p.waterEnabled ? 1.0 : 0.0;
    return true;
  }
  double* f = params_field(const_cast<Params&>(p), name);
  if (!f) return false;
  *value = *f;
  return true;
}

// ---------------------------------------------------------------- uniforms
void write_uniforms(const Params& p, const int32_t dims[3], double dt, double simulation_time, double frame_index,
                    double time_step, bool cinematic, float u[FXU_COUNT], int32_t particle_count, int32_t canvas_width,
                    int32_t canvas_height, int32_t ray_steps) {
  memset(u, 0, sizeof(float) * FXU_COUNT);
  double cells = (double)dims[0] * (double)dims[1] * (double)dims[2];
  u[FXU_GRID_X] = (float)dims[0];
  u[FXU_GRID_Y] = (float)dims[1];
  u[FXU_GRID_Z] = (float)dims[2];
  u[FXU_CELLS] = (float)cells;
  u[FXU_DT] = (float)dt;
  u[FXU_SIM_TIME] = (float)simulation_time;
  u[FXU_FRAME] = (float)frame_index;
  u[FXU_TIME_STEP] = (float)time_step;
  u[FXU_FUEL_RATE] = (float)p.fuelRate;
  u[FXU_OXYGEN_RATE] = (float)p.oxygenRate;
  u[FXU_HEAT_EFFICIENCY] = (float)p.heatEfficiency;
  u[FXU_VORTICITY] = (float)p.vorticity;
  u[FXU_WATER_FLOW] = (float)p.waterFlow;
  u[FXU_SPRAY_ANGLE] = (float)(p.sprayAngle * M_PI / 180.0);
  u[FXU_WATER_ENABLED] = p.waterEnabled ? 1.0f : 0.0f;
  u[FXU_AMBIENT] = 300.0f;
  u[FXU_CAMERA_YAW] = (float)p.cameraYaw;
  u[FXU_CAMERA_PITCH] = (float)p.cameraPitch;
  u[FXU_CAMERA_DISTANCE] = (float)p.cameraDistance;
  u[FXU_ASPECT] = (float)((double)canvas_width / (double)std::max(canvas_height, 1));
  u[FXU_CANVAS_W] = (float)canvas_width;
  u[FXU_CANVAS_H] = (float)canvas_height;
  u[FXU_RAY_STEPS] = (float)std::min(384, ray_steps);
  u[FXU_PARTICLES] = (float)particle_count;
  u[FXU_VIEW_MODE] = (float)p.viewMode;
  u[FXU_AIM_HEIGHT] = (float)p.aimHeight;
  u[FXU_NOZZLE] = (float)p.nozzleType;
  u[FXU_EXPOSURE] = (float)p.exposure;
  u[FXU_BURNER_SIZE] = (float)p.burnerSize;
  u[FXU_BURNER_DEPTH] = (float)p.burnerDepth;
  u[FXU_SOURCE_THICKNESS] = (float)p.sourceThickness;
  u[FXU_FIRE_POWER] = (float)p.firePower;
  u[FXU_DOMAIN_X] = (float)p.domainWidth;
  u[FXU_DOMAIN_Y] = (float)p.domainHeight;
  u[FXU_DOMAIN_Z] = (float)p.domainDepth;
  u[FXU_SOURCE_LIFT] = (float)p.sourceLift;
  u[FXU_BUOYANCY] = (float)p.buoyancyScale;
  u[FXU_PERSISTENCE] = (float)p.flamePersistence;
  u[FXU_VIEW_ZOOM] = (float)p.viewZoom;
  u[FXU_ENTRAINMENT] = (float)p.airEntrainment;
  u[FXU_FRONT_DEFINITION] = (float)p.frontDefinition;
  u[FXU_CINEMATIC] = cinematic ? 1.0f : 0.0f;
  u[FXU_OPTICAL_DETAIL] = (float)p.opticalDetailTarget;
}

FxStep make_step(const float u[FXU_COUNT], const int32_t dims[3], const EmitterConfig& em, uint64_t frame,
                 bool has_liquid) {
  FxStep st;
  memset(&st, 0, sizeof st);
  st.nx = dims[0];
  st.ny = dims[1];
  st.nz = dims[2];
  st.has_liquid = has_liquid ? 1 : 0;
  st.gx = u[FXU_GRID_X];
  st.gy = u[FXU_GRID_Y];
  st.gz = u[FXU_GRID_Z];
  const float dt = u[FXU_DT], time = u[FXU_SIM_TIME];
  st.dt = dt;
  st.time = time;
  // domainExtent / cellMetric, shaders.ts:35-41
  st.ex = fxm_max(u[FXU_DOMAIN_X], 0.001f);
  st.ey = fxm_max(u[FXU_DOMAIN_Y], 0.001f);
  st.ez = fxm_max(u[FXU_DOMAIN_Z], 0.001f);
  st.mx = st.gx / st.ex;
  st.my = st.gy / st.ey;
  st.mz = st.gz / st.ez;
  st.ambient = u[FXU_AMBIENT];
  st.chem_x = u[FXU_FUEL_RATE];
  st.chem_y = u[FXU_OXYGEN_RATE];
  st.chem_z = u[FXU_HEAT_EFFICIENCY];
  st.chem_w = u[FXU_VORTICITY];
  // simulate, shaders.ts:178-187
  st.mix_velocity = fxm_clamp(dt * 0.32f, 0.0f, 0.12f);
  st.mix_temperature = fxm_clamp(dt * 0.18f, 0.0f, 0.08f);
  st.mix_species = fxm_clamp(dt * 0.12f, 0.0f, 0.06f);
  // shaders.ts:190-201
  st.fire_power = fxm_clamp(u[FXU_FIRE_POWER], 0.1f, 4.0f);
  st.power_root = fxm_sqrt(st.fire_power);
  st.thermal_power = fxm_min(st.power_root, 1.0f + 0.22f * (st.power_root - 1.0f));
  st.swirl_tz = time * 2.3f;
  st.swirl_tx = time * 1.9f;
  st.lift_clamp = fxm_clamp(u[FXU_SOURCE_LIFT], 0.0f, 4.0f);
  // shaders.ts:220-221, 253, 256, 270, 284, 288
  st.entrainment = fxm_clamp(u[FXU_ENTRAINMENT], 0.0f, 3.0f);
  st.ambient_oxygen = fxm_max(st.chem_y, 0.001f);
  st.inv_ambient_oxygen = 1.0f / st.ambient_oxygen;
  st.heat_release = 2150.0f + 1450.0f * st.chem_z;
  st.max_dt = fxm_max(dt, 0.0001f);
  st.inv_max_dt = 1.0f / st.max_dt;
  st.vortex_k = 0.00135f * st.chem_w * dt;
  st.turbulence_base = 0.075f + 0.16f * st.chem_w;
  st.tt31 = time * 3.1f;
  st.tt53 = time * 5.3f;
  st.tt41 = time * 4.1f;
  st.tt27 = time * 2.7f;
  st.tt47 = time * 4.7f;
  st.buoyancy_clamp = fxm_clamp(u[FXU_BUOYANCY], 0.0f, 4.0f);
  // spray momentum, shaders.ts:293-296
  {
    float nox = 0.035f * st.ex, noy = 0.50f * st.ey, noz = 0.5f * st.ez;
    float stx = 0.52f * st.ex, sty = u[FXU_AIM_HEIGHT] * st.ey, stz = 0.5f * st.ez;
    float smx = stx - nox, smy = sty - noy, smz = stz - noz;
    float slen = fxm_sqrt(smx * smx + smy * smy + smz * smz);
    st.spray_x = smx / slen;
    st.spray_y = smy / slen;
    st.spray_z = smz / slen;
  }
  // persistence and radiation, shaders.ts:299-304
  st.persistence = fxm_clamp(u[FXU_PERSISTENCE], 0.4f, 3.0f);
  st.ambient4 = fxm_pow(st.ambient, 4.0f);
  st.radiation_denominator = 1005.0f * st.persistence;
  st.inv_radiation_denominator = 1.0f / st.radiation_denominator;
  st.decay_products = fxm_exp(-dt * 0.075f / st.persistence);
  st.decay_vapor = fxm_exp(-dt * 0.14f / st.persistence);
  // emitter, shaders.ts:101-133
  {
    st.em_cx = st.ex * 0.5f + fxm_sin(time * 1.73f) * 0.013f;
    st.em_cy = 0.075f;
    st.em_cz = st.ez * 0.5f + fxm_cos(time * 1.37f) * 0.011f;
    const float baseRadius = (float)std::sqrt(0.02024);
    float radiusX = fxm_min(baseRadius * fxm_clamp(u[FXU_BURNER_SIZE], 0.35f, 3.2f), st.ex * 0.42f);
    float radiusZ = fxm_min(baseRadius * fxm_clamp(u[FXU_BURNER_DEPTH], 0.35f, 3.2f), st.ez * 0.42f);
    float thick = fxm_clamp(u[FXU_SOURCE_THICKNESS], 0.35f, 2.5f);
    st.em_rx2 = radiusX * radiusX;
    st.em_rz2 = radiusZ * radiusZ;
    st.em_vdenom = 0.0022f * thick * thick;
    st.em_rx_safe = fxm_max(radiusX, 0.001f);
    st.em_rz_safe = fxm_max(radiusZ, 0.001f);
    st.em_cinematic = fxm_clamp(u[FXU_CINEMATIC], 0.0f, 1.0f);
    st.et52 = time * 5.2f;
    st.et44 = time * 4.4f;
    st.et17 = time * 1.7f;
    st.et13 = time * 1.3f;
    st.et37 = time * 3.7f;
  }
  st.seeded_count = em.count;
  memcpy(st.seeded, em.data, sizeof st.seeded);
  st.jitter = em.jitter;
  st.velocity_noise = em.velocity_noise;
  st.seed = em.seed;
  st.frame = frame;
  // Jacobi stencil, shaders.ts:386-389, 402
  st.wx = st.mx * st.mx;
  st.wy = st.my * st.my;
  st.wz = st.mz * st.mz;
  st.jacobi_inv = 1.0f / (2.0f * (st.wx + st.wy + st.wz));
  return st;
}

FxDroplets make_droplets(const float u[FXU_COUNT], const int32_t dims[3]) {
  FxDroplets d;
  memset(&d, 0, sizeof d);
  d.count = (uint32_t)u[FXU_PARTICLES];
  d.nx = dims[0];
  d.ny = dims[1];
  d.nz = dims[2];
  d.dt = u[FXU_DT];
  d.frame = u[FXU_FRAME];
  d.gx = u[FXU_GRID_X];
  d.gy = u[FXU_GRID_Y];
  d.gz = u[FXU_GRID_Z];
  d.ex = fxm_max(u[FXU_DOMAIN_X], 0.001f);
  d.ey = fxm_max(u[FXU_DOMAIN_Y], 0.001f);
  d.ez = fxm_max(u[FXU_DOMAIN_Z], 0.001f);
  d.water_flow = u[FXU_WATER_FLOW];
  d.spray_angle = u[FXU_SPRAY_ANGLE];
  d.water_enabled = u[FXU_WATER_ENABLED];
  d.aim = u[FXU_AIM_HEIGHT];
  d.nozzle = u[FXU_NOZZLE];
  return d;
}

// Math.hypot as JS engines evaluate it: the magnitudes scaled by the largest,
// a Kahan-compensated sum of their squares, square root, rescaled.
static double js_hypot(const double* v, int n) {
  double largest = 0;
  bool nan = false;
  for (int i = 0; i < n; ++i) {
    const double a = std::fabs(v[i]);
    if (std::isinf(a)) return INFINITY;
    if (std::isnan(a)) nan = true;
    else largest = std::max(largest, a);
  }
  if (nan) return NAN;
  if (largest == 0) return 0;
  double sum = 0, compensation = 0;
  for (int i = 0; i < n; ++i) {
    const double r = std::fabs(v[i]) / largest;
    const double summand = r * r - compensation;
    const double next = sum + summand;
    compensation = (next - sum) - summand;
    sum = next;
  }
  return std::sqrt(sum) * largest;
}

// writeParticleInitialState, FireXEngine.ts:1976-2034: float64 arithmetic,
// stored as f32 (the reference fills a Float32Array).
void particle_initial_state(const Params& p, int32_t count, bool active, float* out) {
  const double w = p.domainWidth, h = p.domainHeight, depth = p.domainDepth;
  const double aim[2] = {(0.52 - 0.035) * w, (p.aimHeight - 0.5) * h};
  const double aim_length = js_hypot(aim, 2);
  const double bx = aim[0] / aim_length, by = aim[1] / aim_length;
  const double tx = by, ty = -bx;  // tangent in the xy plane
  auto hash = [](double v) {
    const double x = std::sin(v * 12.9898 + 1.317) * 43758.5453;
    return x - std::floor(x);
  };
  const bool spray = p.nozzleType >= 0.5;
  const double spread = spray ? 0.72 : 0.1;
  const double variation = spray ? 0.24 : 0.05;
  const double origin_radius = spray ? 0.006 : 0.0028;
  for (int32_t i = 0; i < count; ++i) {
    const double phase = (double)i / (double)count;
    const double angle = hash(i * 3.771 + 4.91) * M_PI * 2;
    const double radial = std::sqrt(hash(i * 3.771 + 0.17));
    const double cone = std::tan((p.sprayAngle * M_PI) / 180) * radial * spread;
    const double disk_x = tx * std::cos(angle), disk_y = ty * std::cos(angle), disk_z = -std::sin(angle);
    double dir[3] = {bx + disk_x * cone, by + disk_y * cone, disk_z * cone};
    const double length = js_hypot(dir, 3);
    dir[0] /= length;
    dir[1] /= length;
    dir[2] /= length;
    const double speed = (0.82 + p.waterFlow * 0.82) * (1 - variation * 0.5 + hash(i * 3.771 + 9.37) * variation);
    const double age = phase * 0.72;
    float* o = out + (size_t)i * 8;
    o[0] = (float)(0.035 + (disk_x * radial * origin_radius + dir[0] * speed * age) / w);
    o[1] = (float)std::max(0.012, 0.5 + (disk_y * radial * origin_radius + dir[1] * speed * age - 0.29 * age * age) / h);
    o[2] = (float)(0.5 + (disk_z * radial * origin_radius + dir[2] * speed * age) / depth);
    o[3] = (float)(active ? age : -phase * 0.72);
    o[4] = (float)(active ? dir[0] * speed : 0);
    o[5] = (float)(active ? dir[1] * speed - 0.58 * age : 0);
    o[6] = (float)(active ? dir[2] * speed : 0);
    o[7] = (float)(active ? 0.52 + p.waterFlow * 0.88 : 0);
  }
}

// createPeriodicDetailVolume, FireXEngine.ts:311-355: value noise of
// frequencies 4, 8 and 14 over a period of size - 1 texels in r, g, b (alpha
// 255).  The hash is JS 32-bit integer arithmetic (Math.imul, <<, >>>), the
// interpolation float64, Math.round = floor(x + 0.5).
void periodic_detail_volume(int size, uint8_t* out) {
  const int period = size - 1;
  auto hash = [](int x, int y, int z, uint32_t seed) -> double {
    uint32_t v = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u) ^ seed;
    v ^= v << 13;
    v ^= v >> 17;
    v ^= v << 5;
    return (double)((v >> 24) & 255u);
  };
  auto fade = [](double v) { return v * v * (3 - 2 * v); };
  auto mixd = [](double a, double b, double t) { return a + (b - a) * t; };
  auto noise = [&](int x, int y, int z, int frequency, uint32_t seed) -> uint8_t {
    const double px = ((double)x / period) * frequency;
    const double py = ((double)y / period) * frequency;
    const double pz = ((double)z / period) * frequency;
    const int x0 = (int)std::floor(px) % frequency, y0 = (int)std::floor(py) % frequency;
    const int z0 = (int)std::floor(pz) % frequency;
    const int x1 = (x0 + 1) % frequency, y1 = (y0 + 1) % frequency, z1 = (z0 + 1) % frequency;
    const double tx = fade(px - std::floor(px)), ty = fade(py - std::floor(py)), tz = fade(pz - std::floor(pz));
    const double lower = mixd(mixd(hash(x0, y0, z0, seed), hash(x1, y0, z0, seed), tx),
                              mixd(hash(x0, y1, z0, seed), hash(x1, y1, z0, seed), tx), ty);
    const double upper = mixd(mixd(hash(x0, y0, z1, seed), hash(x1, y0, z1, seed), tx),
                              mixd(hash(x0, y1, z1, seed), hash(x1, y1, z1, seed), tx), ty);
    return (uint8_t)std::floor(mixd(lower, upper, tz) + 0.5);
  };
  for (int z = 0; z < size; ++z)
    for (int y = 0; y < size; ++y)
      for (int x = 0; x < size; ++x) {
        uint8_t* o = out + ((size_t)x + (size_t)size * ((size_t)y + (size_t)size * z)) * 4;
        o[0] = noise(x, y, z, 4, 0x51f15e5du);
        o[1] = noise(x, y, z, 8, 0x6d2b79f5u);
        o[2] = noise(x, y, z, 14, 0x1b873593u);
        o[3] = 255;
      }
}

FxRender make_render(const float u[FXU_COUNT], const int32_t dims[3], int32_t width, int32_t height) {
  FxRender r;
  memset(&r, 0, sizeof r);
  r.width = width;
  r.height = height;
  const float steps = u[FXU_RAY_STEPS];
  const uint32_t s = steps > 0.0f ? (uint32_t)steps : 0u;
  r.steps = (int32_t)(s < 384u ? s : 384u);
  r.nx = dims[0];
  r.ny = dims[1];
  r.nz = dims[2];
  r.gx = u[FXU_GRID_X];
  r.gy = u[FXU_GRID_Y];
  r.gz = u[FXU_GRID_Z];
  r.ex = fxm_max(u[FXU_DOMAIN_X], 0.001f);
  r.ey = fxm_max(u[FXU_DOMAIN_Y], 0.001f);
  r.ez = fxm_max(u[FXU_DOMAIN_Z], 0.001f);
  r.yaw = u[FXU_CAMERA_YAW];
  r.pitch = u[FXU_CAMERA_PITCH];
  r.cam_distance = u[FXU_CAMERA_DISTANCE];
  r.aspect = u[FXU_ASPECT];
  r.render_z = steps;
  r.view_mode = u[FXU_VIEW_MODE];
  r.exposure = u[FXU_EXPOSURE];
  r.front_definition = u[FXU_FRONT_DEFINITION];
  r.cinematic = u[FXU_CINEMATIC];
  r.detail_target = u[FXU_OPTICAL_DETAIL];
  r.time = u[FXU_SIM_TIME];
  r.chem_y = u[FXU_OXYGEN_RATE];
  r.chem_z = u[FXU_HEAT_EFFICIENCY];
  r.zoom = u[FXU_VIEW_ZOOM];
  r.canvas_height = u[FXU_CANVAS_H];
  r.nozzle = u[FXU_NOZZLE];
  r.bnx = (dims[0] + kFxBrick - 1) / kFxBrick;
  r.bny = (dims[1] + kFxBrick - 1) / kFxBrick;
  r.bnz = (dims[2] + kFxBrick - 1) / kFxBrick;
  return r;
}

// ------------------------------------------------------------------- clock
int32_t FixedStepClock::advance(double delta_seconds) {
  if (!std::isfinite(delta_seconds)) return 0;
  double bounded = std::max(0.0, std::min(delta_seconds, step_seconds * maximum_steps));
  accumulator += bounded;
  int32_t count = 0;
  while (accumulator >= step_seconds && count < maximum_steps) {
    accumulator -= step_seconds;
    count += 1;
  }
  if (count == maximum_steps && accumulator >= step_seconds) accumulator = std::fmod(accumulator, step_seconds);
  return count;
}

// ------------------------------------------------------------------ memory
bool heavy_grid(const int32_t dims[3]) {
  return (uint64_t)dims[0] * (uint64_t)dims[1] * (uint64_t)dims[2] >= kHeavyCellThreshold;
}

uint64_t device_bytes(const int32_t dims[3], int32_t species_storage, bool f16_pressure, bool compact_f32,
                      bool with_liquid) {
  uint64_t cells = (uint64_t)dims[0] * (uint64_t)dims[1] * (uint64_t)dims[2];
  uint64_t species = species_storage == 0 ? 16 : 8;
  uint64_t pressure = f16_pressure ? 2 : 4;
  uint64_t bytes = cells * (2 * 16 + 2 * species + 2 * pressure + 4 /* div */ + 4 /* reaction */);
  if (!compact_f32) bytes += cells * 2 * pressure;  // residual pair
  if (with_liquid) bytes += cells * 4;
  return bytes + (uint64_t)kReduceScratchBytes + fx_sim_table_floats(dims) * 4;
}

uint64_t particle_device_bytes(int32_t count) { return count > 0 ? (uint64_t)count * (2 * 32 + 16) : 0; }

}  // namespace fx

// =========================================================== C ABI (host only)
using namespace fx;

namespace {

struct Tier {
  const char* name;
  int32_t dims[3];
  int32_t pressure, correction;
  double time_step;
  int32_t cinematic;
  int32_t particles;
  int32_t ray_steps;
};

// FIREX_QUALITY (FireXEngine.ts:187-238) and FIREX_OFFLINE_QUALITY (:240-277).
// The uniform's cinematic flag is set for the cinematic live tier and for
// every offline tier (usesCinematicPresentation, FireXEngine.ts:2819-2821).
const Tier kTiers[] = {
    {"low", {32, 48, 32}, 14, 14, 1.0 / 30.0, 0, 384, 80},
    {"balanced", {40, 64, 40}, 24, 24, 1.0 / 30.0, 0, 640, 112},
    {"high", {64, 96, 64}, 36, 36, 1.0 / 30.0, 0, 1024, 144},
    {"maximum", {80, 120, 80}, 48, 48, 1.0 / 30.0, 0, 1536, 192},
    {"cinematic", {80, 120, 80}, 48, 48, 1.0 / 30.0, 1, 1536, 192},
    {"hd", {128, 192, 128}, 64, 64, 1.0 / 60.0, 1, 2048, 256},
    {"qhd", {160, 240, 160}, 80, 80, 1.0 / 60.0, 1, 3072, 320},
    {"uhd", {192, 288, 192}, 96, 96, 1.0 / 60.0, 1, 4096, 384},
};

const int64_t kGridBytesPerCell = 116;       // gridConfiguration.ts:106
const int64_t kHeavyGridBytesPerCell = 108;  // gridConfiguration.ts:107
const int64_t kParticleBytes = 80;           // gridConfiguration.ts:108
const int64_t kLargestStoragePerCell = 16;   // gridConfiguration.ts:109
const int64_t kDetailTextureBytes = 64 * 64 * 64 * 4;  // gridConfiguration.ts:110

}  // namespace

extern "C" {

const char* firex_last_error(void) { return g_last_error.c_str(); }

int firex_abi_version(void) { return FIREX_ABI_VERSION; }

void firex_config_init(firex_config* c) {
  if (!c) return;
  memset(c, 0, sizeof *c);
  c->dims[0] = 40;
  c->dims[1] = 64;
  c->dims[2] = 40;
  c->pressure_iterations = 24;
  c->correction_iterations = 24;
  c->time_step = 1.0 / 30.0;
  c->maximum_substeps = 1;
  c->species_storage = FIREX_STORAGE_F32;
  c->heavy_mode = FIREX_HEAVY_AUTO;
  c->plan = FIREX_PLAN_REFERENCE;
  c->particle_count = 640;
  c->ray_steps = 112;
}

int firex_config_from_tier(firex_config* c, const char* tier) {
  if (!c || !tier) return set_error(FIREX_ERR_INVALID, "config and tier must be non-null");
  for (const Tier& t : kTiers) {
    if (strcmp(t.name, tier) == 0) {
      memcpy(c->dims, t.dims, sizeof c->dims);
      c->pressure_iterations = t.pressure;
      c->correction_iterations = t.correction;
      c->time_step = t.time_step;
      c->maximum_substeps = 1;
      c->cinematic = t.cinematic;
      c->particle_count = t.particles;
      c->ray_steps = t.ray_steps;
      return FIREX_OK;
    }
  }
  return set_error(FIREX_ERR_INVALID, std::string("unknown Fire-X tier: ") + tier);
}

int firex_checksum_bytes(const void* data, size_t bytes, uint64_t* out) {
  if (!out || (!data && bytes)) return set_error(FIREX_ERR_INVALID, "data and out must be non-null");
  const uint8_t* b = (const uint8_t*)data;
  const uint64_t n = (bytes + 3) / 4;
  uint64_t h = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t w = 0;
    const size_t left = bytes - (size_t)i * 4;
    memcpy(&w, b + i * 4, left < 4 ? left : 4);  // little-endian host
    h += fx_checksum_term(i, w);
  }
  *out = h;
  return FIREX_OK;
}

int firex_estimate_memory(const int32_t dims[3], int64_t particle_count, firex_memory_estimate* out) {
  if (!dims || !out) return set_error(FIREX_ERR_INVALID, "dims and out must be non-null");
  int64_t cells = (int64_t)dims[0] * (int64_t)dims[1] * (int64_t)dims[2];
  int64_t per_cell = (uint64_t)cells >= kHeavyCellThreshold ? kHeavyGridBytesPerCell : kGridBytesPerCell;
  int64_t particles = particle_count > 0 ? particle_count : 0;
  out->cell_count = cells;
  out->grid_bytes = cells * per_cell;
  out->particle_bytes = particles * kParticleBytes;
  out->fixed_bytes = kDetailTextureBytes;
  out->total_bytes = out->grid_bytes + out->particle_bytes + out->fixed_bytes;
  out->largest_storage_buffer_bytes = cells * kLargestStoragePerCell;
  return FIREX_OK;
}

int firex_preflight(const int32_t dims[3], const firex_device_limits* limits, int64_t particle_count,
                    int64_t memory_budget_bytes, uint32_t* issues) {
  if (!dims || !limits || !issues) return set_error(FIREX_ERR_INVALID, "dims, limits and issues must be non-null");
  uint32_t bits = 0;
  for (int i = 0; i < 3; ++i) {
    if (dims[i] <= 0) bits |= FIREX_ISSUE_INVALID_DIMENSION;
    else if (dims[i] % 4 != 0) bits |= FIREX_ISSUE_UNALIGNED_DIMENSION;
    if (dims[i] > limits->max_texture_dimension_3d) bits |= FIREX_ISSUE_TEXTURE_DIMENSION_LIMIT;
  }
  firex_memory_estimate e;
  firex_estimate_memory(dims, particle_count, &e);
  if (e.largest_storage_buffer_bytes > limits->max_storage_buffer_binding_size) bits |= FIREX_ISSUE_STORAGE_BINDING_LIMIT;
  if (e.largest_storage_buffer_bytes > limits->max_buffer_size) bits |= FIREX_ISSUE_BUFFER_SIZE_LIMIT;
  if (memory_budget_bytes >= 0 && e.total_bytes > memory_budget_bytes) bits |= FIREX_ISSUE_MEMORY_BUDGET;
  *issues = bits;
  if (bits) return set_error(FIREX_ERR_UNSUPPORTED, "the grid fails preflight");
  return FIREX_OK;
}

int firex_split_offline_delta(double delta, double maximum_step, double* steps, int32_t cap, int32_t* count) {
  if (!count) return set_error(FIREX_ERR_INVALID, "count must be non-null");
  *count = 0;
  if (!std::isfinite(delta) || delta < 0)
    return set_error(FIREX_ERR_INVALID, "Offline frame delta must be finite and non-negative.");
  if (!std::isfinite(maximum_step) || maximum_step <= 0)
    return set_error(FIREX_ERR_INVALID, "Offline maximum step must be positive and finite.");
  if (delta == 0) return FIREX_OK;
  double n = std::max(1.0, std::ceil(delta / maximum_step - 1e-12));
  if (n > 2147483647.0) return set_error(FIREX_ERR_INVALID, "Offline frame delta needs too many steps.");
  int32_t k = (int32_t)n;
  double step = delta / (double)k;
  for (int32_t i = 0; i < k && i < cap; ++i) steps[i] = step;
  *count = k;
  return FIREX_OK;
}

int32_t firex_fixed_step_clock_advance(double step_seconds, int32_t maximum_steps, double* accumulator,
                                       double delta_seconds) {
  if (!accumulator || !std::isfinite(step_seconds) || !(step_seconds > 0) || maximum_steps < 1) {
    set_error(FIREX_ERR_INVALID, "Fixed step must be finite and positive; maximum step count a positive integer.");
    return -1;
  }
  FixedStepClock c;
  c.step_seconds = step_seconds;
  c.maximum_steps = maximum_steps;
  c.accumulator = *accumulator;
  int32_t n = c.advance(delta_seconds);
  *accumulator = c.accumulator;
  return n;
}

int firex_clamp_parameter(const char* name, double value, double previous, double* out) {
  if (!name || !out) return set_error(FIREX_ERR_INVALID, "name and out must be non-null");
  Params p;
  double* field = params_field(p, name);
  if (!field) return set_error(FIREX_ERR_INVALID, std::string("unknown parameter: ") + name);
  *field = previous;  // the stored value, as it stands before the call
  params_set(p, name, value);
  *out = *field;
  return FIREX_OK;
}

int firex_resolve_offline_dims(const char* tier, const int32_t* override_dims, int32_t out[3]) {
  if (!tier || !out) return set_error(FIREX_ERR_INVALID, "tier and out must be non-null");
  for (const Tier& t : kTiers) {
    if (strcmp(t.name, tier) == 0 && t.time_step == 1.0 / 60.0) {
      for (int i = 0; i < 3; ++i) out[i] = override_dims ? std::max(t.dims[i], override_dims[i]) : t.dims[i];
      return FIREX_OK;
    }
  }
  return set_error(FIREX_ERR_INVALID, std::string("unknown offline tier: ") + tier);
}

}  // extern "C"
