// firex_oracle.cpp — CPU restatement of the Fire-X fixed-step solver.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP solver is
// checked against.  Only the test suite, the smoke check and the CPU-baseline
// leg of the benchmark load it; the product never does.
//
// It restates, kernel by kernel and in the same operation order, the WGSL of
// the reference solver and the host step graph that drives it:
//   * helpers gridDimensions/domainExtent/cellMetric/clampCell/cellIndex:
//     src/sim/firex/shaders.ts:31-51
//   * trilinearFlow/trilinearSpecies:            shaders.ts:61-95
//   * emitter():                                 shaders.ts:101-133
//   * initializeFields:                          shaders.ts:135-149
//   * simulate (advection + sources):            shaders.ts:151-342
//   * computeDivergence:                         shaders.ts:344-364
//   * loadPressure/solvePressureGlobal/Tiled:    shaders.ts:366-459
//   * pressureGradient/projectVelocity:          shaders.ts:461-501
//   * copyFlow:                                  shaders.ts:503-511
//   * f16 pressure storage (solvePressureF16Tiled/projectVelocityF16):
//                                                src/sim/firex/pressureShaders.ts:49-160
//   * uniform block packing (writeUniforms):     src/sim/firex/FireXEngine.ts:2043-2113
//   * field reset (initializeFields host side):  FireXEngine.ts:1933-1974
//   * pass order / ping-pong / warm start:       FireXEngine.ts:2115-2276
//   * parameter defaults and clamps:             FireXEngine.ts:422-451, 937-1032
//   * volume ray marcher of the cinematic scene pass and its periodic detail
//     volume:                                    qualityShaders.ts:198-575,
//                                                FireXEngine.ts:311-355
//   * SPH droplets (density, integration, liquid deposition) and the host
//     particle state:                            shaders.ts:514-825,
//                                                FireXEngine.ts:1976-2034, 2158-2171
//
// Two precisions are instantiated.  State<double> is the FP64 yardstick: every
// operation in double, libm transcendentals, true division.  State<float>
// is the FP32 "reference step": f32 arithmetic in WGSL operation order with
// the f32 built-in layer of fx_math.h standing in for the WGSL built-ins
// (whose f32 results the WGSL spec leaves implementation-defined).
//
// Additions that the reference does not have (documented as such):
//   * seeded emitters (Gaussian blobs with seeded jitter and seeded velocity
//     noise) that replace emitter() when enabled;
//   * species storage quantised to f16 or bf16 (mixed-precision storage);
//   * FP64 residual and conservation sums.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <string>
#include <vector>
#include <algorithm>
#include <type_traits>

#include <omp.h>

#include "../fire_amd/csrc/fx_math.h"

namespace {

// ---------------------------------------------------------------- math policy
template <typename T> struct Math;

template <> struct Math<float> {
  static float exp(float x) { return fxm_exp(x); }
  static float sin(float x) { return fxm_sin(x); }
  static float cos(float x) { return fxm_cos(x); }
  static float pow(float x, float y) { return fxm_pow(x, y); }
  static float sqrt(float x) { return fxm_sqrt(x); }
  static float floor(float x) { return fxm_floor(x); }
};

template <> struct Math<double> {
  static double exp(double x) { return std::exp(x); }
  static double sin(double x) { return std::sin(x); }
  static double cos(double x) { return std::cos(x); }
  static double pow(double x, double y) { return std::pow(x, y); }
  static double sqrt(double x) { return std::sqrt(x); }
  static double floor(double x) { return std::floor(x); }
};

// WGSL min/max/clamp/mix/smoothstep, written as in fx_math.h: the f32 forms
// contract mix to fma(b, t, a (1 - t)) and smoothstep's 3 - 2t to
// fma(-2, t, 3) (fxm_fma); FP64 rounds every operation.
template <typename T> inline T wmin(T a, T b) { return (b < a) ? b : a; }
template <typename T> inline T wmax(T a, T b) { return (a < b) ? b : a; }
template <typename T> inline T wclamp(T x, T lo, T hi) { return wmin(wmax(x, lo), hi); }
inline float wfma(float a, float b, float c) { return fxm_fma(a, b, c); }
inline double wfma(double a, double b, double c) { return a * b + c; }
template <typename T> inline T wmix(T a, T b, T t) { return wfma(b, t, a * (T(1) - t)); }
// x / c for a divisor c that is a constant, a uniform or shared by several
// quotients: the f32 form multiplies by the f32 reciprocal of c (one of the
// results WGSL's 2.5 ULP f32 division admits), as the kernels do; FP64
// divides.
inline float wdivr(float x, float c) { return x * (1.0f / c); }
inline double wdivr(double x, double c) { return x / c; }
template <typename T> inline T wsmooth(T lo, T hi, T x) {
  T t = wclamp(wdivr(x - lo, hi - lo), T(0), T(1));
  return t * t * wfma(T(-2), t, T(3));
}

template <typename T> struct V3 { T x, y, z; };
template <typename T> struct V4 { T x, y, z, w; };

template <typename T> inline V4<T> mix4(const V4<T>& a, const V4<T>& b, T t) {
  return {wmix(a.x, b.x, t), wmix(a.y, b.y, t), wmix(a.z, b.z, t), wmix(a.w, b.w, t)};
}

// ------------------------------------------------------------ f16 / bf16 rounding
// Round-to-nearest-even conversions, used to model reduced-precision storage.
inline float round_to_f16(float v) {
  uint32_t u = fxm_float_to_bits(v);
  uint32_t sign = u & 0x80000000u;
  uint32_t au = u & 0x7fffffffu;
  if (au >= 0x7f800000u) return v;  // inf / nan
  float a = fxm_bits_to_float(au);
  if (a >= 65520.0f) return fxm_bits_to_float(sign | 0x7f800000u);
  float r;
  if (a < 6.103515625e-05f) {
    // f16 subnormal range: quantum 2^-24.
    float q = a * 16777216.0f;             // exact scaling
    float n = std::nearbyint(q);           // ties to even
    r = n / 16777216.0f;
  } else {
    int e;
    std::frexp(a, &e);                      // a = m * 2^e, m in [0.5, 1)
    float quantum = std::ldexp(1.0f, e - 11);  // 11 significant bits
    r = std::nearbyint(a / quantum) * quantum;
  }
  return fxm_bits_to_float(sign | fxm_float_to_bits(r));
}

inline float round_to_bf16(float v) {
  uint32_t u = fxm_float_to_bits(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return fxm_bits_to_float(0x7fc00000u);  // canonical nan
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return fxm_bits_to_float(u & 0xffff0000u);
}

// -------------------------------------------------------------- seeded noise
inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
// Uniform in [0, 1) with 24 bits, from (seed, frame, cell, stream).
inline float seeded_uniform(uint64_t seed, uint64_t frame, uint64_t cell, uint32_t stream) {
  uint64_t k = mix64(seed ^ mix64(frame));
  k = mix64(k ^ (cell * 4ull + stream));
  return (float)(uint32_t)(k >> 40) * 5.9604644775390625e-08f;  // 2^-24
}

// -------------------------------------------------------------------- params
// Engine parameter set, defaults from FireXEngine.ts:422-451 (Inferno).
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

inline double finite_or(double v, double fallback) { return std::isfinite(v) ? v : fallback; }
inline double clampd(double v, double lo, double hi) { return std::max(lo, std::min(hi, v)); }

// FireXEngine.setParameter, FireXEngine.ts:937-1032 (numeric keys).
bool set_parameter(Params& p, const char* name, double v) {
  std::string n(name);
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
  else if (n == "aimHeight") p.aimHeight = clampd(finite_or(v, p.aimHeight), 0.08, 0.82);
  else if (n == "nozzleType") p.nozzleType = finite_or(v, p.nozzleType) >= 0.5 ? 1 : 0;  // FireXEngine.ts:1000-1001
  else if (n == "domainWidth") p.domainWidth = clampd(finite_or(v, p.domainWidth), 0.5, 4);
  else if (n == "domainHeight") p.domainHeight = clampd(finite_or(v, p.domainHeight), 0.5, 6);
  else if (n == "domainDepth") p.domainDepth = clampd(finite_or(v, p.domainDepth), 0.5, 4);
  else if (n == "viewMode") p.viewMode = clampd(std::round(finite_or(v, p.viewMode)), 0, 7);  // FireXEngine.ts:994-995
  else if (n == "cameraYaw") p.cameraYaw = finite_or(v, p.cameraYaw);
  else if (n == "cameraPitch") p.cameraPitch = finite_or(v, p.cameraPitch);
  else if (n == "cameraDistance") p.cameraDistance = finite_or(v, p.cameraDistance);
  else if (n == "exposure") p.exposure = clampd(finite_or(v, p.exposure), 0.25, 3);
  else if (n == "frontDefinition") p.frontDefinition = clampd(finite_or(v, p.frontDefinition), 0.5, 2.5);
  else if (n == "viewZoom") p.viewZoom = clampd(finite_or(v, p.viewZoom), 0.5, 5);
  else if (n == "opticalDetailTarget") {
    double t = finite_or(v, p.opticalDetailTarget);
    p.opticalDetailTarget = t >= 768 ? 1024 : t >= 384 ? 512 : t >= 192 ? 256 : 0;  // FireXEngine.ts:1009-1011
  }
  else return false;
  return true;
}

// --------------------------------------------------------------- emitters
struct SeededEmitter {
  float cx, cy, cz;   // centre in cell units (cell i spans [i, i+1))
  float radius;       // cells; profile exp(-|p - c|^2 / radius^2)
  float amplitude;
};

struct EmitterSet {
  int count = 0;            // 0 = reference analytic emitter
  SeededEmitter e[8];
  float jitter = 0;         // source *= 1 + jitter * (2u - 1)
  float velocityNoise = 0;  // injected velocity noise amplitude
  uint64_t seed = 0;
};

// ------------------------------------------------------------------- state
enum { SPECIES_F32 = 0, SPECIES_F16 = 1, SPECIES_BF16 = 2 };
enum { HEAVY_AUTO = 0, HEAVY_OFF = 1, HEAVY_F16 = 2, HEAVY_F32 = 3 };
constexpr uint64_t kHeavyThreshold = 64ull * 1024 * 1024;  // gridConfiguration.ts:105

template <typename T> struct State {
  int nx = 0, ny = 0, nz = 0;
  size_t n = 0;
  std::vector<V4<T>> flow[2], species[2];
  std::vector<T> pressure[2], corr[2], divergence, reaction;
  std::vector<uint32_t> liquid;
  int speciesIndex = 0, pressureIndex = 0;
  double simulationTime = 0;
  double frameIndex = 0;
  int pressureIterations = 24, correctionIterations = 24;
  double timeStep = 1.0 / 60.0;
  bool cinematic = true;
  Params params;
  EmitterSet emitters;
  int speciesStorage = SPECIES_F32;
  int heavyMode = HEAVY_AUTO;
  std::string log;
  float u[44];  // uniform block as the device sees it

  // ---- uniform-derived values (T precision)
  T gx, gy, gz, dt, time, mx, my, mz, ex, ey, ez;

  size_t idx(int x, int y, int z) const {
    x = std::min(std::max(x, 0), nx - 1);
    y = std::min(std::max(y, 0), ny - 1);
    z = std::min(std::max(z, 0), nz - 1);
    return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
  }

  bool heavy() const {
    if (heavyMode == HEAVY_AUTO) return (uint64_t)n >= kHeavyThreshold;
    return heavyMode == HEAVY_F16 || heavyMode == HEAVY_F32;
  }
  bool f16Pressure() const {
    if (heavyMode == HEAVY_AUTO) return (uint64_t)n >= kHeavyThreshold;  // shader-f16 assumed
    return heavyMode == HEAVY_F16;
  }

  // writeUniforms, FireXEngine.ts:2043-2113 (solver-relevant slots).
  void write_uniforms(double deltaTime) {
    std::memset(u, 0, sizeof u);
    u[0] = (float)nx; u[1] = (float)ny; u[2] = (float)nz; u[3] = (float)n;
    u[4] = (float)deltaTime; u[5] = (float)simulationTime; u[6] = (float)frameIndex; u[7] = (float)timeStep;
    u[8] = (float)params.fuelRate; u[9] = (float)params.oxygenRate;
    u[10] = (float)params.heatEfficiency; u[11] = (float)params.vorticity;
    u[12] = (float)params.waterFlow; u[13] = (float)(params.sprayAngle * M_PI / 180.0);
    u[14] = params.waterEnabled ? 1.0f : 0.0f; u[15] = 300.0f;
    u[16] = (float)params.cameraYaw; u[17] = (float)params.cameraPitch; u[18] = (float)params.cameraDistance;
    u[19] = (float)((double)canvasWidth / (double)std::max(canvasHeight, 1));
    u[20] = (float)canvasWidth; u[21] = (float)canvasHeight; u[22] = (float)std::min(384, raySteps);
    u[23] = (float)particleCount;
    u[24] = (float)params.viewMode; u[25] = (float)params.aimHeight; u[26] = (float)params.nozzleType;
    u[27] = (float)params.exposure;
    u[28] = (float)params.burnerSize; u[29] = (float)params.burnerDepth;
    u[30] = (float)params.sourceThickness; u[31] = (float)params.firePower;
    u[32] = (float)params.domainWidth; u[33] = (float)params.domainHeight; u[34] = (float)params.domainDepth;
    u[36] = (float)params.sourceLift; u[37] = (float)params.buoyancyScale;
    u[38] = (float)params.flamePersistence; u[39] = (float)params.viewZoom;
    u[40] = (float)params.airEntrainment; u[41] = (float)params.frontDefinition;
    u[42] = cinematic ? 1.0f : 0.0f; u[43] = (float)params.opticalDetailTarget;
    derive();
  }

  void derive() {
    gx = T(u[0]); gy = T(u[1]); gz = T(u[2]);
    dt = T(u[4]); time = T(u[5]);
    ex = wmax(T(u[32]), T(0.001)); ey = wmax(T(u[33]), T(0.001)); ez = wmax(T(u[34]), T(0.001));
    mx = gx / ex; my = gy / ey; mz = gz / ez;
  }

  // ------------------------------------------------------------- helpers
  V4<T> trilinear(const std::vector<V4<T>>& f, T px, T py, T pz) const {
    T sx = wclamp(px - T(0.5), T(0), gx - T(1.001));
    T sy = wclamp(py - T(0.5), T(0), gy - T(1.001));
    T sz = wclamp(pz - T(0.5), T(0), gz - T(1.001));
    T flx = Math<T>::floor(sx), fly = Math<T>::floor(sy), flz = Math<T>::floor(sz);
    int bx = (int)flx, by = (int)fly, bz = (int)flz;
    T fx = sx - flx, fy = sy - fly, fz = sz - flz;
    const V4<T>& c000 = f[idx(bx, by, bz)];
    const V4<T>& c100 = f[idx(bx + 1, by, bz)];
    const V4<T>& c010 = f[idx(bx, by + 1, bz)];
    const V4<T>& c110 = f[idx(bx + 1, by + 1, bz)];
    const V4<T>& c001 = f[idx(bx, by, bz + 1)];
    const V4<T>& c101 = f[idx(bx + 1, by, bz + 1)];
    const V4<T>& c011 = f[idx(bx, by + 1, bz + 1)];
    const V4<T>& c111 = f[idx(bx + 1, by + 1, bz + 1)];
    V4<T> z0 = mix4(mix4(c000, c100, fx), mix4(c010, c110, fx), fy);
    V4<T> z1 = mix4(mix4(c001, c101, fx), mix4(c011, c111, fx), fy);
    return mix4(z0, z1, fz);
  }

  // emitter(), shaders.ts:101-133.
  T emitter_reference(T npx, T npy, T npz) const {
    T t = time;
    T px = npx * ex, py = npy * ey, pz = npz * ez;
    T cx = ex * T(0.5) + Math<T>::sin(t * T(1.73)) * T(0.013);
    T cy = T(0.075);
    T cz = ez * T(0.5) + Math<T>::cos(t * T(1.37)) * T(0.011);
    T ox = px - cx, oy = py - cy, oz = pz - cz;
    T baseRadius = T(std::sqrt(0.02024));
    T radiusX = wmin(baseRadius * wclamp(T(u[28]), T(0.35), T(3.2)), ex * T(0.42));
    T radiusZ = wmin(baseRadius * wclamp(T(u[29]), T(0.35), T(3.2)), ez * T(0.42));
    T thick = wclamp(T(u[30]), T(0.35), T(2.5));
    T radial = ox * ox / (radiusX * radiusX) + oz * oz / (radiusZ * radiusZ);
    T vertical = oy * oy / (T(0.0022) * thick * thick);
    T cellular = T(0.76) + T(0.24) * Math<T>::sin(px * T(49) + t * T(5.2)) * Math<T>::sin(pz * T(43) - t * T(4.4));
    T lobeArg = Math<T>::sin(ox / wmax(radiusX, T(0.001)) * T(6.2) + t * T(1.7))
              * Math<T>::sin(oz / wmax(radiusZ, T(0.001)) * T(5.4) - t * T(1.3));
    T lobes = wmix(T(0.48), T(1.18), wsmooth(T(-0.42), T(0.48), lobeArg));
    T texture = wmix(cellular, cellular * lobes, wclamp(T(u[42]), T(0), T(1)));
    T pulse = T(0.84) + T(0.16) * Math<T>::sin(t * T(3.7) + px * T(27) + pz * T(21));
    return Math<T>::exp(-(radial + vertical)) * texture * pulse;
  }

  // Seeded Gaussian emitters (addition; replaces emitter() when enabled).
  T emitter_seeded(T pcx, T pcy, T pcz, size_t cell) const {
    T s = T(0);
    for (int i = 0; i < emitters.count; ++i) {
      const SeededEmitter& e = emitters.e[i];
      T dx = pcx - T(e.cx), dy = pcy - T(e.cy), dz = pcz - T(e.cz);
      T d2 = dx * dx + dy * dy + dz * dz;
      T r = T(e.radius);
      s = s + T(e.amplitude) * Math<T>::exp(-(d2 / (r * r)));
    }
    if (s > T(0) && emitters.jitter > 0.0f) {
      T uj = T(seeded_uniform(emitters.seed, (uint64_t)frameIndex, cell, 0));
      s = s * (T(1) + T(emitters.jitter) * (T(2) * uj - T(1)));
    }
    return s;
  }

  // ---------------------------------------------------------- kernels
  // initializeFields host + kernel, FireXEngine.ts:1933-1974, shaders.ts:135-149.
  void initialize_fields() {
    for (int k = 0; k < 2; ++k) {
      flow[k].assign(n, V4<T>{0, 0, 0, 0});
      species[k].assign(n, V4<T>{0, 0, 0, 0});
      pressure[k].assign(n, T(0));
      corr[k].assign(n, T(0));
    }
    divergence.assign(n, T(0));
    reaction.assign(n, T(0));
    liquid.assign(n, 0u);
    // initialization uniforms: grid, chemistry.y = oxygenRate, water.w = 300;
    // flowOut = flow[0], speciesOut = species[0] (group at FireXEngine.ts:1667-1675).
    T o2 = wmax(T((float)params.oxygenRate), T(0));
    const V4<T> s0 = store_species(V4<T>{0, o2, 0, 0});
    for (size_t i = 0; i < n; ++i) {
      flow[0][i] = V4<T>{0, 0, 0, T(300)};
      species[0][i] = s0;
    }
    particleDensity.assign((size_t)particleCount * 4, T(0));
    write_particle_initial_state(water_active());
    liquidMayContainWater = false;
  }

  void simulate(int sIn) {
    const std::vector<V4<T>>& flowIn = flow[0];
    const std::vector<V4<T>>& specIn = species[sIn];
    std::vector<V4<T>>& flowOut = flow[1];
    std::vector<V4<T>>& specOut = species[1 - sIn];
    const T ambient = T(u[15]);
    const T chemX = T(u[8]), chemY = T(u[9]), chemZ = T(u[10]), chemW = T(u[11]);
    const T fireX = T(u[36]), fireY = T(u[37]), fireZ = T(u[38]);
    const T mixingX = T(u[40]);
    const T controlsY = T(u[25]);
    const bool seeded = emitters.count > 0;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z) {
      for (int y = 0; y < ny; ++y) {
        for (int x = 0; x < nx; ++x) {
          const size_t id = idx(x, y, z);
          T pcx = T(x) + T(0.5), pcy = T(y) + T(0.5), pcz = T(z) + T(0.5);
          T npx = pcx / gx, npy = pcy / gy, npz = pcz / gz;
          const V4<T> orig = flowIn[id];
          T bx = pcx - orig.x * dt * mx;
          T by = pcy - orig.y * dt * my;
          T bz = pcz - orig.z * dt * mz;
          V4<T> fl = trilinear(flowIn, bx, by, bz);
          V4<T> sp = trilinear(specIn, bx, by, bz);
          sp = V4<T>{wmax(sp.x, T(0)), wmax(sp.y, T(0)), wmax(sp.z, T(0)), wmax(sp.w, T(0))};

          const V4<T> L = flowIn[idx(x - 1, y, z)], R = flowIn[idx(x + 1, y, z)];
          const V4<T> B = flowIn[idx(x, y - 1, z)], Tp = flowIn[idx(x, y + 1, z)];
          const V4<T> K = flowIn[idx(x, y, z - 1)], F = flowIn[idx(x, y, z + 1)];
          T vax = wdivr(L.x + R.x + B.x + Tp.x + K.x + F.x, T(6));
          T vay = wdivr(L.y + R.y + B.y + Tp.y + K.y + F.y, T(6));
          T vaz = wdivr(L.z + R.z + B.z + Tp.z + K.z + F.z, T(6));
          T mv = wclamp(dt * T(0.32), T(0), T(0.12));
          fl.x = wmix(fl.x, vax, mv); fl.y = wmix(fl.y, vay, mv); fl.z = wmix(fl.z, vaz, mv);
          T ta = wdivr(L.w + R.w + B.w + Tp.w + K.w + F.w, T(6));
          fl.w = wmix(fl.w, ta, wclamp(dt * T(0.18), T(0), T(0.08)));

          const V4<T> sL = specIn[idx(x - 1, y, z)], sR = specIn[idx(x + 1, y, z)];
          const V4<T> sB = specIn[idx(x, y - 1, z)], sT = specIn[idx(x, y + 1, z)];
          const V4<T> sK = specIn[idx(x, y, z - 1)], sF = specIn[idx(x, y, z + 1)];
          V4<T> sa{wdivr(sL.x + sR.x + sB.x + sT.x + sK.x + sF.x, T(6)),
                   wdivr(sL.y + sR.y + sB.y + sT.y + sK.y + sF.y, T(6)),
                   wdivr(sL.z + sR.z + sB.z + sT.z + sK.z + sF.z, T(6)),
                   wdivr(sL.w + sR.w + sB.w + sT.w + sK.w + sF.w, T(6))};
          sp = mix4(sp, sa, wclamp(dt * T(0.12), T(0), T(0.06)));

          T source = seeded ? emitter_seeded(pcx, pcy, pcz, id) : emitter_reference(npx, npy, npz);
          T firePower = wclamp(T(u[31]), T(0.1), T(4));
          T powerRoot = Math<T>::sqrt(firePower);
          T thermalPower = wmin(powerRoot, T(1) + T(0.22) * (powerRoot - T(1)));
          sp.x = wmax(sp.x, source * chemX * firePower);
          fl.w = wmax(fl.w, ambient + source * T(1180) * thermalPower);
          T swx = Math<T>::sin(time * T(2.3) + npz * T(22));
          T swz = Math<T>::cos(time * T(1.9) + npx * T(25));
          T swk = source * dt * T(0.11) * chemW * powerRoot;
          fl.x = fl.x + swx * swk;
          fl.y = fl.y + T(0) * swk;
          fl.z = fl.z + swz * swk;
          T lift = source * dt * T(0.52) * wclamp(fireX, T(0), T(4)) * powerRoot;
          fl.y += lift;
          if (seeded && source > T(0) && emitters.velocityNoise > 0.0f) {
            T vn = T(emitters.velocityNoise);
            T nxv = (T(2) * T(seeded_uniform(emitters.seed, (uint64_t)frameIndex, id, 1)) - T(1)) * vn;
            T nyv = (T(2) * T(seeded_uniform(emitters.seed, (uint64_t)frameIndex, id, 2)) - T(1)) * vn;
            T nzv = (T(2) * T(seeded_uniform(emitters.seed, (uint64_t)frameIndex, id, 3)) - T(1)) * vn;
            fl.x = fl.x + nxv * lift;
            fl.y = fl.y + nyv * lift;
            fl.z = fl.z + nzv * lift;
          }

          bool ambientBoundary = x == 0 || x == nx - 1 || z == 0 || z == nz - 1 || y == ny - 1;
          if (ambientBoundary) {
            sp.y = wmax(sp.y, chemY);
            fl.w = wmix(fl.w, ambient, T(0.035));
          }

          T liq = T((float)liquid[id]) * T(1.0 / 4096.0);
          T water = wclamp(liq * T(3.4), T(0), T(5));
          T evaporation = water * wsmooth(T(335), T(1040), fl.w) * dt * T(13.5);
          fl.w = wmax(ambient, fl.w - evaporation * T(2080) - water * dt * T(260));
          sp.w = wclamp(sp.w + evaporation * T(2.15), T(0), T(3));

          T entrainment = wclamp(mixingX, T(0), T(3));
          T ambientOxygen = wmax(chemY, T(0.001));
          T oxygenDeficit = wclamp(wdivr(ambientOxygen - sp.y, ambientOxygen), T(0), T(1));
          T fuelCore = wsmooth(T(0.30), T(0.78), sp.x);
          T interfaceWeight = oxygenDeficit * wmix(T(1), T(0.18), fuelCore);
          T speed = Math<T>::sqrt(fl.x * fl.x + fl.y * fl.y + fl.z * fl.z);
          T entrainmentMix = wclamp(
              dt * entrainment * (T(0.12) + T(0.28) * speed + T(0.22) * wsmooth(T(400), T(1500), fl.w)) * interfaceWeight,
              T(0), T(0.08));
          sp.y = wmix(sp.y, ambientOxygen, entrainmentMix);

          T fuelBefore = wmax(sp.x, T(0));
          T oxygenBefore = wmax(sp.y, T(0));
          T availableFuel = wmin(fuelBefore, oxygenBefore * T(0.25));
          T safeFuel = wmax(fuelBefore, T(0.00001));
          T safeOxygen = wmax(oxygenBefore, T(0.00001));
          T activation = Math<T>::exp(-T(15000) / wmax(fl.w, ambient));
          T arrhenius = wclamp(T(2.5e6) * activation * Math<T>::pow(safeFuel, T(-0.30)) * Math<T>::pow(safeOxygen, T(1.30)),
                               T(0), T(80));
          T ignition = wsmooth(T(690), T(930), fl.w);
          T requestedBurn = availableFuel * (T(1) - Math<T>::exp(-arrhenius * ignition * dt)) * Math<T>::exp(-water * T(18));
          T burn = wclamp(requestedBurn, T(0), availableFuel);
          T richness = wclamp(T(1) - oxygenBefore / wmax(fuelBefore * T(4), T(0.00001)), T(0), T(1));
          sp.x = wmax(T(0), fuelBefore - burn);
          sp.y = wmax(T(0), oxygenBefore - burn * T(4));
          sp.z = wclamp(sp.z + burn * (T(1) + richness * T(1.8)), T(0), T(3));
          sp.w = wclamp(sp.w + burn * T(0.62), T(0), T(3));
          fl.w += burn * (T(2150) + T(1450) * chemZ);
          T react = wclamp(wdivr(burn, wmax(dt, T(0.0001))) * T(0.22), T(0), T(4));

          T curlx = T(0.5) * ((Tp.z - B.z) * my - (F.y - K.y) * mz);
          T curly = T(0.5) * ((F.x - K.x) * mz - (R.z - L.z) * mx);
          T curlz = T(0.5) * ((R.y - L.y) * mx - (Tp.x - B.x) * my);
          T gdx = T(0.5) * ((R.w - L.w) * mx) + T(0.00001);
          T gdy = T(0.5) * ((Tp.w - B.w) * my) + T(0.00001);
          T gdz = T(0.5) * ((F.w - K.w) * mz) + T(0.00001);
          T glen = Math<T>::sqrt(gdx * gdx + gdy * gdy + gdz * gdz);
          T dirx = wdivr(gdx, glen), diry = wdivr(gdy, glen), dirz = wdivr(gdz, glen);
          T vk = T(0.00135) * chemW * dt;
          fl.x = fl.x + (diry * curlz - dirz * curly) * vk;
          fl.y = fl.y + (dirz * curlx - dirx * curlz) * vk;
          fl.z = fl.z + (dirx * curly - diry * curlx) * vk;

          T hotPlume = wsmooth(T(520), T(1650), fl.w);
          T heightEnvelope = wsmooth(T(0.045), T(0.15), npy) * (T(1) - wsmooth(T(0.80), T(1.0), npy));
          T ltx = Math<T>::sin(npy * T(31) + npz * T(19) + time * T(3.1)) + T(0.45) * Math<T>::sin(npy * T(67) - time * T(5.3));
          T lty = T(0.12) * Math<T>::sin(npx * T(29) + npz * T(23) - time * T(4.1));
          T ltz = Math<T>::cos(npy * T(27) - npx * T(21) + time * T(2.7)) + T(0.45) * Math<T>::cos(npy * T(59) + time * T(4.7));
          T turbulenceStrength = (T(0.075) + T(0.16) * chemW) * hotPlume * heightEnvelope;
          T tk = turbulenceStrength * dt;
          fl.x = fl.x + ltx * tk;
          fl.y = fl.y + lty * tk;
          fl.z = fl.z + ltz * tk;

          T buoyancy = wclamp(wdivr(fl.w - ambient, T(1500)), T(-0.1), T(1.2));
          fl.y += buoyancy * dt * T(0.72) * wclamp(fireY, T(0), T(4));

          sp.z *= Math<T>::exp(-water * dt * T(4.8));
          T nox = T(0.035) * ex, noy = T(0.50) * ey, noz = T(0.5) * ez;
          T stx = T(0.52) * ex, sty = controlsY * ey, stz = T(0.5) * ez;
          T smx = stx - nox, smy = sty - noy, smz = stz - noz;
          T slen = Math<T>::sqrt(smx * smx + smy * smy + smz * smz);
          smx = smx / slen; smy = smy / slen; smz = smz / slen;
          fl.x = fl.x + smx * water * dt * T(0.82);
          fl.y = fl.y + smy * water * dt * T(0.82);
          fl.z = fl.z + smz * water * dt * T(0.82);

          T persistence = wclamp(fireZ, T(0.4), T(3));
          T radiation = wdivr(T(5.67e-8) * (Math<T>::pow(ambient, T(4)) - Math<T>::pow(wmax(fl.w, T(0)), T(4))),
                              T(1005) * persistence);
          fl.w = wmax(ambient, fl.w + dt * radiation);
          sp.z *= Math<T>::exp(-dt * T(0.075) / persistence);
          sp.w *= Math<T>::exp(-dt * T(0.14) / persistence);

          T outflow = wsmooth(T(0.84), T(0.995), npy);
          sp.x *= T(1) - outflow;
          T outflowDecay = Math<T>::exp(-outflow * dt * T(15) / persistence);
          sp.z *= outflowDecay;
          sp.w *= outflowDecay;
          fl.w = wmix(fl.w, ambient, outflow * T(0.16) / persistence);
          react *= T(1) - outflow;
          if (y == ny - 1) {
            sp.x = T(0); sp.z = T(0); sp.w = T(0);
            fl.w = ambient;
            react = T(0);
          }
          if (x == 0 || x == nx - 1) fl.x = T(0);
          if (z == 0 || z == nz - 1) fl.z = T(0);
          if (y == 0) fl.y = T(0);
          if (y == ny - 1) fl.y = wmax(T(0), fl.y);
          fl.x = wclamp(fl.x, T(-1.4), T(1.4));
          fl.y = wclamp(fl.y, T(-1.4), T(1.4));
          fl.z = wclamp(fl.z, T(-1.4), T(1.4));
          fl.w = wclamp(fl.w, ambient, T(2900));
          sp.x = wclamp(sp.x, T(0), T(3)); sp.y = wclamp(sp.y, T(0), T(3));
          sp.z = wclamp(sp.z, T(0), T(3)); sp.w = wclamp(sp.w, T(0), T(3));
          flowOut[id] = fl;
          specOut[id] = store_species(sp);
          reaction[id] = react;
        }
      }
    }
  }

  // ------------------------------------------------------------- droplets
  // The SPH droplet passes of FIREX_PARTICLE_WGSL (shaders.ts:514-825) and
  // the host particle state (FireXEngine.ts:1976-2034).  Particle records are
  // 8 values (positionAge: position in domain fractions, age; velocityMass:
  // velocity in m/s, mass); density records 4 (density, near density,
  // pressure, near pressure).
  int particleCount = 0;
  std::vector<T> particles[2];
  std::vector<T> particleDensity;
  int particleIndex = 0;
  bool liquidMayContainWater = false, particleFieldActive = false;

  // waterSimulationActive, FireXEngine.ts:2537-2539.
  bool water_active() const { return params.waterEnabled && params.waterFlow > 0.001; }

  // Math.hypot as the JS engine evaluates it: scale by the largest magnitude,
  // Kahan-compensated sum of squares, sqrt, rescale.
  static double js_hypot(const double* v, int n) {
    double mx = 0;
    for (int i = 0; i < n; ++i) {
      double a = std::fabs(v[i]);
      if (std::isinf(a)) return INFINITY;
      if (std::isnan(a)) mx = NAN;
      else if (a > mx) mx = a;
    }
    if (std::isnan(mx)) return NAN;
    if (mx == 0) return 0;
    double sum = 0, comp = 0;
    for (int i = 0; i < n; ++i) {
      double r = std::fabs(v[i]) / mx;
      double summand = r * r - comp;
      double prelim = sum + summand;
      comp = (prelim - sum) - summand;
      sum = prelim;
    }
    return std::sqrt(sum) * mx;
  }

  // writeParticleInitialState, FireXEngine.ts:1976-2034: float64 arithmetic,
  // stored through a Float32Array.
  void write_particle_initial_state(bool active) {
    const int N = particleCount;
    std::vector<float> init((size_t)N * 8, 0.0f);
    const double W = params.domainWidth, H = params.domainHeight, D = params.domainDepth;
    const double aim[2] = {(0.52 - 0.035) * W, (params.aimHeight - 0.5) * H};
    const double aimLength = js_hypot(aim, 2);
    const double base[3] = {aim[0] / aimLength, aim[1] / aimLength, 0};
    const double tangent[2] = {base[1], -base[0]};
    auto hash = [](double v) {
      double h = std::sin(v * 12.9898 + 1.317) * 43758.5453;
      return h - std::floor(h);
    };
    const bool spray = params.nozzleType >= 0.5;
    for (int i = 0; i < N; ++i) {
      const double phase = (double)i / (double)N;
      const double angle = hash(i * 3.771 + 4.91) * M_PI * 2;
      const double radial = std::sqrt(hash(i * 3.771 + 0.17));
      const double spread = spray ? 0.72 : 0.1;
      const double cone = std::tan((params.sprayAngle * M_PI) / 180) * radial * spread;
      const double diskX = tangent[0] * std::cos(angle);
      const double diskY = tangent[1] * std::cos(angle);
      const double diskZ = -std::sin(angle);
      double d[3] = {base[0] + diskX * cone, base[1] + diskY * cone, diskZ * cone};
      const double len = js_hypot(d, 3);
      d[0] /= len;
      d[1] /= len;
      d[2] /= len;
      const double variation = spray ? 0.24 : 0.05;
      const double speed = (0.82 + params.waterFlow * 0.82) * (1 - variation * 0.5 + hash(i * 3.771 + 9.37) * variation);
      const double age = phase * 0.72;
      const double originRadius = spray ? 0.006 : 0.0028;
      float* o = &init[(size_t)i * 8];
      o[0] = (float)(0.035 + (diskX * radial * originRadius + d[0] * speed * age) / W);
      o[1] = (float)std::max(0.012, 0.5 + (diskY * radial * originRadius + d[1] * speed * age - 0.29 * age * age) / H);
      o[2] = (float)(0.5 + (diskZ * radial * originRadius + d[2] * speed * age) / D);
      o[3] = (float)(active ? age : -phase * 0.72);
      o[4] = (float)(active ? d[0] * speed : 0);
      o[5] = (float)(active ? d[1] * speed - 0.58 * age : 0);
      o[6] = (float)(active ? d[2] * speed : 0);
      o[7] = (float)(active ? 0.52 + params.waterFlow * 0.88 : 0);
    }
    for (int k = 0; k < 2; ++k) {
      particles[k].resize(init.size());
      for (size_t j = 0; j < init.size(); ++j) particles[k][j] = T(init[j]);
    }
    particleIndex = 0;
    particleFieldActive = active;
  }

  // Resource rebuild with a new particle count (installResourceSet,
  // FireXEngine.ts:1917-1931, particle part).
  void set_particle_count(int count) {
    particleCount = count;
    particleDensity.assign((size_t)count * 4, T(0));
    liquidMayContainWater = false;
    particleFieldActive = false;
    write_particle_initial_state(water_active());
  }

  // hash11, shaders.ts:550-552.
  T hash11(T v) const {
    T h = Math<T>::sin(v * T(12.9898) + T(1.317)) * T(43758.5453);
    return h - Math<T>::floor(h);
  }
  static T len3(T x, T y, T z) { return Math<T>::sqrt(x * x + y * y + z * z); }
  bool spray_nozzle() const { return T(u[26]) >= T(0.5); }  // params.controls.z

  // spawnParticle, shaders.ts:579-601.
  void spawn(uint32_t id, T cycle, T* o) const {
    const T seed = T((float)id) * T(3.771) + cycle * T(19.173);
    const T radialRandom = Math<T>::sqrt(hash11(seed + T(0.17)));
    const T angularRandom = hash11(seed + T(4.91)) * T(6.2831853);
    const T speedRandom = hash11(seed + T(9.37));
    // nozzleDirection(), shaders.ts:575-578
    T dx = (T(0.52) - T(0.035)) * ex;
    T dy = (wclamp(T(u[25]), T(0.08), T(0.82)) - T(0.50)) * ey;
    T dz = (T(0.5) - T(0.5)) * ez;
    T l = len3(dx, dy, dz);
    dx = dx / l; dy = dy / l; dz = dz / l;
    // tangent = normalize(cross(direction, (0, 0, 1)))
    T tx = dy * T(1) - dz * T(0), ty = dz * T(0) - dx * T(1), tz = dx * T(0) - dy * T(0);
    l = len3(tx, ty, tz);
    tx = tx / l; ty = ty / l; tz = tz / l;
    // bitangent = normalize(cross(direction, tangent))
    T bx = dy * tz - dz * ty, by = dz * tx - dx * tz, bz = dx * ty - dy * tx;
    l = len3(bx, by, bz);
    bx = bx / l; by = by / l; bz = bz / l;
    const T ca = Math<T>::cos(angularRandom), sa = Math<T>::sin(angularRandom);
    const T kx = tx * ca + bx * sa, ky = ty * ca + by * sa, kz = tz * ca + bz * sa;
    const bool spray = spray_nozzle();
    const T nozzleSpread = spray ? T(0.72) : T(0.10);
    const T ang = wclamp(T(u[13]), T(0.005), T(0.78));
    const T cone = Math<T>::sin(ang) / Math<T>::cos(ang) * radialRandom * nozzleSpread;
    T lx = dx + kx * cone, ly = dy + ky * cone, lz = dz + kz * cone;
    l = len3(lx, ly, lz);
    lx = lx / l; ly = ly / l; lz = lz / l;
    const T originRadius = spray ? T(0.006) : T(0.0028);
    const T speedVariation = spray ? T(0.24) : T(0.05);
    const T speed = (T(0.82) + T(u[12]) * T(0.82)) *
                    wmix(T(1) - speedVariation * T(0.5), T(1) + speedVariation * T(0.5), speedRandom);
    o[0] = T(0.035) + kx * radialRandom * originRadius / ex;
    o[1] = T(0.50) + ky * radialRandom * originRadius / ey;
    o[2] = T(0.5) + kz * radialRandom * originRadius / ez;
    o[3] = T(0);
    o[4] = lx * speed; o[5] = ly * speed; o[6] = lz * speed;
    o[7] = T(0.52) + T(u[12]) * T(0.88);
  }

  // computeParticleDensity, shaders.ts:603-637.
  void sph_density(const std::vector<T>& pin) {
    const int N = particleCount;
    const T radius = spray_nozzle() ? T(0.060) : T(0.038);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const T* p = &pin[(size_t)i * 8];
      T* out = &particleDensity[(size_t)i * 4];
      if (p[7] <= T(0.0001)) {
        out[0] = out[1] = out[2] = out[3] = T(0);
        continue;
      }
      T density = T(0), nearDensity = T(0);
      for (int j = 0; j < N; ++j) {
        if (j == i) continue;
        const T* q = &pin[(size_t)j * 8];
        if (q[7] <= T(0.0001)) continue;
        const T distance = len3((p[0] - q[0]) * ex, (p[1] - q[1]) * ey, (p[2] - q[2]) * ez);
        if (distance > T(0.00001) && distance < radius) {
          const T w = T(1) - distance / radius;
          density = density + w * w;
          nearDensity = nearDensity + w * w * w;
        }
      }
      const T restDensity = T(2.7) + T(u[12]) * T(0.9);
      out[0] = density;
      out[1] = nearDensity;
      out[2] = wmax(density - restDensity, T(0)) * T(0.055);
      out[3] = nearDensity * T(0.065);
    }
  }

  // integrateParticles, shaders.ts:639-735.
  void sph_integrate(const std::vector<T>& pin, std::vector<T>& pout) {
    const int N = particleCount;
    const T dt = T(u[4]);
    const bool spray = spray_nozzle();
    const T radius = spray ? T(0.060) : T(0.038);
    const bool respawn = T(u[14]) >= T(0.5) && T(u[12]) > T(0.001);
#pragma omp parallel for schedule(static)
    for (int id = 0; id < N; ++id) {
      const T* p = &pin[(size_t)id * 8];
      T* o = &pout[(size_t)id * 8];
      if (dt <= T(0)) {
        for (int k = 0; k < 8; ++k) o[k] = p[k];
        continue;
      }
      if (p[7] <= T(0.0001)) {
        const T wait = p[3] + dt;
        if (respawn && wait >= T(0)) {
          spawn((uint32_t)id, wait + T(u[6]), o);
        } else {
          o[0] = T(0.035); o[1] = T(0.50); o[2] = T(0.5);
          o[3] = wait > T(0) ? T(-0.72) : wait;
          o[4] = o[5] = o[6] = o[7] = T(0);
        }
        continue;
      }
      const T lifetime = T(0.88) + hash11(T((float)id) * T(2.31) + T(7.0)) * T(0.52);
      const T nextAge = p[3] + dt;
      const bool outside = p[0] < T(-0.03) || p[1] < T(-0.03) || p[2] < T(-0.03) ||
                           p[0] > T(1.03) || p[1] > T(1.03) || p[2] > T(1.03);
      if (nextAge >= lifetime || outside) {
        if (respawn) {
          spawn((uint32_t)id, nextAge + T(u[6]), o);
        } else {
          o[0] = T(0.035); o[1] = T(0.50); o[2] = T(0.5);
          o[3] = -hash11(T((float)id) * T(8.17)) * T(0.72);
          o[4] = o[5] = o[6] = o[7] = T(0);
        }
        continue;
      }
      T vx = p[4], vy = p[5], vz = p[6];
      const T* pd = &particleDensity[(size_t)id * 4];
      T pax = T(0), pay = T(0), paz = T(0), vax = T(0), vay = T(0), vaz = T(0);
      for (int j = 0; j < N; ++j) {
        if (j == id) continue;
        const T* q = &pin[(size_t)j * 8];
        if (q[7] <= T(0.0001)) continue;
        const T sx = (p[0] - q[0]) * ex, sy = (p[1] - q[1]) * ey, sz = (p[2] - q[2]) * ez;
        const T distance = len3(sx, sy, sz);
        if (distance > T(0.00001) && distance < radius) {
          const T w = T(1) - distance / radius;
          const T dx = sx / distance, dy = sy / distance, dz = sz / distance;
          const T* nd = &particleDensity[(size_t)j * 4];
          const T pressure = (pd[2] + nd[2]) * w;
          const T nearPressure = (pd[3] + nd[3]) * w * w;
          const T push = pressure + nearPressure;
          pax = pax + dx * push; pay = pay + dy * push; paz = paz + dz * push;
          const T visc = w * T(0.018);
          vax = vax + (q[4] - vx) * visc; vay = vay + (q[5] - vy) * visc; vaz = vaz + (q[6] - vz) * visc;
        }
      }
      // sampleFlow, shaders.ts:566-569: flow[0] at the cell holding the position
      const int cx = (int)(wclamp(p[0], T(0), T(0.99999)) * gx);
      const int cy = (int)(wclamp(p[1], T(0), T(0.99999)) * gy);
      const int cz = (int)(wclamp(p[2], T(0), T(0.99999)) * gz);
      const V4<T> gas = flow[0][idx(cx, cy, cz)];
      const T heat = wsmooth(T(380), T(1500), gas.w);
      const T drag = T(0.30) + heat * T(0.18);
      T ax = T(0) + (gas.x - vx) * drag, ay = T(-0.58) + (gas.y - vy) * drag, az = T(0) + (gas.z - vz) * drag;
      const T pressureDispersion = spray ? T(1) : T(0.055);
      const T viscousCohesion = spray ? T(1) : T(2.2);
      ax = ax + (pax * pressureDispersion + vax * viscousCohesion);
      ay = ay + (pay * pressureDispersion + vay * viscousCohesion);
      az = az + (paz * pressureDispersion + vaz * viscousCohesion);
      vx = vx + ax * dt; vy = vy + ay * dt; vz = vz + az * dt;
      T nx_ = p[0] + vx * dt / ex, ny_ = p[1] + vy * dt / ey, nz_ = p[2] + vz * dt / ez;
      if (ny_ < T(0.012)) {
        ny_ = T(0.012);
        vy = std::fabs(vy) * T(0.18);
        vx = vx * T(0.72);
        vz = vz * T(0.72);
      }
      if (nx_ < T(0.008) || nx_ > T(0.992)) {
        nx_ = wclamp(nx_, T(0.008), T(0.992));
        vx = vx * T(-0.24);
      }
      if (nz_ < T(0.008) || nz_ > T(0.992)) {
        nz_ = wclamp(nz_, T(0.008), T(0.992));
        vz = vz * T(-0.24);
      }
      const T remainingMass = p[7] * Math<T>::exp(-heat * dt * T(0.34));
      o[0] = nx_; o[1] = ny_; o[2] = nz_; o[3] = nextAge;
      o[4] = wclamp(vx, T(-2), T(2)); o[5] = wclamp(vy, T(-2), T(2)); o[6] = wclamp(vz, T(-2), T(2));
      o[7] = remainingMass;
    }
  }

  // depositCorner, shaders.ts:737-740 (a NaN amount deposits nothing).
  void deposit_corner(int x, int y, int z, T weight, T mass) {
    const T v = wclamp(weight * mass * T(4096), T(0), T(65535));
    liquid[idx(x, y, z)] += (v == v) ? (uint32_t)v : 0u;
  }

  // depositAtPosition, shaders.ts:742-791.
  void deposit_at(T px, T py, T pz, T mass) {
    const T vx = px * gx - T(0.5), vy = py * gy - T(0.5), vz = pz * gz - T(0.5);
    if (!spray_nozzle()) {
      const T flx = Math<T>::floor(vx), fly = Math<T>::floor(vy), flz = Math<T>::floor(vz);
      const int bx = (int)flx, by = (int)fly, bz = (int)flz;
      const T fx = vx - flx, fy = vy - fly, fz = vz - flz;
      const T ix = T(1) - fx, iy = T(1) - fy, iz = T(1) - fz;
      const T m = mass * T(0.20);
      deposit_corner(bx, by, bz, ix * iy * iz, m);
      deposit_corner(bx + 1, by, bz, fx * iy * iz, m);
      deposit_corner(bx, by + 1, bz, ix * fy * iz, m);
      deposit_corner(bx + 1, by + 1, bz, fx * fy * iz, m);
      deposit_corner(bx, by, bz + 1, ix * iy * fz, m);
      deposit_corner(bx + 1, by, bz + 1, fx * iy * fz, m);
      deposit_corner(bx, by + 1, bz + 1, ix * fy * fz, m);
      deposit_corner(bx + 1, by + 1, bz + 1, fx * fy * fz, m);
      return;
    }
    const int cx = (int)std::nearbyint(vx), cy = (int)std::nearbyint(vy), cz = (int)std::nearbyint(vz);
    T kernel[27];
    T total = T(0);
    int k = 0;
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox, ++k) {
          const T distance = len3(vx - T((float)(cx + ox)), vy - T((float)(cy + oy)), vz - T((float)(cz + oz)));
          const T w = wmax(T(0), T(1) - distance * T(0.52));
          kernel[k] = w;
          total = total + w * w;
        }
    k = 0;
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox, ++k)
          deposit_corner(cx + ox, cy + oy, cz + oz, kernel[k] * kernel[k] / wmax(total, T(0.0001)), mass);
  }

  // depositParticles, shaders.ts:793-825.
  void liquid_deposit(const std::vector<T>& current, const std::vector<T>& previous) {
    for (int id = 0; id < particleCount; ++id) {
      const T* p = &current[(size_t)id * 8];
      const T mass = p[7];
      if (mass <= T(0.0001) || p[0] < T(0) || p[1] < T(0) || p[2] < T(0) || p[0] >= T(1) || p[1] >= T(1) ||
          p[2] >= T(1))
        continue;
      const T* q = &previous[(size_t)id * 8];
      const T dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
      const bool canSweep = q[7] > T(0.0001) && p[3] >= q[3] && len3(dx * ex, dy * ey, dz * ez) < T(0.20);
      if (!canSweep) {
        deposit_at(p[0], p[1], p[2], mass);
        continue;
      }
      for (int s = 0; s < 3; ++s) {
        const T f = T(s) / T(2);
        deposit_at(q[0] + dx * f, q[1] + dy * f, q[2] + dz * f, mass / T(3));
      }
    }
  }

  // The water branch at the head of encodeSimulationStep (FireXEngine.ts:2158-2171).
  void droplet_passes() {
    if (water_active()) {
      if (!particleFieldActive) write_particle_initial_state(true);
      std::fill(liquid.begin(), liquid.end(), 0u);
      sph_density(particles[particleIndex]);
      sph_integrate(particles[particleIndex], particles[1 - particleIndex]);
      liquid_deposit(particles[1 - particleIndex], particles[particleIndex]);
      particleIndex = 1 - particleIndex;
      liquidMayContainWater = true;
      log += "clear_liquid;sph_density;sph_integrate;liquid_deposit;";
    } else if (liquidMayContainWater || particleFieldActive) {
      std::fill(liquid.begin(), liquid.end(), 0u);
      write_particle_initial_state(false);
      liquidMayContainWater = false;
      log += "clear_liquid;particle_reset;";
    }
  }

  // ------------------------------------------------------------- rendering
  // The cinematic scene pass's volume ray marcher (volumeFragment,
  // qualityShaders.ts:374-575) over the packed render volumes, FP32 only.
  // Texture filtering is linear interpolation in f32 (texel t = uvw * size
  // - 0.5; x, then y, then z; clamp-to-edge for the render volumes, repeat
  // for the detail volume, rgba8 texels as byte / 255); a pixel (x, y) has
  // uv = ((x + 0.5) / w, 1 - (y + 0.5) / h) and position (x + 0.5, y + 0.5).
  int canvasWidth = 0, canvasHeight = 0, raySteps = 0;

  // createPeriodicDetailVolume, FireXEngine.ts:311-355.
  static std::vector<uint8_t> detail_volume(int size) {
    std::vector<uint8_t> data((size_t)size * size * size * 4);
    const int period = size - 1;
    auto hash = [](int32_t x, int32_t y, int32_t z, int32_t seed) {
      uint32_t v = (uint32_t)(int32_t)((uint32_t)x * 73856093u) ^ (uint32_t)(int32_t)((uint32_t)y * 19349663u) ^
                   (uint32_t)(int32_t)((uint32_t)z * 83492791u) ^ (uint32_t)seed;
      v ^= v << 13;
      v ^= v >> 17;
      v ^= v << 5;
      return (double)((v >> 24) & 255u);
    };
    auto fade = [](double v) { return v * v * (3 - 2 * v); };
    auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
    auto noise = [&](int x, int y, int z, int f, int32_t seed) {
      double px = (double)x / period * f, py = (double)y / period * f, pz = (double)z / period * f;
      int x0 = (int)std::floor(px) % f, y0 = (int)std::floor(py) % f, z0 = (int)std::floor(pz) % f;
      int x1 = (x0 + 1) % f, y1 = (y0 + 1) % f, z1 = (z0 + 1) % f;
      double tx = fade(px - std::floor(px)), ty = fade(py - std::floor(py)), tz = fade(pz - std::floor(pz));
      double lo = lerp(lerp(hash(x0, y0, z0, seed), hash(x1, y0, z0, seed), tx),
                       lerp(hash(x0, y1, z0, seed), hash(x1, y1, z0, seed), tx), ty);
      double hi = lerp(lerp(hash(x0, y0, z1, seed), hash(x1, y0, z1, seed), tx),
                       lerp(hash(x0, y1, z1, seed), hash(x1, y1, z1, seed), tx), ty);
      return (uint8_t)std::floor(lerp(lo, hi, tz) + 0.5);  // Math.round
    };
    for (int z = 0; z < size; ++z)
      for (int y = 0; y < size; ++y)
        for (int x = 0; x < size; ++x) {
          uint8_t* o = &data[((size_t)x + (size_t)size * ((size_t)y + (size_t)size * z)) * 4];
          o[0] = noise(x, y, z, 4, (int32_t)0x51f15e5d);
          o[1] = noise(x, y, z, 8, (int32_t)0x6d2b79f5);
          o[2] = noise(x, y, z, 14, (int32_t)0x1b873593);
          o[3] = 255;
        }
    return data;
  }

  struct Renderer {
    typedef V3<float> F3;
    const State<float>& s;
    const std::vector<V4<float>>& vf;
    const std::vector<V4<float>>& vc;
    const std::vector<V4<float>>& vm;
    const std::vector<uint8_t>& det;
    const float* u;
    int width, height;

    static F3 f3(float x, float y, float z) { return F3{x, y, z}; }
    static F3 plus(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
    static F3 minus(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
    static F3 scale(F3 a, float k) { return f3(a.x * k, a.y * k, a.z * k); }
    static F3 times(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
    static float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    static float len(F3 a) { return fxm_sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
    static F3 unit(F3 a) { float l = len(a); return f3(a.x / l, a.y / l, a.z / l); }
    static F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
    static F3 lerp3(F3 a, F3 b, float t) { return f3(fxm_mix(a.x, b.x, t), fxm_mix(a.y, b.y, t), fxm_mix(a.z, b.z, t)); }
    static V4<float> lerp4(V4<float> a, V4<float> b, float t) {
      return {fxm_mix(a.x, b.x, t), fxm_mix(a.y, b.y, t), fxm_mix(a.z, b.z, t), fxm_mix(a.w, b.w, t)};
    }

    float ext(int a) const { return fxm_max(u[32 + a], 0.001f); }

    V4<float> filtered(const std::vector<V4<float>>& v, float ux, float uy, float uz) const {
      const int n[3] = {s.nx, s.ny, s.nz};
      const float g[3] = {u[0], u[1], u[2]};
      const float c[3] = {ux, uy, uz};
      int i0[3], i1[3];
      float fr[3];
      for (int a = 0; a < 3; ++a) {
        float t = c[a] * g[a] - 0.5f;
        float fl = fxm_floor(t);
        fr[a] = t - fl;
        int i = (int)fl;
        i0[a] = std::min(std::max(i, 0), n[a] - 1);
        i1[a] = std::min(std::max(i + 1, 0), n[a] - 1);
      }
      auto at = [&](int x, int y, int z) { return v[(size_t)x + (size_t)s.nx * ((size_t)y + (size_t)s.ny * z)]; };
      V4<float> a = lerp4(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), fr[0]);
      V4<float> b = lerp4(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), fr[0]);
      V4<float> cc = lerp4(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), fr[0]);
      V4<float> d = lerp4(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), fr[0]);
      return lerp4(lerp4(a, b, fr[1]), lerp4(cc, d, fr[1]), fr[2]);
    }

    // sampleFlow / sampleChemistry / sampleMedia, qualityShaders.ts:262-278
    V4<float> field(const std::vector<V4<float>>& v, F3 p) const {
      float h[3] = {0.5f / u[0], 0.5f / u[1], 0.5f / u[2]};
      float q[3] = {p.x / ext(0), p.y / ext(1), p.z / ext(2)};
      for (int a = 0; a < 3; ++a) q[a] = fxm_clamp(q[a], h[a], 1.0f - h[a]);
      return filtered(v, q[0], q[1], q[2]);
    }

    // sampleDetail, qualityShaders.ts:280-282
    V4<float> detail(F3 p) const {
      const int S = 64;
      float c[3] = {p.x, p.y, p.z};
      int i0[3];
      float fr[3];
      for (int a = 0; a < 3; ++a) {
        float t = c[a] * (float)S - 0.5f;
        float fl = fxm_floor(t);
        fr[a] = t - fl;
        i0[a] = (int)fl;
      }
      auto wrap = [&](int i) { return ((i % S) + S) % S; };
      auto at = [&](int x, int y, int z) {
        const uint8_t* t = &det[((size_t)wrap(x) + (size_t)S * ((size_t)wrap(y) + (size_t)S * wrap(z))) * 4];
        return V4<float>{(float)t[0] / 255.0f, (float)t[1] / 255.0f, (float)t[2] / 255.0f, (float)t[3] / 255.0f};
      };
      int x = i0[0], y = i0[1], z = i0[2];
      V4<float> a = lerp4(at(x, y, z), at(x + 1, y, z), fr[0]);
      V4<float> b = lerp4(at(x, y + 1, z), at(x + 1, y + 1, z), fr[0]);
      V4<float> cc = lerp4(at(x, y, z + 1), at(x + 1, y, z + 1), fr[0]);
      V4<float> d = lerp4(at(x, y + 1, z + 1), at(x + 1, y + 1, z + 1), fr[0]);
      return lerp4(lerp4(a, b, fr[1]), lerp4(cc, d, fr[1]), fr[2]);
    }

    static F3 blackbody(float temp) {
      float t = fxm_clamp((temp - 720.0f) * (1.0f / 1900.0f), 0.0f, 1.0f);
      F3 red = f3(1.0f, 0.035f, 0.002f), orange = f3(1.0f, 0.24f, 0.008f);
      F3 gold = f3(1.0f, 0.68f, 0.12f), white = f3(1.0f, 0.94f, 0.72f);
      return lerp3(lerp3(red, orange, fxm_smoothstep(0.0f, 0.32f, t)), lerp3(gold, white, fxm_smoothstep(0.58f, 1.0f, t)),
                   fxm_smoothstep(0.26f, 0.72f, t));
    }
    static F3 thermal_map(float value) {
      float t = fxm_clamp(value, 0.0f, 1.0f);
      F3 cold = f3(0.015f, 0.025f, 0.12f), blue = f3(0.02f, 0.36f, 1.0f);
      F3 amber = f3(1.0f, 0.34f, 0.01f), white = f3(1.0f, 0.96f, 0.76f);
      return lerp3(lerp3(cold, blue, fxm_smoothstep(0.0f, 0.30f, t)), lerp3(amber, white, fxm_smoothstep(0.62f, 1.0f, t)),
                   fxm_smoothstep(0.24f, 0.72f, t));
    }
    static float aces_curve(float v) {
      return fxm_clamp((v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f), 0.0f, 1.0f);
    }
    F3 aces(F3 c) const {
      float e = fxm_max(u[27], 0.25f);
      F3 x = f3(fxm_max(c.x * e, 0.0f), fxm_max(c.y * e, 0.0f), fxm_max(c.z * e, 0.0f));
      float peak = fxm_max(x.x, fxm_max(x.y, x.z));
      float k = aces_curve(peak) / fxm_max(peak, 0.000001f);
      const float g = (float)(1.0 / 2.2);
      return f3(fxm_pow(fxm_clamp(x.x * k, 0.0f, 1.0f), g), fxm_pow(fxm_clamp(x.y * k, 0.0f, 1.0f), g),
                fxm_pow(fxm_clamp(x.z * k, 0.0f, 1.0f), g));
    }

    // diagnosticSample, qualityShaders.ts:338-380
    V4<float> diagnostic(uint32_t mode, V4<float> fl, V4<float> ch, V4<float> me) const {
      auto out = [](F3 c, float a) { return V4<float>{c.x, c.y, c.z, a}; };
      if (mode == 1u) return out(thermal_map(fl.w), fxm_smoothstep(0.015f, 0.72f, fl.w));
      if (mode == 2u) {
        float st = fxm_smoothstep(0.003f, 0.38f, me.x);
        return out(lerp3(f3(0.45f, 0.01f, 0.0f), f3(1.0f, 0.95f, 0.52f), st), st);
      }
      if (mode == 3u) {
        float fuel = fxm_clamp(ch.x, 0.0f, 1.0f);
        float ao = fxm_max(u[9], 0.001f);
        float deficit = fxm_clamp((ao - ch.y) / ao, 0.0f, 1.0f);
        float overlap = fxm_min(fuel, fxm_clamp(ch.y / ao, 0.0f, 1.0f));
        return out(f3(fuel, overlap * 0.36f, deficit), fxm_max(fuel, fxm_max(deficit * 0.72f, overlap * 0.32f)));
      }
      if (mode == 4u) {
        float st = fxm_smoothstep(0.005f, 0.62f, ch.z);
        return out(lerp3(f3(0.06f, 0.02f, 0.12f), f3(0.74f, 0.26f, 1.0f), st), st);
      }
      if (mode == 5u) {
        float vapor = fxm_smoothstep(0.004f, 0.72f, ch.w), soot = fxm_smoothstep(0.003f, 0.55f, me.y);
        return out(lerp3(f3(0.56f, 0.72f, 0.82f), f3(0.17f, 0.12f, 0.10f), soot), fxm_max(vapor, soot));
      }
      if (mode == 6u) {
        float liquid = fxm_smoothstep(0.002f, 0.32f, me.z);
        float speed = fxm_smoothstep(0.035f, 1.25f, len(f3(fl.x, fl.y, fl.z)));
        F3 dirn = plus(f3(0.5f, 0.5f, 0.5f), scale(unit(plus(f3(fl.x, fl.y, fl.z), f3(0.0001f, 0.0001f, 0.0001f))), 0.5f));
        return out(lerp3(times(dirn, f3(0.25f, 0.31f, 0.36f)), f3(0.78f, 0.91f, 1.0f), liquid),
                   fxm_max(liquid, speed * 0.038f));
      }
      float dv = me.w;
      return out(dv >= 0.0f ? f3(1.0f, 0.18f, 0.045f) : f3(0.12f, 0.42f, 1.0f), fxm_smoothstep(0.01f, 1.2f, std::fabs(dv)));
    }

    F3 background(float uvy, F3 eye, F3 dir) const {
      F3 c = lerp3(f3(0.020f, 0.024f, 0.030f), f3(0.0025f, 0.004f, 0.007f), fxm_clamp(uvy, 0.0f, 1.0f));
      float fd = -eye.y / fxm_min(dir.y, -0.00001f);
      if (dir.y < -0.00001f && fd > 0.0f) {
        F3 fp = plus(eye, scale(dir, fd));
        float dx = fp.x - ext(0) * 0.5f, dz = fp.z - ext(2) * 0.5f;
        float rings = fxm_smoothstep(0.012f, 0.0f, std::fabs(fxm_sqrt(dx * dx + dz * dz) - 0.10f));
        float grid = fxm_pow(fxm_max(std::fabs(fxm_sin(fp.x * 31.4159f)), std::fabs(fxm_sin(fp.z * 31.4159f))), 28.0f);
        c = plus(lerp3(f3(0.018f, 0.020f, 0.022f), f3(0.045f, 0.048f, 0.050f), grid * 0.22f),
                 scale(f3(0.055f, 0.020f, 0.006f), rings));
      }
      return c;
    }

    // volumeFragment, one pixel; returns the f32 colour before the f16 store.
    F3 pixel(int px, int py) const {
      const float uvx = ((float)px + 0.5f) / (float)width, uvy = 1.0f - ((float)py + 0.5f) / (float)height;
      const float sx = uvx * 2.0f - 1.0f, sy = uvy * 2.0f - 1.0f;
      const F3 E = f3(ext(0), ext(1), ext(2));
      // divisions by the extents as multiplies by their f32 reciprocals, as
      // the device kernel makes them
      const F3 iE = f3(1.0f / E.x, 1.0f / E.y, 1.0f / E.z);
      const F3 target = f3(E.x * 0.5f, E.y * 0.40f, E.z * 0.5f);
      const float yaw = u[16], pitch = u[17];
      const F3 orbit = f3(fxm_sin(yaw) * fxm_cos(pitch), fxm_sin(pitch), fxm_cos(yaw) * fxm_cos(pitch));
      const float frame = fxm_max(E.x, fxm_max(E.y, E.z));
      const F3 eye = plus(target, scale(scale(orbit, 0.90f + u[18] * 0.18f), frame));
      const F3 fwd = unit(minus(target, eye));
      const F3 right = unit(cross(fwd, f3(0.0f, 1.0f, 0.0f)));
      const F3 up = unit(cross(right, fwd));
      const float lens = 0.72f / fxm_clamp(u[39], 0.5f, 3.0f);
      const F3 dir = unit(plus(plus(fwd, scale(scale(scale(right, sx), u[19]), lens)), scale(scale(up, sy), lens)));
      float lo[3], hi[3];
      const float o[3] = {eye.x, eye.y, eye.z}, d[3] = {dir.x, dir.y, dir.z}, e3[3] = {E.x, E.y, E.z};
      for (int a = 0; a < 3; ++a) {
        float sd = std::fabs(d[a]) > 0.000001f ? d[a] : 0.000001f;
        float inv = 1.0f / sd;
        float l = -o[a] * inv, h = (e3[a] - o[a]) * inv;
        lo[a] = fxm_min(l, h);
        hi[a] = fxm_max(l, h);
      }
      const float tn = fxm_max(fxm_max(lo[0], lo[1]), lo[2]);
      const float tf = fxm_min(fxm_min(hi[0], hi[1]), hi[2]);
      const F3 bg = background(uvy, eye, dir);
      const float cin = fxm_clamp(u[42], 0.0f, 1.0f);
      if (tn > tf || tf <= 0.0f) return cin >= 0.5f ? bg : aces(bg);
      const float start = fxm_max(tn, 0.0f);
      const uint32_t count = (uint32_t)std::min(384, raySteps_of());
      const float stepLen = (tf - start) / fxm_max((float)count, 1.0f);
      const float hx = fxm_fract(fxm_floor((float)px + 0.5f) * 123.34f), hy = fxm_fract(fxm_floor((float)py + 0.5f) * 456.21f);
      const float jitter = fxm_fract((hx + hy) * (hx + hy + 45.32f));
      float dist = start + jitter * stepLen;
      float trans = 1.0f;
      F3 rad = f3(0.0f, 0.0f, 0.0f);
      const uint32_t view = (uint32_t)fxm_clamp(fxm_rint(u[24]), 0.0f, 7.0f);
      for (uint32_t k = 0; k < 384u; ++k) {
        if (k >= count || dist >= tf || trans < 0.006f) break;
        const F3 pos = plus(eye, scale(dir, dist));
        const V4<float> rawMedia = field(vm, pos);
        if (view == 0u && cin >= 0.5f) {
          float occ = fxm_max(fxm_max(rawMedia.x, rawMedia.y), fxm_max(rawMedia.z, rawMedia.w));
          if (occ < 0.0005f) {
            float nd = dist + stepLen;
            if (nd < tf) {
              V4<float> m2 = field(vm, plus(eye, scale(dir, nd)));
              float occ2 = fxm_max(fxm_max(m2.x, m2.y), fxm_max(m2.z, m2.w));
              dist += occ2 < 0.0005f ? stepLen * 2.0f : stepLen;
            } else {
              dist = tf;
            }
            continue;
          }
        }
        const V4<float> fl = field(vf, pos);
        V4<float> ch = field(vc, pos);
        ch = {fxm_max(ch.x, 0.0f), fxm_max(ch.y, 0.0f), fxm_max(ch.z, 0.0f), fxm_max(ch.w, 0.0f)};
        const V4<float> me{fxm_max(rawMedia.x, 0.0f), fxm_max(rawMedia.y, 0.0f), fxm_max(rawMedia.z, 0.0f), rawMedia.w};
        if (view == 0u) {
          const float temperature = 300.0f + fl.w * 2600.0f;
          const float rr = fxm_max(me.x, 0.0f);
          const float front = fxm_mix(fxm_smoothstep(0.006f, 0.32f, rr), fxm_smoothstep(0.004f, 0.28f, rr), cin);
          const float thermal = fxm_smoothstep(620.0f, 1780.0f, temperature);
          const F3 dp = minus(pos, scale(f3(fl.x, fl.y, fl.z), 0.08f));
          const float da = fxm_sin(dot(dp, f3(27.0f, 35.0f, 13.0f)) + u[5] * 3.4f);
          const float db = fxm_sin(dot(dp, f3(-21.0f, 29.0f, 17.0f)) - u[5] * 4.7f);
          float optical = 0.90f + 0.20f * da * db;
          if (cin > 0.5f && front > 0.002f) {
            const F3 nrm = f3(dp.x * iE.x, dp.y * iE.y, dp.z * iE.z);
            const float req = u[43];
            const float expl = fxm_clamp(req, 64.0f, 1024.0f);
            const F3 cells = req >= 1.0f ? f3(expl, expl, expl) : f3(u[0], u[1], u[2]);
            const F3 tile = f3(fxm_max(cells.x / 64.0f, 0.015625f), fxm_max(cells.y / 64.0f, 0.015625f),
                               fxm_max(cells.z / 64.0f, 0.015625f));
            const F3 toff = scale(f3(0.013f, -0.021f, 0.017f), u[5]);
            const float n1 = detail(plus(times(nrm, tile), toff)).x - 0.5f;
            const F3 c = minus(nrm, f3(0.5f, 0.5f, 0.5f));
            const float ra = c.x * 0.771f + c.z * 0.637f, rb = -c.x * 0.637f + c.z * 0.771f;
            const F3 rot = f3(ra, c.y * 0.909f - rb * 0.417f, c.y * 0.417f + rb * 0.909f);
            const float n2 = detail(minus(plus(scale(times(rot, tile), 0.47f), f3(0.371f, 0.113f, 0.733f)),
                                          scale(f3(toff.y, toff.z, toff.x), 0.61f))).y - 0.5f;
            const float maxCells = fxm_max(cells.x, fxm_max(cells.y, cells.z));
            const float budget = fxm_clamp(u[22], 1.0f, 384.0f);
            const float band1 = fxm_clamp((budget * 2.0f) / maxCells, 0.25f, 1.0f);
            const float band2 = fxm_clamp((budget * 2.0f) / fxm_max(maxCells * 0.47f, 1.0f), 0.35f, 1.0f);
            const float fbm = n1 * (0.55f * band1) + n2 * (0.45f * band2);
            const float mask = fxm_smoothstep(0.025f, 0.24f, front) * (1.0f - fxm_smoothstep(0.72f, 0.96f, front));
            const float breakup = fxm_clamp(1.0f + fbm * 0.44f, 0.84f, 1.16f);
            optical = fxm_mix(1.0f, breakup, mask * 0.85f);
          }
          const float fdef = fxm_clamp(u[41], 0.5f, 2.5f);
          const float fexp = fxm_mix(0.85f + fdef * 0.35f, 0.92f + fdef * 0.16f, cin);
          const float flame = fxm_pow(fxm_clamp(fxm_mix(front * thermal, front * (0.42f + 0.58f * thermal), cin), 0.0f, 1.0f),
                                      fexp) * optical;
          const float soot = fxm_clamp(1.0f - fxm_exp(-fxm_max(me.y, 0.0f) * fxm_mix(1.15f, 0.90f, cin)), 0.0f,
                                       fxm_mix(1.0f, 0.96f, cin));
          const float vapor = 1.0f - fxm_exp(-fxm_max(ch.w, 0.0f) * fxm_mix(0.34f, 0.30f, cin));
          const float rich = fxm_clamp(ch.x / fxm_max(ch.x + ch.y * 0.25f, 0.0001f), 0.0f, 1.0f);
          const float reveal = fxm_mix(1.0f, 1.0f - fxm_smoothstep(0.045f, 0.30f, front) * 0.78f, cin);
          const float sf = flame * fxm_mix(3.20f, 0.88f, cin);
          const float ss = soot * reveal * fxm_mix(1.6f, 1.62f, cin);
          const float sv = vapor * fxm_mix(0.20f, 0.32f, cin);
          const float extinction = fxm_min(sf + ss + sv, fxm_mix(100.0f, 5.5f, cin));
          const float depth = fxm_min(extinction * stepLen * fxm_mix(5.2f, 3.15f, cin), fxm_mix(100.0f, 0.28f, cin));
          const float alpha = 1.0f - fxm_exp(-depth);
          const float hot = soot * thermal;
          const float etemp = fxm_mix(temperature + front * 650.0f,
                                      820.0f + 2150.0f * fxm_pow(fxm_clamp(front * 0.74f + thermal * 0.38f, 0.0f, 1.0f), 0.72f),
                                      cin);
          const float warmth = 1.0f - fxm_smoothstep(0.18f, 0.72f, front * thermal);
          const F3 stdColor = lerp3(blackbody(etemp), f3(1.0f, 0.075f, 0.002f), warmth * 0.68f);
          const F3 amber = lerp3(f3(1.0f, 0.055f, 0.002f), f3(1.0f, 0.43f, 0.018f), fxm_smoothstep(0.025f, 0.30f, flame));
          const F3 cinColor = lerp3(amber, f3(1.0f, 0.91f, 0.60f),
                                    fxm_smoothstep(0.52f, 0.92f, flame) * fxm_smoothstep(980.0f, 2050.0f, temperature));
          const F3 color = lerp3(stdColor, cinColor, cin);
          const float eCin = flame * (1.35f + 3.65f * front) * (0.92f + 0.32f * u[10]) + hot * 0.08f;
          const float eStd = flame * (1.20f + 1.80f * front) * (0.90f + 0.30f * u[10]) + hot * 0.25f;
          const F3 emissive = scale(color, fxm_mix(eStd, eCin, cin));
          const float baseZone = 1.0f - fxm_smoothstep(0.16f, 0.48f, pos.y * iE.y);
          const F3 blue = scale(scale(scale(f3(0.018f, 0.12f, 0.74f), flame), 1.0f - rich),
                                fxm_mix(0.48f, baseZone * 0.34f, cin));
          const F3 toFire = minus(f3(E.x * 0.5f, 0.08f, E.z * 0.5f), pos);
          const float falloff = 1.0f / (1.0f + dot(toFire, toFire) * 8.0f);
          const F3 amb = f3(0.018f, 0.022f, 0.028f);
          const F3 fire = scale(f3(1.0f, 0.24f, 0.035f), falloff);
          const F3 sStd = plus(scale(f3(0.012f, 0.009f, 0.007f), ss), scale(f3(0.12f, 0.15f, 0.17f), sv));
          const F3 sCin = plus(scale(plus(scale(amb, 0.12f), scale(fire, 0.055f)), ss),
                               scale(plus(scale(amb, 0.55f), scale(fire, 0.05f)), sv));
          const F3 scattered = lerp3(sStd, sCin, cin);
          const float m = fxm_max(extinction, 0.0001f);
          const float im = 1.0f / m;
          const F3 src = plus(plus(emissive, blue), scattered);
          rad = plus(rad, scale(f3(src.x * im, src.y * im, src.z * im), trans * alpha));
          trans *= 1.0f - alpha;
        } else {
          const V4<float> fv = diagnostic(view, fl, ch, me);
          const float alpha = 1.0f - fxm_exp(-fv.w * stepLen * 8.5f);
          rad = plus(rad, scale(scale(f3(fv.x, fv.y, fv.z), trans * alpha), 1.35f));
          trans *= 1.0f - alpha;
        }
        dist += stepLen;
      }
      const F3 c = plus(rad, scale(bg, trans));
      return cin >= 0.5f ? c : aces(c);
    }
    int raySteps_of() const {
      float v = u[22];
      return v > 0.0f ? (int)(uint32_t)v : 0;
    }

    // ---- droplet streaks (dropletVertex / dropletFragment, qualityShaders.ts:156-232)
    // The camera of cameraEye / projectWorld (qualityShaders.ts:133-176).
    struct Camera {
      F3 eye, fwd, right, up;
      float lens;
    };
    Camera camera() const {
      const F3 E = f3(ext(0), ext(1), ext(2));
      const F3 target = f3(E.x * 0.5f, E.y * 0.40f, E.z * 0.5f);
      const float yaw = u[16], pitch = u[17];
      const F3 orbit = f3(fxm_sin(yaw) * fxm_cos(pitch), fxm_sin(pitch), fxm_cos(yaw) * fxm_cos(pitch));
      const float frame = fxm_max(E.x, fxm_max(E.y, E.z));
      Camera c;
      c.eye = plus(target, scale(scale(orbit, 0.90f + u[18] * 0.18f), frame));
      c.fwd = unit(minus(target, c.eye));
      c.right = unit(cross(c.fwd, f3(0.0f, 1.0f, 0.0f)));
      c.up = unit(cross(c.right, c.fwd));
      c.lens = 0.72f / fxm_clamp(u[39], 0.5f, 3.0f);
      return c;
    }
    void project(const Camera& c, F3 point, float out[2]) const {
      const F3 rel = minus(point, c.eye);
      const float depth = fxm_max(dot(rel, c.fwd), 0.0001f);
      out[0] = dot(rel, c.right) / (depth * u[19] * c.lens);
      out[1] = dot(rel, c.up) / (depth * c.lens);
    }
    // dropletVertex for droplet i: the quad's centre, unit tangent, radius and
    // half length in NDC (corner (cx, cy) sits at centre + normal * cx *
    // radius + tangent * cy * half).  False when it draws nothing (massless or
    // outside the unit box: the vertex collapses off screen).
    struct Sprite {
      float cx, cy, tx, ty, radius, half, mass;
    };
    bool sprite(const Camera& cam, int i, Sprite& sp) const {
      const float* q = &s.particles[s.particleIndex][(size_t)i * 8];
      if (!(q[7] > 0.0001f && q[0] >= 0.0f && q[1] >= 0.0f && q[2] >= 0.0f && q[0] <= 1.0f && q[1] <= 1.0f &&
            q[2] <= 1.0f))
        return false;
      const F3 world = times(f3(q[0], q[1], q[2]), f3(ext(0), ext(1), ext(2)));
      float head[2], tail[2];
      project(cam, world, head);
      project(cam, minus(world, scale(f3(q[4], q[5], q[6]), 0.018f)), tail);
      const float sx = tail[0] - head[0], sy = tail[1] - head[1];
      const float length = fxm_sqrt(sx * sx + sy * sy);
      const float streak = fxm_min(length, 0.032f);
      const bool along = streak > 0.00001f;
      sp.tx = along ? sx / length : 1.0f;
      sp.ty = along ? sy / length : 0.0f;
      const float pixel_radius = u[26] >= 0.5f ? 1.15f : 1.65f;
      sp.radius = pixel_radius * 2.0f / fxm_max(u[21], 1.0f);
      sp.half = sp.radius + streak * 0.5f;
      sp.cx = head[0];
      sp.cy = head[1];
      sp.mass = q[7];
      return true;
    }
    // dropletFragment at quad coordinates (lx, ly): false when discarded.
    bool droplet_fragment(float lx, float ly, float mass, uint32_t view, float rgba[4]) const {
      const float m = fxm_max(std::fabs(ly) - 0.48f, 0.0f);
      const float d = fxm_sqrt(lx * lx + m * m);
      const float coverage = 1.0f - fxm_smoothstep(0.64f, 1.0f, d);
      if (coverage <= 0.001f) return false;
      const float highlight = fxm_pow(fxm_clamp(1.0f - std::fabs(lx + 0.30f), 0.0f, 1.0f), 5.0f);
      const F3 base = lerp3(f3(0.27f, 0.34f, 0.38f), f3(0.76f, 0.86f, 0.91f), highlight * 0.62f);
      const float boost = view == 6u ? 1.38f : 1.0f;
      const float alpha = coverage * fxm_clamp(0.13f + mass * 0.15f, 0.12f, 0.42f) * boost;
      const F3 linear = scale(base, 0.72f);
      const F3 c = u[42] >= 0.5f ? linear : aces(linear);
      rgba[0] = c.x;
      rgba[1] = c.y;
      rgba[2] = c.z;
      rgba[3] = alpha;
      return true;
    }
    // Draws the streaks over img (width * height rgba, row 0 at the top) in
    // instance order.  A pixel is rasterized for droplet i when its centre's
    // quad coordinates (the fragment's interpolated `local`) lie strictly
    // inside (-1, 1)^2: local = (dot(p - c, normal) / radius, dot(p - c,
    // tangent) / half) with p the pixel centre in NDC.  Blending (src-alpha,
    // one-minus-src-alpha; alpha: one, one-minus-src-alpha) runs in f32 and
    // each result is stored in the target's format: f16 (img holds f16
    // values) or unorm8 (img holds bytes 0..255, read as byte / 255).
    void draw_droplets(std::vector<float>& img, bool unorm8_target) const {
      const uint32_t view = (uint32_t)fxm_clamp(fxm_rint(u[24]), 0.0f, 7.0f);
      if (view != 0u && view != 6u) return;
      const Camera cam = camera();
      for (int i = 0; i < s.particleCount; ++i) {
        Sprite sp;
        if (!sprite(cam, i, sp)) continue;
        const float nx = -sp.ty, ny = sp.tx;
        // a generous pixel window: each corner lies within radius + half of the centre per axis
        const double reach = (double)sp.radius + (double)sp.half;
        const double fx0 = ((double)sp.cx - reach + 1.0) * 0.5 * width - 0.5;
        const double fx1 = ((double)sp.cx + reach + 1.0) * 0.5 * width - 0.5;
        const double fy0 = (1.0 - ((double)sp.cy + reach)) * 0.5 * height - 0.5;
        const double fy1 = (1.0 - ((double)sp.cy - reach)) * 0.5 * height - 0.5;
        const int x0 = (int)std::max(0.0, std::floor(fx0) - 2.0), x1 = (int)std::min(width - 1.0, std::ceil(fx1) + 2.0);
        const int y0 = (int)std::max(0.0, std::floor(fy0) - 2.0), y1 = (int)std::min(height - 1.0, std::ceil(fy1) + 2.0);
        for (int y = y0; y <= y1; ++y)
          for (int x = x0; x <= x1; ++x) {
            const float uvx = ((float)x + 0.5f) / (float)width, uvy = 1.0f - ((float)y + 0.5f) / (float)height;
            const float px = uvx * 2.0f - 1.0f, py = uvy * 2.0f - 1.0f;
            const float dx = px - sp.cx, dy = py - sp.cy;
            const float lx = (dx * nx + dy * ny) / sp.radius, ly = (dx * sp.tx + dy * sp.ty) / sp.half;
            if (!(std::fabs(lx) < 1.0f && std::fabs(ly) < 1.0f)) continue;
            float src[4];
            if (!droplet_fragment(lx, ly, sp.mass, view, src)) continue;
            float* d = &img[((size_t)y * width + x) * 4];
            const float a = src[3], keep = 1.0f - a;
            for (int k = 0; k < 4; ++k) {
              const float dst = unorm8_target ? d[k] / 255.0f : d[k];
              const float v = (k < 3 ? src[k] * a : a) + dst * keep;
              d[k] = unorm8_target ? (float)fxm_rint(fxm_clamp(v, 0.0f, 1.0f) * 255.0f) : round_to_f16(v);
            }
          }
      }
    }
  };

  V4<T> store_species(const V4<T>& s) const;
  T texel(T v) const;  // an rgba16float texel component: f16 rounding (FP32 oracle)

  // packFields / storePackedFields, qualityShaders.ts:37-73, with the
  // presentation divergence refresh that precedes it in view mode 7
  // (FireXEngine.ts:2590-2597): the raymarcher's three rgba16float volumes
  //   flow      (u, v, w, clamp((T - 300) / 2600, 0, 1))
  //   chemistry max(species, 0)
  //   media     (reaction, soot, liquid, display value)
  // from flow[0], species[current], reaction and liquid.
  void pack_fields(std::vector<V4<T>>& pf, std::vector<V4<T>>& pc, std::vector<V4<T>>& pm) {
    // controls.x, the view mode: the current parameter, as the next
    // uniform write (every frame) would hold it
    const T viewMode = T((float)params.viewMode);
    const bool divergenceView = wclamp(T(std::nearbyint(viewMode)), T(0), T(7)) == T(7);
    if (divergenceView) compute_divergence(flow[0]);
    pf.resize(n);
    pc.resize(n);
    pm.resize(n);
    const std::vector<V4<T>>& sp = species[speciesIndex];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const size_t id = idx(x, y, z);
          const V4<T> f = flow[0][id];
          const V4<T> s{wmax(sp[id].x, T(0)), wmax(sp[id].y, T(0)), wmax(sp[id].z, T(0)), wmax(sp[id].w, T(0))};
          const T react = wclamp(reaction[id], T(0), T(4));
          const T richness = wclamp(s.x / wmax(s.x + s.y * T(0.25), T(0.0001)), T(0), T(1));
          const T soot = s.z * wsmooth(T(0.25), T(0.82), richness);
          const T liq = T((float)liquid[id]) * T(3.4f / 4096.0f);
          const bool shell = x == 0 || y == 0 || z == 0 || x + 1 >= nx || y + 1 >= ny || z + 1 >= nz;
          T display = wmax(wmax(react, soot), wmax(liq, wmax(s.w, wsmooth(T(420), T(700), f.w))));
          if (divergenceView) display = shell ? T(0) : divergence[id];
          pf[id] = V4<T>{texel(f.x), texel(f.y), texel(f.z), texel(wclamp((f.w - T(300)) / T(2600), T(0), T(1)))};
          pc[id] = V4<T>{texel(s.x), texel(s.y), texel(s.z), texel(s.w)};
          pm[id] = V4<T>{texel(react), texel(soot), texel(liq), texel(display)};
        }
  }

  // computeDivergence, shaders.ts:344-364.
  void compute_divergence(const std::vector<V4<T>>& f) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const V4<T>& c = f[idx(x, y, z)];
          T left = f[idx(x - 1, y, z)].x;
          T bottom = f[idx(x, y - 1, z)].y;
          T back = f[idx(x, y, z - 1)].z;
          divergence[idx(x, y, z)] = (c.x - left) * mx + (c.y - bottom) * my + (c.z - back) * mz;
        }
  }

  // loadPressure, shaders.ts:366-373 (pressureShaders.ts:49-56 for f16 storage).
  T load_p(const std::vector<T>& p, int x, int y, int z) const {
    if (y >= ny) return T(0);
    return p[idx(x, y, z)];
  }

  // solvePressureGlobal / solvePressureTiled / solvePressureF16Tiled (one sweep).
  // The f32 form multiplies by the f32 reciprocal of 2 (wx + wy + wz), one of
  // the results WGSL's f32 division (2.5 ULP) admits, and contracts the
  // weighted neighbour sum wx sx + wy sy + wz sz into fma(wz, sz, fma(wy, sy,
  // wx sx)), the form a contracting shader compiler emits for it; FP64
  // divides and rounds every operation.
  void jacobi_sweep(const std::vector<T>& pin, std::vector<T>& pout, bool f16Store) {
    T wx = mx * mx, wy = my * my, wz = mz * mz;
    T denom = T(2) * (wx + wy + wz);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          size_t id = idx(x, y, z);
          if (y == ny - 1) { pout[id] = T(0); continue; }
          T nb = jacobi_sum(wx, load_p(pin, x - 1, y, z) + load_p(pin, x + 1, y, z),
                            wy, load_p(pin, x, y - 1, z) + load_p(pin, x, y + 1, z),
                            wz, load_p(pin, x, y, z - 1) + load_p(pin, x, y, z + 1));
          T v = jacobi_divide(nb - divergence[id], denom);
          pout[id] = f16Store ? store_pressure_f16(v) : v;
        }
  }
  T jacobi_sum(T wx, T sx, T wy, T sy, T wz, T sz) const;
  T jacobi_divide(T num, T denom) const;
  T store_pressure_f16(T v) const;

  // pressureGradient + projectVelocity, shaders.ts:461-501.
  void project(const std::vector<V4<T>>& fin, const std::vector<T>& p, std::vector<V4<T>>& fout) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          size_t id = idx(x, y, z);
          auto grad = [&](int cx, int cy, int cz, T& gxo, T& gyo, T& gzo) {
            T c = load_p(p, cx, cy, cz);
            gxo = (load_p(p, cx + 1, cy, cz) - c) * mx;
            gyo = (load_p(p, cx, cy + 1, cz) - c) * my;
            gzo = (load_p(p, cx, cy, cz + 1) - c) * mz;
          };
          T g0, g1, g2;
          grad(x, y, z, g0, g1, g2);
          V4<T> f = fin[id];
          f.x = f.x - g0; f.y = f.y - g1; f.z = f.z - g2;
          if (y == ny - 1) {
            const V4<T>& bf = fin[idx(x, y - 1, z)];
            grad(x, y - 1, z, g0, g1, g2);
            f.x = bf.x - g0; f.y = bf.y - g1; f.z = bf.z - g2;
          }
          if (x == 0 || x == nx - 1) f.x = T(0);
          if (z == 0 || z == nz - 1) f.z = T(0);
          if (y == 0) f.y = T(0);
          if (y == ny - 1) f.y = wmax(T(0), f.y);
          fout[id] = f;
        }
  }

  // encodeSimulationStep, FireXEngine.ts:2115-2276.
  void step(double deltaTime) {
    write_uniforms(deltaTime);
    log.clear();
    droplet_passes();
    log += "simulate;";
    simulate(speciesIndex);
    speciesIndex = 1 - speciesIndex;
    log += "divergence;";
    compute_divergence(flow[1]);
    if (heavy()) {
      if (f16Pressure()) {
        for (int i = 0; i < pressureIterations; ++i) {
          jacobi_sweep(pressure[pressureIndex], pressure[1 - pressureIndex], true);
          pressureIndex = 1 - pressureIndex;
        }
        log += "jacobi_f16x" + std::to_string(pressureIterations) + ";project;";
        project(flow[1], pressure[pressureIndex], flow[0]);
        log += "divergence;clear_corr;";
        compute_divergence(flow[0]);
        std::fill(corr[0].begin(), corr[0].end(), T(0));
        std::fill(corr[1].begin(), corr[1].end(), T(0));
        int ci = 0;
        for (int i = 0; i < correctionIterations; ++i) {
          jacobi_sweep(corr[ci], corr[1 - ci], true);
          ci = 1 - ci;
        }
        log += "jacobi_f16x" + std::to_string(correctionIterations) + ";project;";
        project(flow[0], corr[ci], flow[1]);
      } else {
        // compact f32 fallback: two shared pressure fields, no warm start.
        std::fill(pressure[0].begin(), pressure[0].end(), T(0));
        std::fill(pressure[1].begin(), pressure[1].end(), T(0));
        pressureIndex = 0;
        log += "clear_pressure;";
        for (int i = 0; i < pressureIterations; ++i) {
          jacobi_sweep(pressure[pressureIndex], pressure[1 - pressureIndex], false);
          pressureIndex = 1 - pressureIndex;
        }
        log += "jacobix" + std::to_string(pressureIterations) + ";project;";
        project(flow[1], pressure[pressureIndex], flow[0]);
        log += "divergence;clear_pressure;";
        compute_divergence(flow[0]);
        std::fill(pressure[0].begin(), pressure[0].end(), T(0));
        std::fill(pressure[1].begin(), pressure[1].end(), T(0));
        int ci = 0;
        for (int i = 0; i < correctionIterations; ++i) {
          jacobi_sweep(pressure[ci], pressure[1 - ci], false);
          ci = 1 - ci;
        }
        log += "jacobix" + std::to_string(correctionIterations) + ";project;";
        project(flow[0], pressure[ci], flow[1]);
      }
    } else {
      for (int i = 0; i < pressureIterations; ++i) {
        jacobi_sweep(pressure[pressureIndex], pressure[1 - pressureIndex], false);
        pressureIndex = 1 - pressureIndex;
      }
      log += "jacobix" + std::to_string(pressureIterations) + ";project;";
      project(flow[1], pressure[pressureIndex], flow[0]);
      log += "divergence;clear_corr;";
      compute_divergence(flow[0]);
      std::fill(corr[0].begin(), corr[0].end(), T(0));
      std::fill(corr[1].begin(), corr[1].end(), T(0));
      int ci = 0;
      for (int i = 0; i < correctionIterations; ++i) {
        jacobi_sweep(corr[ci], corr[1 - ci], false);
        ci = 1 - ci;
      }
      log += "jacobix" + std::to_string(correctionIterations) + ";project;";
      project(flow[0], corr[ci], flow[1]);
    }
    log += "copy_flow;";
    flow[0] = flow[1];
    simulationTime += deltaTime;
    frameIndex += 1;
  }
};

template <> V4<float> State<float>::store_species(const V4<float>& s) const {
  if (speciesStorage == SPECIES_F16)
    return {round_to_f16(s.x), round_to_f16(s.y), round_to_f16(s.z), round_to_f16(s.w)};
  if (speciesStorage == SPECIES_BF16)
    return {round_to_bf16(s.x), round_to_bf16(s.y), round_to_bf16(s.z), round_to_bf16(s.w)};
  return s;
}
template <> V4<double> State<double>::store_species(const V4<double>& s) const { return s; }
template <> float State<float>::texel(float v) const { return round_to_f16(v); }
template <> double State<double>::texel(double v) const { return v; }

template <> float State<float>::jacobi_divide(float num, float denom) const {
  return num * (1.0f / denom);
}
template <> double State<double>::jacobi_divide(double num, double denom) const { return num / denom; }
template <> float State<float>::jacobi_sum(float wx, float sx, float wy, float sy, float wz, float sz) const {
  return std::fma(wz, sz, std::fma(wy, sy, wx * sx));
}
template <> double State<double>::jacobi_sum(double wx, double sx, double wy, double sy, double wz,
                                             double sz) const {
  return wx * sx + wy * sy + wz * sz;
}
template <> float State<float>::store_pressure_f16(float v) const { return round_to_f16(v); }
template <> double State<double>::store_pressure_f16(double v) const { return v; }

struct Handle {
  int precision;  // 32 or 64
  State<float> f;
  State<double> d;
};

enum Field {
  F_FLOW = 0, F_SPECIES = 1, F_PRESSURE = 2, F_CORRECTION = 3, F_DIVERGENCE = 4,
  F_REACTION = 5, F_LIQUID = 6, F_FLOW_PREDICTED = 7,
  F_PARTICLES = 11, F_PARTICLE_DENSITY = 12  // per particle: 8 and 4 values
};

template <typename T> size_t field_components(int field) {
  return (field == F_FLOW || field == F_SPECIES || field == F_FLOW_PREDICTED) ? 4 : 1;
}

template <typename T> int read_field(State<T>& s, int field, double* out) {
  size_t n = s.n;
  switch (field) {
    case F_FLOW: case F_FLOW_PREDICTED: {
      const auto& f = s.flow[field == F_FLOW ? 0 : 1];
      for (size_t i = 0; i < n; ++i) { out[4*i] = f[i].x; out[4*i+1] = f[i].y; out[4*i+2] = f[i].z; out[4*i+3] = f[i].w; }
      return 0;
    }
    case F_SPECIES: {
      const auto& f = s.species[s.speciesIndex];
      for (size_t i = 0; i < n; ++i) { out[4*i] = f[i].x; out[4*i+1] = f[i].y; out[4*i+2] = f[i].z; out[4*i+3] = f[i].w; }
      return 0;
    }
    case F_PRESSURE: for (size_t i = 0; i < n; ++i) out[i] = s.pressure[s.pressureIndex][i]; return 0;
    case F_CORRECTION: {
      // last residual solution: correctionIndex = correctionIterations % 2
      // (f32 heavy fallback shares the pressure fields).
      int ci = s.correctionIterations % 2;
      const auto& c = (s.heavy() && !s.f16Pressure()) ? s.pressure[ci] : s.corr[ci];
      for (size_t i = 0; i < n; ++i) out[i] = c[i];
      return 0;
    }
    case F_DIVERGENCE: for (size_t i = 0; i < n; ++i) out[i] = s.divergence[i]; return 0;
    case F_REACTION: for (size_t i = 0; i < n; ++i) out[i] = s.reaction[i]; return 0;
    case F_LIQUID: for (size_t i = 0; i < n; ++i) out[i] = (double)s.liquid[i]; return 0;
    case F_PARTICLES: {
      const auto& q = s.particles[s.particleIndex];
      for (size_t i = 0; i < q.size(); ++i) out[i] = q[i];
      return 0;
    }
    case F_PARTICLE_DENSITY:
      for (size_t i = 0; i < s.particleDensity.size(); ++i) out[i] = s.particleDensity[i];
      return 0;
  }
  return -1;
}

template <typename T> int write_field(State<T>& s, int field, const double* in) {
  size_t n = s.n;
  switch (field) {
    case F_FLOW: case F_FLOW_PREDICTED: {
      auto& f = s.flow[field == F_FLOW ? 0 : 1];
      for (size_t i = 0; i < n; ++i) f[i] = V4<T>{T(in[4*i]), T(in[4*i+1]), T(in[4*i+2]), T(in[4*i+3])};
      return 0;
    }
    case F_SPECIES: {
      auto& f = s.species[s.speciesIndex];
      for (size_t i = 0; i < n; ++i) f[i] = s.store_species(V4<T>{T(in[4*i]), T(in[4*i+1]), T(in[4*i+2]), T(in[4*i+3])});
      return 0;
    }
    case F_PRESSURE: for (size_t i = 0; i < n; ++i) s.pressure[s.pressureIndex][i] = T(in[i]); return 0;
    case F_DIVERGENCE: for (size_t i = 0; i < n; ++i) s.divergence[i] = T(in[i]); return 0;
    case F_REACTION: for (size_t i = 0; i < n; ++i) s.reaction[i] = T(in[i]); return 0;
    case F_LIQUID: for (size_t i = 0; i < n; ++i) s.liquid[i] = (uint32_t)in[i]; return 0;
    case F_PARTICLES: {
      // both particle buffers, as writeParticleInitialState writes them; the
      // written state is the live one (the next water step advances it)
      for (int k = 0; k < 2; ++k)
        for (size_t i = 0; i < s.particles[k].size(); ++i) s.particles[k][i] = T(in[i]);
      s.particleIndex = 0;
      s.particleFieldActive = true;
      return 0;
    }
    case F_PARTICLE_DENSITY: {
      for (size_t i = 0; i < s.particleDensity.size(); ++i) s.particleDensity[i] = T(in[i]);
      return 0;
    }
  }
  return -1;
}

}  // namespace

// =========================================================================== C API
extern "C" {

void* orc_create(int nx, int ny, int nz, int precision, int pressureIterations, int correctionIterations) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || (precision != 32 && precision != 64)) return nullptr;
  Handle* h = new Handle();
  h->precision = precision;
  auto init = [&](auto& s) {
    s.nx = nx; s.ny = ny; s.nz = nz; s.n = (size_t)nx * ny * nz;
    s.pressureIterations = pressureIterations;
    s.correctionIterations = correctionIterations;
    s.initialize_fields();
  };
  if (precision == 32) init(h->f); else init(h->d);
  return h;
}

void orc_destroy(void* p) { delete static_cast<Handle*>(p); }

#define ORC_DISPATCH(h, expr) ((h)->precision == 32 ? ([&](auto& s) { return expr; })((h)->f) \
                                                     : ([&](auto& s) { return expr; })((h)->d))

int orc_set_parameter(void* p, const char* name, double value) {
  Handle* h = static_cast<Handle*>(p);
  return ORC_DISPATCH(h, set_parameter(s.params, name, value) ? 0 : 1);
}

int orc_set_bool(void* p, const char* name, int value) {
  Handle* h = static_cast<Handle*>(p);
  if (std::string(name) == "waterEnabled") { h->f.params.waterEnabled = value != 0; h->d.params.waterEnabled = value != 0; return 0; }
  if (std::string(name) == "cinematic") { h->f.cinematic = value != 0; h->d.cinematic = value != 0; return 0; }
  return 1;
}

double orc_get_parameter(void* p, const char* name) {
  Handle* h = static_cast<Handle*>(p);
  const Params& q = h->precision == 32 ? h->f.params : h->d.params;
  std::string n(name);
  if (n == "burnerSize") return q.burnerSize;
  if (n == "burnerDepth") return q.burnerDepth;
  if (n == "sourceThickness") return q.sourceThickness;
  if (n == "firePower") return q.firePower;
  if (n == "sourceLift") return q.sourceLift;
  if (n == "buoyancyScale") return q.buoyancyScale;
  if (n == "flamePersistence") return q.flamePersistence;
  if (n == "airEntrainment") return q.airEntrainment;
  if (n == "fuelRate") return q.fuelRate;
  if (n == "oxygenRate") return q.oxygenRate;
  if (n == "heatEfficiency") return q.heatEfficiency;
  if (n == "vorticity") return q.vorticity;
  if (n == "aimHeight") return q.aimHeight;
  if (n == "domainWidth") return q.domainWidth;
  if (n == "domainHeight") return q.domainHeight;
  if (n == "domainDepth") return q.domainDepth;
  if (n == "waterFlow") return q.waterFlow;
  if (n == "sprayAngle") return q.sprayAngle;
  if (n == "cameraYaw") return q.cameraYaw;
  if (n == "cameraPitch") return q.cameraPitch;
  if (n == "cameraDistance") return q.cameraDistance;
  if (n == "viewMode") return q.viewMode;
  if (n == "nozzleType") return q.nozzleType;
  if (n == "exposure") return q.exposure;
  if (n == "frontDefinition") return q.frontDefinition;
  if (n == "viewZoom") return q.viewZoom;
  if (n == "opticalDetailTarget") return q.opticalDetailTarget;
  return NAN;
}

// Simulation clock as the uniform block sees it (time.y, time.z): lets a
// test evaluate the step at a chosen time and frame.
void orc_set_time(void* p, double simulationTime, double frameIndex) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, (s.simulationTime = simulationTime, s.frameIndex = frameIndex, 0));
}
void orc_set_time_step(void* p, double ts) { Handle* h = static_cast<Handle*>(p); h->f.timeStep = ts; h->d.timeStep = ts; }
void orc_set_species_storage(void* p, int mode) { static_cast<Handle*>(p)->f.speciesStorage = mode; }
void orc_set_heavy_mode(void* p, int mode) { Handle* h = static_cast<Handle*>(p); h->f.heavyMode = mode; h->d.heavyMode = mode; }

int orc_set_emitters(void* p, int count, const float* data /* count*5: cx cy cz radius amplitude */,
                     float jitter, float velocityNoise, uint64_t seed) {
  if (count < 0 || count > 8) return 1;
  Handle* h = static_cast<Handle*>(p);
  EmitterSet e;
  e.count = count;
  for (int i = 0; i < count; ++i) e.e[i] = SeededEmitter{data[5*i], data[5*i+1], data[5*i+2], data[5*i+3], data[5*i+4]};
  e.jitter = jitter; e.velocityNoise = velocityNoise; e.seed = seed;
  h->f.emitters = e; h->d.emitters = e;
  return 0;
}

void orc_reset(void* p) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, (s.simulationTime = 0, s.frameIndex = 0, s.speciesIndex = 0, s.pressureIndex = 0,
                   s.particleIndex = 0, s.initialize_fields(), 0));
}

// Particle count of the resource set (FireXEngine.ts:187-277 per tier);
// rebuilds the particle state as a resource rebuild does.
int orc_set_particle_count(void* p, int count) {
  if (count < 0) return 1;
  Handle* h = static_cast<Handle*>(p);
  h->f.set_particle_count(count);
  h->d.set_particle_count(count);
  return 0;
}
int orc_particle_count(void* p) { return static_cast<Handle*>(p)->f.particleCount; }
int orc_particle_index(void* p) { Handle* h = static_cast<Handle*>(p); return h->precision == 32 ? h->f.particleIndex : h->d.particleIndex; }

// One droplet pass on the current particle buffers with uniforms for dt:
// 0 computeParticleDensity, 1 integrateParticles (current -> other buffer,
// then flip), 2 depositParticles (current as the new state, other as the
// previous one; adds to the liquid field).
int orc_kernel_droplets(void* p, double dt, int which) {
  Handle* h = static_cast<Handle*>(p);
  return ORC_DISPATCH(h, (s.write_uniforms(dt), [&] {
    const int pi = s.particleIndex;
    if (which == 0) s.sph_density(s.particles[pi]);
    else if (which == 1) { s.sph_integrate(s.particles[pi], s.particles[1 - pi]); s.particleIndex = 1 - pi; }
    else if (which == 2) s.liquid_deposit(s.particles[pi], s.particles[1 - pi]);
    else return 1;
    return 0;
  }()));
}

// Sets the two particle buffers separately (the live one first), for
// single-pass tests of the deposition.
void orc_write_particle_buffers(void* p, const double* current, const double* previous) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, ([&] {
    using T = typename std::remove_reference<decltype(s.particles[0][0])>::type;
    for (size_t i = 0; i < s.particles[0].size(); ++i) {
      s.particles[0][i] = T(current[i]);
      s.particles[1][i] = T(previous[i]);
    }
    s.particleIndex = 0;
    s.particleFieldActive = true;
  }(), 0));
}

// writeParticleInitialState(active) on the current parameters.
void orc_write_particle_initial_state(void* p, int active) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, (s.write_particle_initial_state(active != 0), 0));
}

void orc_step(void* p, double dt, int count) {
  Handle* h = static_cast<Handle*>(p);
  for (int i = 0; i < count; ++i) ORC_DISPATCH(h, (s.step(dt), 0));
}

int orc_read_field(void* p, int field, double* out) {
  Handle* h = static_cast<Handle*>(p);
  return ORC_DISPATCH(h, read_field(s, field, out));
}

int orc_write_field(void* p, int field, const double* in) {
  Handle* h = static_cast<Handle*>(p);
  return ORC_DISPATCH(h, write_field(s, field, in));
}

double orc_simulation_time(void* p) { Handle* h = static_cast<Handle*>(p); return h->precision == 32 ? h->f.simulationTime : h->d.simulationTime; }
int orc_species_index(void* p) { Handle* h = static_cast<Handle*>(p); return h->precision == 32 ? h->f.speciesIndex : h->d.speciesIndex; }
int orc_pressure_index(void* p) { Handle* h = static_cast<Handle*>(p); return h->precision == 32 ? h->f.pressureIndex : h->d.pressureIndex; }

int orc_pass_log(void* p, char* buf, int cap) {
  Handle* h = static_cast<Handle*>(p);
  const std::string& s = h->precision == 32 ? h->f.log : h->d.log;
  int len = (int)s.size();
  if (cap > 0) { int c = std::min(len, cap - 1); std::memcpy(buf, s.data(), c); buf[c] = 0; }
  return len;
}

// The 44-float uniform block of the last step (writeUniforms).
void orc_uniforms(void* p, double dt, float* out44) {
  Handle* h = static_cast<Handle*>(p);
  h->f.write_uniforms(dt);
  std::memcpy(out44, h->f.u, sizeof h->f.u);
}

// ---- single-kernel entry points used by the unit tests -------------------
// Apply one kernel to the state's current buffers, with uniforms for dt.
void orc_kernel_divergence(void* p, double dt, int fromPredicted) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, (s.write_uniforms(dt), s.compute_divergence(s.flow[fromPredicted ? 1 : 0]), 0));
}
// One Jacobi sweep pressure[pi] -> pressure[1-pi], then flip.
void orc_kernel_jacobi(void* p, double dt, int f16) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, (s.write_uniforms(dt), s.jacobi_sweep(s.pressure[s.pressureIndex], s.pressure[1 - s.pressureIndex], f16 != 0),
                   s.pressureIndex = 1 - s.pressureIndex, 0));
}
// projectVelocity(flow[src], pressure[pi]) -> flow[1-src].
void orc_kernel_project(void* p, double dt, int src) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, (s.write_uniforms(dt), s.project(s.flow[src], s.pressure[s.pressureIndex], s.flow[1 - src]), 0));
}
// simulate(flow[0], species[si]) -> flow[1], species[1-si], reaction; flips si.
void orc_kernel_simulate(void* p, double dt) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, (s.write_uniforms(dt), s.simulate(s.speciesIndex), s.speciesIndex = 1 - s.speciesIndex, 0));
}
// trilinearFlow at an arbitrary position (cell units) on flow[0].
void orc_trilinear_flow(void* p, double dt, double px, double py, double pz, double* out4) {
  Handle* h = static_cast<Handle*>(p);
  ORC_DISPATCH(h, (s.write_uniforms(dt), [&] {
    using T = std::remove_reference_t<decltype(s.gx)>;
    V4<T> v = s.trilinear(s.flow[0], T(px), T(py), T(pz));
    out4[0] = v.x; out4[1] = v.y; out4[2] = v.z; out4[3] = v.w;
  }(), 0));
}
// emitter() at a normalised position (reference analytic emitter).
double orc_emitter(void* p, double dt, double npx, double npy, double npz) {
  Handle* h = static_cast<Handle*>(p);
  return ORC_DISPATCH(h, (s.write_uniforms(dt), (double)s.emitter_reference(
      (decltype(s.gx))npx, (decltype(s.gx))npy, (decltype(s.gx))npz)));
}

float orc_round_f16(float v) { return round_to_f16(v); }
float orc_round_bf16(float v) { return round_to_bf16(v); }
float orc_seeded_uniform(uint64_t seed, uint64_t frame, uint64_t cell, uint32_t stream) {
  return seeded_uniform(seed, frame, cell, stream);
}
float orc_f32_exp(float x) { return fxm_exp(x); }
float orc_f32_sin(float x) { return fxm_sin(x); }
float orc_f32_cos(float x) { return fxm_cos(x); }
float orc_f32_pow(float x, float y) { return fxm_pow(x, y); }
float orc_f32_log2(float x) { return fxm_log2(x); }
float orc_f32_exp2(float x) { return fxm_exp2(x); }

// Vectorised form of the f32 built-ins: fn 0 exp, 1 exp2, 2 log2, 3 sin,
// 4 cos, 5 pow(x, y), 6 smoothstep(y0, y1, x) with y = (y0, y1) pairs.
void orc_f32_eval(int fn, const float* x, const float* y, float* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    switch (fn) {
      case 0: out[i] = fxm_exp(x[i]); break;
      case 1: out[i] = fxm_exp2(x[i]); break;
      case 2: out[i] = fxm_log2(x[i]); break;
      case 3: out[i] = fxm_sin(x[i]); break;
      case 4: out[i] = fxm_cos(x[i]); break;
      case 5: out[i] = fxm_pow(x[i], y[i]); break;
      default: out[i] = fxm_smoothstep(y[2 * i], y[2 * i + 1], x[i]); break;
    }
  }
}

// One Jacobi sweep of the FP64 stencil is checked against a dense solve;
// this exports the grid's stencil weights (wx, wy, wz) for a test to build
// the matrix from the same uniforms.
void orc_stencil_weights(void* p, double dt, double* out3) {
  Handle* h = static_cast<Handle*>(p);
  h->d.write_uniforms(dt);
  out3[0] = h->d.mx * h->d.mx;
  out3[1] = h->d.my * h->d.my;
  out3[2] = h->d.mz * h->d.mz;
}

// packFields: three vec4 volumes (flow, chemistry, media), 4 n values each.
int orc_pack_fields(void* p, double* out_flow, double* out_chem, double* out_media) {
  Handle* h = static_cast<Handle*>(p);
  auto run = [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.flow[0][0].x)>::type;
    std::vector<V4<T>> pf, pc, pm;
    s.pack_fields(pf, pc, pm);
    double* outs[3] = {out_flow, out_chem, out_media};
    const std::vector<V4<T>>* ins[3] = {&pf, &pc, &pm};
    for (int k = 0; k < 3; ++k)
      for (size_t i = 0; i < s.n; ++i) {
        const V4<T>& v = (*ins[k])[i];
        outs[k][4 * i] = v.x; outs[k][4 * i + 1] = v.y; outs[k][4 * i + 2] = v.z; outs[k][4 * i + 3] = v.w;
      }
    return 0;
  };
  return h->precision == 32 ? run(h->f) : run(h->d);
}

// The volume ray marcher over freshly packed volumes: out receives width *
// height * 4 values (rgba, f16-rounded as the rgba16float target stores them),
// row 0 at the top.  FP32 oracle only.
int orc_render_volume(void* p, int width, int height, int ray_steps, double* out) {
  Handle* h = static_cast<Handle*>(p);
  if (h->precision != 32 || width < 1 || height < 1 || ray_steps < 1) return 1;
  State<float>& s = h->f;
  std::vector<V4<float>> pf, pc, pm;
  s.pack_fields(pf, pc, pm);
  s.canvasWidth = width;
  s.canvasHeight = height;
  s.raySteps = ray_steps;
  s.write_uniforms(0.0);
  static const std::vector<uint8_t> det = State<float>::detail_volume(64);
  const State<float>::Renderer r{s, pf, pc, pm, det, s.u, width, height};
  std::vector<float> img((size_t)width * height * 4);
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      const V3<float> c = r.pixel(x, y);
      float* o = &img[((size_t)y * width + x) * 4];
      o[0] = round_to_f16(c.x);
      o[1] = round_to_f16(c.y);
      o[2] = round_to_f16(c.z);
      o[3] = 1.0f;
    }
  // the cinematic scene pass draws the droplet streaks after the volume
  // (FireXEngine.ts:2636-2639); the standard path draws them into the canvas
  if (s.u[42] >= 0.5f && s.particleFieldActive && s.particleCount > 0) r.draw_droplets(img, false);
  for (size_t i = 0; i < img.size(); ++i) out[i] = img[i];
  return 0;
}

// The standard presentation's droplet streaks over an rgba8 frame (the canvas
// the standard pipeline draws into, FireXEngine.ts:2673-2690); the renderer's
// uniforms are those of the last orc_render_volume.
static void droplets_over_frame(Handle* h, int width, int height, uint8_t* frame) {
  State<float>& s = h->f;
  if (!(s.particleFieldActive && s.particleCount > 0)) return;
  std::vector<V4<float>> none;
  static const std::vector<uint8_t> det = State<float>::detail_volume(64);
  const State<float>::Renderer r{s, none, none, none, det, s.u, width, height};
  std::vector<float> img((size_t)width * height * 4);
  for (size_t i = 0; i < img.size(); ++i) img[i] = (float)frame[i];
  r.draw_droplets(img, true);
  for (size_t i = 0; i < img.size(); ++i) frame[i] = (uint8_t)img[i];
}

// ---- presentation (FIREX_POST_WGSL, qualityShaders.ts:580-671) ----------
namespace post {
struct Img {
  int w, h;
  std::vector<float> px;  // w * h * 4, row 0 at the top
  float at(int x, int y, int c) const { return px[((size_t)y * w + x) * 4 + c]; }
};
// linear, clamp-to-edge sampling of a 2D image at uv (v = 0: row 0)
static void sample(const Img& im, float u, float v, float out[3]) {
  const float tx = u * (float)im.w - 0.5f, ty = v * (float)im.h - 0.5f;
  const float fx0 = fxm_floor(tx), fy0 = fxm_floor(ty);
  const float fx = tx - fx0, fy = ty - fy0;
  const int x0 = std::min(std::max((int)fx0, 0), im.w - 1), x1 = std::min(std::max((int)fx0 + 1, 0), im.w - 1);
  const int y0 = std::min(std::max((int)fy0, 0), im.h - 1), y1 = std::min(std::max((int)fy0 + 1, 0), im.h - 1);
  for (int c = 0; c < 3; ++c) {
    const float a = fxm_mix(im.at(x0, y0, c), im.at(x1, y0, c), fx);
    const float b = fxm_mix(im.at(x0, y1, c), im.at(x1, y1, c), fx);
    out[c] = fxm_mix(a, b, fy);
  }
}
// bloomFragment + brightExtract, qualityShaders.ts:621-644
static Img bloom(const Img& scene) {
  Img b;
  b.w = std::max((scene.w + 3) / 4, 1);
  b.h = std::max((scene.h + 3) / 4, 1);
  b.px.assign((size_t)b.w * b.h * 4, 0.0f);
  const float rx = 1.0f / fxm_max((float)scene.w, 1.0f) * 5.5f, ry = 1.0f / fxm_max((float)scene.h, 1.0f) * 5.5f;
  // the eight taps of bloomFragment: uv + (a, b), or uv - (a, b)
  struct Tap {
    bool minus;
    float a, b;
  };
  const Tap taps[8] = {{false, rx, 0.0f}, {true, rx, 0.0f},  {false, 0.0f, ry}, {true, 0.0f, ry},
                       {false, rx, ry},   {true, rx, ry},    {false, rx, -ry},  {false, -rx, ry}};
  for (int y = 0; y < b.h; ++y)
    for (int x = 0; x < b.w; ++x) {
      const float u = ((float)x + 0.5f) / (float)b.w, v = ((float)y + 0.5f) / (float)b.h;
      float t[3], c[3];
      sample(scene, u, v, t);
      for (int k = 0; k < 3; ++k) c[k] = t[k] * 0.20f;
      for (const Tap& tp : taps) {
        sample(scene, tp.minus ? u - tp.a : u + tp.a, tp.minus ? v - tp.b : v + tp.b, t);
        for (int k = 0; k < 3; ++k) c[k] = c[k] + t[k] * 0.10f;
      }
      const float lum = c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f;
      const float threshold = 1.15f, knee = 0.35f;
      const float soft = fxm_clamp((lum - threshold + knee) / (2.0f * knee), 0.0f, 1.0f);
      const float contribution = fxm_max(lum - threshold, 0.0f) + soft * soft * knee;
      const float d = fxm_max(lum, 0.0001f);
      for (int k = 0; k < 3; ++k) b.px[((size_t)y * b.w + x) * 4 + k] = round_to_f16(c[k] * contribution / d);
      b.px[((size_t)y * b.w + x) * 4 + 3] = 1.0f;
    }
  return b;
}
static uint8_t unorm8(float c) { return (uint8_t)fxm_rint(fxm_clamp(c, 0.0f, 1.0f) * 255.0f); }
}  // namespace post

// The presentation passes over a given scene image (w * h rgba, row 0 at
// the top): with cinematic != 0, bloomFragment into the quarter-size bloom
// image and compositeFragment (filmic tone map of scene + 0.1 bloom at the
// exposure, then the dither), qualityShaders.ts:613-671; else the scene colour
// itself (the standard pipeline draws straight into the canvas).  frame: w * h
// * 4 bytes, unorm8 = rint(clamp(c, 0, 1) * 255) and alpha 255 (an opaque
// canvas).  bloom_out (may be null): ceil(w/4) * ceil(h/4) * 4 floats.
int orc_present(const float* scene_rgba, int width, int height, float exposure_param, int cinematic,
                uint8_t* out, float* bloom_out) {
  if (width <= 0 || height <= 0 || !scene_rgba || !out) return 1;
  post::Img scene{width, height, std::vector<float>(scene_rgba, scene_rgba + (size_t)width * height * 4)};
  const float exposure = fxm_max(exposure_param, 0.25f);
  post::Img bl;
  if (cinematic) {
    bl = post::bloom(scene);
    if (bloom_out) std::memcpy(bloom_out, bl.px.data(), bl.px.size() * sizeof(float));
  }
  const float gamma = (float)(1.0 / 2.2);
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      float c[3];
      if (cinematic) {
        const float u = ((float)x + 0.5f) / (float)width, v = ((float)y + 0.5f) / (float)height;
        float sc[3], b[3];
        post::sample(scene, u, v, sc);
        post::sample(bl, u, v, b);
        // dither(position.xy) at the fragment centre
        const float px = fxm_fract(((float)x + 0.5f) * 0.06711056f), py = fxm_fract(((float)y + 0.5f) * 0.00583715f);
        const float dither = (fxm_fract(52.9829189f * (px + py)) - 0.5f) / 255.0f;
        for (int k = 0; k < 3; ++k) {
          const float e = fxm_max((sc[k] + b[k] * 0.10f) * exposure, 0.0f);
          const float m = fxm_clamp((e * (2.51f * e + 0.03f)) / (e * (2.43f * e + 0.59f) + 0.14f), 0.0f, 1.0f);
          c[k] = fxm_clamp(fxm_pow(m, gamma) + dither, 0.0f, 1.0f);
        }
      } else {
        for (int k = 0; k < 3; ++k) c[k] = scene.at(x, y, k);
      }
      uint8_t* o = out + ((size_t)y * width + x) * 4;
      o[0] = post::unorm8(c[0]);
      o[1] = post::unorm8(c[1]);
      o[2] = post::unorm8(c[2]);
      o[3] = 255;
    }
  return 0;
}

// A presented frame: the ray march (rgba16f scene image), then the
// presentation passes above.  out: width * height * 4 bytes.  FP32 oracle only.
int orc_render_frame(void* p, int width, int height, int ray_steps, uint8_t* out) {
  Handle* h = static_cast<Handle*>(p);
  std::vector<double> img((size_t)width * height * 4);
  if (orc_render_volume(p, width, height, ray_steps, img.data())) return 1;
  const std::vector<float> scene(img.begin(), img.end());
  const int cinematic = h->f.cinematic ? 1 : 0;
  if (orc_present(scene.data(), width, height, (float)h->f.params.exposure, cinematic, out, nullptr)) return 1;
  if (!cinematic) droplets_over_frame(h, width, height, out);
  return 0;
}

// The 64^3 rgba8 detail volume (createPeriodicDetailVolume), 4 bytes per texel.
void orc_detail_volume(uint8_t* out) {
  const std::vector<uint8_t> d = State<float>::detail_volume(64);
  std::memcpy(out, d.data(), d.size());
}

// Threads of the OpenMP loops (the CPU baseline times 1 thread and all cores).
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
