// firex_droplets.hip — gfx950 kernels of the SPH droplet passes that run at
// the head of a step while water is on (FIREX_PARTICLE_WGSL,
// src/sim/firex/shaders.ts:514-825; dispatched by encodeSimulationStep,
// FireXEngine.ts:2158-2171):
//
//   k_sph_density    computeParticleDensity  shaders.ts:603-637
//   k_sph_integrate  integrateParticles      shaders.ts:639-735
//   k_liquid_deposit depositParticles        shaders.ts:737-825
//
// Particle records are two float4 (positionAge: position in domain
// fractions and age; velocityMass: velocity in m/s and mass), density
// records one float4 (density, near density, pressure, near pressure), as in
// the reference's storage buffers.
//
// The neighbour sums of the density and integration passes are sequential
// folds over particle index in the reference, and f32 addition does not
// associate.  Each particle therefore gets one wave: lane l evaluates
// neighbour base + l of every 64-wide chunk (distance test and the pair
// terms), and the wave then adds the contributing lanes' terms into its
// accumulators one by one in lane order (ballot, then readlane per set bit).
// Lanes that do not contribute add nothing in the reference either, so the
// sums take the same operations in the same order.  A squared-distance guard
// skips the square root and the divisions for chunks with no neighbour
// inside 1.001 radii; it only drops pairs whose rounded distance is at least
// the radius, which the exact test rejects too.
//
// Deposition adds u32 fixed-point amounts with device atomics; integer
// addition commutes, so the liquid field does not depend on the order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "firex_kernels.h"

namespace {

constexpr int kWavesPerBlock = 4;  // particles per density / integration block

__device__ inline float hash11(float v) {
  const float h = fxm_sin(v * 12.9898f + 1.317f) * 43758.5453f;
  return h - fxm_floor(h);
}

__device__ inline float len3(float x, float y, float z) { return fxm_sqrt(x * x + y * y + z * z); }

__device__ inline uint32_t cell_index(int x, int y, int z, const FxDroplets& d) {
  x = min(max(x, 0), d.nx - 1);
  y = min(max(y, 0), d.ny - 1);
  z = min(max(z, 0), d.nz - 1);
  return (uint32_t)x + (uint32_t)d.nx * ((uint32_t)y + (uint32_t)d.ny * (uint32_t)z);
}

__device__ inline float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// spawnParticle, shaders.ts:579-601 (nozzleDirection :575-578, tan as
// sin / cos).
__device__ void spawn(uint32_t id, float cycle, const FxDroplets& d, float4& pa, float4& vm) {
  const float seed = (float)id * 3.771f + cycle * 19.173f;
  const float radialRandom = fxm_sqrt(hash11(seed + 0.17f));
  const float angularRandom = hash11(seed + 4.91f) * 6.2831853f;
  const float speedRandom = hash11(seed + 9.37f);
  float dx = (0.52f - 0.035f) * d.ex;
  float dy = (fxm_clamp(d.aim, 0.08f, 0.82f) - 0.50f) * d.ey;
  float dz = (0.5f - 0.5f) * d.ez;
  float l = len3(dx, dy, dz);
  dx = dx / l;
  dy = dy / l;
  dz = dz / l;
  float tx = dy * 1.0f - dz * 0.0f, ty = dz * 0.0f - dx * 1.0f, tz = dx * 0.0f - dy * 0.0f;
  l = len3(tx, ty, tz);
  tx = tx / l;
  ty = ty / l;
  tz = tz / l;
  float bx = dy * tz - dz * ty, by = dz * tx - dx * tz, bz = dx * ty - dy * tx;
  l = len3(bx, by, bz);
  bx = bx / l;
  by = by / l;
  bz = bz / l;
  const float ca = fxm_cos(angularRandom), sa = fxm_sin(angularRandom);
  const float kx = tx * ca + bx * sa, ky = ty * ca + by * sa, kz = tz * ca + bz * sa;
  const bool spray = d.nozzle >= 0.5f;
  const float nozzleSpread = spray ? 0.72f : 0.10f;
  const float ang = fxm_clamp(d.spray_angle, 0.005f, 0.78f);
  const float cone = fxm_sin(ang) / fxm_cos(ang) * radialRandom * nozzleSpread;
  float lx = dx + kx * cone, ly = dy + ky * cone, lz = dz + kz * cone;
  l = len3(lx, ly, lz);
  lx = lx / l;
  ly = ly / l;
  lz = lz / l;
  const float originRadius = spray ? 0.006f : 0.0028f;
  const float speedVariation = spray ? 0.24f : 0.05f;
  const float speed = (0.82f + d.water_flow * 0.82f) *
                      fxm_mix(1.0f - speedVariation * 0.5f, 1.0f + speedVariation * 0.5f, speedRandom);
  pa = make_float4(0.035f + kx * radialRandom * originRadius / d.ex, 0.50f + ky * radialRandom * originRadius / d.ey,
                   0.5f + kz * radialRandom * originRadius / d.ez, 0.0f);
  vm = make_float4(lx * speed, ly * speed, lz * speed, 0.52f + d.water_flow * 0.88f);
}

// computeParticleDensity: one wave per particle.
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_sph_density(const float4* __restrict__ pin,
                                                                      float4* __restrict__ dens, FxDroplets d) {
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= d.count) return;  // whole waves
  const float4 p = pin[2 * i];
  const float mass = pin[2 * i + 1].w;
  if (mass <= 0.0001f) {
    if (lane == 0) dens[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const float radius = d.nozzle >= 0.5f ? 0.060f : 0.038f;
  const float guard = (radius * 1.001f) * (radius * 1.001f);
  float density = 0.0f, nearDensity = 0.0f;
  for (uint32_t base = 0; base < d.count; base += 64) {
    const uint32_t j = base + (uint32_t)lane;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    bool near = false;
    if (j < d.count && j != i && !(pin[2 * j + 1].w <= 0.0001f)) {
      const float4 q = pin[2 * j];
      sx = (p.x - q.x) * d.ex;
      sy = (p.y - q.y) * d.ey;
      sz = (p.z - q.z) * d.ez;
      near = sx * sx + sy * sy + sz * sz < guard;
    }
    if (__ballot(near) == 0) continue;
    float w2 = 0.0f, w3 = 0.0f;
    bool hit = false;
    if (near) {
      const float distance = len3(sx, sy, sz);
      if (distance > 0.00001f && distance < radius) {
        const float w = 1.0f - distance / radius;
        w2 = w * w;
        w3 = w * w * w;
        hit = true;
      }
    }
    for (uint64_t m = __ballot(hit); m; m &= m - 1) {
      const int l = __builtin_ctzll(m);
      density = density + lane_value(w2, l);
      nearDensity = nearDensity + lane_value(w3, l);
    }
  }
  if (lane == 0) {
    const float restDensity = 2.7f + d.water_flow * 0.9f;
    dens[i] = make_float4(density, nearDensity, fxm_max(density - restDensity, 0.0f) * 0.055f,
                          nearDensity * 0.065f);
  }
}

// integrateParticles: one wave per particle.
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_sph_integrate(const float4* __restrict__ pin,
                                                                        float4* __restrict__ pout,
                                                                        const float4* __restrict__ dens,
                                                                        const float4* __restrict__ flow,
                                                                        FxDroplets d) {
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t id = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (id >= d.count) return;
  const float4 pa = pin[2 * id];
  const float4 vm = pin[2 * id + 1];
  const float dt = d.dt;
  if (dt <= 0.0f) {
    if (lane == 0) {
      pout[2 * id] = pa;
      pout[2 * id + 1] = vm;
    }
    return;
  }
  const bool respawn = d.water_enabled >= 0.5f && d.water_flow > 0.001f;
  if (vm.w <= 0.0001f) {
    const float wait = pa.w + dt;
    float4 oa, ov;
    if (respawn && wait >= 0.0f) {
      spawn(id, wait + d.frame, d, oa, ov);
    } else {
      oa = make_float4(0.035f, 0.50f, 0.5f, wait > 0.0f ? -0.72f : wait);
      ov = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (lane == 0) {
      pout[2 * id] = oa;
      pout[2 * id + 1] = ov;
    }
    return;
  }
  const float lifetime = 0.88f + hash11((float)id * 2.31f + 7.0f) * 0.52f;
  const float nextAge = pa.w + dt;
  const bool outside = pa.x < -0.03f || pa.y < -0.03f || pa.z < -0.03f || pa.x > 1.03f || pa.y > 1.03f ||
                       pa.z > 1.03f;
  if (nextAge >= lifetime || outside) {
    float4 oa, ov;
    if (respawn) {
      spawn(id, nextAge + d.frame, d, oa, ov);
    } else {
      oa = make_float4(0.035f, 0.50f, 0.5f, -hash11((float)id * 8.17f) * 0.72f);
      ov = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (lane == 0) {
      pout[2 * id] = oa;
      pout[2 * id + 1] = ov;
    }
    return;
  }

  const bool spray = d.nozzle >= 0.5f;
  const float radius = spray ? 0.060f : 0.038f;
  const float guard = (radius * 1.001f) * (radius * 1.001f);
  const float4 pd = dens[id];
  float pax = 0.0f, pay = 0.0f, paz = 0.0f, vax = 0.0f, vay = 0.0f, vaz = 0.0f;
  for (uint32_t base = 0; base < d.count; base += 64) {
    const uint32_t j = base + (uint32_t)lane;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    bool near = false;
    if (j < d.count && j != id && !(pin[2 * j + 1].w <= 0.0001f)) {
      const float4 q = pin[2 * j];
      sx = (pa.x - q.x) * d.ex;
      sy = (pa.y - q.y) * d.ey;
      sz = (pa.z - q.z) * d.ez;
      near = sx * sx + sy * sy + sz * sz < guard;
    }
    if (__ballot(near) == 0) continue;
    float tpx = 0.0f, tpy = 0.0f, tpz = 0.0f, tvx = 0.0f, tvy = 0.0f, tvz = 0.0f;
    bool hit = false;
    if (near) {
      const float distance = len3(sx, sy, sz);
      if (distance > 0.00001f && distance < radius) {
        const float w = 1.0f - distance / radius;
        const float dx = sx / distance, dy = sy / distance, dz = sz / distance;
        const float4 nd = dens[j];
        const float4 nv = pin[2 * j + 1];
        const float pressure = (pd.z + nd.z) * w;
        const float nearPressure = (pd.w + nd.w) * w * w;
        const float push = pressure + nearPressure;
        tpx = dx * push;
        tpy = dy * push;
        tpz = dz * push;
        const float visc = w * 0.018f;
        tvx = (nv.x - vm.x) * visc;
        tvy = (nv.y - vm.y) * visc;
        tvz = (nv.z - vm.z) * visc;
        hit = true;
      }
    }
    for (uint64_t m = __ballot(hit); m; m &= m - 1) {
      const int l = __builtin_ctzll(m);
      pax = pax + lane_value(tpx, l);
      pay = pay + lane_value(tpy, l);
      paz = paz + lane_value(tpz, l);
      vax = vax + lane_value(tvx, l);
      vay = vay + lane_value(tvy, l);
      vaz = vaz + lane_value(tvz, l);
    }
  }
  if (lane != 0) return;

  // sampleFlow, shaders.ts:566-569
  const int cx = (int)(fxm_clamp(pa.x, 0.0f, 0.99999f) * d.gx);
  const int cy = (int)(fxm_clamp(pa.y, 0.0f, 0.99999f) * d.gy);
  const int cz = (int)(fxm_clamp(pa.z, 0.0f, 0.99999f) * d.gz);
  const float4 gas = flow[cell_index(cx, cy, cz, d)];
  const float heat = fxm_smoothstep(380.0f, 1500.0f, gas.w);
  const float drag = 0.30f + heat * 0.18f;
  float vx = vm.x, vy = vm.y, vz = vm.z;
  float ax = 0.0f + (gas.x - vx) * drag, ay = -0.58f + (gas.y - vy) * drag, az = 0.0f + (gas.z - vz) * drag;
  const float pressureDispersion = spray ? 1.0f : 0.055f;
  const float viscousCohesion = spray ? 1.0f : 2.2f;
  ax = ax + (pax * pressureDispersion + vax * viscousCohesion);
  ay = ay + (pay * pressureDispersion + vay * viscousCohesion);
  az = az + (paz * pressureDispersion + vaz * viscousCohesion);
  vx = vx + ax * dt;
  vy = vy + ay * dt;
  vz = vz + az * dt;
  float px = pa.x + vx * dt / d.ex, py = pa.y + vy * dt / d.ey, pz = pa.z + vz * dt / d.ez;
  if (py < 0.012f) {
    py = 0.012f;
    vy = fabsf(vy) * 0.18f;
    vx = vx * 0.72f;
    vz = vz * 0.72f;
  }
  if (px < 0.008f || px > 0.992f) {
    px = fxm_clamp(px, 0.008f, 0.992f);
    vx = vx * -0.24f;
  }
  if (pz < 0.008f || pz > 0.992f) {
    pz = fxm_clamp(pz, 0.008f, 0.992f);
    vz = vz * -0.24f;
  }
  const float remainingMass = vm.w * fxm_exp(-heat * dt * 0.34f);
  pout[2 * id] = make_float4(px, py, pz, nextAge);
  pout[2 * id + 1] = make_float4(fxm_clamp(vx, -2.0f, 2.0f), fxm_clamp(vy, -2.0f, 2.0f), fxm_clamp(vz, -2.0f, 2.0f),
                                 remainingMass);
}

// depositCorner, shaders.ts:737-740 (a NaN amount deposits nothing; adding
// zero is skipped).
__device__ inline void deposit_corner(uint32_t* liquid, int x, int y, int z, float weight, float mass,
                                      const FxDroplets& d) {
  const float v = fxm_clamp(weight * mass * 4096.0f, 0.0f, 65535.0f);
  const uint32_t fixed = (v == v) ? (uint32_t)v : 0u;
  if (fixed) atomicAdd(&liquid[cell_index(x, y, z, d)], fixed);
}

// depositAtPosition, shaders.ts:742-791.
__device__ void deposit_at(uint32_t* liquid, float px, float py, float pz, float mass, const FxDroplets& d) {
  const float vx = px * d.gx - 0.5f, vy = py * d.gy - 0.5f, vz = pz * d.gz - 0.5f;
  if (!(d.nozzle >= 0.5f)) {
    const float flx = fxm_floor(vx), fly = fxm_floor(vy), flz = fxm_floor(vz);
    const int bx = (int)flx, by = (int)fly, bz = (int)flz;
    const float fx = vx - flx, fy = vy - fly, fz = vz - flz;
    const float ix = 1.0f - fx, iy = 1.0f - fy, iz = 1.0f - fz;
    const float m = mass * 0.20f;
    deposit_corner(liquid, bx, by, bz, ix * iy * iz, m, d);
    deposit_corner(liquid, bx + 1, by, bz, fx * iy * iz, m, d);
    deposit_corner(liquid, bx, by + 1, bz, ix * fy * iz, m, d);
    deposit_corner(liquid, bx + 1, by + 1, bz, fx * fy * iz, m, d);
    deposit_corner(liquid, bx, by, bz + 1, ix * iy * fz, m, d);
    deposit_corner(liquid, bx + 1, by, bz + 1, fx * iy * fz, m, d);
    deposit_corner(liquid, bx, by + 1, bz + 1, ix * fy * fz, m, d);
    deposit_corner(liquid, bx + 1, by + 1, bz + 1, fx * fy * fz, m, d);
    return;
  }
  const int cx = (int)fxm_rint(vx), cy = (int)fxm_rint(vy), cz = (int)fxm_rint(vz);
  float kernel[27];
  float total = 0.0f;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int ox = k % 3 - 1, oy = (k / 3) % 3 - 1, oz = k / 9 - 1;
    const float distance = len3(vx - (float)(cx + ox), vy - (float)(cy + oy), vz - (float)(cz + oz));
    const float w = fxm_max(0.0f, 1.0f - distance * 0.52f);
    kernel[k] = w;
    total = total + w * w;
  }
  const float norm = fxm_max(total, 0.0001f);
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int ox = k % 3 - 1, oy = (k / 3) % 3 - 1, oz = k / 9 - 1;
    deposit_corner(liquid, cx + ox, cy + oy, cz + oz, kernel[k] * kernel[k] / norm, mass, d);
  }
}

// depositParticles: one thread per particle.
__global__ __launch_bounds__(256) void k_liquid_deposit(const float4* __restrict__ current,
                                                        const float4* __restrict__ previous,
                                                        uint32_t* __restrict__ liquid, FxDroplets d) {
  const uint32_t id = blockIdx.x * 256 + threadIdx.x;
  if (id >= d.count) return;
  const float4 p = current[2 * id];
  const float mass = current[2 * id + 1].w;
  if (mass <= 0.0001f || p.x < 0.0f || p.y < 0.0f || p.z < 0.0f || p.x >= 1.0f || p.y >= 1.0f || p.z >= 1.0f) return;
  const float4 q = previous[2 * id];
  const float qm = previous[2 * id + 1].w;
  const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
  const bool canSweep = qm > 0.0001f && p.w >= q.w && len3(dx * d.ex, dy * d.ey, dz * d.ez) < 0.20f;
  if (!canSweep) {
    deposit_at(liquid, p.x, p.y, p.z, mass, d);
    return;
  }
  for (int s = 0; s < 3; ++s) {
    const float f = (float)s / 2.0f;
    deposit_at(liquid, q.x + dx * f, q.y + dy * f, q.z + dz * f, mass / 3.0f, d);
  }
}

}  // namespace

namespace fx {

void launch_sph_density(const float4* particles, float4* density, const FxDroplets& d, hipStream_t s) {
  if (d.count == 0) return;
  const uint32_t blocks = (d.count + kWavesPerBlock - 1) / kWavesPerBlock;
  k_sph_density<<<blocks, 64 * kWavesPerBlock, 0, s>>>(particles, density, d);
}

void launch_sph_integrate(const float4* particles_in, float4* particles_out, const float4* density,
                          const float4* flow, const FxDroplets& d, hipStream_t s) {
  if (d.count == 0) return;
  const uint32_t blocks = (d.count + kWavesPerBlock - 1) / kWavesPerBlock;
  k_sph_integrate<<<blocks, 64 * kWavesPerBlock, 0, s>>>(particles_in, particles_out, density, flow, d);
}

void launch_liquid_deposit(const float4* current, const float4* previous, uint32_t* liquid, const FxDroplets& d,
                           hipStream_t s) {
  if (d.count == 0) return;
  k_liquid_deposit<<<(d.count + 255) / 256, 256, 0, s>>>(current, previous, liquid, d);
}

}  // namespace fx
