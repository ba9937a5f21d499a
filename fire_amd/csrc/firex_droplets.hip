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
// associate.  Each particle therefore gets one wave (16 per block, which
// stages the neighbours' positions and masses in LDS): lane l evaluates
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

constexpr int kParticlesPerBlock = 16;  // waves (one per particle) of a density / integration block
constexpr int kTile = 4096;             // neighbours staged in LDS at a time (64 KB)
constexpr int kTileIntegrate = 2048;    // the same for the integration pass (72 KB)

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

// Neighbour positions and masses, staged in LDS tile by tile: every wave of
// a block reads each tile, so the block fetches a neighbour once instead of
// once per particle.
struct NeighbourTile {
  float4 pm[kTile];  // (x, y, z, mass)
};

__device__ inline void load_tile(NeighbourTile& t, const float4* __restrict__ pin, uint32_t t0, uint32_t tn) {
  for (uint32_t k = threadIdx.x; k < tn; k += blockDim.x) {
    float4 a = pin[2 * (t0 + k)];
    a.w = pin[2 * (t0 + k) + 1].w;
    t.pm[k] = a;
  }
}

// computeParticleDensity: one wave per particle, kParticlesPerBlock per block.
__global__ __launch_bounds__(64 * kParticlesPerBlock) void k_sph_density(const float4* __restrict__ pin,
                                                                          float4* __restrict__ dens, FxDroplets d) {
  __shared__ NeighbourTile tile;
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t i = blockIdx.x * kParticlesPerBlock + (threadIdx.x >> 6);
  const bool mine = i < d.count;  // whole waves; every wave joins the tile barriers
  const float4 p = mine ? pin[2 * i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float mass = mine ? pin[2 * i + 1].w : 0.0f;
  const bool live = mine && !(mass <= 0.0001f);
  const float radius = d.nozzle >= 0.5f ? 0.060f : 0.038f;
  const float guard = (radius * 1.001f) * (radius * 1.001f);
  float density = 0.0f, nearDensity = 0.0f;
  for (uint32_t t0 = 0; t0 < d.count; t0 += kTile) {
    const uint32_t tn = min((uint32_t)kTile, d.count - t0);
    __syncthreads();
    load_tile(tile, pin, t0, tn);
    __syncthreads();
    if (!live) continue;
    for (uint32_t base = 0; base < tn; base += 64) {
      const uint32_t k = base + (uint32_t)lane;
      float sx = 0.0f, sy = 0.0f, sz = 0.0f;
      bool near = false;
      if (k < tn && t0 + k != i) {
        const float4 q = tile.pm[k];
        if (!(q.w <= 0.0001f)) {
          sx = (p.x - q.x) * d.ex;
          sy = (p.y - q.y) * d.ey;
          sz = (p.z - q.z) * d.ez;
          near = sx * sx + sy * sy + sz * sz < guard;
        }
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
  }
  if (!mine || lane != 0) return;
  if (!live) {
    dens[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const float restDensity = 2.7f + d.water_flow * 0.9f;
  dens[i] = make_float4(density, nearDensity, fxm_max(density - restDensity, 0.0f) * 0.055f, nearDensity * 0.065f);
}

// The integration pass also needs each contributing neighbour's velocity and
// pressures: staged with its position.
struct IntegrateTile {
  float4 pm[kTileIntegrate];  // (x, y, z, mass)
  float4 vp[kTileIntegrate];  // (vx, vy, vz, pressure)
  float np[kTileIntegrate];   // near pressure
};

__device__ inline void load_integrate_tile(IntegrateTile& t, const float4* __restrict__ pin,
                                           const float4* __restrict__ dens, uint32_t t0, uint32_t tn) {
  for (uint32_t k = threadIdx.x; k < tn; k += blockDim.x) {
    float4 a = pin[2 * (t0 + k)];
    float4 v = pin[2 * (t0 + k) + 1];
    const float4 q = dens[t0 + k];
    a.w = v.w;
    v.w = q.z;
    t.pm[k] = a;
    t.vp[k] = v;
    t.np[k] = q.w;
  }
}

// integrateParticles: one wave per particle, kParticlesPerBlock per block.
__global__ __launch_bounds__(64 * kParticlesPerBlock) void k_sph_integrate(const float4* __restrict__ pin,
                                                                            float4* __restrict__ pout,
                                                                            const float4* __restrict__ dens,
                                                                            const float4* __restrict__ flow,
                                                                            FxDroplets d) {
  __shared__ IntegrateTile tile;
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t id = blockIdx.x * kParticlesPerBlock + (threadIdx.x >> 6);
  const bool mine = id < d.count;
  const float4 pa = mine ? pin[2 * id] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 vm = mine ? pin[2 * id + 1] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float dt = d.dt;
  const bool respawn = d.water_enabled >= 0.5f && d.water_flow > 0.001f;
  // what this particle does: 0 copy (dt <= 0), 1 waiting (inactive),
  // 2 expired or outside, 3 integrate
  int mode = 3;
  float lifetime = 0.0f, nextAge = 0.0f;
  if (dt <= 0.0f) {
    mode = 0;
  } else if (vm.w <= 0.0001f) {
    mode = 1;
  } else {
    lifetime = 0.88f + hash11((float)id * 2.31f + 7.0f) * 0.52f;
    nextAge = pa.w + dt;
    const bool outside = pa.x < -0.03f || pa.y < -0.03f || pa.z < -0.03f || pa.x > 1.03f || pa.y > 1.03f ||
                         pa.z > 1.03f;
    if (nextAge >= lifetime || outside) mode = 2;
  }
  const bool live = mine && mode == 3;

  const bool spray = d.nozzle >= 0.5f;
  const float radius = spray ? 0.060f : 0.038f;
  const float guard = (radius * 1.001f) * (radius * 1.001f);
  const float4 pd = live ? dens[id] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float pax = 0.0f, pay = 0.0f, paz = 0.0f, vax = 0.0f, vay = 0.0f, vaz = 0.0f;
  for (uint32_t t0 = 0; t0 < d.count; t0 += kTileIntegrate) {
    const uint32_t tn = min((uint32_t)kTileIntegrate, d.count - t0);
    __syncthreads();
    load_integrate_tile(tile, pin, dens, t0, tn);
    __syncthreads();
    if (!live) continue;
    for (uint32_t base = 0; base < tn; base += 64) {
      const uint32_t k = base + (uint32_t)lane;
      float sx = 0.0f, sy = 0.0f, sz = 0.0f;
      bool near = false;
      if (k < tn && t0 + k != id) {
        const float4 q = tile.pm[k];
        if (!(q.w <= 0.0001f)) {
          sx = (pa.x - q.x) * d.ex;
          sy = (pa.y - q.y) * d.ey;
          sz = (pa.z - q.z) * d.ez;
          near = sx * sx + sy * sy + sz * sz < guard;
        }
      }
      if (__ballot(near) == 0) continue;
      float tpx = 0.0f, tpy = 0.0f, tpz = 0.0f, tvx = 0.0f, tvy = 0.0f, tvz = 0.0f;
      bool hit = false;
      if (near) {
        const float distance = len3(sx, sy, sz);
        if (distance > 0.00001f && distance < radius) {
          const float w = 1.0f - distance / radius;
          const float dx = sx / distance, dy = sy / distance, dz = sz / distance;
          const float4 nv = tile.vp[k];
          const float pressure = (pd.z + nv.w) * w;
          const float nearPressure = (pd.w + tile.np[k]) * w * w;
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
  }
  if (!mine || lane != 0) return;

  if (mode == 0) {
    pout[2 * id] = pa;
    pout[2 * id + 1] = vm;
    return;
  }
  if (mode == 1 || mode == 2) {
    float4 oa, ov;
    const float cycle = mode == 1 ? pa.w + dt : nextAge;  // wait, or the age at expiry
    if (respawn && (mode == 2 || cycle >= 0.0f)) {
      spawn(id, cycle + d.frame, d, oa, ov);
    } else if (mode == 1) {
      oa = make_float4(0.035f, 0.50f, 0.5f, cycle > 0.0f ? -0.72f : cycle);
      ov = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
      oa = make_float4(0.035f, 0.50f, 0.5f, -hash11((float)id * 8.17f) * 0.72f);
      ov = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    pout[2 * id] = oa;
    pout[2 * id + 1] = ov;
    return;
  }

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

// depositParticles: one thread per particle and sweep sample (a droplet
// that is not swept deposits from sample 0 only), so the scattered atomics
// of neighbouring samples are issued by different waves.
__global__ __launch_bounds__(256) void k_liquid_deposit(const float4* __restrict__ current,
                                                        const float4* __restrict__ previous,
                                                        uint32_t* __restrict__ liquid, FxDroplets d) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 3 * d.count) return;
  const uint32_t id = t / 3, sample = t - 3 * id;
  const float4 p = current[2 * id];
  const float mass = current[2 * id + 1].w;
  if (mass <= 0.0001f || p.x < 0.0f || p.y < 0.0f || p.z < 0.0f || p.x >= 1.0f || p.y >= 1.0f || p.z >= 1.0f) return;
  const float4 q = previous[2 * id];
  const float qm = previous[2 * id + 1].w;
  const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
  const bool canSweep = qm > 0.0001f && p.w >= q.w && len3(dx * d.ex, dy * d.ey, dz * d.ez) < 0.20f;
  if (!canSweep) {
    if (sample == 0) deposit_at(liquid, p.x, p.y, p.z, mass, d);
    return;
  }
  const float f = (float)sample / 2.0f;
  deposit_at(liquid, q.x + dx * f, q.y + dy * f, q.z + dz * f, mass / 3.0f, d);
}

}  // namespace

namespace fx {

void launch_sph_density(const float4* particles, float4* density, const FxDroplets& d, hipStream_t s) {
  if (d.count == 0) return;
  const uint32_t blocks = (d.count + kParticlesPerBlock - 1) / kParticlesPerBlock;
  k_sph_density<<<blocks, 64 * kParticlesPerBlock, 0, s>>>(particles, density, d);
}

void launch_sph_integrate(const float4* particles_in, float4* particles_out, const float4* density,
                          const float4* flow, const FxDroplets& d, hipStream_t s) {
  if (d.count == 0) return;
  const uint32_t blocks = (d.count + kParticlesPerBlock - 1) / kParticlesPerBlock;
  k_sph_integrate<<<blocks, 64 * kParticlesPerBlock, 0, s>>>(particles_in, particles_out, density, flow, d);
}

void launch_liquid_deposit(const float4* current, const float4* previous, uint32_t* liquid, const FxDroplets& d,
                           hipStream_t s) {
  if (d.count == 0) return;
  k_liquid_deposit<<<(3 * d.count + 255) / 256, 256, 0, s>>>(current, previous, liquid, d);
}

}  // namespace fx
