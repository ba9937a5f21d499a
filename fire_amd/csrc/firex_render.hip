// firex_render.hip — the volume ray marcher of the cinematic scene pass
// (volumeFragment, src/sim/firex/qualityShaders.ts:374-575, drawn by
// encodeRender, FireXEngine.ts:2607-2640) as a gfx950 compute kernel: one
// thread per pixel, 16 x 16 pixels per block (four 8 x 8 waves, so the rays of
// a wave stay close and share cached texels), writing the rgba16float HDR
// scene image the reference renders into.  The kernel is VALU-bound (texel
// conversion and interpolation): one texel site serves the three volumes at a
// sample position, and each diagnostic view has its own instantiation that
// computes only the components it reads.
//
// The reference samples its three rgba16float render volumes and the 64^3
// rgba8unorm detail volume with the GPU's linear filter, whose weight
// precision is implementation-defined.  Here filtering is arithmetic: texel
// coordinate t = uvw * size - 0.5, i0 = floor(t), f = t - i0, the eight
// texels (clamp-to-edge for the volumes, repeat for the detail volume) mixed
// along x, then y, then z as a * (1 - f) + b * f in f32; rgba8unorm texels
// read as byte / 255.  A pixel's inputs are those of the fullscreen
// triangle: uv = ((x + 0.5) / width, 1 - (y + 0.5) / height) and
// position.xy = (x + 0.5, y + 0.5).  Everything else follows the WGSL in
// operation order with the f32 built-ins of fx_math.h.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "firex_kernels.h"

namespace {

struct V3 {
  float x, y, z;
};

__device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ inline V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ inline V3 divv(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float length3(V3 a) { return fxm_sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ inline V3 normalize3(V3 a) {
  const float l = length3(a);
  return v3(a.x / l, a.y / l, a.z / l);
}
__device__ inline V3 cross3(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline V3 mix3(V3 a, V3 b, float t) {
  return v3(fxm_mix(a.x, b.x, t), fxm_mix(a.y, b.y, t), fxm_mix(a.z, b.z, t));
}
__device__ inline float4 mix4(float4 a, float4 b, float t) {
  return make_float4(fxm_mix(a.x, b.x, t), fxm_mix(a.y, b.y, t), fxm_mix(a.z, b.z, t), fxm_mix(a.w, b.w, t));
}

// A texel of the rgba16float volumes; every texel starts at a multiple of 8
// bytes, so one 8-byte load fetches it (2-byte alignment would let the
// compiler split texel loads into pieces).
struct alignas(8) Half4 {
  _Float16 x, y, z, w;
};

__device__ inline float4 texel(const Half4* __restrict__ vol, uint32_t i) {
  const Half4 h = vol[i];
  return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
}

// Where a sample of the three render volumes falls: the texel of the lower
// corner, the steps to the upper neighbours along x, y, z (0 at a clamped
// edge) and the weights.  The flow, chemistry and media samples at one
// position share it (sampleFlow / sampleChemistry / sampleMedia,
// qualityShaders.ts:262-278: the position clamped half a texel inside the
// volume, then textureSampleLevel with the linear clamp-to-edge sampler).
struct Site {
  uint32_t i000, dx, dy, dz;
  float fx, fy, fz;
};

__device__ inline Site sample_site(V3 position, const FxRender& r) {
  const V3 half_texel = v3(0.5f / r.gx, 0.5f / r.gy, 0.5f / r.gz);
  const V3 q = divv(position, v3(r.ex, r.ey, r.ez));
  const V3 uvw = v3(fxm_clamp(q.x, half_texel.x, 1.0f - half_texel.x), fxm_clamp(q.y, half_texel.y, 1.0f - half_texel.y),
                    fxm_clamp(q.z, half_texel.z, 1.0f - half_texel.z));
  const float tx = uvw.x * r.gx - 0.5f, ty = uvw.y * r.gy - 0.5f, tz = uvw.z * r.gz - 0.5f;
  const float flx = fxm_floor(tx), fly = fxm_floor(ty), flz = fxm_floor(tz);
  const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
  const int x0 = min(max(ix, 0), r.nx - 1), x1 = min(max(ix + 1, 0), r.nx - 1);
  const int y0 = min(max(iy, 0), r.ny - 1), y1 = min(max(iy + 1, 0), r.ny - 1);
  const int z0 = min(max(iz, 0), r.nz - 1), z1 = min(max(iz + 1, 0), r.nz - 1);
  Site s;
  s.i000 = (uint32_t)x0 + (uint32_t)r.nx * ((uint32_t)y0 + (uint32_t)r.ny * (uint32_t)z0);
  s.dx = (uint32_t)(x1 - x0);
  s.dy = y1 != y0 ? (uint32_t)r.nx : 0u;
  s.dz = z1 != z0 ? (uint32_t)r.nx * (uint32_t)r.ny : 0u;
  s.fx = tx - flx;
  s.fy = ty - fly;
  s.fz = tz - flz;
  return s;
}

// The eight texels mixed along x, then y, then z.
__device__ inline float4 sample_volume(const Half4* __restrict__ vol, const Site& s) {
  const uint32_t i = s.i000;
  const float4 c00 = mix4(texel(vol, i), texel(vol, i + s.dx), s.fx);
  const float4 c10 = mix4(texel(vol, i + s.dy), texel(vol, i + s.dy + s.dx), s.fx);
  const float4 c01 = mix4(texel(vol, i + s.dz), texel(vol, i + s.dz + s.dx), s.fx);
  const float4 c11 = mix4(texel(vol, i + s.dz + s.dy), texel(vol, i + s.dz + s.dy + s.dx), s.fx);
  return mix4(mix4(c00, c10, s.fy), mix4(c01, c11, s.fy), s.fz);
}

__device__ inline float4 detail_texel(const uint32_t* __restrict__ detail, int x, int y, int z) {
  const int s = kFxDetailSize;
  x = ((x % s) + s) % s;
  y = ((y % s) + s) % s;
  z = ((z % s) + s) % s;
  const uint32_t v = detail[x + s * (y + s * z)];
  return make_float4((float)(v & 255u) / 255.0f, (float)((v >> 8) & 255u) / 255.0f, (float)((v >> 16) & 255u) / 255.0f,
                     (float)(v >> 24) / 255.0f);
}

// sampleDetail: textureSampleLevel(detailTexture, detailSampler (linear, repeat), position, 0)
__device__ float4 sample_detail(const uint32_t* __restrict__ detail, V3 uvw) {
  const float s = (float)kFxDetailSize;
  const float tx = uvw.x * s - 0.5f, ty = uvw.y * s - 0.5f, tz = uvw.z * s - 0.5f;
  const float flx = fxm_floor(tx), fly = fxm_floor(ty), flz = fxm_floor(tz);
  const float fx = tx - flx, fy = ty - fly, fz = tz - flz;
  const int x0 = (int)flx, y0 = (int)fly, z0 = (int)flz;
  const float4 c00 = mix4(detail_texel(detail, x0, y0, z0), detail_texel(detail, x0 + 1, y0, z0), fx);
  const float4 c10 = mix4(detail_texel(detail, x0, y0 + 1, z0), detail_texel(detail, x0 + 1, y0 + 1, z0), fx);
  const float4 c01 = mix4(detail_texel(detail, x0, y0, z0 + 1), detail_texel(detail, x0 + 1, y0, z0 + 1), fx);
  const float4 c11 = mix4(detail_texel(detail, x0, y0 + 1, z0 + 1), detail_texel(detail, x0 + 1, y0 + 1, z0 + 1), fx);
  return mix4(mix4(c00, c10, fy), mix4(c01, c11, fy), fz);
}

// hash21, qualityShaders.ts:286-289
__device__ inline float hash21(float vx, float vy) {
  const float px = fxm_fract(vx * 123.34f), py = fxm_fract(vy * 456.21f);
  return fxm_fract((px + py) * (px + py + 45.32f));
}

// blackbody, qualityShaders.ts:301-309
__device__ V3 blackbody(float temperature) {
  const float t = fxm_clamp((temperature - 720.0f) / 1900.0f, 0.0f, 1.0f);
  const V3 red = v3(1.0f, 0.035f, 0.002f), orange = v3(1.0f, 0.24f, 0.008f);
  const V3 gold = v3(1.0f, 0.68f, 0.12f), white = v3(1.0f, 0.94f, 0.72f);
  return mix3(mix3(red, orange, fxm_smoothstep(0.0f, 0.32f, t)), mix3(gold, white, fxm_smoothstep(0.58f, 1.0f, t)),
              fxm_smoothstep(0.26f, 0.72f, t));
}

// thermalMap, qualityShaders.ts:311-319
__device__ V3 thermal_map(float value) {
  const float t = fxm_clamp(value, 0.0f, 1.0f);
  const V3 cold = v3(0.015f, 0.025f, 0.12f), blue = v3(0.02f, 0.36f, 1.0f);
  const V3 amber = v3(1.0f, 0.34f, 0.01f), white = v3(1.0f, 0.96f, 0.76f);
  return mix3(mix3(cold, blue, fxm_smoothstep(0.0f, 0.30f, t)), mix3(amber, white, fxm_smoothstep(0.62f, 1.0f, t)),
              fxm_smoothstep(0.24f, 0.72f, t));
}

// acesCurve / aces, qualityShaders.ts:321-336
__device__ inline float aces_curve(float v) {
  return fxm_clamp((v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f), 0.0f, 1.0f);
}

__device__ V3 aces(V3 color, const FxRender& r) {
  const float e = fxm_max(r.exposure, 0.25f);
  const V3 x = v3(fxm_max(color.x * e, 0.0f), fxm_max(color.y * e, 0.0f), fxm_max(color.z * e, 0.0f));
  const float peak = fxm_max(x.x, fxm_max(x.y, x.z));
  const float k = aces_curve(peak) / fxm_max(peak, 0.000001f);
  const float inv_gamma = kFxInverseGamma;
  return v3(fxm_pow(fxm_clamp(x.x * k, 0.0f, 1.0f), inv_gamma), fxm_pow(fxm_clamp(x.y * k, 0.0f, 1.0f), inv_gamma),
            fxm_pow(fxm_clamp(x.z * k, 0.0f, 1.0f), inv_gamma));
}

// diagnosticSample, qualityShaders.ts:338-380
__device__ float4 diagnostic_sample(uint32_t mode, float4 flow, float4 chem, float4 media, const FxRender& r) {
  if (mode == 1u) {
    const float strength = fxm_smoothstep(0.015f, 0.72f, flow.w);
    const V3 c = thermal_map(flow.w);
    return make_float4(c.x, c.y, c.z, strength);
  }
  if (mode == 2u) {
    const float strength = fxm_smoothstep(0.003f, 0.38f, media.x);
    const V3 c = mix3(v3(0.45f, 0.01f, 0.0f), v3(1.0f, 0.95f, 0.52f), strength);
    return make_float4(c.x, c.y, c.z, strength);
  }
  if (mode == 3u) {
    const float fuel = fxm_clamp(chem.x, 0.0f, 1.0f);
    const float ambient_oxygen = fxm_max(r.chem_y, 0.001f);
    const float deficit = fxm_clamp((ambient_oxygen - chem.y) / ambient_oxygen, 0.0f, 1.0f);
    const float overlap = fxm_min(fuel, fxm_clamp(chem.y / ambient_oxygen, 0.0f, 1.0f));
    const float strength = fxm_max(fuel, fxm_max(deficit * 0.72f, overlap * 0.32f));
    return make_float4(fuel, overlap * 0.36f, deficit, strength);
  }
  if (mode == 4u) {
    const float strength = fxm_smoothstep(0.005f, 0.62f, chem.z);
    const V3 c = mix3(v3(0.06f, 0.02f, 0.12f), v3(0.74f, 0.26f, 1.0f), strength);
    return make_float4(c.x, c.y, c.z, strength);
  }
  if (mode == 5u) {
    const float vapor = fxm_smoothstep(0.004f, 0.72f, chem.w);
    const float soot = fxm_smoothstep(0.003f, 0.55f, media.y);
    const V3 c = mix3(v3(0.56f, 0.72f, 0.82f), v3(0.17f, 0.12f, 0.10f), soot);
    return make_float4(c.x, c.y, c.z, fxm_max(vapor, soot));
  }
  if (mode == 6u) {
    const float liquid = fxm_smoothstep(0.002f, 0.32f, media.z);
    const float speed = fxm_smoothstep(0.035f, 1.25f, length3(v3(flow.x, flow.y, flow.z)));
    const V3 n = normalize3(add(v3(flow.x, flow.y, flow.z), v3(0.0001f, 0.0001f, 0.0001f)));
    const V3 direction = add(v3(0.5f, 0.5f, 0.5f), mul(n, 0.5f));
    const V3 c = mix3(mulv(direction, v3(0.25f, 0.31f, 0.36f)), v3(0.78f, 0.91f, 1.0f), liquid);
    return make_float4(c.x, c.y, c.z, fxm_max(liquid, speed * 0.038f));
  }
  const float divergence = media.w;
  const float magnitude = fxm_smoothstep(0.01f, 1.2f, fabsf(divergence));
  const V3 c = divergence >= 0.0f ? v3(1.0f, 0.18f, 0.045f) : v3(0.12f, 0.42f, 1.0f);
  return make_float4(c.x, c.y, c.z, magnitude);
}

// studioBackground, qualityShaders.ts:382-393
__device__ V3 studio_background(float uvy, V3 eye, V3 dir, const FxRender& r) {
  V3 color = mix3(v3(0.020f, 0.024f, 0.030f), v3(0.0025f, 0.004f, 0.007f), fxm_clamp(uvy, 0.0f, 1.0f));
  const float floor_distance = -eye.y / fxm_min(dir.y, -0.00001f);
  if (dir.y < -0.00001f && floor_distance > 0.0f) {
    const V3 fp = add(eye, mul(dir, floor_distance));
    const float ddx = fp.x - r.ex * 0.5f, ddz = fp.z - r.ez * 0.5f;
    const float rings = fxm_smoothstep(0.012f, 0.0f, fabsf(fxm_sqrt(ddx * ddx + ddz * ddz) - 0.10f));
    const float grid = fxm_pow(fxm_max(fabsf(fxm_sin(fp.x * 31.4159f)), fabsf(fxm_sin(fp.z * 31.4159f))), 28.0f);
    color = add(mix3(v3(0.018f, 0.020f, 0.022f), v3(0.045f, 0.048f, 0.050f), grid * 0.22f),
                mul(v3(0.055f, 0.020f, 0.006f), rings));
  }
  return color;
}

// rotateDetailCoordinate, qualityShaders.ts:395-405
__device__ inline V3 rotate_detail(V3 p) {
  const float a = p.x * 0.771f + p.z * 0.637f;
  const float b = -p.x * 0.637f + p.z * 0.771f;
  return v3(a, p.y * 0.909f - b * 0.417f, p.y * 0.417f + b * 0.909f);
}

// volumeFragment, qualityShaders.ts:407-575.  VIEW 1-7 instantiates one
// diagnostic view mode (u32(clamp(round(controls.x), 0, 7))): the compiler
// then drops the texel conversions and interpolations of components that view
// never reads (about 3x faster).  kRuntimeView reads the mode at run time;
// the beauty view uses it, as a beauty-only instantiation measured slower.
constexpr uint32_t kRuntimeView = 8;

template <uint32_t VIEW>
__global__ __launch_bounds__(256) void k_render_volume(const Half4* __restrict__ flow_vol,
                                                       const Half4* __restrict__ chem_vol,
                                                       const Half4* __restrict__ media_vol,
                                                       const uint32_t* __restrict__ detail, Half4* __restrict__ out,
                                                       FxRender r) {
  const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
  const int px = (int)blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
  const int py = (int)blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
  if (px >= r.width || py >= r.height) return;

  const float uvx = ((float)px + 0.5f) / (float)r.width;
  const float uvy = 1.0f - ((float)py + 0.5f) / (float)r.height;
  const float sx = uvx * 2.0f - 1.0f, sy = uvy * 2.0f - 1.0f;
  const V3 extent = v3(r.ex, r.ey, r.ez);
  const V3 target = v3(r.ex * 0.5f, r.ey * 0.40f, r.ez * 0.5f);
  const V3 orbit = v3(fxm_sin(r.yaw) * fxm_cos(r.pitch), fxm_sin(r.pitch), fxm_cos(r.yaw) * fxm_cos(r.pitch));
  const float frame_scale = fxm_max(r.ex, fxm_max(r.ey, r.ez));
  const V3 eye = add(target, mul(mul(orbit, 0.90f + r.cam_distance * 0.18f), frame_scale));
  const V3 forward = normalize3(sub(target, eye));
  const V3 right = normalize3(cross3(forward, v3(0.0f, 1.0f, 0.0f)));
  const V3 up = normalize3(cross3(right, forward));
  const float lens = 0.72f / fxm_clamp(r.zoom, 0.5f, 3.0f);
  const V3 dir = normalize3(add(add(forward, mul(mul(mul(right, sx), r.aspect), lens)), mul(mul(up, sy), lens)));

  // rayBox, qualityShaders.ts:291-299
  const V3 safe = v3(fabsf(dir.x) > 0.000001f ? dir.x : 0.000001f, fabsf(dir.y) > 0.000001f ? dir.y : 0.000001f,
                     fabsf(dir.z) > 0.000001f ? dir.z : 0.000001f);
  const V3 inv = v3(1.0f / safe.x, 1.0f / safe.y, 1.0f / safe.z);
  const V3 lower = mulv(v3(-eye.x, -eye.y, -eye.z), inv);
  const V3 upper = mulv(sub(extent, eye), inv);
  const float near_t = fxm_max(fxm_max(fxm_min(lower.x, upper.x), fxm_min(lower.y, upper.y)), fxm_min(lower.z, upper.z));
  const float far_t = fxm_min(fxm_min(fxm_max(lower.x, upper.x), fxm_max(lower.y, upper.y)), fxm_max(lower.z, upper.z));

  const V3 background = studio_background(uvy, eye, dir, r);
  const float cinematic = fxm_clamp(r.cinematic, 0.0f, 1.0f);
  V3 color;
  if (near_t > far_t || far_t <= 0.0f) {
    color = cinematic >= 0.5f ? background : aces(background, r);
  } else {
    const float start = fxm_max(near_t, 0.0f);
    const uint32_t step_count = (uint32_t)r.steps;
    const float step_length = (far_t - start) / fxm_max((float)step_count, 1.0f);
    const float jitter = hash21(fxm_floor((float)px + 0.5f), fxm_floor((float)py + 0.5f));
    float distance = start + jitter * step_length;
    float transmittance = 1.0f;
    V3 radiance = v3(0.0f, 0.0f, 0.0f);
    const uint32_t view_mode =
        VIEW == kRuntimeView ? (uint32_t)fxm_clamp(fxm_rint(r.view_mode), 0.0f, 7.0f) : VIEW;
    const bool beauty = view_mode == 0u;
    const float front_def = fxm_clamp(r.front_definition, 0.5f, 2.5f);
    for (uint32_t step = 0; step < 384u; ++step) {
      if (step >= step_count || distance >= far_t || transmittance < 0.006f) break;
      const V3 position = add(eye, mul(dir, distance));
      const Site site = sample_site(position, r);
      const float4 raw_media = sample_volume(media_vol, site);
      if (beauty && cinematic >= 0.5f) {
        const float occupancy = fxm_max(fxm_max(raw_media.x, raw_media.y), fxm_max(raw_media.z, raw_media.w));
        if (occupancy < 0.0005f) {
          const float next_distance = distance + step_length;
          if (next_distance < far_t) {
            const float4 nm = sample_volume(media_vol, sample_site(add(eye, mul(dir, next_distance)), r));
            const float next_occupancy = fxm_max(fxm_max(nm.x, nm.y), fxm_max(nm.z, nm.w));
            distance += next_occupancy < 0.0005f ? step_length * 2.0f : step_length;
          } else {
            distance = far_t;
          }
          continue;
        }
      }
      const float4 flow = sample_volume(flow_vol, site);
      float4 chem = sample_volume(chem_vol, site);
      chem = make_float4(fxm_max(chem.x, 0.0f), fxm_max(chem.y, 0.0f), fxm_max(chem.z, 0.0f), fxm_max(chem.w, 0.0f));
      const float4 media = make_float4(fxm_max(raw_media.x, 0.0f), fxm_max(raw_media.y, 0.0f),
                                       fxm_max(raw_media.z, 0.0f), raw_media.w);
      if (beauty) {
        const float temperature = 300.0f + flow.w * 2600.0f;
        const float reaction_rate = fxm_max(media.x, 0.0f);
        const float reaction_front = fxm_mix(fxm_smoothstep(0.006f, 0.32f, reaction_rate),
                                             fxm_smoothstep(0.004f, 0.28f, reaction_rate), cinematic);
        const float thermal = fxm_smoothstep(620.0f, 1780.0f, temperature);
        const V3 detail_position = sub(position, mul(v3(flow.x, flow.y, flow.z), 0.08f));
        const float detail_a = fxm_sin(dot3(detail_position, v3(27.0f, 35.0f, 13.0f)) + r.time * 3.4f);
        const float detail_b = fxm_sin(dot3(detail_position, v3(-21.0f, 29.0f, 17.0f)) - r.time * 4.7f);
        float optical_detail = 0.90f + 0.20f * detail_a * detail_b;
        if (cinematic > 0.5f && reaction_front > 0.002f) {
          const V3 normalized = divv(detail_position, extent);
          const float requested = r.detail_target;
          const float explicit_detail = fxm_clamp(requested, 64.0f, 1024.0f);
          const V3 detail_cells = requested >= 1.0f ? v3(explicit_detail, explicit_detail, explicit_detail)
                                                    : v3(r.gx, r.gy, r.gz);
          const V3 tile_scale = v3(fxm_max(detail_cells.x / 64.0f, 0.015625f), fxm_max(detail_cells.y / 64.0f, 0.015625f),
                                   fxm_max(detail_cells.z / 64.0f, 0.015625f));
          const V3 time_offset = mul(v3(0.013f, -0.021f, 0.017f), r.time);
          const V3 detail_uv = add(mulv(normalized, tile_scale), time_offset);
          const float primary_noise = sample_detail(detail, detail_uv).x - 0.5f;
          const V3 rotated = rotate_detail(sub(normalized, v3(0.5f, 0.5f, 0.5f)));
          const V3 secondary_uv = sub(add(mul(mulv(rotated, tile_scale), 0.47f), v3(0.371f, 0.113f, 0.733f)),
                                      mul(v3(time_offset.y, time_offset.z, time_offset.x), 0.61f));
          const float secondary_noise = sample_detail(detail, secondary_uv).y - 0.5f;
          const float maximum_cells = fxm_max(detail_cells.x, fxm_max(detail_cells.y, detail_cells.z));
          const float ray_budget = fxm_clamp(r.render_z, 1.0f, 384.0f);
          const float primary_band = fxm_clamp((ray_budget * 2.0f) / maximum_cells, 0.25f, 1.0f);
          const float secondary_band = fxm_clamp((ray_budget * 2.0f) / fxm_max(maximum_cells * 0.47f, 1.0f), 0.35f, 1.0f);
          const float fbm = primary_noise * (0.55f * primary_band) + secondary_noise * (0.45f * secondary_band);
          const float edge_mask = fxm_smoothstep(0.025f, 0.24f, reaction_front) *
                                  (1.0f - fxm_smoothstep(0.72f, 0.96f, reaction_front));
          const float breakup = fxm_clamp(1.0f + fbm * 0.44f, 0.84f, 1.16f);
          optical_detail = fxm_mix(1.0f, breakup, edge_mask * 0.85f);
        }
        const float front_exponent = fxm_mix(0.85f + front_def * 0.35f, 0.92f + front_def * 0.16f, cinematic);
        const float cinematic_front = reaction_front * (0.42f + 0.58f * thermal);
        const float standard_front = reaction_front * thermal;
        const float flame_density =
            fxm_pow(fxm_clamp(fxm_mix(standard_front, cinematic_front, cinematic), 0.0f, 1.0f), front_exponent) *
            optical_detail;
        const float soot = fxm_clamp(1.0f - fxm_exp(-fxm_max(media.y, 0.0f) * fxm_mix(1.15f, 0.90f, cinematic)), 0.0f,
                                     fxm_mix(1.0f, 0.96f, cinematic));
        const float vapor = 1.0f - fxm_exp(-fxm_max(chem.w, 0.0f) * fxm_mix(0.34f, 0.30f, cinematic));
        const float richness = fxm_clamp(chem.x / fxm_max(chem.x + chem.y * 0.25f, 0.0001f), 0.0f, 1.0f);
        const float soot_reveal = fxm_mix(1.0f, 1.0f - fxm_smoothstep(0.045f, 0.30f, reaction_front) * 0.78f, cinematic);
        const float flame_sigma = flame_density * fxm_mix(3.20f, 0.88f, cinematic);
        const float soot_sigma = soot * soot_reveal * fxm_mix(1.6f, 1.62f, cinematic);
        const float vapor_sigma = vapor * fxm_mix(0.20f, 0.32f, cinematic);
        const float extinction = fxm_min(flame_sigma + soot_sigma + vapor_sigma, fxm_mix(100.0f, 5.5f, cinematic));
        const float optical_depth =
            fxm_min(extinction * step_length * fxm_mix(5.2f, 3.15f, cinematic), fxm_mix(100.0f, 0.28f, cinematic));
        const float alpha = 1.0f - fxm_exp(-optical_depth);
        const float hot_soot = soot * thermal;
        const float emission_temperature = fxm_mix(
            temperature + reaction_front * 650.0f,
            820.0f + 2150.0f * fxm_pow(fxm_clamp(reaction_front * 0.74f + thermal * 0.38f, 0.0f, 1.0f), 0.72f),
            cinematic);
        const float edge_warmth = 1.0f - fxm_smoothstep(0.18f, 0.72f, reaction_front * thermal);
        const V3 standard_flame = mix3(blackbody(emission_temperature), v3(1.0f, 0.075f, 0.002f), edge_warmth * 0.68f);
        const V3 cinematic_amber =
            mix3(v3(1.0f, 0.055f, 0.002f), v3(1.0f, 0.43f, 0.018f), fxm_smoothstep(0.025f, 0.30f, flame_density));
        const V3 cinematic_flame = mix3(cinematic_amber, v3(1.0f, 0.91f, 0.60f),
                                        fxm_smoothstep(0.52f, 0.92f, flame_density) *
                                            fxm_smoothstep(980.0f, 2050.0f, temperature));
        const V3 flame_color = mix3(standard_flame, cinematic_flame, cinematic);
        const float cinematic_emission =
            flame_density * (1.35f + 3.65f * reaction_front) * (0.92f + 0.32f * r.chem_z) + hot_soot * 0.08f;
        const float standard_emission =
            flame_density * (1.20f + 1.80f * reaction_front) * (0.90f + 0.30f * r.chem_z) + hot_soot * 0.25f;
        const V3 emissive = mul(flame_color, fxm_mix(standard_emission, cinematic_emission, cinematic));
        const float base_zone = 1.0f - fxm_smoothstep(0.16f, 0.48f, position.y / r.ey);
        const V3 blue = mul(mul(mul(v3(0.018f, 0.12f, 0.74f), flame_density), 1.0f - richness),
                            fxm_mix(0.48f, base_zone * 0.34f, cinematic));
        const V3 to_fire = sub(v3(r.ex * 0.5f, 0.08f, r.ez * 0.5f), position);
        const float fire_falloff = 1.0f / (1.0f + dot3(to_fire, to_fire) * 8.0f);
        const V3 ambient_light = v3(0.018f, 0.022f, 0.028f);
        const V3 fire_light = mul(v3(1.0f, 0.24f, 0.035f), fire_falloff);
        const V3 standard_scatter = add(mul(v3(0.012f, 0.009f, 0.007f), soot_sigma), mul(v3(0.12f, 0.15f, 0.17f), vapor_sigma));
        const V3 cinematic_scatter = add(mul(add(mul(ambient_light, 0.12f), mul(fire_light, 0.055f)), soot_sigma),
                                         mul(add(mul(ambient_light, 0.55f), mul(fire_light, 0.05f)), vapor_sigma));
        const V3 scattered = mix3(standard_scatter, cinematic_scatter, cinematic);
        const float inv_ext = fxm_max(extinction, 0.0001f);
        const V3 source = add(add(emissive, blue), scattered);
        const V3 source_radiance = v3(source.x / inv_ext, source.y / inv_ext, source.z / inv_ext);
        radiance = add(radiance, mul(source_radiance, transmittance * alpha));
        transmittance *= 1.0f - alpha;
      } else {
        const float4 fv = diagnostic_sample(view_mode, flow, chem, media, r);
        const float alpha = 1.0f - fxm_exp(-fv.w * step_length * 8.5f);
        radiance = add(radiance, mul(mul(v3(fv.x, fv.y, fv.z), transmittance * alpha), 1.35f));
        transmittance *= 1.0f - alpha;
      }
      distance += step_length;
    }
    const V3 c = add(radiance, mul(background, transmittance));
    color = cinematic >= 0.5f ? c : aces(c, r);
  }
  Half4 o;
  o.x = (_Float16)fx_keep_f32(color.x);
  o.y = (_Float16)fx_keep_f32(color.y);
  o.z = (_Float16)fx_keep_f32(color.z);
  o.w = (_Float16)1.0f;
  out[(uint32_t)py * (uint32_t)r.width + (uint32_t)px] = o;
}

}  // namespace

namespace fx {

void launch_render_volume(const void* packed, const uint32_t* detail, void* out, const FxRender& r, hipStream_t s) {
  const Half4* vol = (const Half4*)packed;
  const uint64_t cells = (uint64_t)r.nx * r.ny * r.nz;
  dim3 grid((uint32_t)((r.width + 15) / 16), (uint32_t)((r.height + 15) / 16));
  const Half4 *f = vol, *c = vol + cells, *m = vol + 2 * cells;
  Half4* o = (Half4*)out;
  // u32(clamp(round(controls.x), 0.0, 7.0)), qualityShaders.ts:436
  switch ((uint32_t)fxm_clamp(fxm_rint(r.view_mode), 0.0f, 7.0f)) {
    case 0: k_render_volume<kRuntimeView><<<grid, 256, 0, s>>>(f, c, m, detail, o, r); break;
    case 1: k_render_volume<1><<<grid, 256, 0, s>>>(f, c, m, detail, o, r); break;
    case 2: k_render_volume<2><<<grid, 256, 0, s>>>(f, c, m, detail, o, r); break;
    case 3: k_render_volume<3><<<grid, 256, 0, s>>>(f, c, m, detail, o, r); break;
    case 4: k_render_volume<4><<<grid, 256, 0, s>>>(f, c, m, detail, o, r); break;
    case 5: k_render_volume<5><<<grid, 256, 0, s>>>(f, c, m, detail, o, r); break;
    case 6: k_render_volume<6><<<grid, 256, 0, s>>>(f, c, m, detail, o, r); break;
    default: k_render_volume<7><<<grid, 256, 0, s>>>(f, c, m, detail, o, r); break;
  }
}

}  // namespace fx
