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
  uint32_t brick;  // the occupancy brick holding texel (x0, y0, z0)
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
  s.brick = ((uint32_t)(z0 / kFxBrick) * (uint32_t)r.bny + (uint32_t)(y0 / kFxBrick)) * (uint32_t)r.bnx +
            (uint32_t)(x0 / kFxBrick);
  return s;
}

// Empty-space skipping by bricks.  bricks[b] is the largest of the four
// media components over the texels of brick b plus one texel beyond its
// upper faces, which holds all eight texels of any sample whose lower corner
// (x0, y0, z0) lies in the brick.  Each f32 mix a (1 - t) + b t of values
// <= M stays <= M (1 + u)^3 (u = 2^-24), so three levels of mixing stay
// below M (1 + 6e-7): a brick whose maximum is below kFxEmptyBrick
// guarantees an interpolated occupancy below 0.0005, which is when the
// beauty view's loop skips the sample, and the loop then runs exactly as if
// it had sampled.
#define kFxEmptyBrick 0.00049995f

__device__ inline float max4(float4 v) { return fxm_max(fxm_max(v.x, v.y), fxm_max(v.z, v.w)); }

__global__ __launch_bounds__(64) void k_render_bricks(const Half4* __restrict__ media, int nx, int ny, int nz,
                                                      int bnx, int bny, float* __restrict__ bricks) {
  const int b = (int)blockIdx.x;
  const int bx = b % bnx, by = (b / bnx) % bny, bz = b / (bnx * bny);
  const int x0 = bx * kFxBrick, y0 = by * kFxBrick, z0 = bz * kFxBrick;
  const int ex = min(x0 + kFxBrick, nx - 1) - x0 + 1, ey = min(y0 + kFxBrick, ny - 1) - y0 + 1;
  const int ez = min(z0 + kFxBrick, nz - 1) - z0 + 1;
  float m = -INFINITY;
  for (int t = (int)threadIdx.x; t < ex * ey * ez; t += 64) {
    const int x = x0 + t % ex, y = y0 + (t / ex) % ey, z = z0 + t / (ex * ey);
    m = fxm_max(m, max4(texel(media, (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)z))));
  }
  for (int o = 32; o > 0; o >>= 1) m = fxm_max(m, __shfl_xor(m, o));
  if (threadIdx.x == 0) bricks[b] = m;
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
  const float t = fxm_clamp((temperature - 720.0f) * (1.0f / 1900.0f), 0.0f, 1.0f);
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
                                                       const uint32_t* __restrict__ detail,
                                                       const float* __restrict__ bricks, Half4* __restrict__ out,
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
    // Divisions by the extents and by a sample's extinction are multiplies
    // by the f32 reciprocal (one of the results WGSL's 2.5 ULP f32 division
    // admits), as in simulate; the FP32 oracle makes the same ones.
    const V3 inv_extent = v3(1.0f / extent.x, 1.0f / extent.y, 1.0f / extent.z);
    for (uint32_t step = 0; step < 384u; ++step) {
      if (step >= step_count || distance >= far_t || transmittance < 0.006f) break;
      const V3 position = add(eye, mul(dir, distance));
      const Site site = sample_site(position, r);
      const bool skipping = beauty && cinematic >= 0.5f;
      float4 raw_media = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      bool empty = skipping && bricks[site.brick] < kFxEmptyBrick;
      if (!empty) {
        raw_media = sample_volume(media_vol, site);
        empty = skipping && max4(raw_media) < 0.0005f;
      }
      if (empty) {
        const float next_distance = distance + step_length;
        if (next_distance < far_t) {
          const Site next = sample_site(add(eye, mul(dir, next_distance)), r);
          bool next_empty = bricks[next.brick] < kFxEmptyBrick;
          if (!next_empty) next_empty = max4(sample_volume(media_vol, next)) < 0.0005f;
          distance += next_empty ? step_length * 2.0f : step_length;
        } else {
          distance = far_t;
        }
        continue;
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
          const V3 normalized = mulv(detail_position, inv_extent);
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
        const float base_zone = 1.0f - fxm_smoothstep(0.16f, 0.48f, position.y * inv_extent.y);
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
        const float rcp_ext = 1.0f / inv_ext;
        const V3 source_radiance = v3(source.x * rcp_ext, source.y * rcp_ext, source.z * rcp_ext);
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

// ---------------------------------------------------------- droplet streaks
// dropletVertex / dropletFragment (qualityShaders.ts:156-232): each live
// droplet draws a camera-facing quad, blended over the target in instance
// order.  Here the draw is a tiled compute rasterizer in four launches:
//   k_sprite_setup  one thread per droplet: the quad (FxSprite) and its
//                   conservative pixel box; counts the 128 x 128 bins it meets
//   k_sprite_scan   bin offsets (exclusive scan of the counts)
//   k_sprite_fill   one block per bin: the ids of the droplets meeting it, in
//                   instance order (a bin whose list would overrun the
//                   capacity is left empty and read as "every droplet")
//   k_sprite_draw   one block per 16 x 16 tile: walks its bin's list in order,
//                   keeps the droplets meeting the tile (ordered compaction in
//                   LDS) and blends each covered pixel in registers.
// Rasterization and blending are defined in f32 arithmetic: a pixel is
// covered when its centre p (the fullscreen pass's NDC of the pixel) has
// quad coordinates local = (dot(p - c, normal) / radius, dot(p - c, tangent)
// / half) strictly inside (-1, 1)^2 (local is the fragment's interpolated
// `local`); the blend (src-alpha, one-minus-src-alpha; alpha: one,
// one-minus-src-alpha) is src * a + dst * (1 - a) in f32, each result stored
// in the target's format (f16 nearest-even, or unorm8 rint(clamp * 255) with
// dst read as byte / 255).

struct Camera {
  V3 eye, fwd, right, up;
  float lens;
};

// cameraEye / projectWorld's basis (qualityShaders.ts:133-176): the same f32
// operations as the ray marcher's camera.
__device__ Camera make_camera(const FxRender& r) {
  const V3 target = v3(r.ex * 0.5f, r.ey * 0.40f, r.ez * 0.5f);
  const V3 orbit = v3(fxm_sin(r.yaw) * fxm_cos(r.pitch), fxm_sin(r.pitch), fxm_cos(r.yaw) * fxm_cos(r.pitch));
  const float frame_scale = fxm_max(r.ex, fxm_max(r.ey, r.ez));
  Camera c;
  c.eye = add(target, mul(mul(orbit, 0.90f + r.cam_distance * 0.18f), frame_scale));
  c.fwd = normalize3(sub(target, c.eye));
  c.right = normalize3(cross3(c.fwd, v3(0.0f, 1.0f, 0.0f)));
  c.up = normalize3(cross3(c.right, c.fwd));
  c.lens = 0.72f / fxm_clamp(r.zoom, 0.5f, 3.0f);
  return c;
}

__device__ inline void project_world(const Camera& c, V3 point, float aspect, float& x, float& y) {
  const V3 rel = sub(point, c.eye);
  const float depth = fxm_max(dot3(rel, c.fwd), 0.0001f);
  x = dot3(rel, c.right) / (depth * aspect * c.lens);
  y = dot3(rel, c.up) / (depth * c.lens);
}

// A pixel box packed as (x0 | x1 << 16, y0 | y1 << 16); x0 > x1 when empty.
__device__ inline bool box_meets(uint2 b, int x0, int y0, int x1, int y1) {
  const int bx0 = (int)(b.x & 0xffffu), bx1 = (int)(b.x >> 16), by0 = (int)(b.y & 0xffffu), by1 = (int)(b.y >> 16);
  return bx0 <= bx1 && bx0 <= x1 && bx1 >= x0 && by0 <= y1 && by1 >= y0;
}

__global__ __launch_bounds__(256) void k_sprite_setup(const float4* __restrict__ particles, int count, FxRender r,
                                                      FxSprite* __restrict__ sprites, uint2* __restrict__ boxes,
                                                      uint32_t* __restrict__ bin_counts, int bins_x) {
  const int i = (int)(blockIdx.x * 256u + threadIdx.x);
  if (i >= count) return;
  const float4 pa = particles[2 * i], vm = particles[2 * i + 1];
  const uint2 empty = make_uint2(1u, 0u);
  if (!(vm.w > 0.0001f && pa.x >= 0.0f && pa.y >= 0.0f && pa.z >= 0.0f && pa.x <= 1.0f && pa.y <= 1.0f &&
        pa.z <= 1.0f)) {
    boxes[i] = empty;
    return;
  }
  const Camera cam = make_camera(r);
  const V3 world = mulv(v3(pa.x, pa.y, pa.z), v3(r.ex, r.ey, r.ez));
  float hx, hy, tx, ty;
  project_world(cam, world, r.aspect, hx, hy);
  project_world(cam, sub(world, mul(v3(vm.x, vm.y, vm.z), 0.018f)), r.aspect, tx, ty);
  const float sx = tx - hx, sy = ty - hy;
  const float length = fxm_sqrt(sx * sx + sy * sy);
  const float streak = fxm_min(length, 0.032f);
  const bool along = streak > 0.00001f;
  FxSprite q;
  q.tx = along ? sx / length : 1.0f;
  q.ty = along ? sy / length : 0.0f;
  const float pixel_radius = r.nozzle >= 0.5f ? 1.15f : 1.65f;
  q.radius = pixel_radius * 2.0f / fxm_max(r.canvas_height, 1.0f);
  q.half = q.radius + streak * 0.5f;
  q.cx = hx;
  q.cy = hy;
  q.mass = vm.w;
  q.pad = 0.0f;
  sprites[i] = q;
  // conservative box: the quad's NDC half extents, one pixel of margin
  const float ex = fabsf(q.ty) * q.radius + fabsf(q.tx) * q.half;
  const float ey = fabsf(q.tx) * q.radius + fabsf(q.ty) * q.half;
  const float w = (float)r.width, h = (float)r.height;
  const float fx0 = floorf((q.cx - ex + 1.0f) * 0.5f * w - 0.5f) - 1.0f;
  const float fx1 = ceilf((q.cx + ex + 1.0f) * 0.5f * w - 0.5f) + 1.0f;
  const float fy0 = floorf((1.0f - (q.cy + ey)) * 0.5f * h - 0.5f) - 1.0f;
  const float fy1 = ceilf((1.0f - (q.cy - ey)) * 0.5f * h - 0.5f) + 1.0f;
  // clamp in f32 first (the quad may project far off screen); a box left of
  // or below the image ends as x1 = -1, one right of or above it as x0 = w
  const int x0 = (int)fminf(fmaxf(fx0, 0.0f), w), x1 = (int)fmaxf(fminf(fx1, w - 1.0f), -1.0f);
  const int y0 = (int)fminf(fmaxf(fy0, 0.0f), h), y1 = (int)fmaxf(fminf(fy1, h - 1.0f), -1.0f);
  if (!(fx0 <= fx1 && fy0 <= fy1) || x0 > x1 || y0 > y1) {
    boxes[i] = empty;
    return;
  }
  boxes[i] = make_uint2((uint32_t)x0 | ((uint32_t)x1 << 16), (uint32_t)y0 | ((uint32_t)y1 << 16));
  for (int by = y0 / kFxSpriteBin; by <= y1 / kFxSpriteBin; ++by)
    for (int bx = x0 / kFxSpriteBin; bx <= x1 / kFxSpriteBin; ++bx) atomicAdd(&bin_counts[by * bins_x + bx], 1u);
}

// Exclusive scan of the bin counts, one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_sprite_scan(const uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ offsets, int bins) {
  __shared__ uint32_t part[1024];
  const int t = (int)threadIdx.x;
  const int per = (bins + 1023) / 1024;
  const int b0 = t * per, b1 = min(b0 + per, bins);
  uint32_t sum = 0;
  for (int b = b0; b < b1; ++b) sum += counts[b];
  part[t] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const uint32_t v = t >= d ? part[t - d] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (int b = b0; b < b1; ++b) {
    offsets[b] = run;
    run += counts[b];
  }
}

// Ordered compaction over a 256-thread block: returns the slot of a hit among
// the block's hits (in thread order) and the block's hit count.
__device__ inline uint32_t block_compact(bool hit, uint32_t* wave_hits, uint32_t& total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t m = __ballot(hit);
  const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wave_hits[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t base = 0;
  total = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t c = wave_hits[k];
    base += k < wave ? c : 0u;
    total += c;
  }
  return base + below;
}

__global__ __launch_bounds__(256) void k_sprite_fill(const uint2* __restrict__ boxes, int count,
                                                     const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ offsets, uint32_t* __restrict__ list,
                                                     uint32_t capacity, int bins_x) {
  __shared__ uint32_t wave_hits[4];
  const int bin = (int)blockIdx.x;
  const uint32_t n = counts[bin], off = offsets[bin];
  if (n == 0 || (uint64_t)off + n > capacity) return;
  const int x0 = (bin % bins_x) * kFxSpriteBin, y0 = (bin / bins_x) * kFxSpriteBin;
  uint32_t run = 0;
  for (int base = 0; base < count; base += 256) {
    const int i = base + (int)threadIdx.x;
    const bool hit = i < count && box_meets(boxes[i], x0, y0, x0 + kFxSpriteBin - 1, y0 + kFxSpriteBin - 1);
    uint32_t total;
    const uint32_t slot = block_compact(hit, wave_hits, total);
    if (hit) list[off + run + slot] = (uint32_t)i;
    run += total;
    __syncthreads();
  }
}

// dropletFragment at quad coordinates (lx, ly): false when discarded.
__device__ inline bool droplet_fragment(float lx, float ly, float mass, float boost, bool linear, const FxRender& r,
                                        V3& color, float& alpha) {
  const float m = fxm_max(fabsf(ly) - 0.48f, 0.0f);
  const float d = fxm_sqrt(lx * lx + m * m);
  const float coverage = 1.0f - fxm_smoothstep(0.64f, 1.0f, d);
  if (coverage <= 0.001f) return false;
  const float highlight = fxm_pow(fxm_clamp(1.0f - fabsf(lx + 0.30f), 0.0f, 1.0f), 5.0f);
  const V3 base = mix3(v3(0.27f, 0.34f, 0.38f), v3(0.76f, 0.86f, 0.91f), highlight * 0.62f);
  alpha = coverage * fxm_clamp(0.13f + mass * 0.15f, 0.12f, 0.42f) * boost;
  const V3 lin = mul(base, 0.72f);
  color = linear ? lin : aces(lin, r);
  return true;
}

__device__ inline float blend_f16(float src, float dst, float keep_a) {
  return (float)(_Float16)fx_keep_f32(src + dst * keep_a);
}

__device__ inline float blend_unorm8(float src, float dst, float keep_a) {
  return fxm_rint(fxm_clamp(src + dst * keep_a, 0.0f, 1.0f) * 255.0f);
}

template <bool UNORM8>
__global__ __launch_bounds__(256) void k_sprite_draw(const FxSprite* __restrict__ sprites, const uint2* __restrict__ boxes,
                                                     const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ list, uint32_t capacity, int count,
                                                     int bins_x, void* __restrict__ target, FxRender r) {
  __shared__ FxSprite tile_sprites[256];
  __shared__ uint32_t wave_hits[4];
  const int tx0 = (int)blockIdx.x * kFxSpriteTile, ty0 = (int)blockIdx.y * kFxSpriteTile;
  const int bin = (ty0 / kFxSpriteBin) * bins_x + tx0 / kFxSpriteBin;
  const uint32_t n = counts[bin];
  if (n == 0) return;
  const uint32_t off = offsets[bin];
  const bool every = (uint64_t)off + n > capacity;  // the bin's list was not written
  const int total = every ? count : (int)n;
  const int px = tx0 + (int)(threadIdx.x & 15u), py = ty0 + (int)(threadIdx.x >> 4);
  const bool on = px < r.width && py < r.height;
  const int tx1 = tx0 + kFxSpriteTile - 1, ty1 = ty0 + kFxSpriteTile - 1;
  // the pixel centre in NDC, as the ray marcher computes it
  const float uvx = ((float)px + 0.5f) / (float)r.width;
  const float uvy = 1.0f - ((float)py + 0.5f) / (float)r.height;
  const float sx = uvx * 2.0f - 1.0f, sy = uvy * 2.0f - 1.0f;
  const uint32_t view = (uint32_t)fxm_clamp(fxm_rint(r.view_mode), 0.0f, 7.0f);
  const float boost = view == 6u ? 1.38f : 1.0f;
  const bool linear = r.cinematic >= 0.5f;
  const size_t pixel = (size_t)(uint32_t)py * (uint32_t)r.width + (uint32_t)px;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f;
  if (on) {
    if (UNORM8) {
      const uint32_t b = ((const uint32_t*)target)[pixel];
      c0 = (float)(b & 255u);
      c1 = (float)((b >> 8) & 255u);
      c2 = (float)((b >> 16) & 255u);
      c3 = (float)(b >> 24);
    } else {
      const Half4 h = ((const Half4*)target)[pixel];
      c0 = (float)h.x;
      c1 = (float)h.y;
      c2 = (float)h.z;
      c3 = (float)h.w;
    }
  }
  bool dirty = false;
  for (int base = 0; base < total; base += 256) {
    const int k = base + (int)threadIdx.x;
    uint32_t id = 0;
    bool hit = false;
    if (k < total) {
      id = every ? (uint32_t)k : list[off + (uint32_t)k];
      hit = box_meets(boxes[id], tx0, ty0, tx1, ty1);
    }
    uint32_t found;
    const uint32_t slot = block_compact(hit, wave_hits, found);
    if (hit) tile_sprites[slot] = sprites[id];
    __syncthreads();
    if (on) {
      for (uint32_t j = 0; j < found; ++j) {
        const FxSprite q = tile_sprites[j];
        const float dx = sx - q.cx, dy = sy - q.cy;
        const float lx = (dx * -q.ty + dy * q.tx) / q.radius, ly = (dx * q.tx + dy * q.ty) / q.half;
        if (!(fabsf(lx) < 1.0f && fabsf(ly) < 1.0f)) continue;
        V3 src;
        float a;
        if (!droplet_fragment(lx, ly, q.mass, boost, linear, r, src, a)) continue;
        const float keep = 1.0f - a;
        if (UNORM8) {
          c0 = blend_unorm8(src.x * a, c0 / 255.0f, keep);
          c1 = blend_unorm8(src.y * a, c1 / 255.0f, keep);
          c2 = blend_unorm8(src.z * a, c2 / 255.0f, keep);
          c3 = blend_unorm8(a, c3 / 255.0f, keep);
        } else {
          c0 = blend_f16(src.x * a, c0, keep);
          c1 = blend_f16(src.y * a, c1, keep);
          c2 = blend_f16(src.z * a, c2, keep);
          c3 = blend_f16(a, c3, keep);
        }
        dirty = true;
      }
    }
    __syncthreads();
  }
  if (!dirty) return;
  if (UNORM8) {
    ((uint32_t*)target)[pixel] = (uint32_t)c0 | ((uint32_t)c1 << 8) | ((uint32_t)c2 << 16) | ((uint32_t)c3 << 24);
  } else {
    Half4 o;
    o.x = (_Float16)c0;
    o.y = (_Float16)c1;
    o.z = (_Float16)c2;
    o.w = (_Float16)c3;
    ((Half4*)target)[pixel] = o;
  }
}

// ------------------------------------------------------------- post passes
// FIREX_POST_WGSL (qualityShaders.ts:580-671): the quarter-resolution bloom
// and the tone-map / composite of the cinematic presentation, and the 8-bit
// store of the canvas.  A post pass pixel (x, y) of a w x h target reads uv =
// ((x + 0.5) / w, (y + 0.5) / h) (the pass flips y once, so v = 0 is the
// scene image's top row); 2D textures are sampled like the volumes: t = uv *
// size - 0.5, the four texels clamped to the edge, mixed along x, then y.

__device__ inline float4 image_texel(const Half4* __restrict__ img, int x, int y, int w) {
  const Half4 h = img[(uint32_t)y * (uint32_t)w + (uint32_t)x];
  return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
}

__device__ float4 sample_image(const Half4* __restrict__ img, int w, int h, float u, float v) {
  const float tx = u * (float)w - 0.5f, ty = v * (float)h - 0.5f;
  const float flx = fxm_floor(tx), fly = fxm_floor(ty);
  const float fx = tx - flx, fy = ty - fly;
  const int ix = (int)flx, iy = (int)fly;
  const int x0 = min(max(ix, 0), w - 1), x1 = min(max(ix + 1, 0), w - 1);
  const int y0 = min(max(iy, 0), h - 1), y1 = min(max(iy + 1, 0), h - 1);
  return mix4(mix4(image_texel(img, x0, y0, w), image_texel(img, x1, y0, w), fx),
              mix4(image_texel(img, x0, y1, w), image_texel(img, x1, y1, w), fx), fy);
}

__device__ inline V3 rgb(float4 c) { return v3(c.x, c.y, c.z); }

// brightExtract, qualityShaders.ts:621-628
__device__ V3 bright_extract(V3 c) {
  const float luminance = dot3(c, v3(0.2126f, 0.7152f, 0.0722f));
  const float threshold = 1.15f, knee = 0.35f;
  const float soft = fxm_clamp((luminance - threshold + knee) / (2.0f * knee), 0.0f, 1.0f);
  const float contribution = fxm_max(luminance - threshold, 0.0f) + soft * soft * knee;
  const float d = fxm_max(luminance, 0.0001f);
  return v3(c.x * contribution / d, c.y * contribution / d, c.z * contribution / d);
}

// bloomFragment, qualityShaders.ts:630-644: nine taps of the scene image.
__global__ __launch_bounds__(256) void k_bloom(const Half4* __restrict__ scene, int w, int h,
                                               Half4* __restrict__ bloom, int bw, int bh) {
  const int x = (int)(blockIdx.x * 16 + (threadIdx.x & 15u)), y = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
  if (x >= bw || y >= bh) return;
  const float u = ((float)x + 0.5f) / (float)bw, v = ((float)y + 0.5f) / (float)bh;
  const float rx = 1.0f / fxm_max((float)w, 1.0f) * 5.5f, ry = 1.0f / fxm_max((float)h, 1.0f) * 5.5f;
  V3 c = mul(rgb(sample_image(scene, w, h, u, v)), 0.20f);
  c = add(c, mul(rgb(sample_image(scene, w, h, u + rx, v + 0.0f)), 0.10f));
  c = add(c, mul(rgb(sample_image(scene, w, h, u - rx, v - 0.0f)), 0.10f));
  c = add(c, mul(rgb(sample_image(scene, w, h, u + 0.0f, v + ry)), 0.10f));
  c = add(c, mul(rgb(sample_image(scene, w, h, u - 0.0f, v - ry)), 0.10f));
  c = add(c, mul(rgb(sample_image(scene, w, h, u + rx, v + ry)), 0.10f));
  c = add(c, mul(rgb(sample_image(scene, w, h, u - rx, v - ry)), 0.10f));
  c = add(c, mul(rgb(sample_image(scene, w, h, u + rx, v + -ry)), 0.10f));
  c = add(c, mul(rgb(sample_image(scene, w, h, u + -rx, v + ry)), 0.10f));
  const V3 b = bright_extract(c);
  Half4 o;
  o.x = (_Float16)fx_keep_f32(b.x);
  o.y = (_Float16)fx_keep_f32(b.y);
  o.z = (_Float16)fx_keep_f32(b.z);
  o.w = (_Float16)1.0f;
  bloom[(uint32_t)y * (uint32_t)bw + (uint32_t)x] = o;
}

// filmic, qualityShaders.ts:654-662
__device__ inline float filmic1(float c, float e) {
  const float x = fxm_max(c * e, 0.0f);
  return fxm_pow(fxm_clamp((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f), kFxInverseGamma);
}

// The 8-bit store of a colour target: round(clamp(c, 0, 1) * 255), ties to even.
__device__ inline uint32_t unorm8(float c) { return (uint32_t)fxm_rint(fxm_clamp(c, 0.0f, 1.0f) * 255.0f); }

// compositeFragment, qualityShaders.ts:664-671 (with dither, :659-662), into
// rgba8 pixels; cinematic = 0 stores the scene image as the standard path's
// canvas receives it.
__global__ __launch_bounds__(256) void k_composite(const Half4* __restrict__ scene, const Half4* __restrict__ bloom,
                                                   int w, int h, int bw, int bh, float exposure, int cinematic,
                                                   uint32_t* __restrict__ frame) {
  const int x = (int)(blockIdx.x * 16 + (threadIdx.x & 15u)), y = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
  if (x >= w || y >= h) return;
  V3 c;
  if (cinematic) {
    const float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
    const V3 sc = rgb(sample_image(scene, w, h, u, v));
    const V3 bl = rgb(sample_image(bloom, bw, bh, u, v));
    const V3 hdr = add(sc, mul(bl, 0.10f));
    const float e = fxm_max(exposure, 0.25f);
    c = v3(filmic1(hdr.x, e), filmic1(hdr.y, e), filmic1(hdr.z, e));
    const float px = fxm_fract(((float)x + 0.5f) * 0.06711056f), py = fxm_fract(((float)y + 0.5f) * 0.00583715f);
    const float d = (fxm_fract(52.9829189f * (px + py)) - 0.5f) / 255.0f;
    c = v3(fxm_clamp(c.x + d, 0.0f, 1.0f), fxm_clamp(c.y + d, 0.0f, 1.0f), fxm_clamp(c.z + d, 0.0f, 1.0f));
  } else {
    c = rgb(image_texel(scene, x, y, w));
  }
  frame[(uint32_t)y * (uint32_t)w + (uint32_t)x] = unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (255u << 24);
}

}  // namespace

namespace fx {

void launch_render_bricks(const void* packed, float* bricks, const FxRender& r, hipStream_t s) {
  const uint64_t cells = (uint64_t)r.nx * r.ny * r.nz;
  const Half4* media = (const Half4*)packed + 2 * cells;
  k_render_bricks<<<(uint32_t)(r.bnx * r.bny * r.bnz), 64, 0, s>>>(media, r.nx, r.ny, r.nz, r.bnx, r.bny, bricks);
}

void launch_render_volume(const void* packed, const uint32_t* detail, const float* bricks, void* out,
                          const FxRender& r, hipStream_t s) {
  const Half4* vol = (const Half4*)packed;
  const uint64_t cells = (uint64_t)r.nx * r.ny * r.nz;
  dim3 grid((uint32_t)((r.width + 15) / 16), (uint32_t)((r.height + 15) / 16));
  const Half4 *f = vol, *c = vol + cells, *m = vol + 2 * cells;
  Half4* o = (Half4*)out;
  // u32(clamp(round(controls.x), 0.0, 7.0)), qualityShaders.ts:436
  switch ((uint32_t)fxm_clamp(fxm_rint(r.view_mode), 0.0f, 7.0f)) {
    case 0: k_render_volume<kRuntimeView><<<grid, 256, 0, s>>>(f, c, m, detail, bricks, o, r); break;
    case 1: k_render_volume<1><<<grid, 256, 0, s>>>(f, c, m, detail, bricks, o, r); break;
    case 2: k_render_volume<2><<<grid, 256, 0, s>>>(f, c, m, detail, bricks, o, r); break;
    case 3: k_render_volume<3><<<grid, 256, 0, s>>>(f, c, m, detail, bricks, o, r); break;
    case 4: k_render_volume<4><<<grid, 256, 0, s>>>(f, c, m, detail, bricks, o, r); break;
    case 5: k_render_volume<5><<<grid, 256, 0, s>>>(f, c, m, detail, bricks, o, r); break;
    case 6: k_render_volume<6><<<grid, 256, 0, s>>>(f, c, m, detail, bricks, o, r); break;
    default: k_render_volume<7><<<grid, 256, 0, s>>>(f, c, m, detail, bricks, o, r); break;
  }
}

uint64_t sprite_work_bytes(int count) {
  const uint64_t n = count > 0 ? (uint64_t)count : 0;
  const uint64_t capacity = 8 * n + 4096;
  auto up = [](uint64_t b) { return (b + 255) / 256 * 256; };
  return up(n * sizeof(FxSprite)) + up(n * 8) + 2 * up((uint64_t)kFxSpriteMaxBins * 4) + up(capacity * 4);
}

void launch_droplet_streaks(const void* particles, int count, const FxRender& r, void* target, int unorm8,
                            void* work, hipStream_t s) {
  const uint32_t view = (uint32_t)fxm_clamp(fxm_rint(r.view_mode), 0.0f, 7.0f);
  if (count <= 0 || (view != 0u && view != 6u)) return;  // dropletFragment discards every fragment
  const uint64_t n = (uint64_t)count, capacity = 8 * n + 4096;
  auto up = [](uint64_t b) { return (b + 255) / 256 * 256; };
  uint8_t* p = (uint8_t*)work;
  FxSprite* sprites = (FxSprite*)p;
  p += up(n * sizeof(FxSprite));
  uint2* boxes = (uint2*)p;
  p += up(n * 8);
  uint32_t* counts = (uint32_t*)p;
  p += up((uint64_t)kFxSpriteMaxBins * 4);
  uint32_t* offsets = (uint32_t*)p;
  p += up((uint64_t)kFxSpriteMaxBins * 4);
  uint32_t* list = (uint32_t*)p;
  const int bins_x = (r.width + kFxSpriteBin - 1) / kFxSpriteBin, bins_y = (r.height + kFxSpriteBin - 1) / kFxSpriteBin;
  const int bins = bins_x * bins_y;
  launch_fill_u32(counts, (uint64_t)bins, 0u, s);
  k_sprite_setup<<<(uint32_t)((count + 255) / 256), 256, 0, s>>>((const float4*)particles, count, r, sprites, boxes,
                                                                  counts, bins_x);
  k_sprite_scan<<<1, 1024, 0, s>>>(counts, offsets, bins);
  k_sprite_fill<<<(uint32_t)bins, 256, 0, s>>>(boxes, count, counts, offsets, list, (uint32_t)capacity, bins_x);
  dim3 tiles((uint32_t)((r.width + kFxSpriteTile - 1) / kFxSpriteTile),
             (uint32_t)((r.height + kFxSpriteTile - 1) / kFxSpriteTile));
  if (unorm8)
    k_sprite_draw<true><<<tiles, 256, 0, s>>>(sprites, boxes, counts, offsets, list, (uint32_t)capacity, count, bins_x,
                                              target, r);
  else
    k_sprite_draw<false><<<tiles, 256, 0, s>>>(sprites, boxes, counts, offsets, list, (uint32_t)capacity, count,
                                               bins_x, target, r);
}

void launch_render_post(const void* scene, void* bloom, int width, int height, float exposure, int cinematic,
                        uint32_t* frame, hipStream_t s) {
  const int bw = (width + 3) / 4 > 1 ? (width + 3) / 4 : 1, bh = (height + 3) / 4 > 1 ? (height + 3) / 4 : 1;
  if (cinematic) {
    dim3 g((uint32_t)((bw + 15) / 16), (uint32_t)((bh + 15) / 16));
    k_bloom<<<g, 256, 0, s>>>((const Half4*)scene, width, height, (Half4*)bloom, bw, bh);
  }
  dim3 g((uint32_t)((width + 15) / 16), (uint32_t)((height + 15) / 16));
  k_composite<<<g, 256, 0, s>>>((const Half4*)scene, (const Half4*)bloom, width, height, bw, bh, exposure, cinematic,
                                frame);
}

}  // namespace fx
