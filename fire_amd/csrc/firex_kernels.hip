// firex_kernels.hip — gfx950 kernels of the Fire-X fixed step.
//
// One pass of the reference (src/sim/firex/shaders.ts) maps to one kernel
// here, except that the FUSED plan adds kernels that merge passes (see
// firex_jacobi.hip).  Every per-cell expression keeps the reference's f32
// operation order, so each kernel is bit-identical to the FP32 oracle; the
// uniform-only subexpressions arrive precomputed in FxStep.
//
// Data layout in HBM (index = x + nx * (y + ny * z), x fastest):
//   flow     float4 (u, v, w, T)             16 B/cell, two buffers
//   species  float4 | 4 x f16 | 4 x bf16     16 or 8 B/cell, two buffers
//   pressure float | f16                     4 or 2 B/cell, two primary + two residual
//   divergence, reaction float; liquid u32   4 B/cell each
#include <hip/hip_runtime.h>

#include "fx_common.h"
#include "firex_kernels.h"

namespace fx {

// ------------------------------------------------------------------ storage
template <int S> struct SpeciesIO;

// 16 bytes at an 8-byte aligned address (two consecutive 8-byte words).
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef u32x4_t u32x4_a8_t __attribute__((aligned(8)));

// Each storage also loads two consecutive words at once (load_pair: the x
// pair of a trilinear corner, one 16-byte load for the 8-byte storages).
template <> struct SpeciesIO<FX_STORAGE_F32> {
  using Word = float4;
  using Bits = float4;
  __device__ static inline float4 load(const Word* p, uint32_t i) { return p[i]; }
  __device__ static inline void store(Word* p, uint32_t i, float4 v) { p[i] = v; }
  __device__ static inline Bits load_bits(const Word* p, uint32_t i) { return p[i]; }
  __device__ static inline float4 unpack(Bits b) { return b; }
  __device__ static inline void load_pair(const Word* p, uint32_t i, float4& a, float4& b) {
    a = p[i];
    b = p[i + 1];
  }
};

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));

template <> struct SpeciesIO<FX_STORAGE_F16> {
  using Word = half4_t;
  using Bits = uint2;
  __device__ static inline float4 unpack(Bits b) {
    half4_t h = __builtin_bit_cast(half4_t, b);
    return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
  }
  __device__ static inline Bits load_bits(const Word* p, uint32_t i) { return __builtin_bit_cast(uint2, p[i]); }
  __device__ static inline float4 load(const Word* p, uint32_t i) { return unpack(load_bits(p, i)); }
  __device__ static inline void load_pair(const Word* p, uint32_t i, float4& a, float4& b) {
    const u32x4_t v = *reinterpret_cast<const u32x4_a8_t*>(p + i);
    a = unpack(make_uint2(v.x, v.y));
    b = unpack(make_uint2(v.z, v.w));
  }
  __device__ static inline void store(Word* p, uint32_t i, float4 v) {
    half4_t h;
    h.x = (_Float16)fx_keep_f32(v.x);
    h.y = (_Float16)fx_keep_f32(v.y);
    h.z = (_Float16)fx_keep_f32(v.z);
    h.w = (_Float16)fx_keep_f32(v.w);
    p[i] = h;
  }
};

template <> struct SpeciesIO<FX_STORAGE_BF16> {
  using Word = ushort4;
  using Bits = uint2;
  __device__ static inline float4 unpack(Bits w) {
    return make_float4(fxm_bits_to_float(w.x << 16), fxm_bits_to_float(w.x & 0xffff0000u),
                       fxm_bits_to_float(w.y << 16), fxm_bits_to_float(w.y & 0xffff0000u));
  }
  __device__ static inline Bits load_bits(const Word* p, uint32_t i) { return __builtin_bit_cast(uint2, p[i]); }
  __device__ static inline float4 load(const Word* p, uint32_t i) { return unpack(load_bits(p, i)); }
  __device__ static inline void load_pair(const Word* p, uint32_t i, float4& a, float4& b) {
    const u32x4_t v = *reinterpret_cast<const u32x4_a8_t*>(p + i);
    a = unpack(make_uint2(v.x, v.y));
    b = unpack(make_uint2(v.z, v.w));
  }
  __device__ static inline void store(Word* p, uint32_t i, float4 v) {
    ushort4 b;
    b.x = fx_float_to_bf16_bits(v.x); b.y = fx_float_to_bf16_bits(v.y);
    b.z = fx_float_to_bf16_bits(v.z); b.w = fx_float_to_bf16_bits(v.w);
    p[i] = b;
  }
};

template <int P> struct PressureIO;
template <> struct PressureIO<0> {
  using Word = float;
  __device__ static inline float load(const Word* p, uint32_t i) { return p[i]; }
  __device__ static inline void store(Word* p, uint32_t i, float v) { p[i] = v; }
};
template <> struct PressureIO<1> {
  using Word = _Float16;
  __device__ static inline float load(const Word* p, uint32_t i) { return (float)p[i]; }
  __device__ static inline void store(Word* p, uint32_t i, float v) { p[i] = (_Float16)fx_keep_f32(v); }
};

// ------------------------------------------------------------ cell mapping
// Cell kernels run 64 x 4 x 1 blocks (one wave per x row of 64 cells).  The
// hardware hands consecutive workgroups to the 8 XCDs in turn; the block id
// is remapped so that each XCD owns one contiguous range of blocks, i.e. one
// z slab, and the slab's halo planes stay in that XCD's L2.
constexpr int kBX = 64, kBY = 4;

__device__ inline uint32_t xcd_remap(uint32_t hw, uint32_t total) {
  uint32_t q = total >> 3, r = total & 7u;
  uint32_t xcd = hw & 7u, slot = hw >> 3;
  uint32_t base = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return base + slot;
}

template <int BY = kBY>
__device__ inline bool cell_coords(const FxStep& st, int& x, int& y, int& z) {
  uint32_t bx = (uint32_t)(st.nx + kBX - 1) / kBX;
  uint32_t by = (uint32_t)(st.ny + BY - 1) / BY;
  uint32_t total = bx * by * (uint32_t)st.nz;
  uint32_t b = xcd_remap(blockIdx.x, total);
  uint32_t ix = b % bx, rest = b / bx;
  uint32_t iy = rest % by, iz = rest / by;
  x = (int)(ix * kBX + threadIdx.x);
  y = (int)(iy * BY + threadIdx.y);
  z = (int)iz;
  return x < st.nx && y < st.ny;
}

__device__ inline uint32_t cidx(const FxStep& st, int x, int y, int z) {
  x = min(max(x, 0), st.nx - 1);
  y = min(max(y, 0), st.ny - 1);
  z = min(max(z, 0), st.nz - 1);
  return (uint32_t)x + (uint32_t)st.nx * ((uint32_t)y + (uint32_t)st.ny * (uint32_t)z);
}

__device__ inline float4 mix4(float4 a, float4 b, float t) {
  return make_float4(fxm_mix(a.x, b.x, t), fxm_mix(a.y, b.y, t), fxm_mix(a.z, b.z, t), fxm_mix(a.w, b.w, t));
}

// The sample point of trilinearFlow / trilinearSpecies (shaders.ts:61-95) is
// clamped to [0, dims - 1.001] per axis, so its floor lies in [0, dims - 2]
// (every grid has at least two cells per axis) and no corner index needs the
// edge clamp: the corners are base, base + 1 (x), + nx (y), + nx ny (z).
// 24-bit multiplies suffice (coordinates and grid sizes are at most 1024).
struct TriCorners {
  uint32_t i000, sy, sz;
  float fx, fy, fz;
};
__device__ inline TriCorners tri_corners(const FxStep& st, float px, float py, float pz) {
  float sx = fxm_clamp(px - 0.5f, 0.0f, st.gx - 1.001f);
  float sy = fxm_clamp(py - 0.5f, 0.0f, st.gy - 1.001f);
  float sz = fxm_clamp(pz - 0.5f, 0.0f, st.gz - 1.001f);
  float flx = fxm_floor(sx), fly = fxm_floor(sy), flz = fxm_floor(sz);
  TriCorners c;
  const uint32_t bx = (uint32_t)(int)flx, by = (uint32_t)(int)fly, bz = (uint32_t)(int)flz;
  c.sy = (uint32_t)st.nx;
  c.sz = __umul24((uint32_t)st.nx, (uint32_t)st.ny);
  c.i000 = bx + __umul24(by, c.sy) + __umul24(bz, c.sz);
  c.fx = sx - flx;
  c.fy = sy - fly;
  c.fz = sz - flz;
  return c;
}
template <typename Pair>
__device__ inline float4 trilinear_pairs(const TriCorners& c, Pair pair) {
  float4 c000, c100, c010, c110, c001, c101, c011, c111;
  pair(c.i000, c000, c100);
  pair(c.i000 + c.sy, c010, c110);
  pair(c.i000 + c.sz, c001, c101);
  pair(c.i000 + c.sz + c.sy, c011, c111);
  float4 z0 = mix4(mix4(c000, c100, c.fx), mix4(c010, c110, c.fx), c.fy);
  float4 z1 = mix4(mix4(c001, c101, c.fx), mix4(c011, c111, c.fx), c.fy);
  return mix4(z0, z1, c.fz);
}

template <typename Load>
__device__ inline float4 trilinear(const FxStep& st, float px, float py, float pz, Load load) {
  float sx = fxm_clamp(px - 0.5f, 0.0f, st.gx - 1.001f);
  float sy = fxm_clamp(py - 0.5f, 0.0f, st.gy - 1.001f);
  float sz = fxm_clamp(pz - 0.5f, 0.0f, st.gz - 1.001f);
  float flx = fxm_floor(sx), fly = fxm_floor(sy), flz = fxm_floor(sz);
  int bx = (int)flx, by = (int)fly, bz = (int)flz;
  float fx = sx - flx, fy = sy - fly, fz = sz - flz;
  float4 c000 = load(cidx(st, bx, by, bz));
  float4 c100 = load(cidx(st, bx + 1, by, bz));
  float4 c010 = load(cidx(st, bx, by + 1, bz));
  float4 c110 = load(cidx(st, bx + 1, by + 1, bz));
  float4 c001 = load(cidx(st, bx, by, bz + 1));
  float4 c101 = load(cidx(st, bx + 1, by, bz + 1));
  float4 c011 = load(cidx(st, bx, by + 1, bz + 1));
  float4 c111 = load(cidx(st, bx + 1, by + 1, bz + 1));
  float4 z0 = mix4(mix4(c000, c100, fx), mix4(c010, c110, fx), fy);
  float4 z1 = mix4(mix4(c001, c101, fx), mix4(c011, c111, fx), fy);
  return mix4(z0, z1, fz);
}

// ------------------------------------------------------- coordinate tables
// Every term of simulate that depends on at most two cell coordinates is
// evaluated once per step by k_sim_tables, with exactly the f32 operations
// the per-cell code would use (so the results are the same bits), and read
// back per cell: the turbulence sines and cosines, the height envelope, the
// outflow terms, the swirl, and the reference emitter's horizontal factors.
// Layout (floats, SimTab offsets): xz4 [nx*nz] float4 (emitter radial,
// texture, pulse, 0); y4 [ny] float4 (height envelope, 1 - outflow, outflow
// decay, outflow mix); x [nx] swz; yv [ny] emitter vertical; z [nz] swx;
// yz [ny*nz] ltx; xz [nx*nz] lty; xy [nx*ny] ltz.
struct SimTab {
  uint32_t xz4, y4, x, yv, z, yz, xz, xy, total;
  __host__ __device__ SimTab(int nx, int ny, int nz) {
    const uint32_t a = (uint32_t)nx, b = (uint32_t)ny, c = (uint32_t)nz;
    xz4 = 0;
    y4 = 4 * a * c;
    x = y4 + 4 * b;
    yv = x + a;
    z = yv + b;
    yz = z + c;
    xz = yz + b * c;
    xy = xz + a * c;
    total = xy + a * b;
  }
};

// normalized cell centres, (i + 0.5) / grid (shaders.ts:153-156)
__device__ inline float np_of(int i, float g) { return ((float)i + 0.5f) / g; }

// emitter(), shaders.ts:101-133, split into its (x, z) factors and its y
// factor; the per-cell value is exp(-(radial + vertical)) * texture * pulse.
__device__ inline float4 emitter_xz(const FxStep& st, float npx, float npz) {
  float px = npx * st.ex, pz = npz * st.ez;
  float ox = px - st.em_cx, oz = pz - st.em_cz;
  float radial = ox * ox / st.em_rx2 + oz * oz / st.em_rz2;
  float cellular = 0.76f + 0.24f * fxm_sin(px * 49.0f + st.et52) * fxm_sin(pz * 43.0f - st.et44);
  float lobeArg = fxm_sin(ox / st.em_rx_safe * 6.2f + st.et17) * fxm_sin(oz / st.em_rz_safe * 5.4f - st.et13);
  float lobes = fxm_mix(0.48f, 1.18f, fxm_smoothstep(-0.42f, 0.48f, lobeArg));
  float texture = fxm_mix(cellular, cellular * lobes, st.em_cinematic);
  float pulse = 0.84f + 0.16f * fxm_sin(st.et37 + px * 27.0f + pz * 21.0f);
  return make_float4(radial, texture, pulse, 0.0f);
}
__device__ inline float emitter_vertical(const FxStep& st, float npy) {
  float oy = npy * st.ey - st.em_cy;
  return oy * oy / st.em_vdenom;
}

__global__ __launch_bounds__(256) void k_sim_tables(float* __restrict__ tab, FxStep st, int reference_emitter) {
  const int nx = st.nx, ny = st.ny, nz = st.nz;
  const SimTab o(nx, ny, nz);
  const uint32_t n_xz = (uint32_t)nx * nz, n_xy = (uint32_t)nx * ny, n_yz = (uint32_t)ny * nz;
  const uint32_t total = n_xz + n_xy + n_yz + (uint32_t)(nx + ny + nz);
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < total; i += gridDim.x * 256u) {
    uint32_t j = i;
    if (j < n_xz) {  // (x, z): lty and the reference emitter's horizontal factors
      const float npx = np_of((int)(j % nx), st.gx), npz = np_of((int)(j / nx), st.gz);
      tab[o.xz + j] = 0.12f * fxm_sin(npx * 29.0f + npz * 23.0f - st.tt41);
      if (reference_emitter) reinterpret_cast<float4*>(tab + o.xz4)[j] = emitter_xz(st, npx, npz);
      continue;
    }
    j -= n_xz;
    if (j < n_xy) {  // (x, y): ltz
      const float npx = np_of((int)(j % nx), st.gx), npy = np_of((int)(j / nx), st.gy);
      tab[o.xy + j] = fxm_cos(npy * 27.0f - npx * 21.0f + st.tt27) + 0.45f * fxm_cos(npy * 59.0f + st.tt47);
      continue;
    }
    j -= n_xy;
    if (j < n_yz) {  // (y, z): ltx
      const float npy = np_of((int)(j % ny), st.gy), npz = np_of((int)(j / ny), st.gz);
      tab[o.yz + j] = fxm_sin(npy * 31.0f + npz * 19.0f + st.tt31) + 0.45f * fxm_sin(npy * 67.0f - st.tt53);
      continue;
    }
    j -= n_yz;
    if (j < (uint32_t)nx) {  // x: swirl z-factor
      tab[o.x + j] = fxm_cos(st.swirl_tx + np_of((int)j, st.gx) * 25.0f);
      continue;
    }
    j -= nx;
    if (j < (uint32_t)ny) {  // y: height envelope, outflow terms, emitter vertical
      const float npy = np_of((int)j, st.gy);
      const float envelope = fxm_smoothstep(0.045f, 0.15f, npy) * (1.0f - fxm_smoothstep(0.80f, 1.0f, npy));
      const float outflow = fxm_smoothstep(0.84f, 0.995f, npy);
      reinterpret_cast<float4*>(tab + o.y4)[j] =
          make_float4(envelope, 1.0f - outflow, fxm_exp(-outflow * st.dt * 15.0f / st.persistence),
                      outflow * 0.16f / st.persistence);
      tab[o.yv + j] = emitter_vertical(st, npy);
      continue;
    }
    j -= ny;  // z: swirl x-factor
    tab[o.z + j] = fxm_sin(st.swirl_tz + np_of((int)j, st.gz) * 22.0f);
  }
}

// A term whose exponent is below -86 adds exactly nothing: fxm_exp returns +0
// there, amplitude * +0 is a zero and s + that zero is s (s starts at +0).
// d2 >= 87 r^2 (the product rounded) guarantees d2 / r^2 rounds above 86, so
// those cells skip the division and the exponential without changing a bit.
__device__ inline float emitter_seeded(const FxStep& st, float pcx, float pcy, float pcz, uint32_t cell) {
  float s = 0.0f;
  for (int i = 0; i < st.seeded_count; ++i) {
    float dx = pcx - st.seeded[i][0], dy = pcy - st.seeded[i][1], dz = pcz - st.seeded[i][2];
    float d2 = dx * dx + dy * dy + dz * dz;
    float r = st.seeded[i][3];
    const float rr = r * r;
    if (d2 < rr * 87.0f) s = s + st.seeded[i][4] * fxm_exp(-(d2 / rr));
  }
  if (s > 0.0f && st.jitter > 0.0f) {
    float uj = fx_seeded_uniform(st.seed, st.frame, cell, 0);
    s = s * (1.0f + st.jitter * (2.0f * uj - 1.0f));
  }
  return s;
}

// ------------------------------------------------------------------ kernels
// initializeFields, shaders.ts:135-149: flow = (0, 0, 0, water.w),
// species = (0, max(chemistry.y, 0), 0, 0).
template <int S>
__global__ __launch_bounds__(256) void k_initialize(float4* __restrict__ flow, typename SpeciesIO<S>::Word* __restrict__ species,
                                                    uint32_t cells, float ambient, float oxygen) {
  uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= cells) return;
  flow[i] = make_float4(0.0f, 0.0f, 0.0f, ambient);
  SpeciesIO<S>::store(species, i, make_float4(0.0f, fxm_max(oxygen, 0.0f), 0.0f, 0.0f));
}

// Fill kernel (used instead of memset so that a captured graph holds no
// memset node).
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint64_t words, uint32_t value) {
  uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * 256u;
  for (; i < words; i += stride) p[i] = value;
}

__global__ __launch_bounds__(256) void k_copy_u128(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * 256u;
  for (; i < n; i += stride) dst[i] = src[i];
}

// simulate, shaders.ts:151-342 (advection, sources, reaction, forces, BCs).
// kLiquid = false (no liquid field, as when the hose never ran) fixes the
// water amount at +0; the terms that then reduce to exact identities are
// skipped: exp(-0) = 1 factors and the zero evaporation.  Per cell, the
// reaction's Arrhenius rate is evaluated only where the ignition weight is
// nonzero, and the radiation only where T is not exactly ambient; the other
// cells take the exact values those terms reduce to (see below), so every
// result keeps its bits, signed zeros included.
// simulate runs one-wave blocks (64 x 1 cells): its waves take different
// times (the Arrhenius rate, radiation and emitter terms are evaluated only
// where they are nonzero), and a one-wave block frees its slot as soon as
// its wave is done (285 -> 274 us at UHD against 64 x 4 blocks).
constexpr int kSimRows = 1;

template <int S, bool kSeeded, bool kLiquid>
__global__ __launch_bounds__(64 * kSimRows) void k_simulate(const float4* __restrict__ flowIn,
                                                  const typename SpeciesIO<S>::Word* __restrict__ specIn,
                                                  float4* __restrict__ flowOut,
                                                  typename SpeciesIO<S>::Word* __restrict__ specOut,
                                                  float* __restrict__ reactionOut,
                                                  const uint32_t* __restrict__ liquid,
                                                  const float* __restrict__ tab, FxStep st) {
  int x, y, z;
  if (!cell_coords<kSimRows>(st, x, y, z)) return;
  const uint32_t id = cidx(st, x, y, z);
  const float dt = st.dt;
  float pcx = (float)x + 0.5f, pcy = (float)y + 0.5f, pcz = (float)z + 0.5f;
  const SimTab to(st.nx, st.ny, st.nz);
  const float4 ty = reinterpret_cast<const float4*>(tab + to.y4)[y];  // envelope, 1 - outflow, decay, mix

  const float4 orig = flowIn[id];
  float bxp = pcx - orig.x * dt * st.mx;
  float byp = pcy - orig.y * dt * st.my;
  float bzp = pcz - orig.z * dt * st.mz;
  const TriCorners tc = tri_corners(st, bxp, byp, bzp);
  float4 fl = trilinear_pairs(tc, [&](uint32_t i, float4& a, float4& b) { a = flowIn[i]; b = flowIn[i + 1]; });
  float4 sp = trilinear_pairs(tc, [&](uint32_t i, float4& a, float4& b) { SpeciesIO<S>::load_pair(specIn, i, a, b); });
  sp = make_float4(fxm_max(sp.x, 0.0f), fxm_max(sp.y, 0.0f), fxm_max(sp.z, 0.0f), fxm_max(sp.w, 0.0f));

  // Six neighbours, clamped at the walls (loadFlow, shaders.ts:170-175).
  const uint32_t sy = (uint32_t)st.nx, sz = __umul24((uint32_t)st.nx, (uint32_t)st.ny);
  const uint32_t iL = x > 0 ? id - 1 : id, iR = x < st.nx - 1 ? id + 1 : id;
  const float4 L = flowIn[iL], R = flowIn[iR];
  const uint32_t iB = y > 0 ? id - sy : id, iT = y < st.ny - 1 ? id + sy : id;
  const uint32_t iK = z > 0 ? id - sz : id, iF = z < st.nz - 1 ? id + sz : id;
  const float4 B = flowIn[iB], Tp = flowIn[iT];
  const float4 K = flowIn[iK], F = flowIn[iF];
  float vax = (L.x + R.x + B.x + Tp.x + K.x + F.x) * (1.0f / 6.0f);
  float vay = (L.y + R.y + B.y + Tp.y + K.y + F.y) * (1.0f / 6.0f);
  float vaz = (L.z + R.z + B.z + Tp.z + K.z + F.z) * (1.0f / 6.0f);
  fl.x = fxm_mix(fl.x, vax, st.mix_velocity);
  fl.y = fxm_mix(fl.y, vay, st.mix_velocity);
  fl.z = fxm_mix(fl.z, vaz, st.mix_velocity);
  float ta = (L.w + R.w + B.w + Tp.w + K.w + F.w) * (1.0f / 6.0f);
  fl.w = fxm_mix(fl.w, ta, st.mix_temperature);

  {
    using SIO = SpeciesIO<S>;
    const float4 sL = SIO::load(specIn, iL), sR = SIO::load(specIn, iR);
    const float4 sB = SIO::load(specIn, iB), sT = SIO::load(specIn, iT);
    const float4 sK = SIO::load(specIn, iK), sF = SIO::load(specIn, iF);
    float4 sa = make_float4((sL.x + sR.x + sB.x + sT.x + sK.x + sF.x) * (1.0f / 6.0f),
                            (sL.y + sR.y + sB.y + sT.y + sK.y + sF.y) * (1.0f / 6.0f),
                            (sL.z + sR.z + sB.z + sT.z + sK.z + sF.z) * (1.0f / 6.0f),
                            (sL.w + sR.w + sB.w + sT.w + sK.w + sF.w) * (1.0f / 6.0f));
    sp = mix4(sp, sa, st.mix_species);
  }

  float source;
  if (kSeeded) {
    source = emitter_seeded(st, pcx, pcy, pcz, id);
  } else {
    const float4 e = reinterpret_cast<const float4*>(tab + to.xz4)[x + st.nx * z];
    source = fxm_exp(-(e.x + tab[to.yv + y])) * e.y * e.z;
  }
  sp.x = fxm_max(sp.x, source * st.chem_x * st.fire_power);
  fl.w = fxm_max(fl.w, st.ambient + source * 1180.0f * st.thermal_power);
  {
    float swx = tab[to.z + z];
    float swz = tab[to.x + x];
    float swk = source * dt * 0.11f * st.chem_w * st.power_root;
    fl.x = fl.x + swx * swk;
    fl.y = fl.y + 0.0f * swk;
    fl.z = fl.z + swz * swk;
  }
  const float lift = source * dt * 0.52f * st.lift_clamp * st.power_root;
  fl.y += lift;
  if (kSeeded && source > 0.0f && st.velocity_noise > 0.0f) {
    float vn = st.velocity_noise;
    float nxv = (2.0f * fx_seeded_uniform(st.seed, st.frame, id, 1) - 1.0f) * vn;
    float nyv = (2.0f * fx_seeded_uniform(st.seed, st.frame, id, 2) - 1.0f) * vn;
    float nzv = (2.0f * fx_seeded_uniform(st.seed, st.frame, id, 3) - 1.0f) * vn;
    fl.x = fl.x + nxv * lift;
    fl.y = fl.y + nyv * lift;
    fl.z = fl.z + nzv * lift;
  }

  if (x == 0 || x == st.nx - 1 || z == 0 || z == st.nz - 1 || y == st.ny - 1) {
    sp.y = fxm_max(sp.y, st.chem_y);
    fl.w = fxm_mix(fl.w, st.ambient, 0.035f);
  }

  const float liq = kLiquid ? (float)liquid[id] * (1.0f / 4096.0f) : 0.0f;
  const float water = kLiquid ? fxm_clamp(liq * 3.4f, 0.0f, 5.0f) : 0.0f;
  if (kLiquid) {
    float evaporation = water * fxm_smoothstep(335.0f, 1040.0f, fl.w) * dt * 13.5f;
    fl.w = fxm_max(st.ambient, fl.w - evaporation * 2080.0f - water * dt * 260.0f);
    sp.w = fxm_clamp(sp.w + evaporation * 2.15f, 0.0f, 3.0f);
  } else {
    // water = +0: evaporation = +0 (smoothstep >= +0), so T - 0 - 0 = T and
    // vapor + 0 (which turns a -0 into +0)
    fl.w = fxm_max(st.ambient, fl.w);
    sp.w = fxm_clamp(sp.w + 0.0f, 0.0f, 3.0f);
  }
  {
    float oxygenDeficit = fxm_clamp((st.ambient_oxygen - sp.y) * st.inv_ambient_oxygen, 0.0f, 1.0f);
    float fuelCore = fxm_smoothstep(0.30f, 0.78f, sp.x);
    float interfaceWeight = oxygenDeficit * fxm_mix(1.0f, 0.18f, fuelCore);
    float speed = fxm_sqrt(fl.x * fl.x + fl.y * fl.y + fl.z * fl.z);
    float entrainmentMix = fxm_clamp(
        dt * st.entrainment * (0.12f + 0.28f * speed + 0.22f * fxm_smoothstep(400.0f, 1500.0f, fl.w)) * interfaceWeight,
        0.0f, 0.08f);
    sp.y = fxm_mix(sp.y, st.ambient_oxygen, entrainmentMix);
  }
  float react;
  {
    float fuelBefore = fxm_max(sp.x, 0.0f);
    float oxygenBefore = fxm_max(sp.y, 0.0f);
    float availableFuel = fxm_min(fuelBefore, oxygenBefore * 0.25f);
    float ignition = fxm_smoothstep(690.0f, 930.0f, fl.w);
    if (ignition != 0.0f) {
      float safeFuel = fxm_max(fuelBefore, 0.00001f);
      float safeOxygen = fxm_max(oxygenBefore, 0.00001f);
      float activation = fxm_exp(-15000.0f / fxm_max(fl.w, st.ambient));
      float arrhenius =
          fxm_clamp(2.5e6f * activation * fxm_pow(safeFuel, -0.30f) * fxm_pow(safeOxygen, 1.30f), 0.0f, 80.0f);
      float requestedBurn = availableFuel * (1.0f - fxm_exp(-arrhenius * ignition * dt));
      if (kLiquid) requestedBurn = requestedBurn * fxm_exp(-water * 18.0f);  // else * exp(-0) = * 1
      float burn = fxm_clamp(requestedBurn, 0.0f, availableFuel);
      float richness = fxm_clamp(1.0f - oxygenBefore / fxm_max(fuelBefore * 4.0f, 0.00001f), 0.0f, 1.0f);
      sp.x = fxm_max(0.0f, fuelBefore - burn);
      sp.y = fxm_max(0.0f, oxygenBefore - burn * 4.0f);
      sp.z = fxm_clamp(sp.z + burn * (1.0f + richness * 1.8f), 0.0f, 3.0f);
      sp.w = fxm_clamp(sp.w + burn * 0.62f, 0.0f, 3.0f);
      fl.w += burn * st.heat_release;
      react = fxm_clamp(burn * st.inv_max_dt * 0.22f, 0.0f, 4.0f);
    } else {
      // Below 690 K the ignition weight is a zero: the finite rate times it
      // is a zero, exp of a zero is exactly 1, so the request is
      // availableFuel * (+0) (times exp(-18 w) > 0), a zero with
      // availableFuel's sign, and every burn-weighted term adds that zero:
      // burn * c = burn for the positive factors c below, and T (far from
      // zero) + burn = T.
      float burn = fxm_clamp(availableFuel * 0.0f, 0.0f, availableFuel);
      sp.x = fxm_max(0.0f, fuelBefore - burn);
      sp.y = fxm_max(0.0f, oxygenBefore - burn);
      sp.z = fxm_clamp(sp.z + burn, 0.0f, 3.0f);
      sp.w = fxm_clamp(sp.w + burn, 0.0f, 3.0f);
      react = fxm_clamp(burn, 0.0f, 4.0f);
    }
  }
  {
    float curlx = 0.5f * ((Tp.z - B.z) * st.my - (F.y - K.y) * st.mz);
    float curly = 0.5f * ((F.x - K.x) * st.mz - (R.z - L.z) * st.mx);
    float curlz = 0.5f * ((R.y - L.y) * st.mx - (Tp.x - B.x) * st.my);
    float gdx = 0.5f * ((R.w - L.w) * st.mx) + 0.00001f;
    float gdy = 0.5f * ((Tp.w - B.w) * st.my) + 0.00001f;
    float gdz = 0.5f * ((F.w - K.w) * st.mz) + 0.00001f;
    float glen = fxm_sqrt(gdx * gdx + gdy * gdy + gdz * gdz);
    const float ilen = 1.0f / glen;
    float dirx = gdx * ilen, diry = gdy * ilen, dirz = gdz * ilen;
    fl.x = fl.x + (diry * curlz - dirz * curly) * st.vortex_k;
    fl.y = fl.y + (dirz * curlx - dirx * curlz) * st.vortex_k;
    fl.z = fl.z + (dirx * curly - diry * curlx) * st.vortex_k;
  }
  {
    float hotPlume = fxm_smoothstep(520.0f, 1650.0f, fl.w);
    float heightEnvelope = ty.x;
    float ltx = tab[to.yz + y + st.ny * z];
    float lty = tab[to.xz + x + st.nx * z];
    float ltz = tab[to.xy + x + st.nx * y];
    float turbulenceStrength = st.turbulence_base * hotPlume * heightEnvelope;
    float tk = turbulenceStrength * dt;
    fl.x = fl.x + ltx * tk;
    fl.y = fl.y + lty * tk;
    fl.z = fl.z + ltz * tk;
  }
  {
    float buoyancy = fxm_clamp((fl.w - st.ambient) * (1.0f / 1500.0f), -0.1f, 1.2f);
    fl.y += buoyancy * dt * 0.72f * st.buoyancy_clamp;
  }
  if (kLiquid) sp.z *= fxm_exp(-water * dt * 4.8f);  // else * exp(-0) = * 1
  fl.x = fl.x + st.spray_x * water * dt * 0.82f;
  fl.y = fl.y + st.spray_y * water * dt * 0.82f;
  fl.z = fl.z + st.spray_z * water * dt * 0.82f;
  {
    // At exactly ambient T the radiation is 5.67e-8 (ambient4 - ambient4) ... = +0
    // (ambient4 is this same fxm_pow of the ambient), and max(ambient,
    // ambient + dt * 0) = ambient: nothing to evaluate.
    if (fl.w != st.ambient) {
      float radiation = 5.67e-8f * (st.ambient4 - fxm_pow(fxm_max(fl.w, 0.0f), 4.0f)) * st.inv_radiation_denominator;
      fl.w = fxm_max(st.ambient, fl.w + dt * radiation);
    }
    sp.z *= st.decay_products;
    sp.w *= st.decay_vapor;
  }
  {
    sp.x *= ty.y;  // 1 - outflow, outflow = smoothstep(0.84, 0.995, npy)
    float outflowDecay = ty.z;
    sp.z *= outflowDecay;
    sp.w *= outflowDecay;
    fl.w = fxm_mix(fl.w, st.ambient, ty.w);
    react *= ty.y;
  }
  if (y == st.ny - 1) {
    sp.x = 0.0f; sp.z = 0.0f; sp.w = 0.0f;
    fl.w = st.ambient;
    react = 0.0f;
  }
  if (x == 0 || x == st.nx - 1) fl.x = 0.0f;
  if (z == 0 || z == st.nz - 1) fl.z = 0.0f;
  if (y == 0) fl.y = 0.0f;
  if (y == st.ny - 1) fl.y = fxm_max(0.0f, fl.y);
  fl.x = fxm_clamp(fl.x, -1.4f, 1.4f);
  fl.y = fxm_clamp(fl.y, -1.4f, 1.4f);
  fl.z = fxm_clamp(fl.z, -1.4f, 1.4f);
  fl.w = fxm_clamp(fl.w, st.ambient, 2900.0f);
  sp = make_float4(fxm_clamp(sp.x, 0.0f, 3.0f), fxm_clamp(sp.y, 0.0f, 3.0f),
                   fxm_clamp(sp.z, 0.0f, 3.0f), fxm_clamp(sp.w, 0.0f, 3.0f));
  flowOut[id] = fl;
  SpeciesIO<S>::store(specOut, id, sp);
  reactionOut[id] = react;
}

// packFields / storePackedFields, qualityShaders.ts:37-73: the raymarcher's
// three rgba16float volumes, stored as 4 x f16 per cell in cell order:
//   flow      (u, v, w, clamp((T - 300) / 2600, 0, 1))
//   chemistry max(species, 0)
//   media     (reaction, soot, liquid, display value)
// With the divergence view (view mode 7) the display value is the
// divergence field (refreshed from flow[0] just before), 0 on the outer
// cell shell.
__device__ inline uint32_t pack_half2(float a, float b) {
  const uint16_t ha = __builtin_bit_cast(uint16_t, (_Float16)fx_keep_f32(a));
  const uint16_t hb = __builtin_bit_cast(uint16_t, (_Float16)fx_keep_f32(b));
  return (uint32_t)ha | ((uint32_t)hb << 16);
}
__device__ inline uint2 pack_half4(float a, float b, float c, float d) {
  return make_uint2(pack_half2(a, b), pack_half2(c, d));
}

template <int S>
__global__ __launch_bounds__(256) void k_pack_fields(const float4* __restrict__ flow,
                                                     const typename SpeciesIO<S>::Word* __restrict__ species,
                                                     const float* __restrict__ reaction,
                                                     const uint32_t* __restrict__ liquid,
                                                     const float* __restrict__ div, uint2* __restrict__ out,
                                                     FxStep st, int divergence_view) {
  int x, y, z;
  if (!cell_coords(st, x, y, z)) return;
  const uint32_t id = cidx(st, x, y, z);
  const uint32_t cells = (uint32_t)st.nx * (uint32_t)st.ny * (uint32_t)st.nz;
  const float4 f = flow[id];
  float4 sp = SpeciesIO<S>::load(species, id);
  sp = make_float4(fxm_max(sp.x, 0.0f), fxm_max(sp.y, 0.0f), fxm_max(sp.z, 0.0f), fxm_max(sp.w, 0.0f));
  const float react = fxm_clamp(reaction[id], 0.0f, 4.0f);
  const float richness = fxm_clamp(sp.x / fxm_max(sp.x + sp.y * 0.25f, 0.0001f), 0.0f, 1.0f);
  const float soot = sp.z * fxm_smoothstep(0.25f, 0.82f, richness);
  const float liq = (float)(liquid ? liquid[id] : 0u) * (3.4f / 4096.0f);
  float display = fxm_max(fxm_max(react, soot),
                          fxm_max(liq, fxm_max(sp.w, fxm_smoothstep(420.0f, 700.0f, f.w))));
  if (divergence_view) {
    const bool shell = x == 0 || y == 0 || z == 0 || x + 1 >= st.nx || y + 1 >= st.ny || z + 1 >= st.nz;
    display = shell ? 0.0f : div[id];
  }
  out[id] = pack_half4(f.x, f.y, f.z, fxm_clamp((f.w - 300.0f) / 2600.0f, 0.0f, 1.0f));
  out[cells + id] = pack_half4(sp.x, sp.y, sp.z, sp.w);
  out[2u * cells + id] = pack_half4(react, soot, liq, display);
}

// Column form (the divergence): a 64 x 4 block owns a 64 x 4 column of
// cells over kColPlanes consecutive planes, and each thread issues every
// load of its column before it computes, so a wave has kColPlanes planes of
// reads in flight instead of one.  The z neighbour of a plane is the
// previous plane's value, already loaded.  Lanes past the x wall and planes
// past the last one work on the clamped cell and store exactly the value
// that cell's own lane or plane stores, to the same address, so no store is
// conditional (a conditional store lets the compiler sink the loads it
// consumes behind the branch, one plane at a time).  Blocks run with the z
// chunk fastest (after x), so the z - 1 plane of a chunk's first plane was
// read by the chunk before it on the same XCD moments earlier and is an L2
// hit, and each XCD owns a range of rows across all planes (8 % faster than
// z-slab order).  The projection keeps one cell per thread: its five
// independent loads per cell already run it at the rate of a plain copy,
// and the column form measured no faster.
constexpr int kColPlanes = 4;

struct ColumnBlock {
  int x, y, z0;
  uint32_t col;  // x (clamped to the grid) + nx * y
};
__device__ inline bool column_block(const FxStep& st, ColumnBlock& c) {
  const uint32_t bx = (uint32_t)(st.nx + kBX - 1) / kBX;
  const uint32_t by = (uint32_t)(st.ny + kBY - 1) / kBY;
  const uint32_t bz = (uint32_t)(st.nz + kColPlanes - 1) / kColPlanes;
  const uint32_t b = xcd_remap(blockIdx.x, bx * by * bz);
  const uint32_t ix = b % bx, rest = b / bx;
  const uint32_t iz = rest % bz, iy = rest / bz;
  c.x = (int)(ix * kBX + threadIdx.x);
  c.y = (int)(iy * kBY + threadIdx.y);
  c.z0 = (int)iz * kColPlanes;
  c.col = (uint32_t)min(c.x, st.nx - 1) + (uint32_t)st.nx * (uint32_t)c.y;
  return c.y < st.ny;  // uniform over a wave (one x row)
}
static inline dim3 column_grid(const FxStep& st) {
  const uint32_t bx = (uint32_t)(st.nx + kBX - 1) / kBX, by = (uint32_t)(st.ny + kBY - 1) / kBY;
  return dim3(bx * by * ((uint32_t)(st.nz + kColPlanes - 1) / kColPlanes));
}

// computeDivergence, shaders.ts:344-364.  The reads at x - 1, y - 1, z - 1
// clamp to the cell itself at the walls, as cidx does.
__global__ __launch_bounds__(256) void k_divergence(const float4* __restrict__ flow, float* __restrict__ div, FxStep st) {
  ColumnBlock cb;
  if (!column_block(st, cb)) return;
  const uint32_t plane = (uint32_t)st.nx * (uint32_t)st.ny;
  const uint32_t dx = min(cb.x, st.nx - 1) > 0 ? 1u : 0u, dy = cb.y > 0 ? (uint32_t)st.nx : 0u;
  uint32_t id[kColPlanes];
  float4 c[kColPlanes];
  float left[kColPlanes], bottom[kColPlanes];
#pragma unroll
  for (int k = 0; k < kColPlanes; ++k) {
    id[k] = cb.col + plane * (uint32_t)min(cb.z0 + k, st.nz - 1);
    c[k] = flow[id[k]];
    left[k] = flow[id[k] - dx].x;
    bottom[k] = flow[id[k] - dy].y;
  }
  float back = flow[cb.col + plane * (uint32_t)max(cb.z0 - 1, 0)].z;
  float d = 0.0f;
#pragma unroll
  for (int k = 0; k < kColPlanes; ++k) {
    // past the last plane: the last plane's value again (k = 0 never is)
    const float dk = (c[k].x - left[k]) * st.mx + (c[k].y - bottom[k]) * st.my + (c[k].z - back) * st.mz;
    d = k == 0 || cb.z0 + k < st.nz ? dk : d;
    div[id[k]] = d;
    back = c[k].z;
  }
}

// loadPressure, shaders.ts:366-373: y >= ny is the open-top Dirichlet zero.
template <int P>
__device__ inline float load_p(const typename PressureIO<P>::Word* p, const FxStep& st, int x, int y, int z) {
  if (y >= st.ny) return 0.0f;
  return PressureIO<P>::load(p, cidx(st, x, y, z));
}

// solvePressureGlobal, shaders.ts:375-403 (one Jacobi sweep).  The division
// by 2 (wx + wy + wz) is a multiply by its f32 reciprocal, and the weighted
// neighbour sum is contracted: fma(wz, sz, fma(wy, sy, wx sx)).
template <int P>
__global__ __launch_bounds__(256) void k_jacobi(const typename PressureIO<P>::Word* __restrict__ pin,
                                                typename PressureIO<P>::Word* __restrict__ pout,
                                                const float* __restrict__ div, FxStep st) {
  int x, y, z;
  if (!cell_coords(st, x, y, z)) return;
  const uint32_t id = cidx(st, x, y, z);
  if (y == st.ny - 1) {
    PressureIO<P>::store(pout, id, 0.0f);
    return;
  }
  const float nb = jacobi_sum(st.wx, load_p<P>(pin, st, x - 1, y, z) + load_p<P>(pin, st, x + 1, y, z),
                              st.wy, load_p<P>(pin, st, x, y - 1, z) + load_p<P>(pin, st, x, y + 1, z),
                              st.wz, load_p<P>(pin, st, x, y, z - 1) + load_p<P>(pin, st, x, y, z + 1));
  PressureIO<P>::store(pout, id, (nb - div[id]) * st.jacobi_inv);
}

// pressureGradient + projectVelocity, shaders.ts:461-501
// (projectVelocityF16, pressureShaders.ts:119-160, for P = 1).
template <int P>
__device__ inline void gradient(const typename PressureIO<P>::Word* p, const FxStep& st, int x, int y, int z,
                                float& g0, float& g1, float& g2) {
  float c = load_p<P>(p, st, x, y, z);
  g0 = (load_p<P>(p, st, x + 1, y, z) - c) * st.mx;
  g1 = (load_p<P>(p, st, x, y + 1, z) - c) * st.my;
  g2 = (load_p<P>(p, st, x, y, z + 1) - c) * st.mz;
}

template <int P>
__global__ __launch_bounds__(256) void k_project(const float4* __restrict__ fin, const typename PressureIO<P>::Word* __restrict__ p,
                                                 float4* __restrict__ fout, FxStep st) {
  int x, y, z;
  if (!cell_coords(st, x, y, z)) return;
  const uint32_t id = cidx(st, x, y, z);
  float g0, g1, g2;
  gradient<P>(p, st, x, y, z, g0, g1, g2);
  float4 f = fin[id];
  f.x = f.x - g0; f.y = f.y - g1; f.z = f.z - g2;
  if (y == st.ny - 1) {
    const float4 bf = fin[cidx(st, x, y - 1, z)];
    gradient<P>(p, st, x, y - 1, z, g0, g1, g2);
    f.x = bf.x - g0; f.y = bf.y - g1; f.z = bf.z - g2;
  }
  if (x == 0 || x == st.nx - 1) f.x = 0.0f;
  if (z == 0 || z == st.nz - 1) f.z = 0.0f;
  if (y == 0) f.y = 0.0f;
  if (y == st.ny - 1) f.y = fxm_max(0.0f, f.y);
  fout[id] = f;
}

// ---------------------------------------------------------------- reductions
// FP64 sums (additions: the reference computes none).  Each block reduces a
// fixed slice in a fixed order and writes one partial; a single block then
// adds the partials in order, so results do not depend on scheduling.
constexpr int kRedThreads = 256;

__device__ inline double block_sum(double v) {
  __shared__ double part[kRedThreads / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) part[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kRedThreads / 64; ++w) s += part[w];
  return s;
}

template <int S>
__global__ __launch_bounds__(kRedThreads) void k_reduce_field(const void* __restrict__ data, int kind, int component,
                                                              uint32_t cells, uint32_t per_block, double* __restrict__ partial) {
  uint32_t begin = blockIdx.x * per_block;
  uint32_t end = min(begin + per_block, cells);
  double acc = 0.0;
  for (uint32_t i = begin + threadIdx.x; i < end; i += kRedThreads) {
    float v;
    if (kind == 0) {  // species (storage S)
      float4 s = SpeciesIO<S>::load((const typename SpeciesIO<S>::Word*)data, i);
      v = component == 0 ? s.x : component == 1 ? s.y : component == 2 ? s.z : s.w;
    } else if (kind == 1) {  // flow component
      float4 f = ((const float4*)data)[i];
      v = component == 0 ? f.x : component == 1 ? f.y : component == 2 ? f.z : f.w;
    } else {  // scalar f32
      v = ((const float*)data)[i];
    }
    acc += (double)v;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Squared residual of the pressure system at non-top cells:
// r = wx (p[x-1] + p[x+1]) + wy (..) + wz (..) - 2 (wx + wy + wz) p - div, in FP64.
template <int P>
__global__ __launch_bounds__(kRedThreads) void k_reduce_residual(const typename PressureIO<P>::Word* __restrict__ p,
                                                                 const float* __restrict__ div, FxStep st,
                                                                 uint32_t per_block, double* __restrict__ partial) {
  uint32_t cells = (uint32_t)st.nx * st.ny * st.nz;
  uint32_t begin = blockIdx.x * per_block;
  uint32_t end = min(begin + per_block, cells);
  double wx = st.wx, wy = st.wy, wz = st.wz;
  double acc = 0.0;
  for (uint32_t i = begin + threadIdx.x; i < end; i += kRedThreads) {
    int x = (int)(i % (uint32_t)st.nx);
    uint32_t r = i / (uint32_t)st.nx;
    int y = (int)(r % (uint32_t)st.ny), z = (int)(r / (uint32_t)st.ny);
    if (y == st.ny - 1) continue;
    double c = load_p<P>(p, st, x, y, z);
    double nb = wx * ((double)load_p<P>(p, st, x - 1, y, z) + (double)load_p<P>(p, st, x + 1, y, z))
              + wy * ((double)load_p<P>(p, st, x, y - 1, z) + (double)load_p<P>(p, st, x, y + 1, z))
              + wz * ((double)load_p<P>(p, st, x, y, z - 1) + (double)load_p<P>(p, st, x, y, z + 1));
    double res = nb - 2.0 * (wx + wy + wz) * c - (double)div[i];
    acc += res * res;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Squared backward divergence of a flow field (computeDivergence in FP64).
__global__ __launch_bounds__(kRedThreads) void k_reduce_flow_divergence(const float4* __restrict__ flow, FxStep st,
                                                                        uint32_t per_block, double* __restrict__ partial) {
  uint32_t cells = (uint32_t)st.nx * st.ny * st.nz;
  uint32_t begin = blockIdx.x * per_block;
  uint32_t end = min(begin + per_block, cells);
  double acc = 0.0;
  for (uint32_t i = begin + threadIdx.x; i < end; i += kRedThreads) {
    int x = (int)(i % (uint32_t)st.nx);
    uint32_t r = i / (uint32_t)st.nx;
    int y = (int)(r % (uint32_t)st.ny), z = (int)(r / (uint32_t)st.ny);
    float4 c = flow[i];
    double d = ((double)c.x - (double)flow[cidx(st, x - 1, y, z)].x) * (double)st.mx
             + ((double)c.y - (double)flow[cidx(st, x, y - 1, z)].y) * (double)st.my
             + ((double)c.z - (double)flow[cidx(st, x, y, z - 1)].z) * (double)st.mz;
    acc += d * d;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void k_reduce_final(const double* __restrict__ partial, uint32_t n,
                                                              double* __restrict__ out) {
  // One block, fixed order: thread t sums partials t, t + 256, ...; then the
  // block sum adds the 256 thread sums in a fixed tree.
  double acc = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += kRedThreads) acc += partial[i];
  double s = block_sum(acc);
  if (threadIdx.x == 0) out[0] = s;
}

// ------------------------------------------------------------------ launchers
static inline dim3 cell_grid(const FxStep& st) {
  uint32_t bx = (uint32_t)(st.nx + kBX - 1) / kBX;
  uint32_t by = (uint32_t)(st.ny + kBY - 1) / kBY;
  return dim3(bx * by * (uint32_t)st.nz);
}

void launch_initialize(int storage, float4* flow, void* species, uint32_t cells, float ambient, float oxygen,
                       hipStream_t s) {
  dim3 g((cells + 255) / 256);
  switch (storage) {
    case FX_STORAGE_F32: k_initialize<FX_STORAGE_F32><<<g, 256, 0, s>>>(flow, (float4*)species, cells, ambient, oxygen); break;
    case FX_STORAGE_F16: k_initialize<FX_STORAGE_F16><<<g, 256, 0, s>>>(flow, (half4_t*)species, cells, ambient, oxygen); break;
    default: k_initialize<FX_STORAGE_BF16><<<g, 256, 0, s>>>(flow, (ushort4*)species, cells, ambient, oxygen); break;
  }
}

void launch_fill_u32(void* p, uint64_t words, uint32_t value, hipStream_t s) {
  if (words == 0) return;
  uint64_t blocks = (words + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  k_fill_u32<<<dim3((uint32_t)blocks), 256, 0, s>>>((uint32_t*)p, words, value);
}

void launch_copy_u128(void* dst, const void* src, uint64_t n16, hipStream_t s) {
  if (n16 == 0) return;
  uint64_t blocks = (n16 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  k_copy_u128<<<dim3((uint32_t)blocks), 256, 0, s>>>((uint4*)dst, (const uint4*)src, n16);
}

template <int S>
static void simulate_s(const SimulateArgs& a, hipStream_t s) {
  using W = typename SpeciesIO<S>::Word;
  const FxStep& st = a.st;
  const uint32_t entries = (uint32_t)st.nx * st.nz + (uint32_t)st.nx * st.ny + (uint32_t)st.ny * st.nz +
                           (uint32_t)(st.nx + st.ny + st.nz);
  uint32_t blocks = (entries + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_sim_tables<<<blocks, 256, 0, s>>>(a.tables, st, st.seeded_count > 0 ? 0 : 1);
  const W* sin_ = (const W*)a.species_in;
  W* sout = (W*)a.species_out;
  const dim3 g((uint32_t)((st.nx + kBX - 1) / kBX) * (uint32_t)((st.ny + kSimRows - 1) / kSimRows) * (uint32_t)st.nz),
      b(kBX, kSimRows);
  const bool liquid = st.has_liquid != 0;
  if (st.seeded_count > 0) {
    if (liquid)
      k_simulate<S, true, true><<<g, b, 0, s>>>(a.flow_in, sin_, a.flow_out, sout, a.reaction, a.liquid, a.tables, st);
    else
      k_simulate<S, true, false><<<g, b, 0, s>>>(a.flow_in, sin_, a.flow_out, sout, a.reaction, a.liquid, a.tables, st);
  } else {
    if (liquid)
      k_simulate<S, false, true><<<g, b, 0, s>>>(a.flow_in, sin_, a.flow_out, sout, a.reaction, a.liquid, a.tables, st);
    else
      k_simulate<S, false, false><<<g, b, 0, s>>>(a.flow_in, sin_, a.flow_out, sout, a.reaction, a.liquid, a.tables, st);
  }
}

size_t sim_table_bytes(const int32_t dims[3]) { return (size_t)fx_sim_table_floats(dims) * 4u; }

void launch_simulate(int storage, const SimulateArgs& a, hipStream_t s) {
  switch (storage) {
    case FX_STORAGE_F32: simulate_s<FX_STORAGE_F32>(a, s); break;
    case FX_STORAGE_F16: simulate_s<FX_STORAGE_F16>(a, s); break;
    default: simulate_s<FX_STORAGE_BF16>(a, s); break;
  }
}

void launch_pack_fields(int storage, const float4* flow, const void* species, const float* reaction,
                        const uint32_t* liquid, const float* div, void* out, const FxStep& st, int divergence_view,
                        hipStream_t s) {
  uint2* o = (uint2*)out;
  switch (storage) {
    case FX_STORAGE_F32:
      k_pack_fields<FX_STORAGE_F32><<<cell_grid(st), dim3(kBX, kBY), 0, s>>>(
          flow, (const float4*)species, reaction, liquid, div, o, st, divergence_view);
      break;
    case FX_STORAGE_F16:
      k_pack_fields<FX_STORAGE_F16><<<cell_grid(st), dim3(kBX, kBY), 0, s>>>(
          flow, (const typename SpeciesIO<FX_STORAGE_F16>::Word*)species, reaction, liquid, div, o, st,
          divergence_view);
      break;
    default:
      k_pack_fields<FX_STORAGE_BF16><<<cell_grid(st), dim3(kBX, kBY), 0, s>>>(
          flow, (const typename SpeciesIO<FX_STORAGE_BF16>::Word*)species, reaction, liquid, div, o, st,
          divergence_view);
      break;
  }
}

void launch_divergence(const float4* flow, float* div, const FxStep& st, hipStream_t s) {
  k_divergence<<<column_grid(st), dim3(kBX, kBY), 0, s>>>(flow, div, st);
}

void launch_jacobi(int f16, const void* pin, void* pout, const float* div, const FxStep& st, hipStream_t s) {
  if (f16)
    k_jacobi<1><<<cell_grid(st), dim3(kBX, kBY), 0, s>>>((const _Float16*)pin, (_Float16*)pout, div, st);
  else
    k_jacobi<0><<<cell_grid(st), dim3(kBX, kBY), 0, s>>>((const float*)pin, (float*)pout, div, st);
}

void launch_project(int f16, const float4* fin, const void* p, float4* fout, const FxStep& st, hipStream_t s) {
  if (f16)
    k_project<1><<<cell_grid(st), dim3(kBX, kBY), 0, s>>>(fin, (const _Float16*)p, fout, st);
  else
    k_project<0><<<cell_grid(st), dim3(kBX, kBY), 0, s>>>(fin, (const float*)p, fout, st);
}

static uint32_t reduce_blocks(uint32_t cells, uint32_t& per_block) {
  uint32_t blocks = 1024;
  per_block = (cells + blocks - 1) / blocks;
  per_block = (per_block + kRedThreads - 1) / kRedThreads * kRedThreads;
  if (per_block == 0) per_block = kRedThreads;
  return (cells + per_block - 1) / per_block;
}

uint32_t reduce_block_count(uint32_t cells) {
  uint32_t per_block;
  return reduce_blocks(cells, per_block);
}

namespace {
__global__ void k_sqrt_f64(double* v) { v[0] = sqrt(v[0]); }
}

void launch_sqrt_f64(double* value, hipStream_t s) { k_sqrt_f64<<<1, 1, 0, s>>>(value); }

void launch_reduce_field(int storage, int kind, int component, const void* data, uint32_t cells, double* partial,
                         double* out, hipStream_t s) {
  uint32_t per_block;
  uint32_t blocks = reduce_blocks(cells, per_block);
  switch (storage) {
    case FX_STORAGE_F32: k_reduce_field<FX_STORAGE_F32><<<blocks, kRedThreads, 0, s>>>(data, kind, component, cells, per_block, partial); break;
    case FX_STORAGE_F16: k_reduce_field<FX_STORAGE_F16><<<blocks, kRedThreads, 0, s>>>(data, kind, component, cells, per_block, partial); break;
    default: k_reduce_field<FX_STORAGE_BF16><<<blocks, kRedThreads, 0, s>>>(data, kind, component, cells, per_block, partial); break;
  }
  k_reduce_final<<<1, kRedThreads, 0, s>>>(partial, blocks, out);
}

void launch_reduce_residual(int f16, const void* p, const float* div, const FxStep& st, double* partial, double* out,
                            hipStream_t s) {
  uint32_t cells = (uint32_t)st.nx * st.ny * st.nz;
  uint32_t per_block;
  uint32_t blocks = reduce_blocks(cells, per_block);
  if (f16)
    k_reduce_residual<1><<<blocks, kRedThreads, 0, s>>>((const _Float16*)p, div, st, per_block, partial);
  else
    k_reduce_residual<0><<<blocks, kRedThreads, 0, s>>>((const float*)p, div, st, per_block, partial);
  k_reduce_final<<<1, kRedThreads, 0, s>>>(partial, blocks, out);
}

void launch_reduce_flow_divergence(const float4* flow, const FxStep& st, double* partial, double* out, hipStream_t s) {
  uint32_t cells = (uint32_t)st.nx * st.ny * st.nz;
  uint32_t per_block;
  uint32_t blocks = reduce_blocks(cells, per_block);
  k_reduce_flow_divergence<<<blocks, kRedThreads, 0, s>>>(flow, st, per_block, partial);
  k_reduce_final<<<1, kRedThreads, 0, s>>>(partial, blocks, out);
}

}  // namespace fx
