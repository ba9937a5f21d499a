// fx_math.h — f32 implementations of the WGSL built-in functions used by the
// Fire-X solver step (exp, exp2, log2, pow, sin, cos, sqrt, floor, fract, min,
// max, clamp, mix, smoothstep).
//
// The reference's kernels (src/sim/firex/shaders.ts:101-342) call WGSL
// built-ins whose f32 results are implementation-defined (the WGSL spec only
// bounds their error).  This header fixes one implementation so that the same
// source gives the same bits on the gfx950 device and on the x86-64 host:
//
//   * exp, exp2: Cody-Waite argument reduction (W. J. Cody and W. Waite,
//     "Software Manual for the Elementary Functions", 1980) by ln 2 / 1,
//     followed by the degree-7 Taylor polynomial of e^r (or 2^r) on
//     |r| <= ln2/2 (|r| <= 1/2), scaled by 2^k built from the exponent bits.
//   * log2: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), then
//     ln m = 2 atanh(s), s = (m - 1) / (m + 1), summed as the odd Taylor
//     series of atanh through s^9.
//   * pow(x, y) = exp2(y * log2(x)), which is the WGSL definition of pow.
//   * sin, cos: Cody-Waite reduction by pi/2 with a three-part split of pi/2
//     (8 + 12 + 24 significant bits), then the Taylor series of sin through
//     r^9 and of cos through r^10 on |r| <= pi/4, selected by quadrant.
//
// Results below 2^-125 flush to +0 so that no denormal is ever produced by
// these functions (the device and the host then agree whatever their
// denormal mode).  Every function uses only IEEE-754 single-precision add,
// subtract, multiply, divide, sqrt and fused multiply-add (fxm_fma: one
// rounding, as v_fma_f32 on the device and the FMA instruction or fmaf on
// the host), rounding to nearest, and integer/bit operations.  The fused
// multiply-adds are written out explicitly -- the Horner steps of the
// polynomials, the Cody-Waite reductions, mix and smoothstep, each the
// contraction a shader compiler may form for that expression -- and no other
// may be formed (build with -ffp-contract=off on both sides).
//
// min/max/clamp are written as compare-and-select so that their treatment
// of signed zeros is identical on every target.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIP__)
#define FXM_FN __host__ __device__ static inline
#else
#define FXM_FN static inline
#endif

FXM_FN float fxm_bits_to_float(uint32_t u) {
  float f;
  memcpy(&f, &u, sizeof f);
  return f;
}

FXM_FN uint32_t fxm_float_to_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, sizeof u);
  return u;
}

// a * b + c rounded once.
FXM_FN float fxm_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

FXM_FN float fxm_min(float a, float b) { return (b < a) ? b : a; }
FXM_FN float fxm_max(float a, float b) { return (a < b) ? b : a; }
FXM_FN float fxm_clamp(float x, float lo, float hi) { return fxm_min(fxm_max(x, lo), hi); }

// WGSL mix(e1, e2, e3) = e1 * (1 - e3) + e2 * e3, contracted as
// fma(e2, e3, e1 * (1 - e3)).
FXM_FN float fxm_mix(float a, float b, float t) { return fxm_fma(b, t, a * (1.0f - t)); }

// WGSL smoothstep(low, high, x): t = clamp((x - low) / (high - low), 0, 1);
// t * t * (3 - 2 t), with 3 - 2 t as fma(-2, t, 3).  The division is a
// multiply by the f32 reciprocal of high - low (one of the results WGSL's
// 2.5 ULP f32 division admits), which folds to a constant where the edges
// are constants.
FXM_FN float fxm_smoothstep(float lo, float hi, float x) {
  float t = fxm_clamp((x - lo) * (1.0f / (hi - lo)), 0.0f, 1.0f);
  return t * t * fxm_fma(-2.0f, t, 3.0f);
}

FXM_FN float fxm_floor(float x) { return __builtin_floorf(x); }
FXM_FN float fxm_rint(float x) { return __builtin_rintf(x); }
FXM_FN float fxm_fract(float x) { return x - __builtin_floorf(x); }
FXM_FN float fxm_sqrt(float x) { return __builtin_sqrtf(x); }

// p * 2^k for integer-valued k in [-125, 128] (k = 128 is split as 2 * 2^127).
FXM_FN float fxm_ldexp(float p, float k) {
  if (k > 127.0f) {
    p = p * 2.0f;
    k = 127.0f;
  }
  int32_t e = (int32_t)k + 127;
  return p * fxm_bits_to_float((uint32_t)e << 23);
}

// e^x.
FXM_FN float fxm_exp(float x) {
  if (x != x) return x;
  if (x > 88.72283f) return fxm_bits_to_float(0x7f800000u);
  if (x < -86.0f) return 0.0f;
  const float kLog2e = 1.44269502f;       // 1/ln 2 rounded to f32
  const float kLn2Hi = 0.693115234375f;   // 0x1.62ep-1, 12 significant bits
  const float kLn2Lo = 3.19461833e-05f;   // ln 2 - kLn2Hi rounded to f32
  float k = fxm_rint(x * kLog2e);
  float r = fxm_fma(-k, kLn2Lo, fxm_fma(-k, kLn2Hi, x));  // k kLn2Hi is exact
  float p = fxm_fma(r, 1.98412701e-04f, 1.38888892e-03f);
  p = fxm_fma(r, p, 8.33333377e-03f);
  p = fxm_fma(r, p, 4.16666679e-02f);
  p = fxm_fma(r, p, 1.66666672e-01f);
  p = fxm_fma(r, p, 0.5f);
  p = fxm_fma(r, p, 1.0f);
  p = fxm_fma(r, p, 1.0f);
  return fxm_ldexp(p, k);
}

// 2^x.
FXM_FN float fxm_exp2(float x) {
  if (x != x) return x;
  if (x >= 128.0f) return fxm_bits_to_float(0x7f800000u);
  if (x < -125.0f) return 0.0f;
  float k = fxm_rint(x);
  float r = x - k;  // exact
  // c_n = (ln 2)^n / n!
  float p = fxm_fma(r, 1.52527336e-05f, 1.54035297e-04f);
  p = fxm_fma(r, p, 1.33335579e-03f);
  p = fxm_fma(r, p, 9.61812865e-03f);
  p = fxm_fma(r, p, 5.55041097e-02f);
  p = fxm_fma(r, p, 2.40226507e-01f);
  p = fxm_fma(r, p, 6.93147182e-01f);
  p = fxm_fma(r, p, 1.0f);
  return fxm_ldexp(p, k);
}

// log2(x) for x > 0; log2(0) = -inf, negative or NaN input gives NaN.
FXM_FN float fxm_log2(float x) {
  if (!(x > 0.0f)) {
    return (x == 0.0f) ? fxm_bits_to_float(0xff800000u) : fxm_bits_to_float(0x7fc00000u);
  }
  uint32_t u = fxm_float_to_bits(x);
  if (u >= 0x7f800000u) return x;  // +inf
  float bias = 0.0f;
  if (u < 0x00800000u) {  // subnormal input: scale into the normal range
    x = x * 16777216.0f;  // 2^24
    u = fxm_float_to_bits(x);
    bias = -24.0f;
  }
  int32_t e = (int32_t)(u >> 23) - 127;
  float m = fxm_bits_to_float((u & 0x007fffffu) | 0x3f800000u);  // [1, 2)
  if (m > 1.41421354f) {
    m = m * 0.5f;
    e += 1;
  }
  float f = m - 1.0f;          // exact
  float s = f / (2.0f + f);
  float z = s * s;
  float poly = fxm_fma(z, 1.11111112e-01f, 1.42857149e-01f);
  poly = fxm_fma(z, poly, 2.00000003e-01f);
  poly = fxm_fma(z, poly, 3.33333343e-01f);
  poly = z * poly;
  float t = fxm_fma(s, poly, s);  // atanh(s)
  float ln_m = t + t;
  return fxm_fma(ln_m, 1.44269502f, (float)e + bias);
}

// WGSL pow(x, y) = exp2(y * log2(x)).
FXM_FN float fxm_pow(float x, float y) { return fxm_exp2(y * fxm_log2(x)); }

// Reduction of x by pi/2: returns r in about [-pi/4, pi/4] and the quadrant.
FXM_FN float fxm_reduce_pio2(float x, int* quadrant) {
  const float kTwoOverPi = 0.636619747f;
  const float kP1 = 1.5703125f;             // 0x1.92p+0
  const float kP2 = 4.83751297e-04f;        // 0x1.fb4p-12
  const float kP3 = 7.54978995e-08f;        // 0x1.4442d2p-24
  float k = fxm_rint(x * kTwoOverPi);
  float r = fxm_fma(-k, kP3, fxm_fma(-k, kP2, fxm_fma(-k, kP1, x)));
  float q = k - 4.0f * fxm_floor(k * 0.25f);  // k mod 4 in {0,1,2,3}
  *quadrant = (int)q;
  return r;
}

FXM_FN float fxm_sin_poly(float r) {
  float z = r * r;
  float p = fxm_fma(z, 2.75573188e-06f, -1.98412701e-04f);
  p = fxm_fma(z, p, 8.33333377e-03f);
  p = fxm_fma(z, p, -1.66666672e-01f);
  return fxm_fma(r * z, p, r);
}

FXM_FN float fxm_cos_poly(float r) {
  float z = r * r;
  float p = fxm_fma(z, -2.75573188e-07f, 2.48015876e-05f);
  p = fxm_fma(z, p, -1.38888892e-03f);
  p = fxm_fma(z, p, 4.16666679e-02f);
  p = fxm_fma(z, p, -0.5f);
  return fxm_fma(z, p, 1.0f);
}

FXM_FN float fxm_sin(float x) {
  if (x != x) return x;
  int q;
  float r = fxm_reduce_pio2(x, &q);
  float s = fxm_sin_poly(r);
  float c = fxm_cos_poly(r);
  float v = (q & 1) ? c : s;
  return (q & 2) ? -v : v;
}

FXM_FN float fxm_cos(float x) {
  if (x != x) return x;
  int q;
  float r = fxm_reduce_pio2(x, &q);
  float s = fxm_sin_poly(r);
  float c = fxm_cos_poly(r);
  float v = (q & 1) ? s : c;
  return ((q + 1) & 2) ? -v : v;
}
