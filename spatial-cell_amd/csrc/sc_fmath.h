// Scalar primitives that must give the same bits on the host (g++) and on
// gfx950 (hipcc): min/max with fixed NaN rules, sin/cos/acos/log/exp2/pow
// written from their series definitions, the Philox4x32-10 counter-based
// generator (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
// SC'11) with its uniform and Box-Muller transforms, and an IEEE
// float -> binary16 round-to-nearest-even conversion.
//
// Every function uses only IEEE-754 +, -, *, /, sqrt, floor and integer
// operations, so with floating-point contraction disabled (-ffp-contract=off on
// both compilers) host and device produce identical results.  The reference
// (spatial_cell_sim/src/math.cuh, dynamics.cuh) calls CUDA's sinf/cosf/acosf/
// powf and cuRAND; those are replaced here by definition.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SC_HD __host__ __device__ __forceinline__
#else
#define SC_HD static inline
#endif

#pragma clang fp contract(off)

namespace scm {

SC_HD uint32_t f2u(float f) { union { float f; uint32_t u; } c; c.f = f; return c.u; }
SC_HD float u2f(uint32_t u) { union { float f; uint32_t u; } c; c.u = u; return c.f; }
SC_HD float qnan() { return u2f(0x7fc00000u); }

// fminf/fmaxf semantics (a NaN operand yields the other operand), written as
// explicit selects so that neither compiler substitutes an instruction with a
// different signed-zero rule.
SC_HD float fmin_(float a, float b) { if (a != a) return b; if (b != b) return a; return (b < a) ? b : a; }
SC_HD float fmax_(float a, float b) { if (a != a) return b; if (b != b) return a; return (b > a) ? b : a; }
SC_HD double dmin_(double a, double b) { if (a != a) return b; if (b != b) return a; return (b < a) ? b : a; }
SC_HD double dmax_(double a, double b) { if (a != a) return b; if (b != b) return a; return (b > a) ? b : a; }
SC_HD float fabs_(float a) { return u2f(f2u(a) & 0x7fffffffu); }
SC_HD double dabs_(double a) { return a < 0.0 ? -a : (a == 0.0 ? 0.0 : a); }
SC_HD float sqrt_(float a) { return __builtin_sqrtf(a); }
SC_HD float floor_(float a) { return __builtin_floorf(a); }

// ---- sin / cos ------------------------------------------------------------
// x = k*pi/2 + r with a three-part Cody-Waite split of pi/2 (first part has 8
// significant bits, second 12), so k*C1 and k*C2 are exact for |k| < 2^12.
// r in [-pi/4, pi/4]; Taylor series of sin to r^9 and cos to r^10 (truncation
// below 3e-9 relative).  Arguments beyond 1e6 return NaN (never produced by
// the simulation: every angle it passes is below 2*pi in magnitude).
#define SC_PIO2_1 0x1.92p+0f
#define SC_PIO2_2 0x1.fb4p-12f
#define SC_PIO2_3 0x1.4442d2p-24f
#define SC_TWO_OVER_PI 0x1.45f306p-1f

SC_HD float sin_poly_(float r) {
  float r2 = r * r;
  float p = -0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f;   // -1/7!, 1/9!
  p = 0x1.111112p-7f + r2 * p;                         // 1/5!
  p = -0x1.555556p-3f + r2 * p;                        // -1/3!
  return r + r * (r2 * p);
}
SC_HD float cos_poly_(float r) {
  float r2 = r * r;
  float p = 0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f;   // 1/8!, -1/10!
  p = -0x1.6c16c2p-10f + r2 * p;                       // -1/6!
  p = 0x1.555556p-5f + r2 * p;                         // 1/4!
  return (1.0f - 0.5f * r2) + (r2 * r2) * p;
}
// quadrant-shifted evaluation: returns sin(x + shift*pi/2)
SC_HD float sincos_q_(float x, int shift) {
  if (!(fabs_(x) <= 1.0e6f)) return qnan();
  float k = floor_(x * SC_TWO_OVER_PI + 0.5f);
  float r = ((x - k * SC_PIO2_1) - k * SC_PIO2_2) - k * SC_PIO2_3;
  int q = ((int)k + shift) & 3;
  float v = (q & 1) ? cos_poly_(r) : sin_poly_(r);
  return (q & 2) ? -v : v;
}
SC_HD float sin_(float x) { return sincos_q_(x, 0); }
SC_HD float cos_(float x) { return sincos_q_(x, 1); }

// ---- acos -------------------------------------------------------------------
// asin(t) for |t| <= 1/2 by its Taylor series sum_n (2n)!/(4^n (n!)^2 (2n+1)) t^(2n+1)
// up to n = 10; acos(x) = pi/2 - asin(x) for |x| <= 1/2, otherwise
// 2*asin(sqrt((1-|x|)/2)) reflected for negative x.  |x| > 1 or NaN -> NaN.
SC_HD float asin_small_(float t) {
  float z = t * t;
  float p = 0x1.12ef3cp-7f;                 // 46189/5505024
  p = 0x1.3fde50p-7f + z * p;              // 12155/1245184
  p = 0x1.7a8788p-7f + z * p;              // 6435/557056
  p = 0x1.c9999ap-7f + z * p;              // 143/10240
  p = 0x1.1c4ec4p-6f + z * p;              // 231/13312
  p = 0x1.6e8ba2p-6f + z * p;              // 63/2816
  p = 0x1.f1c71cp-6f + z * p;              // 35/1152
  p = 0x1.6db6dcp-5f + z * p;              // 5/112
  p = 0x1.333334p-4f + z * p;              // 3/40
  p = 0x1.555556p-3f + z * p;              // 1/6
  return t + t * (z * p);
}
SC_HD float acos_(float x) {
  if (!(x >= -1.0f && x <= 1.0f)) return qnan();
  float ax = fabs_(x);
  if (ax <= 0.5f) return 0x1.921fb6p+0f - asin_small_(x);
  float s = sqrt_((1.0f - ax) * 0.5f);
  float a = 2.0f * asin_small_(s);
  return (x > 0.0f) ? a : 0x1.921fb6p+1f - a;
}

// ---- log / exp2 / pow -------------------------------------------------------
// x = 2^e * m, m in [sqrt(1/2), sqrt(2)); ln(m) = 2*atanh(s), s = (m-1)/(m+1),
// |s| <= 0.1716, series to s^11.  ln2 is split (16 significant bits + rest).
SC_HD float ln_mant_(float x, int* e_out) {
  uint32_t u = f2u(x);
  int e = 0;
  if (u < 0x00800000u) { x = x * 0x1p24f; u = f2u(x); e = -24; }   // subnormal input
  e += (int)(u >> 23) - 127;
  float m = u2f((u & 0x007fffffu) | 0x3f800000u);                   // [1, 2)
  if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = 0x1.745d18p-4f;                // 1/11
  p = 0x1.c71c72p-4f + s2 * p;             // 1/9
  p = 0x1.24924ap-3f + s2 * p;             // 1/7
  p = 0x1.99999ap-3f + s2 * p;             // 1/5
  p = 0x1.555556p-2f + s2 * p;             // 1/3
  *e_out = e;
  return 2.0f * s + 2.0f * s * (s2 * p);
}
SC_HD float log_(float x) {
  if (x != x || x < 0.0f) return qnan();
  if (x == 0.0f) return -u2f(0x7f800000u);
  if (f2u(x) == 0x7f800000u) return x;
  int e;
  float lm = ln_mant_(x, &e);
  float fe = (float)e;
  return fe * 0x1.62e4p-1f + (fe * 0x1.7f7d1cp-20f + lm);
}
SC_HD float log2_(float x) {
  if (x != x || x < 0.0f) return qnan();
  if (x == 0.0f) return -u2f(0x7f800000u);
  if (f2u(x) == 0x7f800000u) return x;
  int e;
  float lm = ln_mant_(x, &e);
  return (float)e + lm * 0x1.715476p+0f;   // log2(e)
}
// 2^t = 2^k * e^(f*ln2), k = round(t), |f| <= 1/2, exponential series to g^7.
SC_HD float exp2_(float t) {
  if (t != t) return t;
  if (t > 128.0f) return u2f(0x7f800000u);
  if (t < -152.0f) return 0.0f;
  float k = floor_(t + 0.5f);
  float f = t - k;
  float g = f * 0x1.62e430p-1f;            // ln2
  float p = 0x1.a01a02p-13f;               // 1/7!
  p = 0x1.6c16c2p-10f + g * p;             // 1/6!
  p = 0x1.111112p-7f + g * p;              // 1/5!
  p = 0x1.555556p-5f + g * p;              // 1/4!
  p = 0x1.555556p-3f + g * p;              // 1/3!
  p = 0.5f + g * p;
  float y = 1.0f + (g + (g * g) * p);
  int ki = (int)k;
  int k1 = ki / 2, k2 = ki - k1;           // each half within the normal exponent range
  return (y * u2f((uint32_t)(k1 + 127) << 23)) * u2f((uint32_t)(k2 + 127) << 23);
}
SC_HD float pow_(float x, float y) {
  if (x == 1.0f) return 1.0f;
  if (x != x || y != y) return qnan();
  if (x == 0.0f) return (y > 0.0f) ? 0.0f : u2f(0x7f800000u);
  if (x < 0.0f) return qnan();
  return exp2_(y * log2_(x));
}

// ---- Philox4x32-10 ----------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

SC_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  for (int i = 0; i < 10; i++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    u32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
// 24 random bits -> (0, 1]; every value is exactly representable.
SC_HD float u01_(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }
// standard normal from two (0,1] uniforms (Box-Muller, cosine branch)
SC_HD float box_muller_(float u1, float u2) {
  return sqrt_(-2.0f * log_(u1)) * cos_(0x1.921fb6p+2f * u2);
}

// ---- float -> binary16, round to nearest even -------------------------------
SC_HD uint16_t f2h_(float f) {
  uint32_t x = f2u(f);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u)
    return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? (0x200u | ((ax >> 13) & 0x3ffu)) : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
  if (ax < 0x38800000u) {
    if (ax < 0x33000000u) return (uint16_t)sign;
    uint32_t e = ax >> 23;
    uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126u - e;
    uint32_t q = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t r = ax - 0x38000000u;
  uint32_t q = r >> 13;
  uint32_t rem = r & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}

}  // namespace scm
