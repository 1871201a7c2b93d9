// ORACLE -- test infrastructure only (parity checker and CPU baseline); no
// product code links this.  Serial restatement of the vector/quaternion math
// of spatial_cell_sim/src/math.cuh plus the helper_math.h operators it relies
// on (componentwise float3/float4 arithmetic, dot, cross, length, normalize,
// clamp), written from their mathematical definitions.  Scalar transcendentals
// come from sc_fmath.h so that host and device round identically.
//
// `real` is float, the arithmetic of the reference and of the device code.
// Built with ORACLE_FP64 defined, every real quantity of the step is a double
// and the scalar functions are the C library's: the same steps in double
// precision, which bounds the rounding error of the float path.  Random
// draws (uniforms, normals) and the constants' values are float in both.
#pragma once
#include <cmath>

#include "../spatial-cell_amd/csrc/sc_fmath.h"

#ifdef ORACLE_FP64
namespace scm {
static inline double sqrt_(double x) { return std::sqrt(x); }
static inline double acos_(double x) { return std::acos(x); }
static inline double sin_(double x) { return std::sin(x); }
static inline double cos_(double x) { return std::cos(x); }
static inline double pow_(double x, double y) { return std::pow(x, y); }
static inline double floor_(double x) { return std::floor(x); }
static inline double fabs_(double x) { return std::fabs(x); }
static inline double fmax_(double a, double b) { return std::fmax(a, b); }
static inline double fmin_(double a, double b) { return std::fmin(a, b); }
// a double beside a float literal: the double overload, exactly
static inline double fmax_(double a, float b) { return std::fmax(a, (double)b); }
static inline double fmax_(float a, double b) { return std::fmax((double)a, b); }
static inline double fmin_(double a, float b) { return std::fmin(a, (double)b); }
static inline double fmin_(float a, double b) { return std::fmin((double)a, b); }
static inline double pow_(double x, float y) { return std::pow(x, (double)y); }
static inline double log_(double x) { return std::log(x); }
static inline double log2_(double x) { return std::log2(x); }
static inline double exp2_(double x) { return std::exp2(x); }
}  // namespace scm
#endif

namespace orc {

#ifdef ORACLE_FP64
using real = double;
#else
using real = float;
#endif

struct float3 { real x, y, z; };
struct float4 { real x, y, z, w; };

static inline float3 mk3(real x, real y, real z) { float3 r = {x, y, z}; return r; }
static inline float4 mk4(real x, real y, real z, real w) { float4 r = {x, y, z, w}; return r; }

// componentwise operators (helper_math.h semantics, division per component)
static inline float3 operator+(float3 a, float3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline float3 operator-(float3 a, float3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline float3 operator-(float3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline float3 operator*(float3 a, real b) { return mk3(a.x * b, a.y * b, a.z * b); }
static inline float3 operator*(real b, float3 a) { return mk3(b * a.x, b * a.y, b * a.z); }
static inline float3 operator/(float3 a, real b) { return mk3(a.x / b, a.y / b, a.z / b); }
static inline void operator+=(float3& a, float3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
static inline void operator*=(float3& a, real b) { a.x *= b; a.y *= b; a.z *= b; }
static inline float4 operator*(float4 a, real b) { return mk4(a.x * b, a.y * b, a.z * b, a.w * b); }
static inline void operator/=(float4& a, real b) { a.x /= b; a.y /= b; a.z /= b; a.w /= b; }

static inline real dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline real dot(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
static inline float3 cross(float3 a, float3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real length(float3 v) { return scm::sqrt_(dot(v, v)); }
static inline real length(float4 v) { return scm::sqrt_(dot(v, v)); }
// normalize = v * (1/|v|); the reciprocal square root is exact here (the CUDA
// sample header uses an approximate rsqrtf, which is replaced by definition)
static inline float3 normalize(float3 v) { real inv = 1.0f / scm::sqrt_(dot(v, v)); return v * inv; }
static inline float4 normalize(float4 v) { real inv = 1.0f / scm::sqrt_(dot(v, v)); return v * inv; }
static inline real clampf(real f, real a, real b) { return scm::fmax_(a, scm::fmin_(f, b)); }
static inline float3 clamp3(float3 v, real a, real b) { return mk3(clampf(v.x, a, b), clampf(v.y, a, b), clampf(v.z, a, b)); }

static const real PI_F = 3.1415926535f;                // constants.cuh:4
static inline float3 VEC_ZERO() { return mk3(0, 0, 0); }
static inline float3 VEC_UP() { return mk3(0, 1, 0); }
static inline float3 VEC_RIGHT() { return mk3(1, 0, 0); }
static inline float3 VEC_FORWARD() { return mk3(0, 0, 1); }
static inline float4 QUAT_ID() { return mk4(0, 0, 0, 1); }

// math.cuh:15-33
static inline real normsq(float3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline real normsq(float4 a) { return a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w; }
static inline real norm(float3 a) { return scm::sqrt_(normsq(a)); }

// math.cuh:90-96: zero cross product -> +z axis
static inline float3 safeCross(float3 a, float3 b) {
  float3 c = cross(a, b);
  if (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f) return VEC_FORWARD();
  return normalize(c);
}
// math.cuh:116-129: Hamilton product
static inline float4 qmul(float4 a, float4 b) {
  real cx = a.y * b.z - a.z * b.y;
  real cy = a.z * b.x - a.x * b.z;
  real cz = a.x * b.y - a.y * b.x;
  real d = a.x * b.x + a.y * b.y + a.z * b.z;
  return mk4(a.x * b.w + b.x * a.w + cx, a.y * b.w + b.y * a.w + cy,
             a.z * b.w + b.z * a.w + cz, a.w * b.w - d);
}
// math.cuh:131-136: clamped angle between vectors
static inline real angle(float3 a, float3 b) {
  return scm::acos_(clampf(dot(a, b) / (norm(a) * norm(b)), -1.0f, 1.0f));
}
// math.cuh:138-141: rotation angle of a quaternion, unclamped
static inline real angleq(float4 a) { return 2.0f * scm::acos_(a.w); }
// math.cuh:143-155
static inline float4 quaternion(float3 axis, real ang) {
  real s = scm::sin_(ang * 0.5f);
  real c = scm::cos_(ang * 0.5f);
  return normalize(mk4(axis.x * s, axis.y * s, axis.z * s, c));
}
// math.cuh:157-160
static inline float4 quaternionFromTo(float3 a, float3 b) { return quaternion(safeCross(a, b), angle(a, b)); }
// math.cuh:171-181
static inline float4 inverse(float4 a) {
  real invNorm = 1.0f / normsq(a);
  return mk4(-a.x * invNorm, -a.y * invNorm, -a.z * invNorm, a.w * invNorm);
}
// math.cuh:220-262: sign flip for the short arc, linear fallback near cos = 1
static inline float4 slerp(float4 a, float4 b, real t) {
  const real eps = 1e-6f;
  real cosOmega = a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
  bool flip = false;
  if (cosOmega < 0.0f) { flip = true; cosOmega = -cosOmega; }
  real s1, s2;
  if (cosOmega > (1.0f - eps)) {
    s1 = 1.0f - t;
    s2 = flip ? -t : t;
  } else {
    real omega = scm::acos_(cosOmega);
    real invSinOmega = 1 / scm::sin_(omega);
    s1 = scm::sin_((1.0f - t) * omega) * invSinOmega;
    s2 = flip ? -scm::sin_(t * omega) * invSinOmega : scm::sin_(t * omega) * invSinOmega;
  }
  return mk4(s1 * a.x + s2 * b.x, s1 * a.y + s2 * b.y, s1 * a.z + s2 * b.z, s1 * a.w + s2 * b.w);
}
// math.cuh:264-277: Shoemake's uniform rotation from three uniforms
static inline float4 random_rotation(real u, real v, real w) {
  real su = scm::sqrt_(u), su1 = scm::sqrt_(1 - u);
  return mk4(su1 * scm::sin_(2 * PI_F * v), su1 * scm::cos_(2 * PI_F * v),
             su * scm::sin_(2 * PI_F * w), su * scm::cos_(2 * PI_F * w));
}
// math.cuh:309-339: rotate a by q (rotation-matrix form)
static inline float3 transform_vector(float3 a, float4 q) {
  real x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  real wx2 = q.w * x2, wy2 = q.w * y2, wz2 = q.w * z2;
  real xx2 = q.x * x2, xy2 = q.x * y2, xz2 = q.x * z2;
  real yy2 = q.y * y2, yz2 = q.y * z2, zz2 = q.z * z2;
  return mk3(a.x * (1.0f - yy2 - zz2) + a.y * (xy2 - wz2) + a.z * (xz2 + wy2),
             a.x * (xy2 + wz2) + a.y * (1.0f - xx2 - zz2) + a.z * (yz2 - wx2),
             a.x * (xz2 - wy2) + a.y * (yz2 + wx2) + a.z * (1.0f - xx2 - yy2));
}

}  // namespace orc
