// Vector and quaternion operations used by the gfx950 kernels
// (spatial_cell_sim/src/math.cuh:15-339 and the componentwise operators of
// CUDA's helper_math.h it relies on).  Expression order follows the reference
// term by term; with contraction disabled every result is reproducible on any
// IEEE-754 target.
#pragma once
#include "sc_fmath.h"

#pragma clang fp contract(off)

namespace scv {

struct v3 { float x, y, z; };
struct v4 { float x, y, z, w; };

SC_HD v3 m3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
SC_HD v4 m4(float x, float y, float z, float w) { v4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }

SC_HD v3 add(v3 a, v3 b) { return m3(a.x + b.x, a.y + b.y, a.z + b.z); }
SC_HD v3 sub(v3 a, v3 b) { return m3(a.x - b.x, a.y - b.y, a.z - b.z); }
SC_HD v3 neg(v3 a) { return m3(-a.x, -a.y, -a.z); }
SC_HD v3 mul(v3 a, float b) { return m3(a.x * b, a.y * b, a.z * b); }
SC_HD v3 smul(float b, v3 a) { return m3(b * a.x, b * a.y, b * a.z); }
SC_HD v3 div(v3 a, float b) { return m3(a.x / b, a.y / b, a.z / b); }
SC_HD v4 mul4(v4 a, float b) { return m4(a.x * b, a.y * b, a.z * b, a.w * b); }
SC_HD v4 div4(v4 a, float b) { return m4(a.x / b, a.y / b, a.z / b, a.w / b); }

SC_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SC_HD float dot4(v4 a, v4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
SC_HD v3 cross(v3 a, v3 b) { return m3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
SC_HD float len(v3 a) { return scm::sqrt_(dot(a, a)); }
SC_HD float len4(v4 a) { return scm::sqrt_(dot4(a, a)); }
SC_HD v3 nrm(v3 a) { float inv = 1.0f / scm::sqrt_(dot(a, a)); return mul(a, inv); }
SC_HD v4 nrm4(v4 a) { float inv = 1.0f / scm::sqrt_(dot4(a, a)); return mul4(a, inv); }
SC_HD float clampf(float f, float a, float b) { return scm::fmax_(a, scm::fmin_(f, b)); }

#define SCV_PI 3.1415926535f
SC_HD v3 up3() { return m3(0.0f, 1.0f, 0.0f); }
SC_HD v4 qid() { return m4(0.0f, 0.0f, 0.0f, 1.0f); }

// |a|^2 with the reference's own summation (math.cuh:15-18)
SC_HD float normsq(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
SC_HD float normsq4(v4 a) { return a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w; }
SC_HD float norm(v3 a) { return scm::sqrt_(normsq(a)); }

SC_HD v3 safe_cross(v3 a, v3 b) {
  v3 c = cross(a, b);
  if (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f) return m3(0.0f, 0.0f, 1.0f);
  return nrm(c);
}
SC_HD v4 qmul(v4 a, v4 b) {
  float cx = a.y * b.z - a.z * b.y;
  float cy = a.z * b.x - a.x * b.z;
  float cz = a.x * b.y - a.y * b.x;
  float d = a.x * b.x + a.y * b.y + a.z * b.z;
  return m4(a.x * b.w + b.x * a.w + cx, a.y * b.w + b.y * a.w + cy, a.z * b.w + b.z * a.w + cz, a.w * b.w - d);
}
SC_HD float angle_vv(v3 a, v3 b) { return scm::acos_(clampf(dot(a, b) / (norm(a) * norm(b)), -1.0f, 1.0f)); }
SC_HD v4 quat_axis_angle(v3 axis, float ang) {
  float s = scm::sin_(ang * 0.5f);
  float c = scm::cos_(ang * 0.5f);
  return nrm4(m4(axis.x * s, axis.y * s, axis.z * s, c));
}
SC_HD v4 quat_from_to(v3 a, v3 b) { return quat_axis_angle(safe_cross(a, b), angle_vv(a, b)); }
SC_HD v4 qinv(v4 a) {
  float inv = 1.0f / normsq4(a);
  return m4(-a.x * inv, -a.y * inv, -a.z * inv, a.w * inv);
}
SC_HD v4 slerp(v4 a, v4 b, float t) {
  float c = a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
  bool flip = false;
  if (c < 0.0f) { flip = true; c = -c; }
  float s1, s2;
  if (c > (1.0f - 1e-6f)) {
    s1 = 1.0f - t;
    s2 = flip ? -t : t;
  } else {
    float om = scm::acos_(c);
    float inv = 1 / scm::sin_(om);
    s1 = scm::sin_((1.0f - t) * om) * inv;
    s2 = flip ? -scm::sin_(t * om) * inv : scm::sin_(t * om) * inv;
  }
  return m4(s1 * a.x + s2 * b.x, s1 * a.y + s2 * b.y, s1 * a.z + s2 * b.z, s1 * a.w + s2 * b.w);
}
SC_HD v4 shoemake(float u, float v, float w) {
  float su = scm::sqrt_(u), su1 = scm::sqrt_(1 - u);
  return m4(su1 * scm::sin_(2 * SCV_PI * v), su1 * scm::cos_(2 * SCV_PI * v),
            su * scm::sin_(2 * SCV_PI * w), su * scm::cos_(2 * SCV_PI * w));
}
SC_HD v3 rotate(v3 a, v4 q) {
  float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  float wx2 = q.w * x2, wy2 = q.w * y2, wz2 = q.w * z2;
  float xx2 = q.x * x2, xy2 = q.x * y2, xz2 = q.x * z2;
  float yy2 = q.y * y2, yz2 = q.y * z2, zz2 = q.z * z2;
  return m3(a.x * (1.0f - yy2 - zz2) + a.y * (xy2 - wz2) + a.z * (xz2 + wy2),
            a.x * (xy2 + wz2) + a.y * (1.0f - xx2 - zz2) + a.z * (yz2 - wx2),
            a.x * (xz2 - wy2) + a.y * (yz2 + wx2) + a.z * (1.0f - xx2 - yy2));
}
// R(q) * (0, 1, 0): the three products that survive for a = +y
SC_HD v3 rotate_up(v4 q) { return rotate(up3(), q); }

}  // namespace scv
