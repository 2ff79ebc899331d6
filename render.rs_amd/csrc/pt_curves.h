// Curve segments (render.rs src/geometry/curves.rs): rounded cones — a
// sphere swept from p0 (radius r0) to p1 (radius r1), caps at both ends —
// intersected in f64 from the f32 record and the f32 object-space ray.
//
// A hair segment is thin next to its distance from the ray origin, and the
// quadric's constant term cancels terms of order |o - p0|^2 |p1 - p0|^2
// down to |p1 - p0|^2 r^2: (|o - p0| / r)^2 ~ 1e6 for a 5 mm strand seen
// from 5 m, more than f32 keeps. The test therefore runs in f64 (half the
// f32 rate on CDNA4, and a small share of a traversal's work), where the
// f32 inputs' differences are exact and the roots keep ~1e-10 relative
// error, matching render.rs's own f64 arithmetic.
#pragma once
#include <hip/hip_runtime.h>

#include "pt_device.h"
#include "pt_types.h"

namespace rrhip {
namespace pt {

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3of(V3 a) { return {(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ D3 d3of(const float* a) { return {(double)a[0], (double)a[1], (double)a[2]}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Cap sphere (curves.rs:171-189): the first root, or the second when the
// first is at or behind 1e-9.
__device__ __forceinline__ bool cap_hit(D3 o, D3 d, D3 c, double r, double tmax, double& t) {
  const D3 oc = o - c;
  const double a = dotd(d, d), b = dotd(oc, d), cc = dotd(oc, oc) - r * r;
  const double disc = b * b - a * cc;
  if (disc < 0.0) return false;
  const double s = sqrt(disc);
  double x = (-b - s) / a;
  if (x <= 1e-9) x = (-b + s) / a;
  if (x <= 1e-9 || x >= tmax) return false;
  t = x;
  return true;
}

// curves.rs:112-169 (after Quilez): the body's quadratic, whose root
// (-k1 - sqrt(h)) / k2 is accepted for 1e-9 < t < tmax and 0 < y < d2
// with y = (p1 - p0) . (o + t d - p0); else the start cap where y <= 0,
// then the end cap below it where y >= d2; a zero-length segment is a
// sphere of the larger radius about p0. part: 0 body, 1 start cap or
// sphere, 2 end cap (for the normal).
__device__ __forceinline__ bool cone_hit(const CurveRec& c, D3 o, D3 d, double tmax, double& t, int& part) {
  const D3 pa = d3of(c.p0), pb = d3of(c.p1);
  const double ra = (double)c.r0, rb = (double)c.r1;
  const D3 ba = pb - pa, oa = o - pa;
  const double m0 = dotd(ba, ba);
  if (m0 < 1e-24) {
    part = 1;
    return cap_hit(o, d, pa, fmax(ra, rb), tmax, t);
  }
  const double rr = ra - rb;
  const double m1 = dotd(ba, oa), m2 = dotd(ba, d), m3 = dotd(d, oa), m5 = dotd(oa, oa), dd = dotd(d, d);
  const double d2 = m0 - rr * rr;
  const double k2 = d2 * dd - m2 * m2;
  const double k1 = d2 * m3 - m1 * m2 + m2 * rr * ra;
  const double k0 = d2 * m5 - m1 * m1 + m1 * rr * ra * 2.0 - m0 * ra * ra;
  if (fabs(k2) > 1e-24) {
    const double h = k1 * k1 - k0 * k2;
    if (h >= 0.0) {
      const double x = (-k1 - sqrt(h)) / k2;
      const double y = m1 + x * m2;
      if (x > 1e-9 && x < tmax && y > 0.0 && y < d2) {
        t = x;
        part = 0;
        return true;
      }
    }
  }
  bool found = false;
  double x;
  if (cap_hit(o, d, pa, ra, tmax, x) && m1 + x * m2 <= 0.0) {
    t = x;
    part = 1;
    found = true;
  }
  if (cap_hit(o, d, pb, rb, found ? t : tmax, x) && m1 + x * m2 >= d2) {
    t = x;
    part = 2;
    found = true;
  }
  return found;
}

// Outward object-space normal of a hit at parametric t (curves.rs:147,
// 188): the body's d2 q - (ba . q) ba with q = hit - p0, or the cap
// sphere's radial direction.
__device__ __forceinline__ D3 cone_normal(const CurveRec& c, D3 o, D3 d, double t, int part) {
  const D3 pa = d3of(c.p0), pb = d3of(c.p1);
  const D3 q = o + d * t - pa;
  D3 n;
  if (part == 0) {
    const D3 ba = pb - pa;
    const double rr = (double)c.r0 - (double)c.r1;
    const double d2 = dotd(ba, ba) - rr * rr;
    n = q * d2 - ba * dotd(ba, q);
  } else {
    n = part == 1 ? q : o + d * t - pb;
  }
  const double l = sqrt(dotd(n, n));
  return l > 1e-300 ? n * (1.0 / l) : D3{0.0, 0.0, 1.0};
}

// Conservative f32 rejection ahead of the f64 test: every surface point
// of the segment lies within max(r0, r1) / cos(alpha) of its axis line
// (sin(alpha) = |r0 - r1| / |p1 - p0|; the cone's widest cross-section
// within the accepted body range, the caps within max(r0, r1)), so a ray
// whose line passes farther from the axis line misses. The distance is
// |w . n| / |n| with w = p0 - o, n = d x (p1 - p0); the slack gamma(32)
// (|p0| + |o| + r) |d| |p1 - p0| covers its f32 rounding (w's subtraction,
// the cross and dot products, |n|). Parallel rays (n ~ 0) and zero-length
// or fully tapered segments always pass on to the f64 test.
__device__ __forceinline__ bool curve_may_hit(const CurveRec& c, V3 o, V3 d) {
  const V3 p0 = v3(c.p0);
  const V3 a = v3(c.p1) - p0;
  const float aa = dot(a, a), rr = c.r0 - c.r1, den = aa - rr * rr;
  if (!(aa > 1e-12f) || !(den > 0.f)) return true;
  const V3 n = cross(d, a);
  const float rmax = fmaxf(c.r0, c.r1) * sqrtf(aa / den) * 1.0001f;
  const float slack = err_gamma(32) * (len(p0) + len(o) + rmax) * len(d) * sqrtf(aa);
  return fabsf(dot(p0 - o, n)) <= rmax * len(n) + slack;
}

// Closest-hit test inside the BLAS traversal: t as f32, strictly below tmax.
__device__ __forceinline__ bool curve_seg(const CurveRec& c, float ox, float oy, float oz, float dx, float dy,
                                          float dz, float tmax, float& t) {
  if (!curve_may_hit(c, v3(ox, oy, oz), v3(dx, dy, dz))) return false;
  double x;
  int part;
  if (!cone_hit(c, D3{ox, oy, oz}, D3{dx, dy, dz}, (double)tmax, x, part)) return false;
  const float tf = (float)x;
  if (!(tf < tmax)) return false;
  t = tf;
  return true;
}

// Unit vector <-> two floats in [-1, 1] (octahedral map); the path
// buffers carry a curve hit's tangent this way in two spare lanes.
__device__ __forceinline__ float2 oct_encode(V3 v) {
  const float s = fabsf(v.x) + fabsf(v.y) + fabsf(v.z);
  float px = v.x / s, py = v.y / s;
  if (v.z < 0.f) {
    const float qx = (1.f - fabsf(py)) * (px >= 0.f ? 1.f : -1.f);
    const float qy = (1.f - fabsf(px)) * (py >= 0.f ? 1.f : -1.f);
    px = qx;
    py = qy;
  }
  return make_float2(px, py);
}
__device__ __forceinline__ V3 oct_decode(float px, float py) {
  V3 v = v3(px, py, 1.f - fabsf(px) - fabsf(py));
  if (v.z < 0.f) {
    const float qx = (1.f - fabsf(py)) * (px >= 0.f ? 1.f : -1.f);
    const float qy = (1.f - fabsf(px)) * (py >= 0.f ? 1.f : -1.f);
    v.x = qx;
    v.y = qy;
  }
  return normalize(v);
}
// The lane value of hits without a tangent (outside the map's range).
constexpr float kNoTangent = 4.f;

}  // namespace pt
}  // namespace rrhip
