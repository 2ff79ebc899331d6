// Geometric core of the ray-query kernels, shared by host and device
// code: the conservative slab test, the error-bounded f32 triangle
// prefilter and the f64 triangle test (see trace.hip for the argument
// that these reproduce the f64 CPU decisions).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "layout.h"

#define RRHIP_HD __host__ __device__

namespace rrhip {

constexpr float kSlabUp = 1.0f + 4.76837158203125e-7f;  // 1 + 2^-21
constexpr float kErrK = 64.0f * 5.9604644775390625e-8f;  // 64 * 2^-24

RRHIP_HD inline float safe_rcp(float d) {
  const float s = fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d;
  return 1.0f / s;  // correctly rounded
}

// Smallest float >= x.
RRHIP_HD inline float f32_round_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  return f;
}

struct Ray32 {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
  float ix, iy, iz;
};

// Conservative f32 slab test of both children of a node.
RRHIP_HD inline void slab2(const Node2& n, const Ray32& r, float tmax, bool& h0, bool& h1,
                                      float& t0, float& t1) {
  float a0 = (n.c0_lox - r.ox) * r.ix, b0 = (n.c0_hix - r.ox) * r.ix;
  float a1 = (n.c0_loy - r.oy) * r.iy, b1 = (n.c0_hiy - r.oy) * r.iy;
  float a2 = (n.c0_loz - r.oz) * r.iz, b2 = (n.c0_hiz - r.oz) * r.iz;
  float tn = fmaxf(fmaxf(fminf(a0, b0), fminf(a1, b1)), fmaxf(fminf(a2, b2), r.tmin));
  float tf = fminf(fminf(fmaxf(a0, b0), fmaxf(a1, b1)), fmaxf(a2, b2)) * kSlabUp;
  h0 = tn <= fminf(tf, tmax);
  t0 = tn;
  float c0 = (n.c1_lox - r.ox) * r.ix, d0 = (n.c1_hix - r.ox) * r.ix;
  float c1 = (n.c1_loy - r.oy) * r.iy, d1 = (n.c1_hiy - r.oy) * r.iy;
  float c2 = (n.c1_loz - r.oz) * r.iz, d2 = (n.c1_hiz - r.oz) * r.iz;
  float un = fmaxf(fmaxf(fminf(c0, d0), fminf(c1, d1)), fmaxf(fminf(c2, d2), r.tmin));
  float uf = fminf(fminf(fmaxf(c0, d0), fmaxf(c1, d1)), fmaxf(c2, d2)) * kSlabUp;
  h1 = un <= fminf(uf, tmax);
  t1 = un;
}

// f32 Moller-Trumbore with a forward error bound. Returns false only when
// the f64 test of mesh.rs:117-142 certainly rejects the triangle for any
// t in (tmin, tmax].
RRHIP_HD inline bool may_hit(const Ray32& r, const TriRec& tr, float tmax) {
  const float e1x = tr.v1x - tr.v0x, e1y = tr.v1y - tr.v0y, e1z = tr.v1z - tr.v0z;
  const float e2x = tr.v2x - tr.v0x, e2y = tr.v2y - tr.v0y, e2z = tr.v2z - tr.v0z;
  const float tvx = r.ox - tr.v0x, tvy = r.oy - tr.v0y, tvz = r.oz - tr.v0z;
  // p = d x e2
  const float p1 = r.dy * e2z, p2 = r.dz * e2y, p3 = r.dz * e2x;
  const float p4 = r.dx * e2z, p5 = r.dx * e2y, p6 = r.dy * e2x;
  const float px = p1 - p2, py = p3 - p4, pz = p5 - p6;
  const float pax = fabsf(p1) + fabsf(p2), pay = fabsf(p3) + fabsf(p4), paz = fabsf(p5) + fabsf(p6);
  const float det = e1x * px + e1y * py + e1z * pz;
  const float mdet = fabsf(e1x) * pax + fabsf(e1y) * pay + fabsf(e1z) * paz;
  const float eD = kErrK * mdet;
  if (fabsf(det) <= eD) return true;  // orientation uncertain: let f64 decide
  const float s = det > 0.0f ? 1.0f : -1.0f;
  const float ad = fabsf(det);
  const float U = s * (tvx * px + tvy * py + tvz * pz);
  const float eU = kErrK * (fabsf(tvx) * pax + fabsf(tvy) * pay + fabsf(tvz) * paz);
  if (U < -eU) return false;
  // q = tv x e1
  const float q1 = tvy * e1z, q2 = tvz * e1y, q3 = tvz * e1x;
  const float q4 = tvx * e1z, q5 = tvx * e1y, q6 = tvy * e1x;
  const float qx = q1 - q2, qy = q3 - q4, qz = q5 - q6;
  const float qax = fabsf(q1) + fabsf(q2), qay = fabsf(q3) + fabsf(q4), qaz = fabsf(q5) + fabsf(q6);
  const float V = s * (r.dx * qx + r.dy * qy + r.dz * qz);
  const float eV = kErrK * (fabsf(r.dx) * qax + fabsf(r.dy) * qay + fabsf(r.dz) * qaz);
  if (V < -eV) return false;
  if (U + V - ad > eU + eV + eD) return false;
  const float T = s * (e2x * qx + e2y * qy + e2z * qz);
  const float eT = kErrK * (fabsf(e2x) * qax + fabsf(e2y) * qay + fabsf(e2z) * qaz);
  if (r.tmin >= 0.0f && T < -eT) return false;
  if (T > tmax * (ad + eD) + eT) return false;
  return true;
}

// mesh.rs:117-142 in f64, same operation order, no contraction. Accepts
// tmin < t <= tmax.
RRHIP_HD inline bool hit_f64(const Ray32& r, const TriRec& tr, double tmin, double tmax,
                                     double& t_out, double& u_out, double& v_out) {
#pragma clang fp contract(off)
  const double ox = r.ox, oy = r.oy, oz = r.oz, dx = r.dx, dy = r.dy, dz = r.dz;
  const double v0x = tr.v0x, v0y = tr.v0y, v0z = tr.v0z;
  const double e1x = (double)tr.v1x - v0x, e1y = (double)tr.v1y - v0y, e1z = (double)tr.v1z - v0z;
  const double e2x = (double)tr.v2x - v0x, e2y = (double)tr.v2y - v0y, e2z = (double)tr.v2z - v0z;
  const double px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
  const double det = e1x * px + e1y * py + e1z * pz;
  if (fabs(det) < 1e-14) return false;
  const double inv = 1.0 / det;
  const double tvx = ox - v0x, tvy = oy - v0y, tvz = oz - v0z;
  const double u = (tvx * px + tvy * py + tvz * pz) * inv;
  if (!(u >= 0.0 && u <= 1.0)) return false;
  const double qx = tvy * e1z - tvz * e1y, qy = tvz * e1x - tvx * e1z, qz = tvx * e1y - tvy * e1x;
  const double v = (dx * qx + dy * qy + dz * qz) * inv;
  if (v < 0.0 || u + v > 1.0) return false;
  const double t = (e2x * qx + e2y * qy + e2z * qz) * inv;
  if (!(t > tmin && t <= tmax)) return false;
  t_out = t;
  u_out = u;
  v_out = v;
  return true;
}

}  // namespace rrhip
