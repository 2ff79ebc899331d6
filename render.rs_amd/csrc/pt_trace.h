// Scene traversal of the path integrator: loose objects (linear, as
// render.rs's Scene::objects) then the two-level instance hierarchy
// (Scene::intersect, src/scene/mod.rs:105-146; Instance::intersect,
// src/geometry/mesh.rs:308-381). The world ray is mapped into each
// instance's object space WITHOUT renormalising the direction, so t is the
// same parameter at both levels and one f32 bound culls both.
//
// One traversal stack per lane in LDS ([entry][thread]): TLAS node
// indices, instance entries (top bit) for TLAS leaves still to enter, and
// the current instance's BLAS nodes above `base`.
#pragma once
#include <hip/hip_runtime.h>

#include "pt_curves.h"
#include "pt_device.h"
#include "trace_math.h"

namespace rrhip {
namespace pt {

// Stack entry of a TLAS leaf still to enter: flag, count - 1 (bits 28-30),
// first instance (bits 0-27). One entry per leaf keeps the TLAS phase at
// one entry per level (two at the last), so a lane needs at most
// tlas_levels + blas_levels + 2 entries.
constexpr uint32_t kInstEntry = 0x80000000u;
constexpr uint32_t kInstFirst = 0x0FFFFFFFu;

struct MeshHit {
  float t;
  uint32_t inst;  // index into insts
  uint32_t slot;  // BLAS triangle record
  float u, v;
};

// Fetches of one traversal, for the profiling kernels: 64-byte nodes
// (TLAS and BLAS), 48-byte BLAS triangle records, instances entered (a
// 64-byte instance record and its 16-byte mesh record).
struct TraceCount {
  uint32_t nodes = 0, tris = 0, insts = 0;
};

__device__ __forceinline__ V3 xform_point(const float* m, V3 p) {
  return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
          m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
__device__ __forceinline__ V3 xform_vec(const float* m, V3 v) {
  return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
          m[8] * v.x + m[9] * v.y + m[10] * v.z};
}
// Transpose of the upper 3x3 (matrix.rs:222-227, with the inverse).
__device__ __forceinline__ V3 xform_normal(const float* m, V3 n) {
  return {m[0] * n.x + m[4] * n.y + m[8] * n.z, m[1] * n.x + m[5] * n.y + m[9] * n.z,
          m[2] * n.x + m[6] * n.y + m[10] * n.z};
}

__device__ __forceinline__ Ray32 make_ray32(V3 o, V3 d) {
  Ray32 r;
  r.ox = o.x;
  r.oy = o.y;
  r.oz = o.z;
  r.tmin = 0.f;
  r.dx = d.x;
  r.dy = d.y;
  r.dz = d.z;
  r.tmax = INFINITY;
  r.ix = safe_rcp(d.x);
  r.iy = safe_rcp(d.y);
  r.iz = safe_rcp(d.z);
  return r;
}

// mesh.rs:117-142 in f32 (1e-9 < t < tmax). Returns t, u, v.
__device__ __forceinline__ bool mesh_tri(const TriRec& tr, const Ray32& r, float tmax, float& t, float& u,
                                         float& v) {
  const V3 v0 = v3(tr.v0x, tr.v0y, tr.v0z);
  const V3 e1 = v3(tr.v1x, tr.v1y, tr.v1z) - v0, e2 = v3(tr.v2x, tr.v2y, tr.v2z) - v0;
  const V3 d = v3(r.dx, r.dy, r.dz);
  const V3 p = cross(d, e2);
  const float det = dot(e1, p);
  if (fabsf(det) < 1e-14f) return false;
  const float inv = 1.f / det;
  const V3 tv = v3(r.ox, r.oy, r.oz) - v0;
  u = dot(tv, p) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  const V3 q = cross(tv, e1);
  v = dot(d, q) * inv;
  if (v < 0.f || u + v > 1.f) return false;
  t = dot(e2, q) * inv;
  return t > 1e-9f && t < tmax;
}

// A deforming mesh's triangle at shutter time t > 0 (mesh.rs:96-111): each
// vertex lerped from the rest record a towards the shutter-close record b.
__device__ __forceinline__ TriRec lerp_tri(const TriRec& a, const TriRec& b, float t) {
  TriRec r = a;
  r.v0x = a.v0x + (b.v0x - a.v0x) * t;
  r.v0y = a.v0y + (b.v0y - a.v0y) * t;
  r.v0z = a.v0z + (b.v0z - a.v0z) * t;
  r.v1x = a.v1x + (b.v1x - a.v1x) * t;
  r.v1y = a.v1y + (b.v1y - a.v1y) * t;
  r.v1z = a.v1z + (b.v1z - a.v1z) * t;
  r.v2x = a.v2x + (b.v2x - a.v2x) * t;
  r.v2y = a.v2y + (b.v2y - a.v2y) * t;
  r.v2z = a.v2z + (b.v2z - a.v2z) * t;
  return r;
}

// The record instance I's traversal tests at slot k: the rest record, or
// for a deforming mesh at time > 0 the lerped one.
template <bool MOTION>
__device__ __forceinline__ TriRec tri_at(const PtSceneView& S, bool deform, uint32_t k, float time) {
  const TriRec a = S.btris[k];
  if (MOTION && deform && time > 0.f) return lerp_tri(a, S.btris1[k], time);
  return a;
}

// World-to-object transform of instance I at shutter time `time`
// (mesh.rs:294-305): its stored inverse, or for a moving instance the
// inverse of the element-wise lerp of its two transforms (identity when
// singular, as Matrix4::inverse().unwrap_or).
__device__ __forceinline__ void instance_inverse(const PtSceneView& S, const PtInstance& I, float time, float inv[12]) {
  if (I.motion == 0xFFFFFFFFu || !(time > 0.f)) {
#pragma unroll
    for (int k = 0; k < 12; k++) inv[k] = I.inv[k];
    return;
  }
  const PtInstMotion& M = S.inst_motion[I.motion];
  float m[12];
#pragma unroll
  for (int k = 0; k < 12; k++) m[k] = M.fwd0[k] + (M.fwd1[k] - M.fwd0[k]) * time;
  const float a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
  const float A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  const float det = a * A + b * B + c * C;
  if (!(fabsf(det) > 1e-30f)) {
#pragma unroll
    for (int k = 0; k < 12; k++) inv[k] = (k % 5 == 0) ? 1.f : 0.f;
    return;
  }
  const float r = 1.f / det;
  inv[0] = A * r;
  inv[1] = -(b * i - c * h) * r;
  inv[2] = (b * f - c * e) * r;
  inv[4] = B * r;
  inv[5] = (a * i - c * g) * r;
  inv[6] = -(a * f - c * d) * r;
  inv[8] = C * r;
  inv[9] = -(a * h - b * g) * r;
  inv[10] = (a * e - b * d) * r;
#pragma unroll
  for (int row = 0; row < 3; row++)
    inv[4 * row + 3] = -(inv[4 * row] * m[3] + inv[4 * row + 1] * m[7] + inv[4 * row + 2] * m[11]);
}

// Closest (ANY = false) or any (ANY = true) instance hit with t < tmax, for
// a ray at shutter time `time` (MOTION: the scene has moving instances;
// without it the time is ignored and the code is the static traversal).
// CURVES: some instances place curve sets, whose BLAS leaves index curve
// segment records (mesh.rs:324-343, 397-402).
template <int STRIDE, bool ANY, bool MOTION, bool COUNT = false, bool CURVES = false>
__device__ __forceinline__ bool trace_instances(const PtSceneView& S, V3 o, V3 d, float time, float tmax,
                                                uint32_t* stk, MeshHit& hit, TraceCount* tc = nullptr) {
  if (S.num_insts == 0) return false;
  const Ray32 world = make_ray32(o, d);
  Ray32 r = world;
  bool found = false;
  int sp = 0, base = -1;
  uint32_t node = 0, inst = 0;
  bool have = true;
  bool curve = false;   // the instance being traversed places a curve set
  bool deform = false;  // MOTION: the instance being traversed deforms its mesh
  for (;;) {
    if (have) {
      if (COUNT) tc->nodes++;
      const Node2 nd = S.nodes[node];
      bool h0, h1;
      float t0, t1;
      slab2(nd, r, tmax, h0, h1, t0, t1);
      const bool in0 = h0 && nd.cnt0 == 0, in1 = h1 && nd.cnt1 == 0;
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const bool hc = c == 0 ? h0 : h1;
        const uint32_t cnt = c == 0 ? nd.cnt0 : nd.cnt1;
        const uint32_t ref = c == 0 ? nd.ref0 : nd.ref1;
        if (!hc || cnt == 0) continue;
        if (base < 0) {
          stk[(sp++) * STRIDE] = kInstEntry | ((cnt - 1) << 28) | ref;
        } else {
          for (uint32_t k = ref; k < ref + cnt; k++) {
            float t, u = 0.f, v = 0.f;
            if (COUNT) tc->tris++;
            if (CURVES && curve ? curve_seg(S.csegs[k], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, tmax, t)
                                : mesh_tri(tri_at<MOTION>(S, deform, k, time), r, tmax, t, u, v)) {
              hit.t = t;
              hit.inst = inst;
              hit.slot = k;
              hit.u = u;
              hit.v = v;
              found = true;
              if (ANY) return true;
              tmax = t;
            }
          }
        }
      }
      if (in0 && in1) {
        const bool first0 = t0 <= t1;
        stk[(sp++) * STRIDE] = first0 ? nd.ref1 : nd.ref0;
        node = first0 ? nd.ref0 : nd.ref1;
      } else if (in0) {
        node = nd.ref0;
      } else if (in1) {
        node = nd.ref1;
      } else {
        have = false;
      }
      continue;
    }
    if (base >= 0 && sp == base) {  // this instance's BLAS is done
      base = -1;
      r = world;
    }
    if (sp == 0) break;
    const uint32_t e = stk[(--sp) * STRIDE];
    if (e & kInstEntry) {
      inst = e & kInstFirst;
      const uint32_t more = (e >> 28) & 7u;  // instances of the leaf after this one
      if (more) stk[(sp++) * STRIDE] = kInstEntry | ((more - 1) << 28) | (inst + 1);
      if (COUNT) tc->insts++;
      const PtInstance I = S.insts[inst];
      const uint32_t root = S.meshes[I.mesh].root;
      if (root == 0xFFFFFFFFu) continue;
      if (CURVES) curve = S.meshes[I.mesh].curves != 0;
      if (MOTION) deform = I.deform != 0;
      if (!MOTION || I.motion == 0xFFFFFFFFu) {
        r = make_ray32(xform_point(I.inv, o), xform_vec(I.inv, d));
      } else {
        float inv[12];
        instance_inverse(S, I, time, inv);
        r = make_ray32(xform_point(inv, o), xform_vec(inv, d));
      }
      base = sp;
      node = root;
    } else {
      node = e;
    }
    have = true;
  }
  return found;
}

// Shading normal and material of an instance hit (mesh.rs:199-220, 355-356,
// 427-439): interpolated or geometric local normal to world space through
// the inverse transpose, flipped to face the ray; front_face from the
// unflipped normal.
template <bool MOTION>
__device__ __forceinline__ void instance_normal(const PtSceneView& S, const MeshHit& h, V3 d, float time, V3& n,
                                                bool& front, uint32_t& material) {
  const PtInstance I = S.insts[h.inst];
  const TriRec tr = tri_at<MOTION>(S, I.deform != 0, h.slot, time);
  const PtMeshInfo mi = S.meshes[I.mesh];
  V3 ln;
  if (mi.nrm_base != 0xFFFFFFFFu && !(MOTION && I.deform)) {  // a deforming mesh: geometric normals
    const uint32_t* ix = S.mesh_idx + mi.idx_base + 3 * (size_t)tr.prim;
    const float* nb = S.mesh_nrm + 3 * (size_t)mi.nrm_base;
    const float w = 1.f - h.u - h.v;
    ln = normalize(v3(nb + 3 * (size_t)ix[0]) * w + v3(nb + 3 * (size_t)ix[1]) * h.u + v3(nb + 3 * (size_t)ix[2]) * h.v);
  } else {
    const V3 v0 = v3(tr.v0x, tr.v0y, tr.v0z);
    ln = normalize(cross(v3(tr.v1x, tr.v1y, tr.v1z) - v0, v3(tr.v2x, tr.v2y, tr.v2z) - v0));
  }
  if (!MOTION || I.motion == 0xFFFFFFFFu) {
    n = normalize(xform_normal(I.inv, ln));
  } else {
    float inv[12];
    instance_inverse(S, I, time, inv);
    n = normalize(xform_normal(inv, ln));
  }
  front = dot(n, d) < 0.f;
  if (!front) n = -n;
  material = I.material;
}

// Hit point and push of an instance hit (see Hit): the object-space point
// v0 + e1 u + e2 v on the plane mesh_tri intersects (its edges e = v - v0
// rounded as there; gamma(3) of the summed magnitudes), mapped by the
// instance's object-to-world transform with the transform's own rounding
// (Pharr et al., PBRT 3rd ed., eq. 3.10-3.11), plus gamma(4) of the world
// magnitudes for the world-to-object mapping the next ray's traversal
// applies to its origin. The geometric normal is the cross product of the
// mapped edges.
template <bool MOTION>
__device__ __forceinline__ void instance_spawn(const PtSceneView& S, const MeshHit& h, V3 d, float time, V3& p,
                                               V3& e) {
  const PtInstance& I = S.insts[h.inst];
  const bool lerped = MOTION && I.deform && time > 0.f;
  const TriRec tr = tri_at<MOTION>(S, I.deform != 0, h.slot, time);
  float m[12];
  if (!MOTION || I.motion == 0xFFFFFFFFu || !(time > 0.f)) {
    const float* f = S.inst_fwd + 12 * (size_t)h.inst;
#pragma unroll
    for (int k = 0; k < 12; k++) m[k] = f[k];
  } else {
    const PtInstMotion& M = S.inst_motion[I.motion];
#pragma unroll
    for (int k = 0; k < 12; k++) m[k] = M.fwd0[k] + (M.fwd1[k] - M.fwd0[k]) * time;
  }
  const V3 v0 = v3(tr.v0x, tr.v0y, tr.v0z);
  const V3 e1 = v3(tr.v1x, tr.v1y, tr.v1z) - v0, e2 = v3(tr.v2x, tr.v2y, tr.v2z) - v0;
  const V3 b = e1 * h.u, c = e2 * h.v;
  const V3 po = v0 + b + c;
  V3 eo = (vabs(v0) + vabs(b) + vabs(c)) * err_gamma(3);
  if (lerped) {  // each lerped vertex a + (b - a) t is within gamma(3) (2|a| + |b|) of the exact one
    const TriRec ra = S.btris[h.slot], rb = S.btris1[h.slot];
    const V3 va = vabs(v3(ra.v0x, ra.v0y, ra.v0z)) + vabs(v3(ra.v1x, ra.v1y, ra.v1z)) + vabs(v3(ra.v2x, ra.v2y, ra.v2z));
    const V3 vb = vabs(v3(rb.v0x, rb.v0y, rb.v0z)) + vabs(v3(rb.v1x, rb.v1y, rb.v1z)) + vabs(v3(rb.v2x, rb.v2y, rb.v2z));
    eo = eo + (va * 2.f + vb) * err_gamma(3);
  }
  p = xform_point(m, po);
  const float am[12] = {fabsf(m[0]), fabsf(m[1]), fabsf(m[2]),  fabsf(m[3]), fabsf(m[4]),  fabsf(m[5]),
                        fabsf(m[6]), fabsf(m[7]), fabsf(m[8]), fabsf(m[9]), fabsf(m[10]), fabsf(m[11])};
  const V3 tabs = v3(am[3], am[7], am[11]);
  const V3 err = xform_vec(am, eo) * (1.f + err_gamma(3)) + xform_point(am, vabs(po)) * err_gamma(3) +
                 (vabs(p) + tabs) * err_gamma(4);
  const V3 ng = normalize(cross(xform_vec(m, e1), xform_vec(m, e2)));
  e = bound_push(ng, err, d);
}

// A curve-set instance hit (mesh.rs:324-343, 427-439): the segment test
// repeated in f64 on the same object-space ray for its normal, mapped to
// world space through the inverse transpose and flipped to face the ray;
// the world tangent through the shutter-open transform (mesh.rs:334); with
// SPAWN the hit point o + t d (f64, rounded) mapped to world space and its
// push, bounded like instance_spawn with gamma(2) of the object-space
// magnitudes for the rounded point and the root.
template <bool MOTION, bool SPAWN>
__device__ __forceinline__ void curve_instance_hit(const PtSceneView& S, const MeshHit& h, V3 o, V3 d, float time,
                                                   Hit& best) {
  const PtInstance I = S.insts[h.inst];
  const CurveRec c = S.csegs[h.slot];
  float inv[12];
  if (!MOTION || I.motion == 0xFFFFFFFFu) {
#pragma unroll
    for (int k = 0; k < 12; k++) inv[k] = I.inv[k];
  } else {
    instance_inverse(S, I, time, inv);
  }
  const Ray32 r = make_ray32(xform_point(inv, o), xform_vec(inv, d));
  const D3 lo = {r.ox, r.oy, r.oz}, ld = {r.dx, r.dy, r.dz};
  double t;
  int part;
  if (!cone_hit(c, lo, ld, INFINITY, t, part)) {  // a grazing root lost on the repeat: the traced t
    t = (double)h.t;
    part = 0;
  }
  const D3 nl = cone_normal(c, lo, ld, t, part);
  V3 n = normalize(xform_normal(inv, v3((float)nl.x, (float)nl.y, (float)nl.z)));
  best.front = dot(n, d) < 0.f;
  if (!best.front) n = -n;
  best.n = n;
  best.material = I.material;
  const V3 axis = v3(c.p1[0] - c.p0[0], c.p1[1] - c.p0[1], c.p1[2] - c.p0[2]);
  const V3 tl = dot(axis, axis) > 1e-24f ? axis : v3(0.f, 0.f, 1.f);
  best.tan = normalize(xform_vec(S.inst_fwd + 12 * (size_t)h.inst, tl));
  if (SPAWN) {
    float m[12];
    if (!MOTION || I.motion == 0xFFFFFFFFu || !(time > 0.f)) {
      const float* f = S.inst_fwd + 12 * (size_t)h.inst;
#pragma unroll
      for (int k = 0; k < 12; k++) m[k] = f[k];
    } else {
      const PtInstMotion& M = S.inst_motion[I.motion];
#pragma unroll
      for (int k = 0; k < 12; k++) m[k] = M.fwd0[k] + (M.fwd1[k] - M.fwd0[k]) * time;
    }
    const D3 pd = lo + ld * t;
    const V3 po = v3((float)pd.x, (float)pd.y, (float)pd.z);
    const V3 eo = (vabs(po) + vabs(v3(r.dx, r.dy, r.dz)) * (float)t) * err_gamma(2);
    best.p = xform_point(m, po);
    const float am[12] = {fabsf(m[0]), fabsf(m[1]), fabsf(m[2]),  fabsf(m[3]), fabsf(m[4]),  fabsf(m[5]),
                          fabsf(m[6]), fabsf(m[7]), fabsf(m[8]), fabsf(m[9]), fabsf(m[10]), fabsf(m[11])};
    const V3 tabs = v3(am[3], am[7], am[11]);
    const V3 err = xform_vec(am, eo) * (1.f + err_gamma(3)) + xform_point(am, vabs(po)) * err_gamma(3) +
                   (vabs(best.p) + tabs) * err_gamma(4);
    best.e = bound_push(n, err, d);
  }
}

// Geometry classes of a trace instantiation: loose quadrics (their
// solvers) and curve sets (the segment test). GEO = 0 is the lean
// triangle / sphere traversal.
constexpr int kGeoQuadrics = 1, kGeoCurves = 2, kGeoSt = 4;

// Closest hit over loose objects then instances (GEO: the geometry
// classes above). SPAWN: also the winning hit's point and push (Hit::p,
// Hit::e). Hit::tan is the world tangent of a curve hit, zero otherwise.
template <int STRIDE, bool MOTION, int GEO, bool SPAWN = false, bool COUNT = false>
__device__ __forceinline__ bool scene_closest(const PtSceneView& S, V3 o, V3 d, float time, uint32_t* stk,
                                              Hit& best, TraceCount* tc = nullptr) {
  constexpr bool QUAD = (GEO & kGeoQuadrics) != 0, CURVES = (GEO & kGeoCurves) != 0;
  bool found = closest_hit<QUAD>(S, o, d, best);
  best.tan = v3(0.f, 0.f, 0.f);
  best.inst = 0xFFFFFFFFu;
  MeshHit mh;
  if (trace_instances<STRIDE, false, MOTION, COUNT, CURVES>(S, o, d, time, found ? best.t : INFINITY, stk, mh,
                                                            tc)) {
    best.t = mh.t;
    best.inst = mh.inst;
    best.slot = mh.slot;
    best.bu = mh.u;
    best.bv = mh.v;
    if (CURVES && S.meshes[S.insts[mh.inst].mesh].curves) {
      curve_instance_hit<MOTION, SPAWN>(S, mh, o, d, time, best);
      return true;
    }
    instance_normal<MOTION>(S, mh, d, time, best.n, best.front, best.material);
    if (SPAWN) instance_spawn<MOTION>(S, mh, d, time, best.p, best.e);
    return true;
  }
  if (SPAWN && found) loose_spawn(S, o, d, best);
  return found;
}

template <int STRIDE, bool MOTION, int GEO, bool COUNT = false>
__device__ __forceinline__ bool scene_occluded(const PtSceneView& S, V3 o, V3 d, float time, float tmax,
                                               uint32_t* stk, TraceCount* tc = nullptr) {
  if (occluded<(GEO & kGeoQuadrics) != 0>(S, o, d, tmax)) return true;
  MeshHit mh;
  return trace_instances<STRIDE, true, MOTION, COUNT, (GEO & kGeoCurves) != 0>(S, o, d, time, tmax, stk, mh, tc);
}

}  // namespace pt
}  // namespace rrhip
