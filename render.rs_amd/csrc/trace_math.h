// Geometric core of the ray-query kernels, shared by host and device
// code: the conservative slab test, the error-bounded f32 triangle
// prefilter and the f64 triangle test (see trace.hip for the argument
// that these reproduce the f64 CPU decisions).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "layout.h"
#include "trace.h"

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

// ---- per-lane traversal state machine (shared by the kernels and by host
// checks of the control flow) ------------------------------------------------

constexpr uint32_t kDone = 0xFFFFFFFFu;     // traversal finished
constexpr uint32_t kNeedPop = 0xFFFFFFFEu;  // next action: pop the stack
constexpr uint32_t kLeafBit = 0x80000000u;  // stack entry is a leaf range
constexpr uint32_t kLeafFirstMask = 0x0FFFFFFFu;  // leaf entry: first triangle slot (28 bits)

RRHIP_HD inline Ray32 load_ray(const RayRec* __restrict__ rays, uint32_t i) {
  const RayRec q = rays[i];
  Ray32 r;
  r.ox = q.ox;
  r.oy = q.oy;
  r.oz = q.oz;
  r.tmin = q.tmin;
  r.dx = q.dx;
  r.dy = q.dy;
  r.dz = q.dz;
  r.tmax = q.tmax;
  r.ix = safe_rcp(r.dx);
  r.iy = safe_rcp(r.dy);
  r.iz = safe_rcp(r.dz);
  return r;
}

// Per-lane traversal state.
struct Lane {
  Ray32 r;
  double best_t;
  float tmax_f;  // f32 culling bound, never below best_t
  uint32_t best;
  double best_u, best_v;
  bool occ;
  uint32_t node;        // interior node to visit, kNeedPop or kDone
  uint32_t lref, lcnt;  // pending leaf range (lcnt == 0: none)
  uint32_t mask4;       // four-wide: children of `node` still to visit
  int sp;
  uint32_t n_nodes, n_tris;
};

RRHIP_HD inline void lane_start(Lane& L, const RayRec* __restrict__ rays, uint32_t id) {
  L.r = load_ray(rays, id);
  L.best_t = (double)L.r.tmax;
  L.tmax_f = L.r.tmax;
  L.best = kNoHit;
  L.best_u = 0.0;
  L.best_v = 0.0;
  L.occ = false;
  L.node = 0;
  L.mask4 = 0xFu;
  L.lcnt = 0;
  L.lref = 0;
  L.sp = 0;
  L.n_nodes = 0;
  L.n_tris = 0;
}

template <bool ANY>
RRHIP_HD inline void test_triangle(Lane& L, const TriRec& tr) {
  if (!may_hit(L.r, tr, L.tmax_f)) return;
  double tt, uu, vv;
  if (!hit_f64(L.r, tr, (double)L.r.tmin, L.best_t, tt, uu, vv)) return;
  if (ANY) {
    if (tt < L.best_t) L.occ = true;  // the any-hit bound is exclusive (mesh.rs:141)
    return;
  }
  // Strictly closer wins; an equal t only replaces a recorded hit with a
  // higher triangle id (the ray's own tmax stays exclusive).
  if (tt < L.best_t || (tt == L.best_t && L.best != kNoHit && tr.prim < L.best)) {
    L.best_t = tt;
    L.best = tr.prim;
    L.best_u = uu;
    L.best_v = vv;
    L.tmax_f = f32_round_up(tt);
  }
}

template <bool ANY, bool COUNT>
RRHIP_HD inline void lane_finish(const Lane& L, uint32_t id, HitRec* __restrict__ hits,
                                            uint8_t* __restrict__ occluded,
                                            uint32_t* __restrict__ counts) {
  if (ANY) {
    occluded[id] = L.occ ? 1 : 0;
  } else {
    HitRec h;
    if (L.best != kNoHit) {
      h.t = (float)L.best_t;
      h.prim = L.best;
      h.u = (float)L.best_u;
      h.v = (float)L.best_v;
    } else {
      h.t = INFINITY;
      h.prim = kNoHit;
      h.u = 0.0f;
      h.v = 0.0f;
    }
    hits[id] = h;
  }
  if (COUNT) {
    counts[2 * (size_t)id] = L.n_nodes;
    counts[2 * (size_t)id + 1] = L.n_tris;
  }
}

// One interior step: a pop when the lane needs one, then the visit of the
// interior node it yields (a popped leaf entry becomes the pending leaf
// instead). Leaf children become the pending leaf (or a stacked leaf
// entry); interior children continue nearest-first. At most one entry is
// pushed per visit, so the stack never holds more entries than the tree
// has levels.
template <int STRIDE, bool COUNT>
RRHIP_HD inline void interior_step(Lane& L, const Node2* __restrict__ nodes,
                                              uint32_t* stk) {
  if (L.node == kNeedPop) {
    if (L.sp == 0) {
      L.node = kDone;
      return;
    }
    L.sp--;
    const uint32_t e = stk[L.sp * STRIDE];
    if (e & kLeafBit) {
      L.lref = e & kLeafFirstMask;
      L.lcnt = ((e >> 28) & 7u) + 1u;
      return;
    }
    L.node = e;
  }
  const Node2 nd = nodes[L.node];
  if (COUNT) L.n_nodes++;
  bool h0, h1;
  float t0, t1;
  slab2(nd, L.r, L.tmax_f, h0, h1, t0, t1);
  const bool leaf0 = h0 && nd.cnt0 > 0, leaf1 = h1 && nd.cnt1 > 0;
  const bool in0 = h0 && nd.cnt0 == 0, in1 = h1 && nd.cnt1 == 0;
  if (leaf0) {
    L.lref = nd.ref0;
    L.lcnt = nd.cnt0;
  }
  if (leaf1) {
    if (!leaf0) {
      L.lref = nd.ref1;
      L.lcnt = nd.cnt1;
    } else {
      stk[L.sp * STRIDE] = kLeafBit | ((nd.cnt1 - 1u) << 28) | nd.ref1;
      L.sp++;
    }
  }
  if (in0 && in1) {
    const bool first0 = t0 <= t1;
    stk[L.sp * STRIDE] = first0 ? nd.ref1 : nd.ref0;
    L.sp++;
    L.node = first0 ? nd.ref0 : nd.ref1;
  } else if (in0) {
    L.node = nd.ref0;
  } else if (in1) {
    L.node = nd.ref1;
  } else {
    L.node = kNeedPop;
  }
}

// One four-wide interior step (a pop when needed, then the visit of the
// popped or current node). Children still in L.mask4
// are slab-tested (the arithmetic of slab2); the nearest hit leaf becomes
// the pending leaf and the nearest hit interior child the next node
// (min reductions over register values: no sort, no dynamic indexing);
// when hit children remain the node is pushed back with their mask, to be
// re-tested against the then tighter bound when popped, so each pop takes
// the nearest remaining children again.
template <int STRIDE, bool COUNT>
RRHIP_HD inline void interior_step4(Lane& L, const Node4* __restrict__ nodes, uint32_t* stk) {
  if (L.node == kNeedPop) {
    if (L.sp == 0) {
      L.node = kDone;
      return;
    }
    L.sp--;
    const uint32_t e = stk[L.sp * STRIDE];
    L.node = e & kNode4Mask;
    L.mask4 = e >> 27;
  }
  const uint32_t self = L.node;
  const Node4 nd = nodes[self];
  if (COUNT) L.n_nodes++;
  const Ray32& r = L.r;
  uint32_t hit = 0;
  float best_leaf_t = INFINITY, best_node_t = INFINITY;
  int best_leaf = -1, best_node = -1;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float a0 = (nd.lox[k] - r.ox) * r.ix, b0 = (nd.hix[k] - r.ox) * r.ix;
    const float a1 = (nd.loy[k] - r.oy) * r.iy, b1 = (nd.hiy[k] - r.oy) * r.iy;
    const float a2 = (nd.loz[k] - r.oz) * r.iz, b2 = (nd.hiz[k] - r.oz) * r.iz;
    const float tn = fmaxf(fmaxf(fminf(a0, b0), fminf(a1, b1)), fmaxf(fminf(a2, b2), r.tmin));
    const float tf = fminf(fminf(fmaxf(a0, b0), fmaxf(a1, b1)), fmaxf(a2, b2)) * kSlabUp;
    const bool h = ((L.mask4 >> k) & 1u) && nd.cnt[k] != kEmpty4 && tn <= fminf(tf, L.tmax_f);
    hit |= (h ? 1u : 0u) << k;
    const bool leaf = nd.cnt[k] > 0;
    if (h && leaf && (best_leaf < 0 || tn < best_leaf_t)) {
      best_leaf = k;
      best_leaf_t = tn;
    }
    if (h && !leaf && (best_node < 0 || tn < best_node_t)) {
      best_node = k;
      best_node_t = tn;
    }
  }
  uint32_t next = kNeedPop, lref = 0, lcnt = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (k == best_leaf) {
      lref = nd.ref[k];
      lcnt = nd.cnt[k];
    }
    if (k == best_node) next = nd.ref[k];
  }
  if (best_leaf >= 0) {
    L.lref = lref;
    L.lcnt = lcnt;
    hit &= ~(1u << best_leaf);
  }
  if (best_node >= 0) hit &= ~(1u << best_node);
  if (hit) {
    stk[L.sp * STRIDE] = (hit << 27) | self;
    L.sp++;
  }
  L.node = next;
  L.mask4 = 0xFu;
}

}  // namespace rrhip
