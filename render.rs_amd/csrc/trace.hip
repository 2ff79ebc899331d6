// Closest-hit and any-hit BVH traversal for gfx950 (CDNA4).
//
// Semantics follow render.rs's CPU path: `Bvh::traverse` / `Bvh::any_hit`
// (src/accel/bvh.rs:141-245) over `Mesh::intersect_triangle`
// (src/geometry/mesh.rs:113-142). Hit decisions are made so the returned
// triangle ids and t values are the ones that f64 code computes:
//
//  * node boxes are the exact f32 extents of the f32 vertices, and the
//    f32 slab test widens t_far by (1 + 2^-21) >= 1 + 2*gamma_3 (Ize 2013,
//    "Robust BVH Ray Traversal"), so no box the f64 ray enters is culled;
//  * each triangle first gets an f32 Moller-Trumbore evaluation with a
//    forward error bound (64 ulp of the summed product magnitudes); only a
//    candidate the bound cannot rule out is re-tested in f64 with the
//    operation order of mesh.rs:117-142 and FMA contraction off, which
//    yields the same bits as the f64 CPU code;
//  * equal-t candidates resolve to the lower triangle id.
//
// Layout: one ray per lane (wave64), traversal stack of node indices in
// LDS laid out [depth][lane] so every lane hits its own bank.
#include <hip/hip_runtime.h>

#include "trace.h"
#include "trace_math.h"

namespace rrhip {
namespace {

__device__ __forceinline__ Node2 load_node(const Node2* __restrict__ nodes, uint32_t i) {
  const float4* p = reinterpret_cast<const float4*>(nodes + i);
  float4 a = p[0], b = p[1], c = p[2], d = p[3];
  Node2 n;
  n.c0_lox = a.x;
  n.c0_hix = a.y;
  n.c0_loy = a.z;
  n.c0_hiy = a.w;
  n.c1_lox = b.x;
  n.c1_hix = b.y;
  n.c1_loy = b.z;
  n.c1_hiy = b.w;
  n.c0_loz = c.x;
  n.c0_hiz = c.y;
  n.c1_loz = c.z;
  n.c1_hiz = c.w;
  n.ref0 = __float_as_uint(d.x);
  n.ref1 = __float_as_uint(d.y);
  n.cnt0 = __float_as_uint(d.z);
  n.cnt1 = __float_as_uint(d.w);
  return n;
}

__device__ __forceinline__ TriRec load_tri(const TriRec* __restrict__ tris, uint32_t i) {
  const float4* p = reinterpret_cast<const float4*>(tris + i);
  float4 a = p[0], b = p[1], c = p[2];
  TriRec t;
  t.v0x = a.x;
  t.v0y = a.y;
  t.v0z = a.z;
  t.v1x = a.w;
  t.v1y = b.x;
  t.v1z = b.y;
  t.v2x = b.z;
  t.v2y = b.w;
  t.v2z = c.x;
  t.prim = __float_as_uint(c.y);
  return t;
}

template <int BLOCK, int STACK, bool ANY, bool COUNT>
__global__ __launch_bounds__(BLOCK) void trace_kernel(const Node2* __restrict__ nodes,
                                                      const TriRec* __restrict__ tris,
                                                      const RayRec* __restrict__ rays,
                                                      HitRec* __restrict__ hits,
                                                      uint8_t* __restrict__ occluded,
                                                      uint32_t* __restrict__ counts, uint32_t n) {
  __shared__ uint32_t s_stack[STACK * BLOCK];
  const uint32_t lane = threadIdx.x;
  const uint32_t i = blockIdx.x * BLOCK + lane;
  if (i >= n) return;  // no block-level synchronisation below
  uint32_t* stk = s_stack + lane;

  const float4* rp = reinterpret_cast<const float4*>(rays + i);
  const float4 ra = rp[0], rb = rp[1];
  Ray32 r;
  r.ox = ra.x;
  r.oy = ra.y;
  r.oz = ra.z;
  r.tmin = ra.w;
  r.dx = rb.x;
  r.dy = rb.y;
  r.dz = rb.z;
  r.tmax = rb.w;
  r.ix = safe_rcp(r.dx);
  r.iy = safe_rcp(r.dy);
  r.iz = safe_rcp(r.dz);

  const double tmin_d = (double)r.tmin;
  double best_t = (double)r.tmax;
  float tmax_f = r.tmax;  // f32 bound for culling, never below best_t
  uint32_t best = kNoHit;
  double best_u = 0.0, best_v = 0.0;
  bool occ = false;
  uint32_t n_nodes = 0, n_tris = 0;

  int sp = 0;
  uint32_t node = 0;
  for (;;) {
    const Node2 nd = load_node(nodes, node);
    if (COUNT) n_nodes++;
    bool h0, h1;
    float t0, t1;
    slab2(nd, r, tmax_f, h0, h1, t0, t1);
    // Leaves are tested on the spot; interior children go on the stack.
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const bool hc = c == 0 ? h0 : h1;
      const uint32_t cnt = c == 0 ? nd.cnt0 : nd.cnt1;
      const uint32_t ref = c == 0 ? nd.ref0 : nd.ref1;
      if (hc && cnt > 0) {
        for (uint32_t k = ref; k < ref + cnt; k++) {
          const TriRec tr = load_tri(tris, k);
          if (COUNT) n_tris++;
          if (!may_hit(r, tr, tmax_f)) continue;
          double tt, uu, vv;
          if (!hit_f64(r, tr, tmin_d, best_t, tt, uu, vv)) continue;
          if (ANY) {
            if (tt < best_t) {  // the any-hit bound is exclusive (mesh.rs:141)
              occ = true;
              break;
            }
            continue;
          }
          // Strictly closer wins; an equal t only replaces a recorded hit
          // with a higher triangle id (the ray's own tmax stays exclusive).
          if (tt < best_t || (tt == best_t && best != kNoHit && tr.prim < best)) {
            best_t = tt;
            best = tr.prim;
            best_u = uu;
            best_v = vv;
            tmax_f = f32_round_up(tt);
          }
        }
      }
      if (ANY && occ) break;
    }
    if (ANY && occ) break;
    const bool in0 = h0 && nd.cnt0 == 0;
    const bool in1 = h1 && nd.cnt1 == 0;
    if (in0 && in1) {
      const bool first0 = t0 <= t1;
      stk[sp * BLOCK] = first0 ? nd.ref1 : nd.ref0;
      sp++;
      node = first0 ? nd.ref0 : nd.ref1;
    } else if (in0) {
      node = nd.ref0;
    } else if (in1) {
      node = nd.ref1;
    } else {
      if (sp == 0) break;
      sp--;
      node = stk[sp * BLOCK];
    }
  }

  if (ANY) {
    occluded[i] = occ ? 1 : 0;
  } else {
    HitRec h;
    if (best != kNoHit) {
      h.t = (float)best_t;
      h.prim = best;
      h.u = (float)best_u;
      h.v = (float)best_v;
    } else {
      h.t = __builtin_inff();
      h.prim = kNoHit;
      h.u = 0.0f;
      h.v = 0.0f;
    }
    hits[i] = h;
  }
  if (COUNT) {
    counts[2 * (size_t)i] = n_nodes;
    counts[2 * (size_t)i + 1] = n_tris;
  }
}

constexpr int kBlock = 128;
constexpr int kStack = kMaxStack;

template <bool ANY, bool COUNT>
hipError_t launch_one(const DeviceBvh& bvh, const RayRec* rays, uint32_t n, HitRec* hits,
                      uint8_t* occ, uint32_t* counts, hipStream_t stream) {
  const uint32_t blocks = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL((trace_kernel<kBlock, kStack, ANY, COUNT>), dim3(blocks), dim3(kBlock), 0,
                     stream, bvh.nodes, bvh.tris, rays, hits, occ, counts, n);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_trace(const DeviceBvh& bvh, const RayRec* rays, uint64_t n, TraceMode mode,
                        HitRec* hits, uint8_t* occluded, uint32_t* counts, hipStream_t stream) {
  const uint64_t kChunk = 1ull << 30;
  for (uint64_t off = 0; off < n; off += kChunk) {
    const uint32_t m = (uint32_t)((n - off) < kChunk ? (n - off) : kChunk);
    const RayRec* r = rays + off;
    HitRec* h = hits ? hits + off : nullptr;
    uint8_t* o = occluded ? occluded + off : nullptr;
    uint32_t* c = counts ? counts + 2 * off : nullptr;
    hipError_t e;
    if (mode == TraceMode::kAny) {
      e = c ? launch_one<true, true>(bvh, r, m, h, o, c, stream)
            : launch_one<true, false>(bvh, r, m, h, o, c, stream);
    } else {
      e = c ? launch_one<false, true>(bvh, r, m, h, o, c, stream)
            : launch_one<false, false>(bvh, r, m, h, o, c, stream);
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace rrhip
