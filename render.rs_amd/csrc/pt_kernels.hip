// Wavefront path integrator kernels for gfx950.
//
// The CPU path's loop body (render.rs src/raytracer/pt/mod.rs:449-990:
// intersect -> miss/dome -> emission with MIS -> NEE through the light BVH
// -> BSDF continuation -> Russian roulette) is split into stages that run
// over queues of live paths, so a wave only ever holds paths at the same
// stage and dead paths drop out of the launch instead of idling lanes
// (render.rs's own wavefront scheduler, kernel_wf.metal:1-20, does the
// same with one extend and one shade kernel; here shadow rays are a third
// stage with an any-hit query).
//
// Queue appends are wave-aggregated (one atomic per wave). Results do not
// depend on queue order: every path owns its state, and pixels are
// accumulated by the resolve kernel in sample order, so renders are
// bit-reproducible.
#include <hip/hip_runtime.h>

#include "pt_device.h"
#include "pt_hair.h"
#include "pt_kernels.h"
#include "pt_pattern.h"
#include "pt_trace.h"

namespace rrhip {
namespace {

using namespace pt;

constexpr int kBlock = 256;
constexpr uint32_t kMaxGrid = 4096;
// Traversal kernels: 128 lanes per block, each with a 48-entry LDS stack.
constexpr int kTraceBlock = 128;
constexpr int kTraceStack = 48;

constexpr uint32_t kFlagFromCamera = 1u << 8;
constexpr uint32_t kFlagTrain = 1u << 9;   // radiance-cache training path
constexpr uint32_t kFlagSuffix = 1u << 10; // training path past its termination vertex
// Transparent scenes: medium-stack depth (bits 11-14, 0..8) and the count of
// hull crossings / cutout pass-throughs (bits 15-19, 0..16).
constexpr uint32_t kStackShift = 11, kSkipShift = 15, kStackSkipMask = 0x1FFu << 11;
constexpr uint32_t kMaxSkips = 16;  // pt/mod.rs:32 MAX_PRESENCE_SKIPS
// Film mode: the first real hit's AOVs are recorded (bit 20); the first
// continuation has been sampled (bit 21) and was specular-like (bit 22),
// pt/mod.rs:457-470, 959-961.
constexpr uint32_t kFlagAux = 1u << 20, kFlagFirstSet = 1u << 21, kFlagFirstSpec = 1u << 22;
constexpr uint32_t kFlagKeep = kFlagTrain | kFlagSuffix | kFlagAux | kFlagFirstSet | kFlagFirstSpec;

// pt/mod.rs:33-69 MediumStack: up to 8 medium indices, 8 bits each,
// innermost at depth - 1; the atmosphere is the implicit bottom.
__device__ __forceinline__ int stack_top(uint2 st, uint32_t depth, int atmosphere) {
  if (depth == 0) return atmosphere;
  const uint32_t i = depth - 1;
  return (int)(((i < 4 ? st.x : st.y) >> (8 * (i & 3))) & 0xFFu);
}

__device__ __forceinline__ void stack_push(uint2& st, uint32_t& depth, uint32_t m) {
  if (depth >= 8) return;
  const uint32_t sh = 8 * (depth & 3);
  if (depth < 4)
    st.x = (st.x & ~(0xFFu << sh)) | (m << sh);
  else
    st.y = (st.y & ~(0xFFu << sh)) | (m << sh);
  depth++;
}

__device__ __forceinline__ void stack_pop(uint32_t& depth) {
  if (depth > 0) depth--;
}

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Appends for every lane with `pred`; returns the lane's slot (or ~0u).
// All lanes of the wave that are active must call it together.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return 0xFFFFFFFFu;
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, (int)leader);
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  return pred ? base + rank : 0xFFFFFFFFu;
}

__device__ __forceinline__ Pcg32 load_rng(const uint4& u) {
  Pcg32 r;
  r.state = (uint64_t)u.x | ((uint64_t)u.y << 32);
  r.inc = (uint64_t)u.z | ((uint64_t)u.w << 32);
  return r;
}
__device__ __forceinline__ uint4 store_rng(const Pcg32& r) {
  return make_uint4((uint32_t)r.state, (uint32_t)(r.state >> 32), (uint32_t)r.inc, (uint32_t)(r.inc >> 32));
}

// Camera ray of (pixel, sample) into path slot `slot`.
__device__ __forceinline__ void init_path(const PtSceneView& S, const PathBuffers& P, uint32_t slot, uint32_t pixel,
                                          uint64_t sample, uint32_t flags) {
  const uint32_t x = pixel % S.cam.width, y = pixel / S.cam.width;
  // sampler.rs:24-32 per-(pixel, sample) stream; pt/mod.rs:340-356
  Pcg32 rng = Pcg32::for_pixel_sample(pixel, sample);
  V3 o, d;
  camera_ray(S.cam, x, y, rng, o, d);
  // shutter time (after the filter and lens draws) and camera motion,
  // pt/mod.rs:351-355, camera.rs:134-146
  float time = 0.f;
  if (S.motion) {
    time = rnd(rng);
    if (S.cam.has_motion_inv && time > 0.f) {
      const float* M = S.cam.motion_inv;
      float m[12];
#pragma unroll
      for (int k = 0; k < 12; k++) m[k] = ((k % 5 == 0) ? 1.f : 0.f) + (M[k] - ((k % 5 == 0) ? 1.f : 0.f)) * time;
      o = xform_point(m, o);
      d = normalize(xform_vec(m, d));
    }
  }
  P.o[slot] = make_float4(o.x, o.y, o.z, 0.f);
  P.d[slot] = make_float4(d.x, d.y, d.z, 0.f);
  P.beta[slot] = make_float4(1.f, 1.f, 1.f, time);
  P.L[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
  P.prev[slot] = make_float4(o.x, o.y, o.z, 0.f);
  P.rng[slot] = store_rng(rng);
  P.flags[slot] = flags;
  P.q0[slot] = slot;
  if (P.mstk) P.mstk[slot] = make_uint2(0u, 0u);
  if (P.ld) {
    P.ld[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.aux0[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.aux1[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.aux_id[slot] = 0u;
  }
}

// The first `count` u32 draws of the per-(pixel, sample) streams, seeded
// exactly as init_path seeds a camera path: out[i * count + k].
__global__ __launch_bounds__(kBlock) void pcg32_draws_kernel(const uint64_t* __restrict__ pixel,
                                                            const uint64_t* __restrict__ sample, uint32_t n,
                                                            uint32_t count, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Pcg32 rng = Pcg32::for_pixel_sample(pixel[i], sample[i]);
  for (uint32_t k = 0; k < count; k++) out[(size_t)i * count + k] = rng.next_u32();
}

// sample0_dev (when set) supplies the pass's first sample index from device
// memory, so a captured pass graph can be replayed for any sample range.
__global__ __launch_bounds__(kBlock) void raygen_kernel(PtSceneView S, PathBuffers P, PtParams prm, uint32_t total,
                                                       NrcPathBuffers N, const uint64_t* __restrict__ sample0_dev) {
  if (sample0_dev) prm.sample0 = *sample0_dev;
  const uint32_t all = total + N.ntrain;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    P.counts[0] = all;
    P.counts[1] = 0;
    P.counts[2] = 0;
    // N.counts[1] (training samples of the previous pass) is left alone: the
    // previous pass's optimiser steps may still be reading it, and this
    // pass's scan overwrites it before its own steps.
    if (N.counts) N.counts[0] = 0;
  }
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < all; slot += stride) {
    uint32_t pixel, flags = kFlagFromCamera;
    uint64_t sample;
    if (slot < total) {
      pixel = slot % prm.npix;
      sample = prm.sample0 + slot / prm.npix;
      if (prm.live) pixel = prm.live[pixel];
    } else {
      // Training path i of this pass: a hashed pixel and a PCG32 stream of
      // its own (top bit set: never a rendering sample index).
      const uint32_t i = slot - total;
      const uint64_t tag = ((prm.sample0 << 22) | i) & 0x7FFFFFFFFFFFFFFFull;
      pixel = (uint32_t)(Pcg32::mix(tag ^ 0x6e72637472616e69ull) % prm.npix);
      if (prm.live) pixel = prm.live[pixel];
      sample = 0x8000000000000000ull | tag;
      flags |= kFlagTrain;
      N.tn[i] = 0;
    }
    init_path(S, P, slot, pixel, sample, flags);
  }
}

// Closest hit of the path in `slot` into hit_n / hit_m (MOTION: at the
// path's shutter time; GEO: the geometry classes the scene holds).
template <bool MOTION, int GEO, bool COUNT = false>
__device__ __forceinline__ void extend_one(const PtSceneView& S, const PathBuffers& P, uint32_t slot, uint32_t* stk,
                                           TraceCount* tc = nullptr) {
  const float4 o4 = P.o[slot], d4 = P.d[slot];
  const float time = MOTION ? P.beta[slot].w : 0.f;
  Hit h;
  if (scene_closest<kTraceBlock, MOTION, GEO, true, COUNT>(S, v3(o4.x, o4.y, o4.z), v3(d4.x, d4.y, d4.z), time,
                                                             stk, h, tc)) {
    P.hit_n[slot] = make_float4(h.n.x, h.n.y, h.n.z, h.t);
    P.hit_m[slot] = h.material | (h.front ? 0x80000000u : 0u);
    const V3 front = offset_origin(h.p, h.e), back = offset_origin(h.p, -h.e);
    float2 tw = make_float2(kNoTangent, kNoTangent);
    if ((GEO & kGeoCurves) && dot(h.tan, h.tan) > 0.5f) tw = oct_encode(h.tan);
    P.hit_front[slot] = make_float4(front.x, front.y, front.z, tw.x);
    P.hit_back[slot] = make_float4(back.x, back.y, back.z, tw.y);
    if constexpr ((GEO & kGeoSt) != 0) {  // pattern scenes: texture coordinates of the hit
      const HitSt hs = hit_st<MOTION>(S, h, time);
      P.hit_st[slot] = make_float4(hs.s, hs.t, hs.density, 0.f);
      P.hit_gs[slot] = make_float4(hs.gs.x, hs.gs.y, hs.gs.z, 0.f);
      P.hit_gt[slot] = make_float4(hs.gt.x, hs.gt.y, hs.gt.z, 0.f);
    }
  } else {
    P.hit_n[slot] = make_float4(0.f, 0.f, 0.f, -1.f);
    P.hit_m[slot] = 0;
  }
}

// STACK: LDS entries per lane, the smallest size >= the scene's bound
// (fewer bytes per block, more resident waves).
// Profiling kernels: a wave's traversal fetches summed into prof[0..2].
__device__ __forceinline__ void add_counts(unsigned long long* prof, const TraceCount& tc) {
  uint32_t v[3] = {tc.nodes, tc.tris, tc.insts};
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
  if ((threadIdx.x & 63u) == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) atomicAdd(prof + k, (unsigned long long)v[k]);
  }
}

template <int STACK, bool MOTION, int GEO, bool COUNT = false>
__global__ __launch_bounds__(kTraceBlock) void extend_kernel(PtSceneView S, PathBuffers P, const uint32_t* __restrict__ q) {
  __shared__ uint32_t s_stack[STACK * kTraceBlock];
  uint32_t* stk = s_stack + threadIdx.x;
  const uint32_t n = P.counts[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  TraceCount tc;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    extend_one<MOTION, GEO, COUNT>(S, P, q[i], stk, &tc);
  if (COUNT) add_counts(P.prof, tc);
}

// pt/mod.rs:1000-1023
__device__ __forceinline__ float light_pdf_solid_angle(const PtLight& L, V3 origin, V3 hitp) {
  if (L.kind == kLightRect || L.kind == kLightDisk) {
    const V3 d = hitp - origin;
    const float dist2 = dot(d, d);
    const float cos_light = fabsf(dot(normalize(d), v3(L.n)));
    if (cos_light < 1e-9f || L.area <= 0.f) return 0.f;
    return dist2 / (cos_light * L.area);
  }
  if (L.kind == kLightSphere) {
    const V3 c = v3(L.p) - origin;
    const float dist2 = dot(c, c);
    const float sin2 = fminf(L.radius * L.radius / dist2, 1.f);
    const float cos_max = sqrtf(fmaxf(1.f - sin2, 0.f));
    const float sa = 2.f * kPi * (1.f - cos_max);
    return sa < 1e-12f ? 0.f : 1.f / sa;
  }
  return 0.f;
}

struct ShadowOut {
  V3 o, d;
  float tmax;
  float time;  // the path's shutter time
  V3 c;        // contribution before visibility (f cos L weight / pdf)
  int medium;  // transparent scenes: the medium at the ray's origin (the walk applies its transmittance)
  uint2 seed;  // transparent scenes: random stream of the blocker walk
  bool diff;   // film mode: the contribution belongs to the diffuse AOV
};

// Evaluates the BSDF towards world direction wi: f * cos(wi) and the pdf
// (the `eval` closure of pt/mod.rs:838-845).
__device__ __forceinline__ void eval_world(const PtMaterial& m, const Frame& fr, V3 wo_l, float eta, V3 wi,
                                           V3& f, float& pdf) {
  const V3 wi_l = to_local(fr, wi);
  if (wi_l.z <= 0.f) {
    f = v3(0.f, 0.f, 0.f);
    pdf = 0.f;
    return;
  }
  bsdf_eval(m, wo_l, wi_l, eta, f, pdf);
  f = f * wi_l.z;
}

// sample_light + area_light_contribution (pt/mod.rs:1032-1209), without the
// visibility test, which becomes the returned shadow ray from `origin()`
// (evaluated last, to keep it out of the BSDF evaluation's registers; with
// error-bounded origins p itself already lies outside its surface's
// rounding box). `eval(wi, f, pdf)` gives f * cos (or the phase value) and
// the pdf of the BSDF (or phase) sampling strategy for MIS. render.rs stops
// shadow rays 1e-3 short of the light; with BOUNDED the segment also stops
// short by the sampled light point's own rounding bound (projected on the
// ray through the light's cosine).
template <bool BOUNDED, class Eval>
__device__ __forceinline__ bool sample_light(const PtSceneView& S, const PtLight& L, V3 p, V3 org,
                                             const Eval& eval, float pick, Pcg32& rng, ShadowOut& out) {
  V3 f;
  float bsdf_pdf;
  const V3 rad = v3(L.radiance);
  constexpr bool bounded = BOUNDED;
  switch (L.kind) {
    case kLightPoint: {
      const V3 to = v3(L.p) - p;
      const float dist2 = fmaxf(dot(to, to), 1e-12f);
      const V3 wi = normalize(to);
      eval(wi, f, bsdf_pdf);
      if (maxc(f) <= 0.f) return false;
      out.o = org;
      out.d = wi;
      out.tmax = len(to) - 1e-3f;
      out.c = f * rad / (dist2 * pick);
      return true;
    }
    case kLightDistant: {
      const V3 base = -v3(L.p);
      V3 wi = base;
      if (L.radius > 1e-5f) {
        const float u = rnd(rng), v = rnd(rng);
        const float cos_t = 1.f - u * (1.f - L.cos_max);
        const float sin_t = sqrtf(fmaxf(1.f - cos_t * cos_t, 0.f));
        float sp, cp;
        sincosf(2.f * kPi * v, &sp, &cp);
        wi = to_world(make_frame(base), v3(sin_t * cp, sin_t * sp, cos_t));
      }
      eval(wi, f, bsdf_pdf);
      if (maxc(f) <= 0.f) return false;
      out.o = org;
      out.d = wi;
      out.tmax = 1e8f;
      out.c = f * rad / pick;
      return true;
    }
    case kLightRect:
    case kLightDisk: {
      const float u = rnd(rng), v = rnd(rng);
      V3 a, b;
      if (L.kind == kLightRect) {
        a = v3(L.e1) * u;
        b = v3(L.e2) * v;
      } else {
        const float r = sqrtf(u);
        float sn, cs;
        sincosf(2.f * kPi * v, &sn, &cs);
        a = v3(L.e1) * (r * cs);
        b = v3(L.e2) * (r * sn);
      }
      const V3 sp = v3(L.p) + a + b;
      const V3 to = sp - p;
      const float dist2 = fmaxf(dot(to, to), 1e-12f);
      const V3 wi = normalize(to);
      const float cos_light = fabsf(dot(wi, v3(L.n)));  // double-sided emitter
      if (cos_light < 1e-9f || L.area <= 0.f) return false;
      eval(wi, f, bsdf_pdf);
      if (maxc(f) <= 0.f) return false;
      const float pdf_sa = pick * dist2 / (cos_light * L.area);
      const float w = power_heuristic(pdf_sa, bsdf_pdf);
      out.o = org;
      out.d = wi;
      out.tmax = len(to) - 1e-3f;
      if (bounded) out.tmax -= L.err / fmaxf(cos_light, 1e-2f);
      out.c = f * rad * (w / pdf_sa);
      return true;
    }
    case kLightSphere: {
      const V3 tc = v3(L.p) - p;
      const float dist2 = dot(tc, tc);
      const float r = L.radius;
      if (dist2 <= r * r * 1.0001f) return false;
      const float sin2 = fminf(r * r / dist2, 1.f);
      const float cos_max = sqrtf(fmaxf(1.f - sin2, 0.f));
      const float u = rnd(rng), v = rnd(rng);
      const float cos_t = 1.f - u * (1.f - cos_max);
      const float sin_t = sqrtf(fmaxf(1.f - cos_t * cos_t, 0.f));
      float sn, cs;
      sincosf(2.f * kPi * v, &sn, &cs);
      const V3 wi = to_world(make_frame(normalize(tc)), v3(sin_t * cs, sin_t * sn, cos_t));
      eval(wi, f, bsdf_pdf);
      if (maxc(f) <= 0.f) return false;
      const float dist = sqrtf(dist2) * cos_t - sqrtf(fmaxf(r * r - dist2 * sin_t * sin_t, 0.f));
      const float pdf_sa = pick / fmaxf(2.f * kPi * (1.f - cos_max), 1e-12f);
      const float w = power_heuristic(pdf_sa, bsdf_pdf);
      out.o = org;
      out.d = wi;
      out.tmax = dist - 1e-3f;
      if (bounded) out.tmax -= L.err + sqrtf(dist2) * err_gamma(8);
      out.c = f * rad * (w / pdf_sa);
      return true;
    }
    case kLightDome: {  // pt/mod.rs:1138-1172
      const float u = rnd(rng), v = rnd(rng);
      V3 wi, lrad = rad;
      float pdf_d = kInv4Pi;
      if (S.env_rgb) {  // importance-sampled map texel
        V3 texel;
        env_sample(S, u, v, wi, texel, pdf_d);
        if (!(pdf_d > 0.f)) return false;
        lrad = texel * rad;
      } else {  // constant dome: uniform sphere, matching dome_pdf
        const float z = 1.f - 2.f * u;
        const float rr = sqrtf(fmaxf(1.f - z * z, 0.f));
        float sn, cs;
        sincosf(2.f * kPi * v, &sn, &cs);
        wi = v3(rr * cs, rr * sn, z);
      }
      const float pdf_sa = pdf_d * pick;
      eval(wi, f, bsdf_pdf);
      if (maxc(f) <= 0.f) return false;
      const float w = power_heuristic(pdf_sa, bsdf_pdf);
      out.o = org;
      out.d = wi;
      out.tmax = 1e8f;
      out.c = f * lrad * (w / pdf_sa);
      return true;
    }
    default:
      return false;
  }
}

// With NRC, the area-spread heuristic of Mueller et al. 2021 ("Real-time
// neural radiance caching for path tracing", sec. 4): a(x1..xn) =
// (sum_{i=2..n} sqrt(|x_{i-1} - x_i|^2 / (p(w_i) |cos theta_i|)))^2 against
// c * a0, a0 = |x0 - x1|^2 / (4 pi |cos theta_1|).
// One path's shading step (the per-lane part of shade_kernel): sets
// cont (continue in the next queue), shadow (+ so), shadow2 (+ se, an
// equiangular connection in a medium), query (+ nq, qbeta).
// TR: the scene is transparent (medium stack, hulls, cutouts, shadow walks).
// FILM: also the diffuse split of the radiance and the first real hit's
// AOVs (pt/mod.rs:457-470, 559-573, 690-700, 959-961).
// FIXED: render.rs's fixed push of rays off surfaces (S.ray_offset > 0)
// instead of error-bounded origins.
// HAIR: the scene has hair fibers; their curve hits scatter with the fiber
// BSDF (pt/mod.rs:738-800).
// PAT: the scene has pattern-bound materials: paths carry a ray cone (its
// spread beyond one pixel in d.w, its width in prev.w, pt/mod.rs:477-480),
// and a hit on such a material shades with the fields its patterns give
// there (pt/mod.rs:685-689, 389-440).
template <bool NRC, bool TR, bool FILM = false, bool FIXED = false, bool HAIR = false, bool PAT = false>
__device__ __forceinline__ void shade_one(const PtSceneView& S, const PathBuffers& P, const PtParams& prm,
                                          const NrcPathBuffers& N, uint32_t slot, bool& cont, bool& shadow,
                                          bool& shadow2, ShadowOut& se, bool& query, ShadowOut& so, NrcQuery& nq,
                                          V3& qbeta) {
    const float4 o4 = P.o[slot], d4 = P.d[slot], b4 = P.beta[slot], l4 = P.L[slot], pv4 = P.prev[slot];
    const float4 hn = P.hit_n[slot];
    uint32_t flags = P.flags[slot];
    Pcg32 rng = load_rng(P.rng[slot]);
    V3 beta = v3(b4.x, b4.y, b4.z), Lr = v3(l4.x, l4.y, l4.z);
    const V3 ro = v3(o4.x, o4.y, o4.z), rd = v3(d4.x, d4.y, d4.z);
    // Surface hit point. With error-bounded origins it is the hit rebuilt
    // from the surface and pushed outside its rounding box on the ray's
    // side, which is also where reflected and shadow rays start; rays that
    // cross the surface start from its twin on the far side. render.rs's
    // fixed push takes the point along the ray.
    constexpr bool bounded = !FIXED;
    const auto hit_point = [&]() {
      if (!bounded) return ro + rd * hn.w;
      const float4 p4 = P.hit_front[slot];
      return v3(p4.x, p4.y, p4.z);
    };
    const auto far_side = [&]() {
      const float4 b4 = P.hit_back[slot];
      return v3(b4.x, b4.y, b4.z);
    };
    const float prev_pdf = o4.w;
    const V3 prev_origin = v3(pv4.x, pv4.y, pv4.z);
    const bool from_camera = (flags & kFlagFromCamera) != 0;
    uint32_t depth = flags & 0xFFu;
    float a0 = d4.w, asum = pv4.w;
    // Transparent scenes: the medium stack and the skip count
    uint2 mst = make_uint2(0u, 0u);
    uint32_t sdepth = 0, skips = 0;
    if (TR) {
      mst = P.mstk[slot];
      sdepth = (flags >> kStackShift) & 0xFu;
      skips = (flags >> kSkipShift) & 0x1Fu;
    }
    const int cur = NRC ? -1 : stack_top(mst, sdepth, S.atmosphere);  // the segment's medium
    // Film: radiance of this path that goes to the diffuse AOV (all of it
    // until a first continuation is sampled specular-like); aux_set: the
    // first real hit is recorded; travel: distance to it so far.
    V3 Ld = v3(0.f, 0.f, 0.f);
    const bool diff = FILM && !(flags & kFlagFirstSpec);
    bool aux_set = true;
    float travel = 0.f;
    if (FILM) {
      const float4 ld4 = P.ld[slot];
      Ld = v3(ld4.x, ld4.y, ld4.z);
      aux_set = (flags & kFlagAux) != 0;
      if (!aux_set) travel = P.aux1[slot].w;
    }
    const auto emit = [&](V3 c) {
      Lr = Lr + c;
      if (FILM && diff) Ld = Ld + c;
    };
    constexpr bool walk = TR;  // shadow rays pick up media transmittance on their walk
    bool passed = false;                   // hull crossing or cutout pass-through
    PtMaterial mres;                       // PAT: the hit's material with its patterns applied
    // The medium along this segment (pt/mod.rs:486-633): one light pick
    // per segment; for a point light an equiangular single-scatter
    // connection (pt/mod.rs:501-551); distance sampling, and at a real
    // collision NEE with the phase function (MIS-paired with the
    // equiangular strategy for point lights) and an HG continuation
    // instead of the surface.
    bool in_medium = false, absorbed = false;
    if (!NRC && cur >= 0) {  // caches are never attached to scenes with media
      const PtMedium& md = S.media[cur];
      const float t_limit = hn.w < 0.f ? 1e8f : hn.w;
      const float t_seg = fminf(t_limit, md.max_distance);
      uint32_t seg_li = 0;
      float seg_pick = 0.f;
      bool has_pick = false;
      if (S.num_lights > 0) has_pick = ls_sample(S, ro + rd * fminf(t_seg * 0.5f, 50.f), rnd(rng), seg_li, seg_pick);
      const bool point_pick = has_pick && P.so2 && S.lights[seg_li].kind == kLightPoint;
      if (point_pick && t_seg > 1e-6f) {
        const PtLight& L = S.lights[seg_li];
        const V3 lp = v3(L.p);
        float pdf_eq;
        const float t_eq = equiangular_sample(ro, rd, t_seg, lp, rnd(rng), pdf_eq);
        if (pdf_eq > 1e-12f) {
          const V3 p = ro + rd * t_eq;
          const float rho = density_at(md, p);
          if (rho > 0.f) {
            const V3 tr = medium_transmittance(md, ro, rd, t_eq, rng);
            const V3 to_l = lp - p;
            const float dist2 = fmaxf(dot(to_l, to_l), 1e-12f);
            const V3 wi = normalize(to_l);
            const float ph = hg_phase(md.g, dot(rd, wi));  // hg_phase(g, -wo.wi), wo = -rd
            const float dist = sqrtf(dist2) - 1e-3f;
            const V3 vis = walk ? v3(1.f, 1.f, 1.f) : medium_transmittance(md, p, wi, dist, rng);
            if (maxc(vis) > 0.f) {
              se.medium = cur;
              const float pdf_d = distance_pdf_approx(md, t_eq, t_seg);
              const float w_eq = pdf_eq / (pdf_eq + pdf_d);
              se.o = p;
              se.d = wi;
              se.tmax = dist;
              se.c = beta * v3(md.sigma_s) * tr * vis * v3(L.radiance) *
                     (rho * ph * w_eq / (dist2 * pdf_eq * seg_pick));
              shadow2 = maxc(se.c) > 0.f;
            }
          }
        }
      }
      const MediumEvent ev = sample_distance(md, ro, rd, t_limit, beta, rng);
      beta = beta * ev.weight;
      if (ev.scatter) {
        in_medium = true;
        const V3 p = ro + rd * ev.t;
        if (FILM && !aux_set) {  // pt/mod.rs:559-570
          aux_set = true;
          flags |= kFlagAux;
          const float st[3] = {md.sigma_t[0], md.sigma_t[1], md.sigma_t[2]};
          P.aux0[slot] = make_float4(md.sigma_s[0] / fmaxf(st[0], 1e-9f), md.sigma_s[1] / fmaxf(st[1], 1e-9f),
                                     md.sigma_s[2] / fmaxf(st[2], 1e-9f), travel + ev.t);
          P.aux1[slot] = make_float4(-rd.x, -rd.y, -rd.z, 0.f);
          P.aux_id[slot] = 0u;
        }
        if (FILM) flags |= kFlagFirstSet;  // a medium scatter is a diffuse-like first continuation
        if (maxc(v3(md.emission)) > 0.f) emit(beta * v3(md.emission));
        const V3 wo = -rd;
        if (has_pick) {
          const auto eval = [&](V3 wi, V3& f, float& pdf) {
            const float ph = hg_phase(md.g, -dot(wo, wi));
            f = v3(ph, ph, ph);
            pdf = ph;
          };
          if (sample_light<bounded>(S, S.lights[seg_li], p, p, eval, seg_pick, rng, so)) {
            so.c = beta * so.c;
            so.medium = cur;
            if (!walk) so.c = so.c * medium_transmittance(md, so.o, so.d, so.tmax, rng);
            if (point_pick) {  // pt/mod.rs:583-598
              const float pdf_d = distance_pdf_approx(md, ev.t, t_seg);
              const float pdf_e = equiangular_pdf(ro, rd, t_seg, v3(S.lights[seg_li].p), ev.t);
              so.c = so.c * (pdf_d / (pdf_d + pdf_e));
            }
            shadow = maxc(so.c) > 0.f;
          }
        }
        float pdf;
        const float u1 = rnd(rng), u2 = rnd(rng);
        const V3 wi = hg_sample(md.g, wo, u1, u2, pdf);
        if (PAT) a0 += 0.4f;  // the cone widens after a scatter (pt/mod.rs:622)
        depth++;
        bool alive = true;
        if (depth >= prm.rr_start) {
          const float qq = fminf(maxc(beta), 0.95f);
          if (qq <= 0.f || rnd(rng) > qq) {
            alive = false;
          } else {
            beta = beta / qq;
          }
        }
        cont = alive && depth < prm.max_depth;
        P.o[slot] = make_float4(p.x, p.y, p.z, pdf);
        P.d[slot] = make_float4(wi.x, wi.y, wi.z, a0);
        P.prev[slot] = make_float4(p.x, p.y, p.z, asum);
        flags = depth | (flags & kFlagKeep);
      } else {
        absorbed = !(maxc(beta) > 0.f);
      }
    }
    if (in_medium || absorbed) {
      // handled above
    } else if (hn.w < 0.f) {
      // Miss: dome (with MIS) or background (pt/mod.rs:643-659).
      if (S.dome >= 0) {
        float w = 1.f;
        if (!from_camera) w = power_heuristic(prev_pdf, dome_pdf(S, rd) * ls_pmf(S, prev_origin, (uint32_t)S.dome));
        emit(beta * dome_radiance(S, S.lights[S.dome], rd) * w);
      } else {
        emit(beta * v3(S.background));
      }
    } else {
      const uint32_t hm = P.hit_m[slot];
      PtMaterial m = S.materials[hm & 0x7FFFFFFFu];
      const bool entering = (hm >> 31) != 0;
      const V3 p = hit_point();
      if constexpr (PAT) {
        // The cone grows by this segment at every surface the path does not
        // cross as a hull; bound fields follow their patterns here.
        if (!(TR && m.hull && skips < kMaxSkips)) {
          asum += hn.w * (S.pixel_spread + a0);
          if (m.bind) {
            const PatCtx pc = pat_ctx(P.hit_st[slot], P.hit_gs[slot], P.hit_gt[slot], p, v3(hn.x, hn.y, hn.z), rd,
                                      asum);
            pat_resolve_material(S, hm & 0x7FFFFFFFu, m.bind, pc, m);
          }
        }
        mres = m;
      }
      if (FILM && !aux_set) {
        if (TR && m.hull && skips < kMaxSkips) {
          travel += hn.w;  // a hull crossing is looked through (pt/mod.rs:671)
          P.aux1[slot].w = travel;
        } else {  // the first real hit, cutouts included (pt/mod.rs:688-700)
          aux_set = true;
          flags |= kFlagAux;
          const V3 ng = v3(hn.x, hn.y, hn.z);
          const V3 nf = dot(ng, rd) <= 0.f ? ng : -ng;
          P.aux0[slot] = make_float4(m.film_albedo[0], m.film_albedo[1], m.film_albedo[2], travel + hn.w);
          P.aux1[slot] = make_float4(nf.x, nf.y, nf.z, 0.f);
          P.aux_id[slot] = m.id;
        }
      }
      // Volume hull crossings toggle the medium (pt/mod.rs:662-683); a
      // cutout passes the ray with probability 1 - presence
      // (pt/mod.rs:702-710). Neither counts as a bounce.
      if (TR) {
        if (m.hull && skips < kMaxSkips) {
          passed = true;
          if (entering)
            stack_push(mst, sdepth, (uint32_t)m.interior);
          else
            stack_pop(sdepth);
        } else if (m.presence < 1.f && rnd(rng) >= m.presence && skips < kMaxSkips) {
          passed = true;
        }
        if (passed) {
          skips++;
          const V3 org = bounded ? far_side() : p + rd * S.ray_offset;
          P.o[slot] = make_float4(org.x, org.y, org.z, prev_pdf);
          if (PAT) P.prev[slot] = make_float4(pv4.x, pv4.y, pv4.z, asum);
          cont = true;
        }
      }
    }
    if (passed) {
      // the ray continues unchanged past the surface
    } else if (hn.w >= 0.f && !in_medium && !absorbed) {
      const uint32_t hm = P.hit_m[slot];
      const PtMaterial m = PAT ? mres : S.materials[hm & 0x7FFFFFFFu];
      const bool entering = (hm >> 31) != 0;
      const V3 p = hit_point();
      const V3 ng = v3(hn.x, hn.y, hn.z);
      const V3 wo = -rd;
      const V3 nrm = dot(ng, wo) >= 0.f ? ng : -ng;
      const float eta = entering ? m.glass_ior : 1.f / m.glass_ior;
      // Emitter hit (pt/mod.rs:723-736)
      const V3 em = v3(m.emission);
      if (maxc(em) > 0.f) {
        float w = 1.f;
        if (!from_camera && m.area_light >= 0) {
          const PtLight& L = S.lights[m.area_light];
          const float pick = ls_pmf(S, prev_origin, (uint32_t)m.area_light);
          w = power_heuristic(prev_pdf, light_pdf_solid_angle(L, prev_origin, p) * pick);
        }
        emit(beta * em * w);
      }
      bool scatter = true;
      if constexpr (HAIR) {
        // A hair fiber's curve hit (the w lanes hold its tangent): NEE with
        // the fiber BSDF's value and pdf, whose shadow rays leave from the
        // viewer's side, then a sampled continuation leaving from the side
        // of the normal it points to (pt/mod.rs:738-800). The surface
        // lobes below are skipped.
        const float4 t4 = P.hit_front[slot];
        if (m.hair && t4.w < 2.f) {
          scatter = false;
          const FiberFrame ff = fiber_frame(oct_decode(t4.w, P.hit_back[slot].w), ng, wo);
          const HairGeom g = hair_geom(m, fiber_local(ff, wo), ff.h);
          if (S.num_lights > 0) {
            uint32_t li;
            float pick;
            if (ls_sample(S, p, rnd(rng), li, pick)) {
              const auto eval = [&](V3 wi, V3& f, float& pdf) { hair_eval(m, g, fiber_local(ff, wi), f, pdf); };
              if (sample_light<bounded>(S, S.lights[li], p, bounded ? p : p + ng * S.ray_offset, eval, pick, rng,
                                        so)) {
                so.c = beta * so.c;
                so.medium = cur;
                if (!NRC && cur >= 0 && !walk)
                  so.c = so.c * medium_transmittance(S.media[cur], so.o, so.d, so.tmax, rng);
                shadow = maxc(so.c) > 0.f;
              }
            }
          }
          const float u[4] = {rnd(rng), rnd(rng), rnd(rng), rnd(rng)};
          V3 wf, fv;
          float pv;
          if (hair_sample(m, g, u, wf, fv, pv)) {
            const V3 wi = normalize(fiber_world(ff, wf));
            beta = beta * fv / pv;
            if (PAT) a0 += 0.4f;  // pt/mod.rs:780
            if (FILM && !(flags & kFlagFirstSet)) flags |= kFlagFirstSet;
            const bool out = dot(wi, ng) >= 0.f;
            const V3 org = bounded ? (out ? p : far_side()) : p + ng * (out ? S.ray_offset : -S.ray_offset);
            depth++;
            bool alive = true;
            if (depth >= prm.rr_start) {
              const float qq = fminf(maxc(beta), 0.95f);
              if (qq <= 0.f || rnd(rng) > qq) {
                alive = false;
              } else {
                beta = beta / qq;
              }
            }
            cont = alive && depth < prm.max_depth;
            P.o[slot] = make_float4(org.x, org.y, org.z, pv);
            P.d[slot] = make_float4(wi.x, wi.y, wi.z, a0);
            P.prev[slot] = make_float4(p.x, p.y, p.z, asum);
            flags = depth | (flags & kFlagKeep);
          }
        }
      }
      if constexpr (NRC) {
        const float cosv = fmaxf(fabsf(dot(ng, rd)), 1e-4f);
        const float t2 = hn.w * hn.w;
        if (from_camera) {
          a0 = t2 / (4.f * kPi * cosv);
          asum = 0.f;
        } else {
          asum += sqrtf(t2 / fmaxf(prev_pdf * cosv, 1e-30f));
        }
        const bool wide = asum * asum > N.spread * a0;
        const bool train = (flags & kFlagTrain) != 0;
        bool record = false;
        if (!train) {
          query = wide && m.has_lobes;
        } else if (!(flags & kFlagSuffix)) {
          if (wide) {
            flags |= kFlagSuffix;
            asum = 0.f;
          }
          record = m.has_lobes != 0;
        } else {
          query = wide && m.has_lobes;
          record = !query && m.has_lobes;
        }
        if (query || record) {
          nq.pos[0] = p.x;
          nq.pos[1] = p.y;
          nq.pos[2] = p.z;
          nq.roughness = m.nrc_rough;
          nq.dir[0] = wo.x;
          nq.dir[1] = wo.y;
          nq.dir[2] = wo.z;
          nq.pad0 = 0.f;
          nq.normal[0] = nrm.x;
          nq.normal[1] = nrm.y;
          nq.normal[2] = nrm.z;
          nq.pad1 = 0.f;
          nq.albedo[0] = m.nrc_albedo[0];
          nq.albedo[1] = m.nrc_albedo[1];
          nq.albedo[2] = m.nrc_albedo[2];
          nq.pad2 = 0.f;
        }
        if (record && depth < N.max_vertices) {
          const uint32_t tp = slot - N.train_base;
          const size_t r = (size_t)tp * N.max_vertices + depth;
          N.tq[r] = nq;
          N.tb[r] = make_float4(beta.x, beta.y, beta.z, 0.f);
          N.tl[r] = make_float4(Lr.x, Lr.y, Lr.z, 0.f);
          N.tn[tp] = depth + 1;
        }
        qbeta = beta;
        scatter = !query;
      }
      const Frame fr = make_frame(nrm);
      const V3 wo_l = to_local(fr, wo);
      // NEE through the light BVH (pt/mod.rs:924-956)
      if (scatter && S.num_lights > 0 && m.has_lobes) {
        uint32_t li;
        float pick;
        if (ls_sample(S, p, rnd(rng), li, pick)) {
          const auto eval = [&](V3 wi, V3& f, float& pdf) { eval_world(m, fr, wo_l, eta, wi, f, pdf); };
          if (sample_light<bounded>(S, S.lights[li], p, bounded ? p : p + nrm * S.ray_offset, eval, pick, rng, so)) {
            so.c = beta * so.c;
            so.medium = cur;
            // shadow rays through a medium carry its transmittance (or
            // pick it up on their blocker walk)
            if (!NRC && cur >= 0 && !walk)
              so.c = so.c * medium_transmittance(S.media[cur], so.o, so.d, so.tmax, rng);
            shadow = maxc(so.c) > 0.f;
          }
        }
      }
      // BSDF continuation (pt/mod.rs:958-987)
      BsdfSample bs;
      if (scatter && bsdf_sample(m, wo_l, eta, rng, bs)) {
        const V3 wi = to_world(fr, bs.wi);
        beta = beta * bs.f * (fabsf(bs.wi.z) / bs.pdf);
        if (PAT) a0 += fminf(1.f / (1.f + bs.pdf), 0.4f);  // pt/mod.rs:967
        if (FILM && !(flags & kFlagFirstSet)) flags |= kFlagFirstSet | (bs.specular ? kFlagFirstSpec : 0u);
        const V3 org = bounded ? (bs.transmitted ? far_side() : p)
                               : p + nrm * (bs.transmitted ? -S.ray_offset : S.ray_offset);
        const V3 nd = normalize(wi);
        if (TR && bs.transmitted && m.interior >= 0) {  // pt/mod.rs:969-975
          if (entering)
            stack_push(mst, sdepth, (uint32_t)m.interior);
          else
            stack_pop(sdepth);
        }
        depth++;
        bool alive = true;
        if (depth >= prm.rr_start) {
          const float qq = fminf(maxc(beta), 0.95f);
          if (qq <= 0.f || rnd(rng) > qq) {
            alive = false;
          } else {
            beta = beta / qq;
          }
        }
        cont = alive && depth < prm.max_depth;
        P.o[slot] = make_float4(org.x, org.y, org.z, bs.pdf);
        P.d[slot] = make_float4(nd.x, nd.y, nd.z, a0);
        P.prev[slot] = make_float4(p.x, p.y, p.z, asum);
        flags = depth | (flags & kFlagKeep);  // from_camera cleared
      }
    }
    if (TR && (shadow || shadow2)) {
      so.seed = make_uint2(rng.next_u32(), rng.next_u32());
      se.seed = so.seed;
    }
    so.time = b4.w;
    se.time = b4.w;
    so.diff = diff;
    se.diff = diff;
    if (FILM) P.ld[slot] = make_float4(Ld.x, Ld.y, Ld.z, 0.f);
    if (TR) {
      P.mstk[slot] = mst;
      flags = (flags & ~kStackSkipMask) | (sdepth << kStackShift) | (skips << kSkipShift);
    }
    P.beta[slot] = make_float4(beta.x, beta.y, beta.z, b4.w);  // w: shutter time
    P.L[slot] = make_float4(Lr.x, Lr.y, Lr.z, 0.f);
    P.rng[slot] = store_rng(rng);
    P.flags[slot] = flags;
}

// Shadow record idx of path `slot`: the first ray is `a` (or `b` alone);
// with both, `b` goes to the second arrays. sc.w: bit 0 = the record has
// a second ray, bit 1 (FILM) = its contributions go to the diffuse AOV as
// well; transparent scenes: d.w holds the medium at the origin, srng the
// walk's seed.
template <bool TR, bool FILM = false>
__device__ __forceinline__ void put_shadow(const PathBuffers& P, uint32_t idx, uint32_t slot, bool s1,
                                           const ShadowOut& a, bool s2, const ShadowOut& b) {
  const ShadowOut& f = s1 ? a : b;
  const bool both = s1 && s2;
  // transparent scenes: d.w holds the medium at the origin, srng the walk's seed
  constexpr bool walk = TR;
  P.sq[idx] = slot;
  P.so[idx] = make_float4(f.o.x, f.o.y, f.o.z, f.tmax);
  P.sd[idx] = make_float4(f.d.x, f.d.y, f.d.z, walk ? (float)f.medium : 0.f);
  P.sc[idx] = make_float4(f.c.x, f.c.y, f.c.z, (both ? 1.f : 0.f) + (FILM && f.diff ? 2.f : 0.f));
  if (walk) P.srng[idx] = f.seed;
  if (P.stime) P.stime[idx] = f.time;
  if (both) {
    P.so2[idx] = make_float4(b.o.x, b.o.y, b.o.z, b.tmax);
    P.sd2[idx] = make_float4(b.d.x, b.d.y, b.d.z, walk ? (float)b.medium : 0.f);
    P.sc2[idx] = make_float4(b.c.x, b.c.y, b.c.z, 0.f);
  }
}

// Appends a lane's shadow record (all lanes of the wave call it).
template <bool TR, bool FILM = false>
__device__ __forceinline__ void push_shadow(const PathBuffers& P, bool act, uint32_t slot, bool s1, const ShadowOut& a,
                                            bool s2, const ShadowOut& b) {
  const bool any = act && (s1 || s2);
  const uint32_t idx = wave_append(&P.counts[2], any);
  if (!any) return;
  put_shadow<TR, FILM>(P, idx, slot, s1, a, s2, b);
}

// The shade pass over the queue q: each path's step, then the wave-wide
// appends of its shadow records, continuation and cache query.
template <bool NRC, bool TR, bool FILM, bool FIXED, bool HAIR, bool PAT = false>
__device__ __forceinline__ void shade_pass(const PtSceneView& S, const PathBuffers& P, const PtParams& prm,
                                           const uint32_t* __restrict__ q, uint32_t* __restrict__ qnext,
                                           const NrcPathBuffers& N) {
  const uint32_t n = P.counts[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  // Block-uniform loop bound so every lane reaches the wave-wide appends.
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const bool act = i < n;
    bool cont = false, shadow = false, shadow2 = false, query = false;
    ShadowOut so, se;
    NrcQuery nq;
    V3 qbeta;
    uint32_t slot = 0;
    if (act) {
      slot = q[i];
      shade_one<NRC, TR, FILM, FIXED, HAIR, PAT>(S, P, prm, N, slot, cont, shadow, shadow2, se, query, so, nq,
                                                 qbeta);
    }
    push_shadow<TR, FILM>(P, act, slot, shadow, so, shadow2, se);
    const uint32_t nidx = wave_append(&P.counts[1], act && cont);
    if (act && cont) qnext[nidx] = slot;
    if constexpr (NRC) {
      const uint32_t qidx = wave_append(&N.counts[0], act && query);
      if (act && query) {
        N.q[qidx] = nq;
        N.q_slot[qidx] = slot;
        N.q_beta[qidx] = make_float4(qbeta.x, qbeta.y, qbeta.z, 0.f);
      }
    }
  }
}

// Three waves per SIMD (<= 168 VGPRs): shading is latency-bound on
// scattered path-state loads, so occupancy matters more than registers.
template <bool NRC, bool TR, bool FILM = false, bool FIXED = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void shade_kernel(
    PtSceneView S, PathBuffers P, PtParams prm, const uint32_t* __restrict__ q, uint32_t* __restrict__ qnext,
    NrcPathBuffers N) {
  shade_pass<NRC, TR, FILM, FIXED, false>(S, P, prm, q, qnext, N);
}

// Scenes with hair fibers: the fiber BSDF's state does not fit 168
// registers next to the surface lobes', so two waves per SIMD.
template <bool TR, bool FILM, bool FIXED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void shade_hair_kernel(
    PtSceneView S, PathBuffers P, PtParams prm, const uint32_t* __restrict__ q, uint32_t* __restrict__ qnext,
    NrcPathBuffers N) {
  shade_pass<false, TR, FILM, FIXED, true>(S, P, prm, q, qnext, N);
}

// Scenes with pattern-bound materials (and hair fibers, if any): the
// pattern evaluator's value stack and the texture filters next to both
// BSDFs, at two waves per SIMD (218-243 VGPRs without the fiber BSDF, so
// no smaller variant; a three-wave cap spills 140 VGPRs).
template <bool TR, bool FILM, bool FIXED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void shade_pattern_kernel(
    PtSceneView S, PathBuffers P, PtParams prm, const uint32_t* __restrict__ q, uint32_t* __restrict__ qnext,
    NrcPathBuffers N) {
  shade_pass<false, TR, FILM, FIXED, true, true>(S, P, prm, q, qnext, N);
}

// Fractional visibility along a shadow ray in a transparent scene
// (pt/mod.rs:1238-1306): closest hits in turn; each segment multiplies the
// transmittance of the medium on top of the stack (the origin's medium
// over the atmosphere); hulls push / pop their media, cutouts let 1 -
// presence through, anything else blocks; at most 16 blockers.
template <bool MOTION, int GEO, bool COUNT = false>
__device__ __forceinline__ V3 walk_visibility(const PtSceneView& S, V3 o, V3 d, float time, float remaining,
                                              int medium, Pcg32& rng, uint32_t* stk, TraceCount* tc = nullptr) {
  uint2 st = make_uint2(0u, 0u);
  uint32_t depth = 0;
  if (medium >= 0 && medium != S.atmosphere) stack_push(st, depth, (uint32_t)medium);
  V3 vis = v3(1.f, 1.f, 1.f);
  for (uint32_t i = 0; i < kMaxSkips; i++) {
    Hit h;
    const bool hit = scene_closest<kTraceBlock, MOTION, GEO, true, COUNT>(S, o, d, time, stk, h, tc);
    const float seg = fminf(hit ? h.t : remaining, remaining);
    const int cur = stack_top(st, depth, S.atmosphere);
    if (cur >= 0) {
      vis = vis * medium_transmittance(S.media[cur], o, d, seg, rng);
      if (maxc(vis) < 1e-4f) return v3(0.f, 0.f, 0.f);
    }
    if (!hit || h.t >= remaining) return vis;
    const PtMaterial& m = S.materials[h.material];
    if (m.hull) {
      if (h.front)
        stack_push(st, depth, (uint32_t)m.interior);
      else
        stack_pop(depth);
    } else {
      float presence = m.presence;
      if constexpr ((GEO & kGeoSt) != 0) {  // bound presence at the blocker
        if (m.bind) {
          const HitSt hs = hit_st<MOTION>(S, h, time);
          const PatCtx pc = pat_ctx(make_float4(hs.s, hs.t, hs.density, 0.f), make_float4(0.f, 0.f, 0.f, 0.f),
                                    make_float4(0.f, 0.f, 0.f, 0.f), h.p, h.n, d, 0.f);
          presence = pat_presence(S, m.bind, pc, presence);
        }
      }
      if (presence >= 1.f) return v3(0.f, 0.f, 0.f);
      vis = vis * (1.f - presence);
      if (maxc(vis) < 1e-4f) return v3(0.f, 0.f, 0.f);
    }
    if (S.ray_offset > 0.f) {
      o = o + d * (h.t + S.ray_offset);
      remaining -= h.t + S.ray_offset;
    } else {  // past the blocker: its point pushed to the far side
      const V3 on = offset_origin(h.p, -h.e);
      remaining -= dot(on - o, d);
      o = on;
    }
  }
  return vis;
}

// Shadow record i: adds each unoccluded ray's contribution to its path
// (the first ray, then the second when the record has one); with WALK the
// rays walk their blockers (transparent scenes), else one any-hit test.
template <bool WALK, bool MOTION, int GEO, bool COUNT = false>
__device__ __forceinline__ void shadow_one_t(const PtSceneView& S, const PathBuffers& P, uint32_t i, uint32_t* stk,
                                             TraceCount* tc = nullptr) {
  const float4 o = P.so[i], d = P.sd[i];
  float4 c = P.sc[i];
  const uint32_t cw = (uint32_t)c.w;  // bit 0: second ray, bit 1: diffuse AOV
  const bool two = (cw & 1u) != 0;
  const float time = MOTION ? P.stime[i] : 0.f;
  float4 c2 = make_float4(0.f, 0.f, 0.f, 0.f);
  bool v1, v2 = false;
  if constexpr (WALK) {
    const uint2 sd = P.srng[i];
    const uint64_t seed = ((uint64_t)sd.y << 32) | sd.x;
    Pcg32 r1 = Pcg32::make(seed, 1);
    const V3 t1 = walk_visibility<MOTION, GEO, COUNT>(S, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), time, o.w, (int)d.w,
                                                      r1, stk, tc);
    c = make_float4(c.x * t1.x, c.y * t1.y, c.z * t1.z, c.w);
    v1 = maxc(t1) > 0.f;
    if (two) {
      const float4 o2 = P.so2[i], d2 = P.sd2[i];
      Pcg32 r2 = Pcg32::make(seed, 2);
      const V3 t2 = walk_visibility<MOTION, GEO, COUNT>(S, v3(o2.x, o2.y, o2.z), v3(d2.x, d2.y, d2.z), time, o2.w,
                                                        (int)d2.w, r2, stk, tc);
      c2 = P.sc2[i];
      c2 = make_float4(c2.x * t2.x, c2.y * t2.y, c2.z * t2.z, 0.f);
      v2 = maxc(t2) > 0.f;
    }
  } else {
    v1 = !scene_occluded<kTraceBlock, MOTION, GEO, COUNT>(S, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), time, o.w, stk, tc);
    if (two) {
      const float4 o2 = P.so2[i], d2 = P.sd2[i];
      v2 = !scene_occluded<kTraceBlock, MOTION, GEO, COUNT>(S, v3(o2.x, o2.y, o2.z), v3(d2.x, d2.y, d2.z), time, o2.w,
                                                            stk, tc);
      c2 = P.sc2[i];
    }
  }
  if (!v1 && !v2) return;
  const uint32_t slot = P.sq[i];
  const auto add = [&](float4* buf) {
    float4 l = buf[slot];
    if (v1) {
      l.x += c.x;
      l.y += c.y;
      l.z += c.z;
    }
    if (v2) {
      l.x += c2.x;
      l.y += c2.y;
      l.z += c2.z;
    }
    buf[slot] = l;
  };
  add(P.L);
  if (P.ld && (cw & 2u)) add(P.ld);
}

template <bool QUAD>
__device__ __forceinline__ void shadow_one(const PtSceneView& S, const PathBuffers& P, uint32_t i, uint32_t* stk) {
  if (S.transparent)
    S.motion ? shadow_one_t<true, true, QUAD ? 3 : 0>(S, P, i, stk) : shadow_one_t<true, false, QUAD ? 3 : 0>(S, P, i, stk);
  else
    S.motion ? shadow_one_t<false, true, QUAD ? 3 : 0>(S, P, i, stk) : shadow_one_t<false, false, QUAD ? 3 : 0>(S, P, i, stk);
}

template <int STACK, bool WALK, bool MOTION, int GEO, bool COUNT = false>
__global__ __launch_bounds__(kTraceBlock) void shadow_kernel(PtSceneView S, PathBuffers P) {
  __shared__ uint32_t s_stack[STACK * kTraceBlock];
  uint32_t* stk = s_stack + threadIdx.x;
  const uint32_t n = P.counts[2];
  const uint32_t stride = gridDim.x * blockDim.x;
  TraceCount tc;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    shadow_one_t<WALK, MOTION, GEO, COUNT>(S, P, i, stk, &tc);
  if (COUNT) add_counts(P.prof + 3, tc);
}

// Launch of extend / shadow for the scene's stack size, transparency,
// motion and quadrics; with P.prof set, the counting variants.
template <int STACK, bool MOTION, int GEO, bool COUNT>
void launch_trace_counted(const PtSceneView& S, const PathBuffers& P, const uint32_t* q, uint32_t gt, bool extend,
                          hipStream_t st) {
  constexpr int SG = GEO & ~kGeoSt;  // shadow rays need texture coordinates only for a bound presence
  if (extend)
    hipLaunchKernelGGL((extend_kernel<STACK, MOTION, GEO, COUNT>), dim3(gt), dim3(kTraceBlock), 0, st, S, P, q);
  else if (S.transparent && (GEO & kGeoSt) && (S.patterns & 2u))
    hipLaunchKernelGGL((shadow_kernel<STACK, true, MOTION, GEO, COUNT>), dim3(gt), dim3(kTraceBlock), 0, st, S, P);
  else if (S.transparent)
    hipLaunchKernelGGL((shadow_kernel<STACK, true, MOTION, SG, COUNT>), dim3(gt), dim3(kTraceBlock), 0, st, S, P);
  else
    hipLaunchKernelGGL((shadow_kernel<STACK, false, MOTION, SG, COUNT>), dim3(gt), dim3(kTraceBlock), 0, st, S, P);
}

template <int STACK, bool MOTION, int GEO>
void launch_trace_kernels(const PtSceneView& S, const PathBuffers& P, const uint32_t* q, uint32_t gt, bool extend,
                          hipStream_t st) {
  if (P.prof)
    launch_trace_counted<STACK, MOTION, GEO, true>(S, P, q, gt, extend, st);
  else
    launch_trace_counted<STACK, MOTION, GEO, false>(S, P, q, gt, extend, st);
}

// Scenes with pattern-bound materials take full-stack extension kernels
// that also record each hit's texture coordinates, with the geometry
// classes chosen as below (their shadow rays take the kernels below).
// Scenes with loose quadrics take the full-stack kernels with the quadric
// solvers (and the curve test, for scenes with both), scenes with curve
// sets alone the full-stack kernels with the curve test only (the f64
// quartic's registers stay out of them); one instantiation per motion
// flag. Every other scene keeps the lean kernels.
void launch_trace(const PtSceneView& S, const PathBuffers& P, const uint32_t* q, uint32_t gt, bool extend,
                  hipStream_t st) {
  const bool small = S.stack_need <= 32;
  constexpr int kAll = kGeoQuadrics | kGeoCurves;
  if (S.patterns) {  // hits also record their texture coordinates
    if (S.num_quadrics && S.motion)
      launch_trace_kernels<kTraceStack, true, kAll | kGeoSt>(S, P, q, gt, extend, st);
    else if (S.num_quadrics)
      launch_trace_kernels<kTraceStack, false, kAll | kGeoSt>(S, P, q, gt, extend, st);
    else if (S.curves && S.motion)
      launch_trace_kernels<kTraceStack, true, kGeoCurves | kGeoSt>(S, P, q, gt, extend, st);
    else if (S.curves)
      launch_trace_kernels<kTraceStack, false, kGeoCurves | kGeoSt>(S, P, q, gt, extend, st);
    else if (S.motion)
      launch_trace_kernels<kTraceStack, true, kGeoSt>(S, P, q, gt, extend, st);
    else
      launch_trace_kernels<kTraceStack, false, kGeoSt>(S, P, q, gt, extend, st);
  } else if (S.num_quadrics && S.motion)
    launch_trace_kernels<kTraceStack, true, kAll>(S, P, q, gt, extend, st);
  else if (S.num_quadrics)
    launch_trace_kernels<kTraceStack, false, kAll>(S, P, q, gt, extend, st);
  else if (S.curves && S.motion)
    launch_trace_kernels<kTraceStack, true, kGeoCurves>(S, P, q, gt, extend, st);
  else if (S.curves)
    launch_trace_kernels<kTraceStack, false, kGeoCurves>(S, P, q, gt, extend, st);
  else if (small && S.motion)
    launch_trace_kernels<32, true, 0>(S, P, q, gt, extend, st);
  else if (small)
    launch_trace_kernels<32, false, 0>(S, P, q, gt, extend, st);
  else if (S.motion)
    launch_trace_kernels<kTraceStack, true, 0>(S, P, q, gt, extend, st);
  else
    launch_trace_kernels<kTraceStack, false, 0>(S, P, q, gt, extend, st);
}

__global__ void rotate_kernel(PathBuffers P) {
  P.stats[0] += P.counts[0];
  P.stats[1] += P.counts[2];
  P.counts[0] = P.counts[1];
  P.counts[1] = 0;
  P.counts[2] = 0;
}

// Adds column col's pass samples (slots slot_base + k * npix + col) into
// its pixel's accumulator in sample order, with the firefly clamp
// (pt/mod.rs:88-95): sum of rgb, sum of rgb squared, count, and the sum of
// squared luminance that adaptive sampling reads (pt/mod.rs:314-318).
__device__ __forceinline__ void resolve_one(const PathBuffers& P, const PtParams& prm, double* accum, uint32_t col,
                                            uint32_t slot_base) {
  double s[3] = {0, 0, 0}, q[3] = {0, 0, 0}, ql = 0.0;
  for (uint32_t k = 0; k < prm.spp_pass; k++) {
    const float4 l4 = P.L[slot_base + (size_t)k * prm.npix + col];
    double l[3] = {l4.x, l4.y, l4.z};
    double lm = 0.2126 * l[0] + 0.7152 * l[1] + 0.0722 * l[2];
    if (prm.firefly_clamp > 0.f && lm > (double)prm.firefly_clamp) {
      const double sc = (double)prm.firefly_clamp / lm;
      l[0] *= sc;
      l[1] *= sc;
      l[2] *= sc;
      lm = 0.2126 * l[0] + 0.7152 * l[1] + 0.0722 * l[2];
    }
    for (int c = 0; c < 3; c++) {
      s[c] += l[c];
      q[c] += l[c] * l[c];
    }
    ql += lm * lm;
  }
  const uint32_t px = prm.live ? prm.live[col] : col;
  double* a = accum + 8 * (size_t)px;
  for (int c = 0; c < 3; c++) {
    a[c] += s[c];
    a[3 + c] += q[c];
  }
  a[6] += (double)prm.spp_pass;
  a[7] += ql;
}

__global__ __launch_bounds__(kBlock) void resolve_kernel(PathBuffers P, PtParams prm, double* __restrict__ accum) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < prm.npix; c += stride) resolve_one(P, prm, accum, c, 0);
}

// Film accumulators of column col's pass samples, in sample order
// (pt/mod.rs:200-221): the diffuse split with the sample's firefly scale,
// albedo and normal sums, the depth sum and count over samples with a
// first real hit, and the id vote (kFilmIds slots of id and count, then
// the count of samples whose id found no slot).
__device__ __forceinline__ void resolve_film_one(const PathBuffers& P, const PtParams& prm, double* film,
                                                 uint32_t col) {
  const uint32_t px = prm.live ? prm.live[col] : col;
  double* a = film + (size_t)kFilmStride * px;
  double ds[3] = {0, 0, 0}, al[3] = {0, 0, 0}, nm[3] = {0, 0, 0}, dep = 0.0, hits = 0.0, over = 0.0;
  uint32_t ids[kFilmIds], cnt[kFilmIds];
#pragma unroll
  for (int j = 0; j < (int)kFilmIds; j++) {
    ids[j] = (uint32_t)a[11 + j];
    cnt[j] = (uint32_t)a[11 + kFilmIds + j];
  }
  for (uint32_t k = 0; k < prm.spp_pass; k++) {
    const size_t slot = (size_t)k * prm.npix + col;
    const float4 l4 = P.L[slot], d4 = P.ld[slot], x0 = P.aux0[slot], x1 = P.aux1[slot];
    const uint32_t id = P.aux_id[slot];
    double sc = 1.0;
    const double lm = 0.2126 * (double)l4.x + 0.7152 * (double)l4.y + 0.0722 * (double)l4.z;
    if (prm.firefly_clamp > 0.f && lm > (double)prm.firefly_clamp) sc = (double)prm.firefly_clamp / lm;
    ds[0] += (double)d4.x * sc;
    ds[1] += (double)d4.y * sc;
    ds[2] += (double)d4.z * sc;
    al[0] += x0.x;
    al[1] += x0.y;
    al[2] += x0.z;
    nm[0] += x1.x;
    nm[1] += x1.y;
    nm[2] += x1.z;
    if (x0.w > 0.f) {
      dep += x0.w;
      hits += 1.0;
    }
    bool placed = false;
#pragma unroll
    for (int j = 0; j < (int)kFilmIds; j++) {
      if (!placed && cnt[j] != 0 && ids[j] == id) {
        cnt[j]++;
        placed = true;
      }
    }
#pragma unroll
    for (int j = 0; j < (int)kFilmIds; j++) {
      if (!placed && cnt[j] == 0) {
        ids[j] = id;
        cnt[j] = 1;
        placed = true;
      }
    }
    if (!placed) over += 1.0;
  }
  for (int c = 0; c < 3; c++) {
    a[c] += ds[c];
    a[3 + c] += al[c];
    a[6 + c] += nm[c];
  }
  a[9] += dep;
  a[10] += hits;
#pragma unroll
  for (int j = 0; j < (int)kFilmIds; j++) {
    a[11 + j] = (double)ids[j];
    a[11 + kFilmIds + j] = (double)cnt[j];
  }
  a[11 + 2 * kFilmIds] += over;
}

__global__ __launch_bounds__(kBlock) void resolve_film_kernel(PathBuffers P, PtParams prm, double* __restrict__ film) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < prm.npix; c += stride) resolve_film_one(P, prm, film, c);
}

// ---- batched jobs: one set of launches over the path slots of many scenes ----

// Last job whose begin (slot_begin or pix_begin) is <= x.
template <bool PIX>
__device__ __forceinline__ uint32_t find_job(const PtJob* jobs, uint32_t nj, uint32_t x) {
  uint32_t lo = 0, hi = nj;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((PIX ? jobs[mid].pix_begin : jobs[mid].slot_begin) <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void raygen_batch_kernel(const PtJob* __restrict__ jobs, uint32_t nj, PathBuffers P,
                                                             uint32_t total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    P.counts[0] = total;
    P.counts[1] = 0;
    P.counts[2] = 0;
  }
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < total; slot += stride) {
    const uint32_t j = find_job<false>(jobs, nj, slot);
    const PtJob& J = jobs[j];
    const uint32_t local = slot - J.slot_begin;
    init_path(J.S, P, slot, local % J.prm.npix, J.prm.sample0 + local / J.prm.npix, kFlagFromCamera);
    P.job[slot] = j;
  }
}

template <bool QUAD>
__global__ __launch_bounds__(kTraceBlock) void extend_batch_kernel(const PtJob* __restrict__ jobs, PathBuffers P,
                                                                  const uint32_t* __restrict__ q) {
  __shared__ uint32_t s_stack[kTraceStack * kTraceBlock];
  uint32_t* stk = s_stack + threadIdx.x;
  const uint32_t n = P.counts[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t slot = q[i];
    const PtSceneView& S = jobs[P.job[slot]].S;
    if (S.motion)
      extend_one<true, QUAD ? 3 : 0>(S, P, slot, stk);
    else
      extend_one<false, QUAD ? 3 : 0>(S, P, slot, stk);
  }
}

__global__ __launch_bounds__(kBlock) void shade_batch_kernel(const PtJob* __restrict__ jobs, PathBuffers P,
                                                            const uint32_t* __restrict__ q, uint32_t* __restrict__ qnext) {
  const uint32_t n = P.counts[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  const NrcPathBuffers N{};
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const bool act = i < n;
    bool cont = false, shadow = false, shadow2 = false, query = false;
    ShadowOut so, se;
    NrcQuery nq;
    V3 qbeta;
    uint32_t slot = 0;
    if (act) {
      slot = q[i];
      const PtJob& J = jobs[P.job[slot]];
      const bool fixed = J.S.ray_offset > 0.f;
      if (J.S.transparent && fixed)
        shade_one<false, true, false, true>(J.S, P, J.prm, N, slot, cont, shadow, shadow2, se, query, so, nq, qbeta);
      else if (J.S.transparent)
        shade_one<false, true>(J.S, P, J.prm, N, slot, cont, shadow, shadow2, se, query, so, nq, qbeta);
      else if (fixed)
        shade_one<false, false, false, true>(J.S, P, J.prm, N, slot, cont, shadow, shadow2, se, query, so, nq, qbeta);
      else
        shade_one<false, false>(J.S, P, J.prm, N, slot, cont, shadow, shadow2, se, query, so, nq, qbeta);
    }
    if (P.srng)
      push_shadow<true>(P, act, slot, shadow, so, shadow2, se);
    else
      push_shadow<false>(P, act, slot, shadow, so, shadow2, se);
    const uint32_t nidx = wave_append(&P.counts[1], act && cont);
    if (act && cont) qnext[nidx] = slot;
  }
}

template <bool QUAD>
__global__ __launch_bounds__(kTraceBlock) void shadow_batch_kernel(const PtJob* __restrict__ jobs, PathBuffers P) {
  __shared__ uint32_t s_stack[kTraceStack * kTraceBlock];
  uint32_t* stk = s_stack + threadIdx.x;
  const uint32_t n = P.counts[2];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    shadow_one<QUAD>(jobs[P.job[P.sq[i]]].S, P, i, stk);
}

__global__ __launch_bounds__(kBlock) void resolve_batch_kernel(const PtJob* __restrict__ jobs, uint32_t nj, PathBuffers P,
                                                              uint32_t total_pix) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total_pix; g += stride) {
    const PtJob& J = jobs[find_job<true>(jobs, nj, g)];
    resolve_one(P, J.prm, J.accum, g - J.pix_begin, J.slot_begin);
  }
}

// ---- batched jobs in one persistent launch ---------------------------------------
// Every lane owns one path slot and pulls work items — one pixel of one
// job with all of its samples — from a device counter until the items run
// out; items are numbered job by job (pix_begin) in the order the host
// chose (shortest job first). A pixel's samples run one after another
// through the same device functions the wavefront kernels call (camera
// ray, closest hit, shading, shadow records evaluated in place), and are
// added in sample order into the pixel's accumulator row. The last pixel
// of a job records the device clock, so per-job completion is measured
// inside the one launch. There are no barriers: a lane exits when the
// counter passes the last item.

// Samples of one (pixel, sample) path through the bounce loop in slot
// `slot`; returns its radiance (P.L) and counts rays and shadow records.
// One bounce of the path in slot `slot` (closest hit, shading, its shadow
// record evaluated in place); returns whether the path continues and
// counts rays and shadow records. TR: the job is transparent (its own
// scene's flag, as its session shades it); MOTION / QUAD: some job of the
// batch moves / holds quadrics (the kernel is instantiated for the
// batch's union of them; static and quadric-free scenes take the same
// paths through those templates).
template <bool TR, bool MOTION, bool QUAD>
__device__ __forceinline__ bool persist_bounce(const PtJob& J, const PathBuffers& P, uint32_t slot, uint32_t* stk,
                                               unsigned long long& ext, unsigned long long& shd) {
  const NrcPathBuffers N{};
  extend_one<MOTION, QUAD ? 3 : 0>(J.S, P, slot, stk);
  ext++;
  bool cont = false, shadow = false, shadow2 = false, query = false;
  ShadowOut so, se;
  NrcQuery nq;
  V3 qbeta;
  shade_one<false, TR>(J.S, P, J.prm, N, slot, cont, shadow, shadow2, se, query, so, nq, qbeta);
  if (shadow || shadow2) {
    put_shadow<TR>(P, slot, slot, shadow, so, shadow2, se);
    shadow_one_t<TR, MOTION, QUAD ? 3 : 0>(J.S, P, slot, stk);
    shd++;
  }
  return cont;
}

// A lane runs one flat loop of bounces: when its path ends it adds the
// sample to its pixel's sums and starts the pixel's next sample in the same
// iteration, and when the pixel is done it writes the pixel's accumulator
// row and pulls the next item, so no lane waits for the others' paths.
template <bool TR, bool MOTION, bool QUAD>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(2))) void batch_persistent_kernel(
    const PtJob* __restrict__ jobs, uint32_t nj, PathBuffers P, uint32_t total_items, uint32_t* __restrict__ next_item,
    uint32_t* __restrict__ job_left, unsigned long long* __restrict__ job_clock) {
  __shared__ uint32_t s_stack[kTraceStack * kTraceBlock];
  uint32_t* stk = s_stack + threadIdx.x;
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (threadIdx.x == 0) atomicMin(job_clock + nj, (unsigned long long)wall_clock64());
  unsigned long long ext = 0, shd = 0;
  uint32_t j = 0, pixel = 0, k = 0, b = 0, rounds = 0;
  bool have_item = false, have_path = false;
  double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0}, ql = 0.0;
  for (;;) {
    if (!have_path) {
      if (have_item && k == jobs[j].prm.spp_pass) {  // the pixel is done
        const PtJob& J = jobs[j];
        double* a = J.accum + 8 * (size_t)pixel;
        for (int c = 0; c < 3; c++) {
          a[c] += sum[c];
          a[3 + c] += sq[c];
        }
        a[6] += (double)J.prm.spp_pass;
        a[7] += ql;
        if (atomicSub(job_left + j, 1u) == 1u) job_clock[j] = wall_clock64();
        have_item = false;
      }
      if (!have_item) {
        const uint32_t item = atomicAdd(next_item, 1u);
        if (item >= total_items) break;
        j = find_job<true>(jobs, nj, item);
        pixel = item - jobs[j].pix_begin;
        k = 0;
        have_item = true;
        for (int c = 0; c < 3; c++) sum[c] = sq[c] = 0.0;
        ql = 0.0;
      }
      const PtJob& J = jobs[j];
      init_path(J.S, P, slot, pixel, J.prm.sample0 + k, kFlagFromCamera);
      rounds = J.prm.max_depth + (TR && J.S.transparent ? kPtMaxSkips : 0);
      b = 0;
      have_path = true;
    }
    const PtJob& J = jobs[j];
    const bool cont = TR && J.S.transparent ? persist_bounce<true, MOTION, QUAD>(J, P, slot, stk, ext, shd)
                                            : persist_bounce<false, MOTION, QUAD>(J, P, slot, stk, ext, shd);
    if (!cont || ++b >= rounds) {
      // the sample as resolve_one adds it (firefly clamp on luminance)
      const float4 l4 = P.L[slot];
      double l[3] = {l4.x, l4.y, l4.z};
      double lm = 0.2126 * l[0] + 0.7152 * l[1] + 0.0722 * l[2];
      if (J.prm.firefly_clamp > 0.f && lm > (double)J.prm.firefly_clamp) {
        const double sc = (double)J.prm.firefly_clamp / lm;
        l[0] *= sc;
        l[1] *= sc;
        l[2] *= sc;
        lm = 0.2126 * l[0] + 0.7152 * l[1] + 0.0722 * l[2];
      }
      for (int c = 0; c < 3; c++) {
        sum[c] += l[c];
        sq[c] += l[c] * l[c];
      }
      ql += lm * lm;
      k++;
      have_path = false;
    }
  }
  // rays traced (per wave, one atomic); every lane of the wave has left the
  // loop here
  for (int off = 32; off > 0; off >>= 1) {
    ext += __shfl_xor(ext, off);
    shd += __shfl_xor(shd, off);
  }
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(P.stats + 0, ext);
    atomicAdd(P.stats + 1, shd);
  }
}

// ---- shading-class sort of the live queue (before shade) ------------------------
// Key: 0 = miss, else the hit material's sort_key (1 + active-lobe mask).
// Counting sort over fixed contiguous block ranges: per-block histograms,
// one scan over (class, block), then each block scatters its range.
constexpr uint32_t kSortKeys = 33;
constexpr uint32_t kSortBlocks = 512;

__device__ __forceinline__ uint32_t sort_key(const PtSceneView& S, const PathBuffers& P, uint32_t slot) {
  if (P.hit_n[slot].w < 0.f) return 0;
  return S.materials[P.hit_m[slot] & 0x7FFFFFFFu].sort_key;
}

__device__ __forceinline__ void block_range(uint32_t n, uint32_t& b0, uint32_t& b1) {
  const uint32_t chunk = (n + kSortBlocks - 1) / kSortBlocks;
  b0 = min(n, blockIdx.x * chunk);
  b1 = min(n, b0 + chunk);
}

__global__ __launch_bounds__(kBlock) void sort_count_kernel(PtSceneView S, PathBuffers P, const uint32_t* __restrict__ q) {
  __shared__ uint32_t h[kSortKeys];
  for (uint32_t i = threadIdx.x; i < kSortKeys; i += blockDim.x) h[i] = 0;
  __syncthreads();
  uint32_t b0, b1;
  block_range(P.counts[0], b0, b1);
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) atomicAdd(&h[sort_key(S, P, q[i])], 1u);
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kSortKeys; k += blockDim.x) P.sort_hist[k * kSortBlocks + blockIdx.x] = h[k];
}

// Exclusive scan of the (class, block) counts in place.
__global__ __launch_bounds__(1024) void sort_scan_kernel(PathBuffers P) {
  __shared__ uint32_t part[1024];
  constexpr uint32_t n = kSortKeys * kSortBlocks;
  const uint32_t t = threadIdx.x;
  const uint32_t chunk = (n + 1023) / 1024;
  const uint32_t b = t * chunk, e = min(n, b + chunk);
  uint32_t s = 0;
  for (uint32_t i = b; i < e; i++) s += P.sort_hist[i];
  part[t] = s;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
  for (uint32_t i = b; i < e; i++) {
    const uint32_t c = P.sort_hist[i];
    P.sort_hist[i] = run;
    run += c;
  }
}

__global__ __launch_bounds__(kBlock) void sort_scatter_kernel(PtSceneView S, PathBuffers P, const uint32_t* __restrict__ q) {
  __shared__ uint32_t cur[kSortKeys];
  for (uint32_t k = threadIdx.x; k < kSortKeys; k += blockDim.x) cur[k] = P.sort_hist[k * kSortBlocks + blockIdx.x];
  __syncthreads();
  uint32_t b0, b1;
  block_range(P.counts[0], b0, b1);
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const uint32_t slot = q[i];
    P.qs[atomicAdd(&cur[sort_key(S, P, slot)], 1u)] = slot;
  }
}

uint32_t grid_for(uint64_t n, int block = kBlock) {
  const uint64_t b = (n + block - 1) / block;
  return (uint32_t)(b < 1 ? 1 : (b > kMaxGrid ? kMaxGrid : b));
}

}  // namespace

__global__ void set_u64_kernel(uint64_t* dst, uint64_t v) { *dst = v; }

hipError_t pt_launch_raygen(const PtSceneView& S, const PathBuffers& P, const PtParams& prm, uint32_t total,
                            const NrcPathBuffers& N, const uint64_t* sample0_dev, hipStream_t st) {
  hipLaunchKernelGGL(raygen_kernel, dim3(grid_for((uint64_t)total + N.ntrain)), dim3(kBlock), 0, st, S, P, prm, total,
                     N, sample0_dev);
  return hipGetLastError();
}

hipError_t pt_launch_set_u64(uint64_t* dst, uint64_t v, hipStream_t st) {
  hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, dst, v);
  return hipGetLastError();
}

// The shade kernel for the pass's cache, film, transparency and hair, with
// render.rs's fixed origin push (FIXED) or error-bounded origins (caches
// are never attached to scenes with hair).
template <bool FIXED>
void launch_shade(const PtSceneView& S, const PathBuffers& P, const PtParams& prm, const uint32_t* qshade,
                  uint32_t* qnext, const NrcPathBuffers& N, uint32_t g, hipStream_t st) {
  const dim3 G(g), B(kBlock);
  if (S.patterns) {
    if (P.ld && S.transparent)
      hipLaunchKernelGGL((shade_pattern_kernel<true, true, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
    else if (P.ld)
      hipLaunchKernelGGL((shade_pattern_kernel<false, true, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
    else if (S.transparent)
      hipLaunchKernelGGL((shade_pattern_kernel<true, false, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
    else
      hipLaunchKernelGGL((shade_pattern_kernel<false, false, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
  } else if (S.hair) {
    if (P.ld && S.transparent)
      hipLaunchKernelGGL((shade_hair_kernel<true, true, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
    else if (P.ld)
      hipLaunchKernelGGL((shade_hair_kernel<false, true, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
    else if (S.transparent)
      hipLaunchKernelGGL((shade_hair_kernel<true, false, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
    else
      hipLaunchKernelGGL((shade_hair_kernel<false, false, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
  } else if (N.q) {
    hipLaunchKernelGGL((shade_kernel<true, false, false, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
  } else if (P.ld && S.transparent) {
    hipLaunchKernelGGL((shade_kernel<false, true, true, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
  } else if (P.ld) {
    hipLaunchKernelGGL((shade_kernel<false, false, true, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
  } else if (S.transparent) {
    hipLaunchKernelGGL((shade_kernel<false, true, false, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
  } else {
    hipLaunchKernelGGL((shade_kernel<false, false, false, FIXED>), G, B, 0, st, S, P, prm, qshade, qnext, N);
  }
}

hipError_t pt_launch_bounce(const PtSceneView& S, const PathBuffers& P, const PtParams& prm, const uint32_t* q,
                            uint32_t* qnext, uint32_t capacity, const NrcPathBuffers& N, hipStream_t st,
                            const hipEvent_t* ev) {
  const uint32_t g = grid_for(capacity);
  const uint32_t gt = grid_for(capacity, kTraceBlock);
  if (ev) (void)hipEventRecord(ev[0], st);
  launch_trace(S, P, q, gt, true, st);
  if (ev) (void)hipEventRecord(ev[1], st);
  const uint32_t* qshade = q;
  if (P.qs) {  // group the queue by shading class so shade waves run one BSDF kind
    hipLaunchKernelGGL(sort_count_kernel, dim3(kSortBlocks), dim3(kBlock), 0, st, S, P, q);
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(1024), 0, st, P);
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(kSortBlocks), dim3(kBlock), 0, st, S, P, q);
    qshade = P.qs;
  }
  if (S.ray_offset > 0.f)
    launch_shade<true>(S, P, prm, qshade, qnext, N, g, st);
  else
    launch_shade<false>(S, P, prm, qshade, qnext, N, g, st);
  if (ev) (void)hipEventRecord(ev[2], st);
  launch_trace(S, P, nullptr, gt, false, st);
  if (ev) (void)hipEventRecord(ev[3], st);
  hipLaunchKernelGGL(rotate_kernel, dim3(1), dim3(1), 0, st, P);
  return hipGetLastError();
}

hipError_t pt_launch_batch_pass(const PtJob* jobs, uint32_t nj, const PathBuffers& P, uint32_t total,
                                 uint32_t total_pix, uint32_t bounces, bool quadrics, hipStream_t st) {
  const uint32_t g = grid_for(total);
  const uint32_t gt = grid_for(total, kTraceBlock);
  hipLaunchKernelGGL(raygen_batch_kernel, dim3(g), dim3(kBlock), 0, st, jobs, nj, P, total);
  for (uint32_t b = 0; b < bounces; b++) {
    const uint32_t* q = (b & 1) ? P.q1 : P.q0;
    uint32_t* qn = (b & 1) ? P.q0 : P.q1;
    if (quadrics)
      hipLaunchKernelGGL(extend_batch_kernel<true>, dim3(gt), dim3(kTraceBlock), 0, st, jobs, P, q);
    else
      hipLaunchKernelGGL(extend_batch_kernel<false>, dim3(gt), dim3(kTraceBlock), 0, st, jobs, P, q);
    hipLaunchKernelGGL(shade_batch_kernel, dim3(g), dim3(kBlock), 0, st, jobs, P, q, qn);
    if (quadrics)
      hipLaunchKernelGGL(shadow_batch_kernel<true>, dim3(gt), dim3(kTraceBlock), 0, st, jobs, P);
    else
      hipLaunchKernelGGL(shadow_batch_kernel<false>, dim3(gt), dim3(kTraceBlock), 0, st, jobs, P);
    hipLaunchKernelGGL(rotate_kernel, dim3(1), dim3(1), 0, st, P);
  }
  hipLaunchKernelGGL(resolve_batch_kernel, dim3(grid_for(total_pix)), dim3(kBlock), 0, st, jobs, nj, P, total_pix);
  return hipGetLastError();
}

namespace {
// The persistent kernel of a batch's flags (flags bit 0 transparent, 1
// motion, 2 quadrics).
using PersistFn = void (*)(const PtJob*, uint32_t, PathBuffers, uint32_t, uint32_t*, uint32_t*, unsigned long long*);
PersistFn persist_fn(uint32_t flags) {
  switch (flags & 7u) {
    case 0: return batch_persistent_kernel<false, false, false>;
    case 1: return batch_persistent_kernel<true, false, false>;
    case 2: return batch_persistent_kernel<false, true, false>;
    case 3: return batch_persistent_kernel<true, true, false>;
    case 4: return batch_persistent_kernel<false, false, true>;
    case 5: return batch_persistent_kernel<true, false, true>;
    case 6: return batch_persistent_kernel<false, true, true>;
    default: return batch_persistent_kernel<true, true, true>;
  }
}
}  // namespace

hipError_t pt_launch_batch_persistent(const PtJob* jobs, uint32_t nj, const PathBuffers& P, uint32_t total_items,
                                      uint32_t lanes, uint32_t flags, uint32_t* next_item, uint32_t* job_left,
                                      unsigned long long* job_clock, hipStream_t st) {
  hipLaunchKernelGGL(persist_fn(flags), dim3(lanes / kTraceBlock), dim3(kTraceBlock), 0, st, jobs, nj, P,
                     total_items, next_item, job_left, job_clock);
  return hipGetLastError();
}

uint32_t pt_batch_persistent_lanes(uint32_t flags) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_fn(flags), kTraceBlock, 0) != hipSuccess)
    return 0;
  if (per_cu <= 0 || cus <= 0) return 0;
  return (uint32_t)(per_cu * cus) * kTraceBlock;
}

namespace {
// One fiber-BSDF query per thread (rrhip_hair_bsdf).
__global__ __launch_bounds__(kBlock) void hair_bsdf_kernel(PtMaterial m, uint32_t n, const float* __restrict__ wo,
                                                          const float* __restrict__ wi, const float* __restrict__ h,
                                                          const float* __restrict__ u, float* __restrict__ f,
                                                          float* __restrict__ pdf, float* __restrict__ ws,
                                                          float* __restrict__ fs, float* __restrict__ pdfs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const HairGeom g = hair_geom(m, v3(wo + 3 * (size_t)i), h[i]);
  V3 fv;
  float pv;
  hair_eval(m, g, v3(wi + 3 * (size_t)i), fv, pv);
  f[3 * (size_t)i] = fv.x;
  f[3 * (size_t)i + 1] = fv.y;
  f[3 * (size_t)i + 2] = fv.z;
  pdf[i] = pv;
  const float uu[4] = {u[4 * (size_t)i], u[4 * (size_t)i + 1], u[4 * (size_t)i + 2], u[4 * (size_t)i + 3]};
  V3 w, sf;
  float sp;
  const bool ok = hair_sample(m, g, uu, w, sf, sp);
  ws[3 * (size_t)i] = w.x;
  ws[3 * (size_t)i + 1] = w.y;
  ws[3 * (size_t)i + 2] = w.z;
  fs[3 * (size_t)i] = sf.x;
  fs[3 * (size_t)i + 1] = sf.y;
  fs[3 * (size_t)i + 2] = sf.z;
  pdfs[i] = ok ? sp : 0.f;
}
}  // namespace

hipError_t pt_launch_hair_bsdf(const PtMaterial& m, uint32_t n, const float* wo, const float* wi, const float* h,
                               const float* u, float* f, float* pdf, float* ws, float* fs, float* pdfs,
                               hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(hair_bsdf_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, m, n, wo, wi, h, u, f,
                     pdf, ws, fs, pdfs);
  return hipGetLastError();
}

hipError_t pt_launch_pcg32_draws(const uint64_t* pixel, const uint64_t* sample, uint32_t n, uint32_t count,
                                 uint32_t* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(pcg32_draws_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, pixel, sample, n,
                     count, out);
  return hipGetLastError();
}

hipError_t pt_launch_resolve(const PathBuffers& P, const PtParams& prm, double* accum, double* film,
                             hipStream_t st) {
  hipLaunchKernelGGL(resolve_kernel, dim3(grid_for(prm.npix)), dim3(kBlock), 0, st, P, prm, accum);
  if (film) hipLaunchKernelGGL(resolve_film_kernel, dim3(grid_for(prm.npix)), dim3(kBlock), 0, st, P, prm, film);
  return hipGetLastError();
}

}  // namespace rrhip

namespace rrhip {
namespace {

// relMSE partial sums (f64): per pixel and channel (x - r)^2 / (r^2 + eps)
// with x, r the per-pixel means of two accumulators.
__global__ __launch_bounds__(256) void relmse_kernel(const double* __restrict__ a, const double* __restrict__ r,
                                                    uint32_t npix, double eps, double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += stride) {
    const double* pa = a + 8 * (size_t)p;
    const double* pr = r + 8 * (size_t)p;
    const double na = pa[6] > 0.0 ? pa[6] : 1.0, nr = pr[6] > 0.0 ? pr[6] : 1.0;
    for (int c = 0; c < 3; c++) {
      const double x = pa[c] / na, ref = pr[c] / nr;
      const double d = x - ref;
      acc += d * d / (ref * ref + eps);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

}  // namespace

hipError_t pt_launch_relmse(const double* accum, const double* ref_accum, uint32_t npix, double eps,
                            double* partial, uint32_t blocks, hipStream_t st) {
  hipLaunchKernelGGL(relmse_kernel, dim3(blocks), dim3(256), 0, st, accum, ref_accum, npix, eps, partial);
  return hipGetLastError();
}

namespace {

// Per-block partial sums of the accumulators' sample counts.
__global__ __launch_bounds__(256) void count_sum_kernel(const double* __restrict__ a, uint32_t npix,
                                                        double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += stride) acc += a[8 * (size_t)p + 6];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adaptive stopping rule of pt/mod.rs:319-325: after the 32-sample warm-up
// a pixel stops once the 95 % confidence half-width of its mean luminance,
// 1.96 sqrt(var / n) with var the unbiased sample variance, is below
// tol * max(mean, 0.05). The moments come from the accumulator's sums
// (luminance of the rgb sum, sum of squared luminance, count), evaluated
// without fused operations so that the host can repeat the decision.
__device__ bool pixel_converged(const double* a, double tol) {
#pragma clang fp contract(off)
  const double n = a[6];
  if (n < 32.0) return false;
  const double sum = 0.2126 * a[0] + 0.7152 * a[1] + 0.0722 * a[2];
  const double mean = sum / n;
  const double var = fmax((a[7] - sum * mean) / (n - 1.0), 0.0);
  return 1.96 * sqrt(var / n) < tol * fmax(mean, 0.05);
}

constexpr uint32_t kPruneBlock = 256;

// Column c of the live list `in` (null: column = pixel) keeps sampling.
__device__ __forceinline__ bool keep_column(const double* accum, const uint32_t* in, uint32_t c, double tol,
                                            uint32_t* px) {
  *px = in ? in[c] : c;
  return !pixel_converged(accum + 8 * (size_t)*px, tol);
}

__global__ __launch_bounds__(kPruneBlock) void prune_count_kernel(const double* __restrict__ accum,
                                                                 const uint32_t* __restrict__ in, uint32_t n,
                                                                 double tol, uint32_t* __restrict__ bcount) {
  const uint32_t c = blockIdx.x * kPruneBlock + threadIdx.x;
  uint32_t px = 0;
  const bool k = c < n && keep_column(accum, in, c, tol, &px);
  const int cnt = __syncthreads_count(k);
  if (threadIdx.x == 0) bcount[blockIdx.x] = (uint32_t)cnt;
}

// Exclusive prefix of the nb block counts in place; the total into *total.
__global__ __launch_bounds__(1024) void prune_scan_kernel(uint32_t* __restrict__ bcount, uint32_t nb,
                                                         uint32_t* __restrict__ total) {
  __shared__ uint32_t part[1024];
  const uint32_t per = (nb + 1023) / 1024;
  const uint32_t b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; b++) sum += bcount[b];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - sum;
  for (uint32_t b = b0; b < b1; b++) {
    const uint32_t c = bcount[b];
    bcount[b] = run;
    run += c;
  }
  if (threadIdx.x == 1023) *total = part[1023];
}

// Writes the pixels that keep sampling, in column order.
__global__ __launch_bounds__(kPruneBlock) void prune_scatter_kernel(const double* __restrict__ accum,
                                                                   const uint32_t* __restrict__ in, uint32_t n,
                                                                   double tol, const uint32_t* __restrict__ boff,
                                                                   uint32_t* __restrict__ out) {
  __shared__ uint32_t wsum[kPruneBlock / 64];
  const uint32_t c = blockIdx.x * kPruneBlock + threadIdx.x;
  uint32_t px = 0;
  const bool k = c < n && keep_column(accum, in, c, tol, &px);
  const uint64_t m = __ballot(k);
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t base = boff[blockIdx.x];
  for (uint32_t i = 0; i < w; i++) base += wsum[i];
  if (k) out[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = px;
}

}  // namespace

hipError_t pt_launch_count_sum(const double* accum, uint32_t npix, double* partial, uint32_t blocks,
                               hipStream_t st) {
  hipLaunchKernelGGL(count_sum_kernel, dim3(blocks), dim3(256), 0, st, accum, npix, partial);
  return hipGetLastError();
}

hipError_t pt_launch_adaptive_prune(const double* accum, const uint32_t* in, uint32_t n, double tol,
                                    uint32_t* bcount, uint32_t* out, uint32_t* nlive, hipStream_t st) {
  const uint32_t nb = (n + kPruneBlock - 1) / kPruneBlock;
  if (nb == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(prune_count_kernel, dim3(nb), dim3(kPruneBlock), 0, st, accum, in, n, tol, bcount);
  hipLaunchKernelGGL(prune_scan_kernel, dim3(1), dim3(1024), 0, st, bcount, nb, nlive);
  hipLaunchKernelGGL(prune_scatter_kernel, dim3(nb), dim3(kPruneBlock), 0, st, accum, in, n, tol, bcount, out);
  return hipGetLastError();
}

}  // namespace rrhip

namespace rrhip {
namespace {

// Target of training record r (path radiance `lf` at the end): the
// radiance scattered at the vertex, (L_end - L_at_vertex) / throughput,
// with the pixel estimator's firefly clamp on its luminance. Records whose
// throughput has a channel below 1e-6 carry no usable target.
__device__ __forceinline__ bool nrc_target(const NrcPathBuffers& N, size_t r, float4 lf, float clamp, float t[3]) {
  const float4 b = N.tb[r], l = N.tl[r];
  const float bb[3] = {b.x, b.y, b.z}, ll[3] = {l.x, l.y, l.z}, ff[3] = {lf.x, lf.y, lf.z};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (!(bb[c] > 1e-6f)) return false;
    t[c] = fmaxf((ff[c] - ll[c]) / bb[c], 0.f);
    if (!isfinite(t[c])) return false;
  }
  const float lm = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
  if (clamp > 0.f && lm > clamp) {
    const float s = clamp / lm;
    t[0] *= s;
    t[1] *= s;
    t[2] *= s;
  }
  return true;
}

__global__ __launch_bounds__(256) void nrc_count_kernel(PathBuffers P, NrcPathBuffers N, float clamp,
                                                       uint32_t* __restrict__ cnt) {
  const uint32_t tp = blockIdx.x * blockDim.x + threadIdx.x;
  if (tp >= N.ntrain) return;
  const uint32_t nv = N.tn[tp];
  const float4 lf = P.L[N.train_base + tp];
  uint32_t c = 0;
  for (uint32_t v = 0; v < nv; v++) {
    float t[3];
    c += nrc_target(N, (size_t)tp * N.max_vertices + v, lf, clamp, t) ? 1u : 0u;
  }
  cnt[tp] = c;
}

// Exclusive prefix sum of cnt[0, n) in place (one block of 1024 threads,
// each over a contiguous chunk); the clamped total into N.counts[1].
__global__ __launch_bounds__(1024) void nrc_scan_kernel(uint32_t* __restrict__ cnt, uint32_t n, uint32_t cap,
                                                       NrcPathBuffers N, unsigned long long* __restrict__ stats) {
  __shared__ uint32_t part[1024];
  const uint32_t t = threadIdx.x;
  const uint32_t chunk = (n + 1023) / 1024;
  const uint32_t b = t * chunk, e = min(n, b + chunk);
  uint32_t s = 0;
  for (uint32_t i = b; i < e; i++) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - s;  // exclusive
  for (uint32_t i = b; i < e; i++) {
    const uint32_t c = cnt[i];
    cnt[i] = run;
    run += c;
  }
  if (t == 1023) {
    N.counts[1] = min(part[1023], cap);
    stats[2] += N.counts[0];           // cache queries of the pass
    stats[3] += min(part[1023], cap);  // training samples of the pass
  }
}

__global__ __launch_bounds__(256) void nrc_scatter_kernel(PathBuffers P, NrcPathBuffers N, float clamp,
                                                         const uint32_t* __restrict__ off, uint32_t cap,
                                                         NrcQuery* __restrict__ bq, float* __restrict__ bt) {
  const uint32_t tp = blockIdx.x * blockDim.x + threadIdx.x;
  if (tp >= N.ntrain) return;
  const uint32_t nv = N.tn[tp];
  const float4 lf = P.L[N.train_base + tp];
  uint32_t o = off[tp];
  for (uint32_t v = 0; v < nv && o < cap; v++) {
    const size_t r = (size_t)tp * N.max_vertices + v;
    float t[3];
    if (!nrc_target(N, r, lf, clamp, t)) continue;
    bq[o] = N.tq[r];
    bt[3 * (size_t)o + 0] = t[0];
    bt[3 * (size_t)o + 1] = t[1];
    bt[3 * (size_t)o + 2] = t[2];
    o++;
  }
}

__global__ void nrc_set_batch_kernel(const NrcPathBuffers N, uint32_t k, uint32_t batch, uint32_t* __restrict__ n_out) {
  const uint32_t total = N.counts[1];
  const uint64_t first = (uint64_t)k * batch;
  *n_out = total > first ? (uint32_t)min<uint64_t>(total - first, batch) : 0u;
}

}  // namespace

hipError_t pt_launch_nrc_samples(const PathBuffers& P, const NrcPathBuffers& N, float clamp, uint32_t* scratch,
                                 uint32_t cap, NrcQuery* bq, float* bt, hipStream_t st) {
  if (N.ntrain == 0) return hipSuccess;
  const uint32_t g = (N.ntrain + 255) / 256;
  hipLaunchKernelGGL(nrc_count_kernel, dim3(g), dim3(256), 0, st, P, N, clamp, scratch);
  hipLaunchKernelGGL(nrc_scan_kernel, dim3(1), dim3(1024), 0, st, scratch, N.ntrain, cap, N, P.stats);
  hipLaunchKernelGGL(nrc_scatter_kernel, dim3(g), dim3(256), 0, st, P, N, clamp, scratch, cap, bq, bt);
  return hipGetLastError();
}

hipError_t pt_launch_nrc_set_batch(const NrcPathBuffers& N, uint32_t k, uint32_t batch, uint32_t* n_out,
                                   hipStream_t st) {
  hipLaunchKernelGGL(nrc_set_batch_kernel, dim3(1), dim3(1), 0, st, N, k, batch, n_out);
  return hipGetLastError();
}

}  // namespace rrhip
