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
#include "pt_kernels.h"
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

__global__ __launch_bounds__(kBlock) void raygen_kernel(PtSceneView S, PathBuffers P, PtParams prm, uint32_t total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    P.counts[0] = total;
    P.counts[1] = 0;
    P.counts[2] = 0;
  }
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < total; slot += stride) {
    const uint32_t pixel = slot % prm.npix;
    const uint32_t s = slot / prm.npix;
    const uint32_t x = pixel % S.cam.width, y = pixel / S.cam.width;
    // sampler.rs:24-32 per-(pixel, sample) stream; pt/mod.rs:340-356
    Pcg32 rng = Pcg32::for_pixel_sample(pixel, prm.sample0 + s);
    V3 o, d;
    camera_ray(S.cam, x, y, rng, o, d);
    P.o[slot] = make_float4(o.x, o.y, o.z, 0.f);
    P.d[slot] = make_float4(d.x, d.y, d.z, 0.f);
    P.beta[slot] = make_float4(1.f, 1.f, 1.f, 0.f);
    P.L[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.prev[slot] = make_float4(o.x, o.y, o.z, 0.f);
    P.rng[slot] = store_rng(rng);
    P.flags[slot] = kFlagFromCamera;
    P.q0[slot] = slot;
  }
}

__global__ __launch_bounds__(kTraceBlock) void extend_kernel(PtSceneView S, PathBuffers P, const uint32_t* __restrict__ q) {
  __shared__ uint32_t s_stack[kTraceStack * kTraceBlock];
  uint32_t* stk = s_stack + threadIdx.x;
  const uint32_t n = P.counts[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t slot = q[i];
    const float4 o4 = P.o[slot], d4 = P.d[slot];
    Hit h;
    if (scene_closest<kTraceBlock>(S, v3(o4.x, o4.y, o4.z), v3(d4.x, d4.y, d4.z), stk, h)) {
      P.hit_n[slot] = make_float4(h.n.x, h.n.y, h.n.z, h.t);
      P.hit_m[slot] = h.material | (h.front ? 0x80000000u : 0u);
    } else {
      P.hit_n[slot] = make_float4(0.f, 0.f, 0.f, -1.f);
      P.hit_m[slot] = 0;
    }
  }
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
  V3 c;  // contribution before visibility (f cos L weight / pdf)
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
// visibility test, which becomes the returned shadow ray.
__device__ __forceinline__ bool sample_light(const PtSceneView& S, const PtLight& L, V3 p, V3 n,
                                             const PtMaterial& m, const Frame& fr, V3 wo_l, float eta,
                                             float pick, Pcg32& rng, ShadowOut& out) {
  V3 f;
  float bsdf_pdf;
  const V3 rad = v3(L.radiance);
  const V3 org = p + n * kRayOffset;
  switch (L.kind) {
    case kLightPoint: {
      const V3 to = v3(L.p) - p;
      const float dist2 = fmaxf(dot(to, to), 1e-12f);
      const V3 wi = normalize(to);
      eval_world(m, fr, wo_l, eta, wi, f, bsdf_pdf);
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
      eval_world(m, fr, wo_l, eta, wi, f, bsdf_pdf);
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
      V3 sp;
      if (L.kind == kLightRect) {
        sp = v3(L.p) + v3(L.e1) * u + v3(L.e2) * v;
      } else {
        const float r = sqrtf(u);
        float sn, cs;
        sincosf(2.f * kPi * v, &sn, &cs);
        sp = v3(L.p) + v3(L.e1) * (r * cs) + v3(L.e2) * (r * sn);
      }
      const V3 to = sp - p;
      const float dist2 = fmaxf(dot(to, to), 1e-12f);
      const V3 wi = normalize(to);
      const float cos_light = fabsf(dot(wi, v3(L.n)));  // double-sided emitter
      if (cos_light < 1e-9f || L.area <= 0.f) return false;
      eval_world(m, fr, wo_l, eta, wi, f, bsdf_pdf);
      if (maxc(f) <= 0.f) return false;
      const float pdf_sa = pick * dist2 / (cos_light * L.area);
      const float w = power_heuristic(pdf_sa, bsdf_pdf);
      out.o = org;
      out.d = wi;
      out.tmax = len(to) - 1e-3f;
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
      eval_world(m, fr, wo_l, eta, wi, f, bsdf_pdf);
      if (maxc(f) <= 0.f) return false;
      const float dist = sqrtf(dist2) * cos_t - sqrtf(fmaxf(r * r - dist2 * sin_t * sin_t, 0.f));
      const float pdf_sa = pick / fmaxf(2.f * kPi * (1.f - cos_max), 1e-12f);
      const float w = power_heuristic(pdf_sa, bsdf_pdf);
      out.o = org;
      out.d = wi;
      out.tmax = dist - 1e-3f;
      out.c = f * rad * (w / pdf_sa);
      return true;
    }
    case kLightDome: {
      const float u = rnd(rng), v = rnd(rng);
      const float z = 1.f - 2.f * u;
      const float rr = sqrtf(fmaxf(1.f - z * z, 0.f));
      float sn, cs;
      sincosf(2.f * kPi * v, &sn, &cs);
      const V3 wi = v3(rr * cs, rr * sn, z);
      const float pdf_sa = kInv4Pi * pick;
      eval_world(m, fr, wo_l, eta, wi, f, bsdf_pdf);
      if (maxc(f) <= 0.f) return false;
      const float w = power_heuristic(pdf_sa, bsdf_pdf);
      out.o = org;
      out.d = wi;
      out.tmax = 1e8f;
      out.c = f * rad * (w / pdf_sa);
      return true;
    }
    default:
      return false;
  }
}

__global__ __launch_bounds__(kBlock) void shade_kernel(PtSceneView S, PathBuffers P, PtParams prm,
                                                      const uint32_t* __restrict__ q, uint32_t* __restrict__ qnext) {
  const uint32_t n = P.counts[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  // Block-uniform loop bound so every lane reaches the wave-wide appends.
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const bool act = i < n;
    bool cont = false, shadow = false;
    ShadowOut so;
    uint32_t slot = 0;
    if (act) {
      slot = q[i];
      const float4 o4 = P.o[slot], d4 = P.d[slot], b4 = P.beta[slot], l4 = P.L[slot], pv4 = P.prev[slot];
      const float4 hn = P.hit_n[slot];
      uint32_t flags = P.flags[slot];
      Pcg32 rng = load_rng(P.rng[slot]);
      V3 beta = v3(b4.x, b4.y, b4.z), Lr = v3(l4.x, l4.y, l4.z);
      const V3 ro = v3(o4.x, o4.y, o4.z), rd = v3(d4.x, d4.y, d4.z);
      const float prev_pdf = o4.w;
      const V3 prev_origin = v3(pv4.x, pv4.y, pv4.z);
      const bool from_camera = (flags & kFlagFromCamera) != 0;
      uint32_t depth = flags & 0xFFu;
      if (hn.w < 0.f) {
        // Miss: dome (with MIS) or background (pt/mod.rs:643-659).
        if (S.dome >= 0) {
          float w = 1.f;
          if (!from_camera) w = power_heuristic(prev_pdf, kInv4Pi * ls_pmf(S, prev_origin, (uint32_t)S.dome));
          Lr = Lr + beta * v3(S.lights[S.dome].radiance) * w;
        } else {
          Lr = Lr + beta * v3(S.background);
        }
      } else {
        const uint32_t hm = P.hit_m[slot];
        const PtMaterial m = S.materials[hm & 0x7FFFFFFFu];
        const bool entering = (hm >> 31) != 0;
        const V3 p = ro + rd * hn.w;
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
          Lr = Lr + beta * em * w;
        }
        const Frame fr = make_frame(nrm);
        const V3 wo_l = to_local(fr, wo);
        // NEE through the light BVH (pt/mod.rs:924-956)
        if (S.num_lights > 0 && m.has_lobes) {
          uint32_t li;
          float pick;
          if (ls_sample(S, p, rnd(rng), li, pick)) {
            if (sample_light(S, S.lights[li], p, nrm, m, fr, wo_l, eta, pick, rng, so)) {
              so.c = beta * so.c;
              shadow = maxc(so.c) > 0.f;
            }
          }
        }
        // BSDF continuation (pt/mod.rs:958-987)
        BsdfSample bs;
        if (bsdf_sample(m, wo_l, eta, rng, bs)) {
          const V3 wi = to_world(fr, bs.wi);
          beta = beta * bs.f * (fabsf(bs.wi.z) / bs.pdf);
          const V3 org = p + nrm * (bs.transmitted ? -kRayOffset : kRayOffset);
          const V3 nd = normalize(wi);
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
          P.d[slot] = make_float4(nd.x, nd.y, nd.z, 0.f);
          P.prev[slot] = make_float4(p.x, p.y, p.z, 0.f);
          flags = depth;  // from_camera cleared
        }
      }
      P.beta[slot] = make_float4(beta.x, beta.y, beta.z, 0.f);
      P.L[slot] = make_float4(Lr.x, Lr.y, Lr.z, 0.f);
      P.rng[slot] = store_rng(rng);
      P.flags[slot] = flags;
    }
    const uint32_t sidx = wave_append(&P.counts[2], act && shadow);
    if (act && shadow) {
      P.sq[sidx] = slot;
      P.so[sidx] = make_float4(so.o.x, so.o.y, so.o.z, so.tmax);
      P.sd[sidx] = make_float4(so.d.x, so.d.y, so.d.z, 0.f);
      P.sc[sidx] = make_float4(so.c.x, so.c.y, so.c.z, 0.f);
    }
    const uint32_t nidx = wave_append(&P.counts[1], act && cont);
    if (act && cont) qnext[nidx] = slot;
  }
}

__global__ __launch_bounds__(kTraceBlock) void shadow_kernel(PtSceneView S, PathBuffers P) {
  __shared__ uint32_t s_stack[kTraceStack * kTraceBlock];
  uint32_t* stk = s_stack + threadIdx.x;
  const uint32_t n = P.counts[2];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float4 o = P.so[i], d = P.sd[i];
    if (scene_occluded<kTraceBlock>(S, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), o.w, stk)) continue;
    const uint32_t slot = P.sq[i];
    const float4 c = P.sc[i];
    float4 l = P.L[slot];
    l.x += c.x;
    l.y += c.y;
    l.z += c.z;
    P.L[slot] = l;
  }
}

__global__ void rotate_kernel(PathBuffers P) {
  P.stats[0] += P.counts[0];
  P.stats[1] += P.counts[2];
  P.counts[0] = P.counts[1];
  P.counts[1] = 0;
  P.counts[2] = 0;
}

__global__ __launch_bounds__(kBlock) void resolve_kernel(PathBuffers P, PtParams prm, double* __restrict__ accum) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < prm.npix; px += stride) {
    double s[3] = {0, 0, 0}, q[3] = {0, 0, 0};
    for (uint32_t k = 0; k < prm.spp_pass; k++) {
      const float4 l4 = P.L[(size_t)k * prm.npix + px];
      double l[3] = {l4.x, l4.y, l4.z};
      // pt/mod.rs:88-95 firefly clamp on the sample's luminance
      const double lm = 0.2126 * l[0] + 0.7152 * l[1] + 0.0722 * l[2];
      if (prm.firefly_clamp > 0.f && lm > (double)prm.firefly_clamp) {
        const double sc = (double)prm.firefly_clamp / lm;
        l[0] *= sc;
        l[1] *= sc;
        l[2] *= sc;
      }
      for (int c = 0; c < 3; c++) {
        s[c] += l[c];
        q[c] += l[c] * l[c];
      }
    }
    double* a = accum + 8 * (size_t)px;
    for (int c = 0; c < 3; c++) {
      a[c] += s[c];
      a[3 + c] += q[c];
    }
    a[6] += (double)prm.spp_pass;
  }
}

uint32_t grid_for(uint64_t n, int block = kBlock) {
  const uint64_t b = (n + block - 1) / block;
  return (uint32_t)(b < 1 ? 1 : (b > kMaxGrid ? kMaxGrid : b));
}

}  // namespace

hipError_t pt_launch_raygen(const PtSceneView& S, const PathBuffers& P, const PtParams& prm, uint32_t total,
                            hipStream_t st) {
  hipLaunchKernelGGL(raygen_kernel, dim3(grid_for(total)), dim3(kBlock), 0, st, S, P, prm, total);
  return hipGetLastError();
}

hipError_t pt_launch_bounce(const PtSceneView& S, const PathBuffers& P, const PtParams& prm, const uint32_t* q,
                            uint32_t* qnext, uint32_t capacity, hipStream_t st) {
  const uint32_t g = grid_for(capacity);
  const uint32_t gt = grid_for(capacity, kTraceBlock);
  hipLaunchKernelGGL(extend_kernel, dim3(gt), dim3(kTraceBlock), 0, st, S, P, q);
  hipLaunchKernelGGL(shade_kernel, dim3(g), dim3(kBlock), 0, st, S, P, prm, q, qnext);
  hipLaunchKernelGGL(shadow_kernel, dim3(gt), dim3(kTraceBlock), 0, st, S, P);
  hipLaunchKernelGGL(rotate_kernel, dim3(1), dim3(1), 0, st, P);
  return hipGetLastError();
}

hipError_t pt_launch_resolve(const PathBuffers& P, const PtParams& prm, double* accum, hipStream_t st) {
  hipLaunchKernelGGL(resolve_kernel, dim3(grid_for(prm.npix)), dim3(kBlock), 0, st, P, prm, accum);
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

}  // namespace rrhip
