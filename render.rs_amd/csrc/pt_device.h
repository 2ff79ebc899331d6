// Device functions of the wavefront path integrator (f32 arithmetic; the
// light-BVH branch probabilities in f64).
//
// Each function follows the render.rs CPU routine it cites so that, given
// the same PCG32 draws, the GPU makes the same decisions up to f32
// rounding; draws are consumed in the CPU path's order
// (src/raytracer/pt/mod.rs:449-990).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pt_types.h"
#include "trace_math.h"
#include "sampling.h"

namespace rrhip {
namespace pt {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInv4Pi = 0.0795774715459476678f;
constexpr float kObjEps = 1e-6f;      // math/mod.rs:9 (object hits need t > EPSILON)

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3(const float* a) { return V3{a[0], a[1], a[2]}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
// vector.rs:31-38: divide only when the length exceeds EPSILON.
__device__ __forceinline__ V3 normalize(V3 a) {
  const float l = len(a);
  return l > 1e-6f ? a / l : a;
}
__device__ __forceinline__ float maxc(V3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
__device__ __forceinline__ float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// Uniform in [0, 1) with 24 bits.
__device__ __forceinline__ float rnd(Pcg32& r) { return r.next_f32(); }

// ---- error-bounded ray origins -------------------------------------------------
// The method of Pharr, Jakob and Humphreys, "Physically Based Rendering"
// (3rd ed., 2016), sec. 3.9 "Managing Rounding Error": a hit point
// computed in f32 carries a componentwise bound on its distance from the
// exact surface; a ray leaving the surface starts outside that box, pushed
// along the geometric normal by dot(|n|, bound) and then one ulp further
// away from the surface in every component the push moved. render.rs moves
// origins a fixed 1e-4 instead (pt/mod.rs:30), which falls below one ulp
// once coordinates exceed ~1e3.

// gamma(n) = n u / (1 - n u), u = 2^-24: the relative error bound of n
// chained f32 operations.
__device__ __forceinline__ constexpr float err_gamma(int n) {
  return (float)n * 5.9604645e-8f / (1.f - (float)n * 5.9604645e-8f);
}
__device__ __forceinline__ V3 vabs(V3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }

// The next representable float above / below x (x finite).
__device__ __forceinline__ float float_up(float x) {
  if (x == 0.f) return 1.4e-45f;
  const int b = __float_as_int(x);
  return __int_as_float(x > 0.f ? b + 1 : b - 1);
}
__device__ __forceinline__ float float_down(float x) {
  if (x == 0.f) return -1.4e-45f;
  const int b = __float_as_int(x);
  return __int_as_float(x > 0.f ? b - 1 : b + 1);
}

// Origin of a ray leaving hit point p on the side of offset vector `off`
// (the bound's push along the geometric normal, or its negation).
__device__ __forceinline__ V3 offset_origin(V3 p, V3 off) {
  V3 o = p + off;
  o.x = off.x > 0.f ? float_up(o.x) : (off.x < 0.f ? float_down(o.x) : o.x);
  o.y = off.y > 0.f ? float_up(o.y) : (off.y < 0.f ? float_down(o.y) : o.y);
  o.z = off.z > 0.f ? float_up(o.z) : (off.z < 0.f ? float_down(o.z) : o.z);
  return o;
}

// Push of a hit point with error bound `err` along unit geometric normal
// ng, oriented to the side direction d came from.
__device__ __forceinline__ V3 bound_push(V3 ng, V3 err, V3 d) {
  const float k = dot(vabs(ng), err);
  return ng * (dot(ng, d) < 0.f ? k : -k);
}

// bxdf.rs:16-39
struct Frame {
  V3 t, b, n;
};
__device__ __forceinline__ Frame make_frame(V3 n) {
  const V3 t0 = fabsf(n.x) > 0.9f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
  Frame f;
  f.n = n;
  f.t = normalize(cross(n, t0));
  f.b = cross(n, f.t);
  return f;
}
__device__ __forceinline__ V3 to_local(const Frame& f, V3 w) { return {dot(w, f.t), dot(w, f.b), dot(w, f.n)}; }
__device__ __forceinline__ V3 to_world(const Frame& f, V3 w) { return f.t * w.x + f.b * w.y + f.n * w.z; }

// ---- BSDF (bxdf.rs) ----------------------------------------------------------

__device__ __forceinline__ float pow5(float x) {
  const float x2 = x * x;
  return x2 * x2 * x;
}

// bxdf.rs:63-73
__device__ __forceinline__ float fresnel_dielectric(float cos_i, float eta) {
  cos_i = clampf(cos_i, 0.f, 1.f);
  const float sin2_t = (1.f - cos_i * cos_i) / (eta * eta);
  if (sin2_t >= 1.f) return 1.f;
  const float cos_t = sqrtf(1.f - sin2_t);
  const float r_par = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  const float r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  return 0.5f * (r_par * r_par + r_perp * r_perp);
}

// bxdf.rs:77-101. The denominator cos²(a² - 1) + 1 is evaluated as
// sin² + cos² a² with sin² = (1 - cos)(1 + cos) (at least 0 for a
// normalised h that rounds past 1): in f32, a² - 1 loses a² to rounding
// for small roughness and becomes exactly -1 below roughness ~0.013 (a² is
// 6e-10 at render.rs's floor of 0.005), which made D infinite for h on the
// normal.
__device__ __forceinline__ float ggx_d(V3 h, float alpha) {
  if (h.z <= 0.f) return 0.f;
  const float a2 = alpha * alpha;
  const float d = fmaf(h.z * h.z, a2, fmaxf((1.f - h.z) * (1.f + h.z), 0.f));
  return a2 / (kPi * d * d);
}
__device__ __forceinline__ float ggx_lambda(V3 w, float alpha) {
  const float cos2 = w.z * w.z;
  if (cos2 <= 0.f) return 0.f;
  const float tan2 = fmaxf(1.f - cos2, 0.f) / cos2;
  return (sqrtf(1.f + alpha * alpha * tan2) - 1.f) * 0.5f;
}
__device__ __forceinline__ float ggx_g1(V3 w, float alpha) { return 1.f / (1.f + ggx_lambda(w, alpha)); }
__device__ __forceinline__ float ggx_g2(V3 wo, V3 wi, float alpha) {
  return 1.f / (1.f + ggx_lambda(wo, alpha) + ggx_lambda(wi, alpha));
}

// bxdf.rs:104-121 (Heitz 2018 VNDF)
__device__ __forceinline__ V3 ggx_sample_vndf(V3 wo, float alpha, float u1, float u2) {
  const V3 v = normalize(v3(alpha * wo.x, alpha * wo.y, wo.z));
  const float lensq = v.x * v.x + v.y * v.y;
  const V3 t1 = lensq > 1e-12f ? v3(-v.y, v.x, 0.f) / sqrtf(lensq) : v3(1.f, 0.f, 0.f);
  const V3 t2 = cross(v, t1);
  const float r = sqrtf(u1);
  float sp, cp;
  sincosf(2.f * kPi * u2, &sp, &cp);
  const float p1 = r * cp;
  float p2 = r * sp;
  const float s = 0.5f * (1.f + v.z);
  p2 = (1.f - s) * sqrtf(fmaxf(1.f - p1 * p1, 0.f)) + s * p2;
  const V3 nh = t1 * p1 + t2 * p2 + v * sqrtf(fmaxf(1.f - p1 * p1 - p2 * p2, 0.f));
  return normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(nh.z, 1e-9f)));
}

// bxdf.rs:124-130
__device__ __forceinline__ float ggx_pdf_h(V3 wo, V3 h, float alpha) {
  const float odh = dot(wo, h);
  if (odh <= 0.f || wo.z <= 0.f) return 0.f;
  return ggx_g1(wo, alpha) * odh * ggx_d(h, alpha) / wo.z;
}

// bxdf.rs:173-179 (per-channel Schlick)
__device__ __forceinline__ V3 eval_spec(V3 f0, V3 f90, float alpha, V3 wo, V3 wi) {
  const V3 h = normalize(wo + wi);
  const float d = ggx_d(h, alpha);
  const float g = ggx_g2(wo, wi, alpha);
  const float m = pow5(clampf(1.f - fmaxf(dot(wo, h), 0.f), 0.f, 1.f));
  const V3 f = f0 + (f90 - f0) * m;
  return f * (d * g / fmaxf(4.f * wo.z * wi.z, 1e-9f));
}

// bxdf.rs:181-185
__device__ __forceinline__ float spec_pdf(V3 wo, V3 wi, float alpha) {
  const V3 h = normalize(wo + wi);
  const float odh = fmaxf(dot(wo, h), 1e-9f);
  return ggx_pdf_h(wo, h, alpha) / (4.f * odh);
}

// bxdf.rs:145-165 (Oren-Nayar; scale folds diffuse_gain * under_layer / pi)
__device__ __forceinline__ V3 eval_diffuse(const PtMaterial& m, V3 wo, V3 wi) {
  const float sin_o = sqrtf(fmaxf(1.f - wo.z * wo.z, 0.f));
  const float sin_i = sqrtf(fmaxf(1.f - wi.z * wi.z, 0.f));
  const float cos_dphi =
      (sin_o > 1e-6f && sin_i > 1e-6f) ? clampf((wo.x * wi.x + wo.y * wi.y) / (sin_o * sin_i), -1.f, 1.f) : 0.f;
  float sin_a, tan_b;
  if (wo.z < wi.z) {
    sin_a = sin_o;
    tan_b = sin_i / fmaxf(wi.z, 1e-6f);
  } else {
    sin_a = sin_i;
    tan_b = sin_o / fmaxf(wo.z, 1e-6f);
  }
  return v3(m.diffuse) * (m.on_a + m.on_b * fmaxf(cos_dphi, 0.f) * sin_a * tan_b);
}

// bxdf.rs:167-171
__device__ __forceinline__ V3 eval_fuzz(const PtMaterial& m, V3 wo, V3 wi) {
  const float a = 1.f - wi.z, b = 1.f - wo.z;
  const float rim = 0.5f * (a * a * a * a + b * b * b * b);
  return v3(m.fuzz) * rim;
}

// bxdf.rs:241-299. Returns f (no cosine) and the composite pdf.
__device__ __forceinline__ void bsdf_eval(const PtMaterial& m, V3 wo, V3 wi, float eta, V3& f, float& pdf) {
  f = v3(0.f, 0.f, 0.f);
  pdf = 0.f;
  if (!m.has_lobes) return;
  const float wd = m.w[0], ws = m.w[1], wc = m.w[2], wf = m.w[3], wg = m.w[4];
  if (wi.z > 0.f) {
    const float cos_pdf = wi.z / kPi;
    if (wd > 0.f) {
      f = f + eval_diffuse(m, wo, wi);
      pdf += wd * cos_pdf;
    }
    if (wf > 0.f) {
      f = f + eval_fuzz(m, wo, wi);
      pdf += wf * cos_pdf;
    }
    if (ws > 0.f) {
      f = f + eval_spec(v3(m.spec_f0), v3(m.spec_f90), m.spec_alpha, wo, wi);
      pdf += ws * spec_pdf(wo, wi, m.spec_alpha);
    }
    if (wc > 0.f) {
      f = f + eval_spec(v3(m.coat_f0, m.coat_f0, m.coat_f0), v3(m.coat_f90, m.coat_f90, m.coat_f90),
                        m.coat_alpha, wo, wi);
      pdf += wc * spec_pdf(wo, wi, m.coat_alpha);
    }
    if (wg > 0.f) {
      const V3 h = normalize(wo + wi);
      const float fres = fresnel_dielectric(fmaxf(dot(wo, h), 0.f), eta);
      const float d = ggx_d(h, m.glass_alpha);
      const float g = ggx_g2(wo, wi, m.glass_alpha);
      const float v = m.glass_scale * fres * d * g / fmaxf(4.f * wo.z * wi.z, 1e-9f);
      f = f + v3(v, v, v);
      pdf += wg * fres * spec_pdf(wo, wi, m.glass_alpha);
    }
    return;
  }
  if (wg <= 0.f) return;
  // bxdf.rs:193-232 (PBRT-3 microfacet transmission)
  V3 h = normalize(wo + wi * eta);
  if (h.z < 0.f) h = -h;
  const float odh = dot(wo, h), idh = dot(wi, h);
  const float fres = fresnel_dielectric(fmaxf(odh, 0.f), eta);
  if (odh <= 0.f || idh >= 0.f) return;
  const float d = ggx_d(h, m.glass_alpha);
  const float g = ggx_g2(wo, v3(wi.x, wi.y, -wi.z), m.glass_alpha);
  const float sqrt_denom = odh + eta * idh;
  if (fabsf(sqrt_denom) < 1e-9f) return;
  const float factor = 1.f / (eta * eta);
  const float v = m.glass_scale * (1.f - fres) * d * g * factor * fabsf(idh * odh / (wi.z * wo.z)) * (eta * eta) /
                  (sqrt_denom * sqrt_denom);
  f = v3(m.refraction) * v;
  const float dwh_dwi = fabsf(eta * eta * idh) / (sqrt_denom * sqrt_denom);
  pdf = wg * (1.f - fres) * ggx_pdf_h(wo, h, m.glass_alpha) * dwh_dwi;
}

struct BsdfSample {
  V3 wi, f;
  float pdf;
  bool transmitted;
  bool specular;  // a lobe other than diffuse / fuzz was picked (bxdf.rs:315)
};

// bxdf.rs:309-374 (same draw order)
__device__ __forceinline__ bool bsdf_sample(const PtMaterial& m, V3 wo, float eta, Pcg32& rng, BsdfSample& s) {
  if (!m.has_lobes) return false;
  const float wd = m.w[0], ws = m.w[1], wc = m.w[2], wf = m.w[3], wg = m.w[4];
  const float pick = rnd(rng);
  bool transmitted = false;
  V3 wi;
  if (pick < wd + wf) {
    const float u = rnd(rng), v = rnd(rng);
    const float r = sqrtf(u);
    float sp, cp;
    sincosf(2.f * kPi * v, &sp, &cp);
    wi = v3(r * cp, r * sp, sqrtf(fmaxf(1.f - u, 0.f)));
  } else if (pick < wd + wf + ws || pick < wd + wf + ws + wc) {
    const float alpha = pick < wd + wf + ws ? m.spec_alpha : m.coat_alpha;
    const float u1 = rnd(rng), u2 = rnd(rng);
    const V3 h = ggx_sample_vndf(wo, alpha, u1, u2);
    wi = h * (2.f * dot(wo, h)) - wo;  // (-wo).reflect(h)
    if (wi.z <= 0.f) return false;
  } else if (wg > 0.f) {
    const float u1 = rnd(rng), u2 = rnd(rng);
    const V3 h = ggx_sample_vndf(wo, m.glass_alpha, u1, u2);
    const float fres = fresnel_dielectric(fmaxf(dot(wo, h), 0.f), eta);
    if (rnd(rng) < fres) {
      wi = h * (2.f * dot(wo, h)) - wo;
      if (wi.z <= 0.f) return false;
    } else {
      const float cos_i = dot(wo, h);
      const float sin2_t = (1.f - cos_i * cos_i) / (eta * eta);
      if (sin2_t >= 1.f) return false;
      const float cos_t = sqrtf(1.f - sin2_t);
      wi = normalize((-wo) / eta + h * (cos_i / eta - cos_t));
      if (wi.z >= 0.f) return false;
      transmitted = true;
    }
  } else {
    return false;
  }
  V3 f;
  float pdf;
  bsdf_eval(m, wo, wi, eta, f, pdf);
  if (!(pdf > 0.f)) return false;
  s.wi = wi;
  s.f = f;
  s.pdf = pdf;
  s.transmitted = transmitted;
  s.specular = pick >= wd + wf;
  return true;
}

__device__ __forceinline__ float power_heuristic(float a, float b) {  // pt/mod.rs:992-996
  const float a2 = a * a, b2 = b * b;
  return (a2 + b2 <= 0.f) ? 0.f : a2 / (a2 + b2);
}

// ---- dome environment map (envmap.rs), nearest texel, CDFs in f64 ---------------

// Texel of direction d (unit): u = atan2(x, -z) / 2pi wrapped, v = acos(y) / pi
// (envmap.rs:84-91).
__device__ __forceinline__ void env_texel(const PtSceneView& S, V3 d, uint32_t& x, uint32_t& y) {
  float u = atan2f(d.x, -d.z) * (0.5f / kPi);
  u -= floorf(u);
  const float v = fminf(fmaxf(acosf(fminf(fmaxf(d.y, -1.f), 1.f)) / kPi, 0.f), 1.f);
  x = min((uint32_t)(u * (float)S.env_w), S.env_w - 1);
  y = min((uint32_t)(v * (float)S.env_h), S.env_h - 1);
}

__device__ __forceinline__ V3 env_rgb(const PtSceneView& S, uint32_t x, uint32_t y) {
  const float* c = S.env_rgb + 3 * ((size_t)y * S.env_w + x);
  return v3(c[0], c[1], c[2]);
}

// Solid-angle pdf of sampling d (envmap.rs:134-148): the texel's share of
// the sin-weighted luminance over the 2 pi^2 sin(theta) Jacobian. The row's
// sin(theta) cancels (it is >= sin(pi / 2^29) > 1e-9 for every row the
// scene builder accepts, so the reference's zero test never fires), which
// leaves lum * W * H / (2 pi^2 total) with the factor precomputed in f64.
__device__ __forceinline__ float env_pdf(const PtSceneView& S, V3 d) {
  uint32_t x, y;
  env_texel(S, d, x, y);
  return (float)((double)lum(env_rgb(S, x, y)) * S.env_pdf_k);
}

// Largest i in [0, n) with c[i] <= t (the row / column the reference's
// binary search lands on), at most n - 1.
__device__ __forceinline__ uint32_t cdf_find(const double* c, uint32_t n, double t) {
  uint32_t lo = 0, hi = n;  // c[lo] <= t holds for lo = 0 (c[0] = 0 <= t)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (c[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Importance-sampled direction, its texel radiance and pdf (envmap.rs:
// 99-131): the row from the marginal CDF, the column from the row's
// conditional CDF, the texel centre's direction.
__device__ __forceinline__ void env_sample(const PtSceneView& S, float u1, float u2, V3& dir, V3& rad, float& pdf) {
  const uint32_t y = min(cdf_find(S.env_marg, S.env_h + 1, (double)u1 * S.env_total), S.env_h - 1);
  const double row_total = fmax(S.env_marg[y + 1] - S.env_marg[y], 1e-12);
  const double* cond = S.env_cond + (size_t)y * (S.env_w + 1);
  const uint32_t x = min(cdf_find(cond, S.env_w + 1, (double)u2 * row_total), S.env_w - 1);
  const float phi = ((float)x + 0.5f) / (float)S.env_w * 2.f * kPi;
  const float theta = ((float)y + 0.5f) / (float)S.env_h * kPi;
  float sp, cp, st, ct;
  sincosf(phi, &sp, &cp);
  sincosf(theta, &st, &ct);
  dir = v3(st * sp, ct, -st * cp);
  rad = env_rgb(S, x, y);
  pdf = env_pdf(S, dir);
}

// Dome radiance along d and the pdf the dome's light sampling has for it
// (pt/mod.rs:375-387).
__device__ __forceinline__ V3 dome_radiance(const PtSceneView& S, const PtLight& L, V3 d) {
  if (!S.env_rgb) return v3(L.radiance);
  uint32_t x, y;
  env_texel(S, d, x, y);
  return env_rgb(S, x, y) * v3(L.radiance);
}
__device__ __forceinline__ float dome_pdf(const PtSceneView& S, V3 d) {
  return S.env_rgb ? env_pdf(S, d) : kInv4Pi;
}

// ---- light BVH (light_sampler.rs), branch probabilities in f64 ---------------

__device__ __forceinline__ double node_importance(const LightNode& n, double px, double py, double pz) {
  const double cx = ((double)n.min[0] + (double)n.max[0]) * 0.5;
  const double cy = ((double)n.min[1] + (double)n.max[1]) * 0.5;
  const double cz = ((double)n.min[2] + (double)n.max[2]) * 0.5;
  const double dx = px - cx, dy = py - cy, dz = pz - cz;
  const double d2 = dx * dx + dy * dy + dz * dz;
  const double rx = (double)(n.max[0] - n.min[0]) * 0.5;
  const double ry = (double)(n.max[1] - n.min[1]) * 0.5;
  const double rz = (double)(n.max[2] - n.min[2]) * 0.5;
  const double r2 = rx * rx + ry * ry + rz * rz;
  return (double)n.power / fmax(fmax(d2, r2), 1e-6);
}

// light_sampler.rs:150-208
__device__ __forceinline__ bool ls_sample(const PtSceneView& S, V3 p, float uf, uint32_t& light, float& pmf_out) {
  const bool has_f = S.num_lnodes > 0, has_i = S.num_inf > 0;
  if (!has_f && !has_i) return false;
  double u = uf, pmf = 1.0;
  bool inf;
  if (has_f && has_i) {
    if (u < S.p_infinite) {
      u /= S.p_infinite;
      pmf *= S.p_infinite;
      inf = true;
    } else {
      u = (u - S.p_infinite) / (1.0 - S.p_infinite);
      pmf *= 1.0 - S.p_infinite;
      inf = false;
    }
  } else {
    inf = has_i;
  }
  if (inf) {
    const double target = u * S.inf_total;
    double prev = 0.0;
    for (uint32_t k = 0; k < S.num_inf; k++) {
      const double acc = S.inf_cdf[k];
      if (target <= acc || k == S.num_inf - 1) {
        light = S.inf_lights[k];
        pmf_out = (float)(pmf * (acc - prev) / fmax(S.inf_total, 1e-12));
        return true;
      }
      prev = acc;
    }
    return false;
  }
  const double px = p.x, py = p.y, pz = p.z;
  uint32_t node = 0;
  for (;;) {
    const LightNode nd = S.lnodes[node];
    if (nd.b == 0xFFFFFFFFu) {
      light = nd.a;
      pmf_out = (float)pmf;
      return true;
    }
    const double ia = node_importance(S.lnodes[nd.a], px, py, pz);
    const double ib = node_importance(S.lnodes[nd.b], px, py, pz);
    const double total = ia + ib;
    const double pa = total > 1e-30 ? ia / total : 0.5;
    if (u < pa) {
      u = fmin(u / pa, 1.0 - 1e-12);
      pmf *= pa;
      node = nd.a;
    } else {
      u = fmin((u - pa) / (1.0 - pa), 1.0 - 1e-12);
      pmf *= 1.0 - pa;
      node = nd.b;
    }
  }
}

// light_sampler.rs:212-238
__device__ __forceinline__ float ls_pmf(const PtSceneView& S, V3 p, uint32_t light) {
  const bool has_f = S.num_lnodes > 0, has_i = S.num_inf > 0;
  for (uint32_t k = 0; k < S.num_inf; k++) {
    if (S.inf_lights[k] == light) {
      const double prev = k ? S.inf_cdf[k - 1] : 0.0;
      const double group = has_f ? S.p_infinite : 1.0;
      return (float)(group * (S.inf_cdf[k] - prev) / fmax(S.inf_total, 1e-12));
    }
  }
  const uint32_t leaf = S.light_leaf[light];
  if (leaf == 0xFFFFFFFFu) return 0.f;
  double pmf = has_i ? 1.0 - S.p_infinite : 1.0;
  const double px = p.x, py = p.y, pz = p.z;
  uint32_t node = leaf;
  while (S.lnodes[node].parent != 0xFFFFFFFFu) {
    const uint32_t parent = S.lnodes[node].parent;
    const LightNode pn = S.lnodes[parent];
    const double ia = node_importance(S.lnodes[pn.a], px, py, pz);
    const double ib = node_importance(S.lnodes[pn.b], px, py, pz);
    const double total = ia + ib;
    const double mine = node == pn.a ? ia : ib;
    pmf *= total > 1e-30 ? mine / total : 0.5;
    node = parent;
  }
  return (float)pmf;
}

// ---- geometry -------------------------------------------------------------------

struct Hit {
  float t;
  V3 n;  // triangles: flipped to face the ray; spheres: outward
  uint32_t material;
  bool front;
  uint32_t obj;  // loose object: kind (kObjTri / kObjSphere / kObjQuadric) << 30 | index
  // Filled by scene_closest for the winning hit: the hit point rebuilt from
  // the surface (barycentrics, sphere reprojection) and the push that puts
  // an origin outside its rounding-error box, towards the ray's side.
  V3 p, e;
  V3 tan;  // curve hits: the world tangent (zero otherwise)
  // Instance hits: the instance (~0u on loose objects), BLAS record and
  // barycentrics (texture coordinates are rebuilt from them).
  uint32_t inst, slot;
  float bu, bv;
};
constexpr uint32_t kObjTri = 0, kObjSphere = 1, kObjQuadric = 2, kObjIndex = 0x3FFFFFFFu;

// extra.rs:673-720 (world-space, unit direction): t > EPSILON
__device__ __forceinline__ bool hit_triangle(const PtTriangle& tr, V3 o, V3 d, float tmax, Hit& h) {
  const V3 e1 = v3(tr.e1), e2 = v3(tr.e2);
  const V3 pv = cross(d, e2);
  const float det = dot(e1, pv);
  if (fabsf(det) < 1e-12f) return false;
  const float inv = 1.f / det;
  const V3 tv = o - v3(tr.v0);
  const float u = dot(tv, pv) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  const V3 qv = cross(tv, e1);
  const float v = dot(d, qv) * inv;
  if (v < 0.f || u + v > 1.f) return false;
  const float t = dot(e2, qv) * inv;
  if (!(t > kObjEps && t < tmax)) return false;
  V3 n = v3(tr.n);
  const bool front = dot(n, d) < 0.f;
  h.t = t;
  h.n = front ? n : -n;
  h.material = tr.material;
  h.front = front;
  return true;
}

// sphere.rs:38-109 (full sphere): near root, else far root, > EPSILON.
// Roots from the geometric form of Haines et al., "Precision Improvements
// for Ray/Sphere Intersection" (Ray Tracing Gems, 2019).
__device__ __forceinline__ bool sphere_roots(const PtSphere& s, V3 o, V3 d, float& t0, float& t1) {
  const V3 f = o - v3(s.c[0], s.c[1], s.c[2]);
  const float b = -dot(f, d);
  const V3 l = f + d * b;
  const float disc = s.r * s.r - dot(l, l);
  if (disc < 0.f) return false;
  const float c = dot(f, f) - s.r * s.r;
  const float q = b + copysignf(sqrtf(disc), b);
  t0 = c / q;
  t1 = q;
  if (t0 > t1) {
    const float tmp = t0;
    t0 = t1;
    t1 = tmp;
  }
  return true;
}

__device__ __forceinline__ bool hit_sphere(const PtSphere& s, V3 o, V3 d, float tmax, Hit& h) {
  float t0, t1;
  if (!sphere_roots(s, o, d, t0, t1)) return false;
  float t;
  if (t0 > kObjEps)
    t = t0;
  else if (t1 > kObjEps)
    t = t1;
  else
    return false;
  if (!(t < tmax)) return false;
  const V3 n = normalize((o + d * t) - v3(s.c[0], s.c[1], s.c[2]));
  h.t = t;
  h.n = n;
  h.material = s.material;
  h.front = dot(d, n) < 0.f;
  return true;
}

// ---- quadrics (sphere.rs, cylinder.rs, cone.rs, extra.rs Disk / Paraboloid) ----

__device__ __forceinline__ V3 q_point(const float* m, V3 p) {
  return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
__device__ __forceinline__ V3 q_vec(const float* m, V3 v) {
  return v3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
__device__ __forceinline__ V3 q_normal(const float* m, V3 n) {  // transpose of the inverse's 3x3
  return v3(m[0] * n.x + m[4] * n.y + m[8] * n.z, m[1] * n.x + m[5] * n.y + m[9] * n.z,
            m[2] * n.x + m[6] * n.y + m[10] * n.z);
}

// phi in [0, 360) from object x / y accepted when <= thetamax (extra.rs:44-55).
__device__ __forceinline__ bool theta_ok(float x, float y, float thetamax) {
  if (thetamax >= 360.f) return true;
  float phi = atan2f(y, x) * (180.f / kPi);
  if (phi < 0.f) phi += 360.f;
  return phi <= thetamax;
}

// Roots of a t^2 + b t + c as the reference names them: rm = (-b - s) / 2a,
// rp = (-b + s) / 2a, computed without cancellation. False if disc < 0.
__device__ __forceinline__ bool quad_roots(float a, float b, float c, float& rm, float& rp) {
  const float disc = b * b - 4.f * a * c;
  if (disc < 0.f) return false;
  const float s = sqrtf(disc);
  const float q = -0.5f * (b + copysignf(s, b));
  if (q == 0.f) {
    rm = rp = 0.f;
    return true;
  }
  const float r1 = q / a, r2 = c / q;  // r1 = rm when b >= 0, else rp
  if (b >= 0.f) {
    rm = r1;
    rp = r2;
  } else {
    rm = r2;
    rp = r1;
  }
  return true;
}

// Real roots of a3 t^3 + a2 t^2 + a1 t + a0 (ascending; a3 ~ 0 falls back to
// the quadratic / linear case): Cardano with one real root, the
// trigonometric form with three.
__device__ inline int cubic_roots(double a3, double a2, double a1, double a0, double* r) {
  if (fabs(a3) < 1e-12) {
    if (fabs(a2) < 1e-12) {
      if (fabs(a1) < 1e-12) return 0;
      r[0] = -a0 / a1;
      return 1;
    }
    const double disc = a1 * a1 - 4.0 * a2 * a0;
    if (disc < 0.0) return 0;
    const double s = sqrt(disc);
    const double x = (-a1 - s) / (2.0 * a2), y = (-a1 + s) / (2.0 * a2);
    r[0] = fmin(x, y);
    r[1] = fmax(x, y);
    return 2;
  }
  const double a = a2 / a3, b = a1 / a3, c = a0 / a3;
  const double p = b - a * a / 3.0, q = 2.0 * a * a * a / 27.0 - a * b / 3.0 + c, shift = -a / 3.0;
  const double disc = q * q / 4.0 + p * p * p / 27.0;
  if (disc > 0.0) {
    const double s = sqrt(disc);
    r[0] = cbrt(-q / 2.0 + s) + cbrt(-q / 2.0 - s) + shift;
    return 1;
  }
  const double m = 2.0 * sqrt(fmax(-p / 3.0, 0.0));
  if (m < 1e-30) {
    r[0] = shift;
    return 1;
  }
  const double th = acos(fmin(fmax(3.0 * q / (p * m), -1.0), 1.0)) / 3.0;
  for (int k = 0; k < 3; k++) r[k] = m * cos(th - 2.0 * 3.14159265358979323846 * k / 3.0) + shift;
  for (int i = 0; i < 3; i++)  // ascending
    for (int j = i + 1; j < 3; j++)
      if (r[j] < r[i]) {
        const double tmp = r[i];
        r[i] = r[j];
        r[j] = tmp;
      }
  return 3;
}

// Real roots of a4 t^4 + ... + a0 in f64 (ascending, near-duplicates
// merged): Ferrari's reduction to a depressed quartic, split into two
// quadratics through the largest positive root of the resolvent cubic (or
// a biquadratic when the odd term vanishes), each root polished by Newton
// steps and kept when its residual is small (extra.rs solve_quartic).
__device__ inline int quartic_roots(double a4, double a3, double a2, double a1, double a0, double* out) {
  if (fabs(a4) < 1e-12) return cubic_roots(a3, a2, a1, a0, out);
  const double b = a3 / a4, c = a2 / a4, d = a1 / a4, e = a0 / a4;
  const double p = c - 3.0 * b * b / 8.0;
  const double q = d - b * c / 2.0 + b * b * b / 8.0;
  const double r = e - b * d / 4.0 + b * b * c / 16.0 - 3.0 * b * b * b * b / 256.0;
  const double shift = -b / 4.0;
  double cand[4];
  int n = 0;
  if (fabs(q) < 1e-12) {
    const double disc = p * p - 4.0 * r;
    if (disc >= 0.0) {
      const double s = sqrt(disc);
      const double y2s[2] = {(-p - s) / 2.0, (-p + s) / 2.0};
      for (int i = 0; i < 2; i++)
        if (y2s[i] >= 0.0) {
          const double y = sqrt(y2s[i]);
          cand[n++] = y + shift;
          cand[n++] = -y + shift;
        }
    }
  } else {
    double z3[3];
    const int nz = cubic_roots(1.0, 2.0 * p, p * p - 4.0 * r, -q * q, z3);
    double z = -1.0;
    for (int i = 0; i < nz; i++)
      if (z3[i] > 1e-12 && z3[i] > z) z = z3[i];
    if (z > 0.0) {
      const double m = sqrt(z), nn = q / (2.0 * m);
      for (int sgn = 0; sgn < 2; sgn++) {
        const double bq = sgn == 0 ? m : -m;
        const double cq = (p + z) / 2.0 + (sgn == 0 ? -nn : nn);
        const double disc = bq * bq - 4.0 * cq;
        if (disc >= 0.0) {
          const double s = sqrt(disc);
          cand[n++] = (-bq - s) / 2.0 + shift;
          cand[n++] = (-bq + s) / 2.0 + shift;
        }
      }
    }
  }
  int m = 0;
  for (int i = 0; i < n; i++) {
    double t = cand[i];
    for (int it = 0; it < 8; it++) {
      const double f = (((a4 * t + a3) * t + a2) * t + a1) * t + a0;
      const double df = ((4.0 * a4 * t + 3.0 * a3) * t + 2.0 * a2) * t + a1;
      if (fabs(df) < 1e-30) break;
      const double step = f / df;
      t -= step;
      if (fabs(step) < 1e-12 * fmax(fabs(t), 1.0)) break;
    }
    const double res = (((a4 * t + a3) * t + a2) * t + a1) * t + a0;
    if (isfinite(t) && fabs(res) < 1e-6 * (1.0 + t * t * t * t * fabs(a4))) out[m++] = t;
  }
  for (int i = 0; i < m; i++)  // ascending, then merge near-duplicates
    for (int j = i + 1; j < m; j++)
      if (out[j] < out[i]) {
        const double tmp = out[i];
        out[i] = out[j];
        out[j] = tmp;
      }
  int k = 0;
  for (int i = 0; i < m; i++)
    if (k == 0 || fabs(out[i] - out[k - 1]) >= 1e-9) out[k++] = out[i];
  return k;
}

// The quadric's hit as render.rs computes it: the object-space ray with its
// direction normalised, the kind's root order and windows, t = the world
// distance to the hit, the normal through the inverse transpose (not
// flipped); front face only the sphere decides (sphere.rs:105), the others
// keep Intersection::new's `true`. Accepted when t < tmax.
__device__ __forceinline__ bool hit_quadric(const PtQuadric& q, V3 o, V3 d, float tmax, Hit& h) {
  const V3 lo = q_point(q.inv, o);
  const V3 ld = normalize(q_vec(q.inv, d));
  float t = 0.f;
  V3 p, ln;
  switch (q.kind) {
    case 0: {  // sphere.rs:38-75: nearest root > EPSILON, then the z / theta windows
      float rm, rp;
      if (!quad_roots(dot(ld, ld), 2.f * dot(lo, ld), dot(lo, lo) - q.radius * q.radius, rm, rp)) return false;
      if (rm > kObjEps)
        t = rm;
      else if (rp > kObjEps)
        t = rp;
      else
        return false;
      p = lo + ld * t;
      if (p.z < q.zmin || p.z > q.zmax || !theta_ok(p.x, p.y, q.thetamax)) return false;
      ln = normalize(p);
      break;
    }
    case 1: {  // cylinder.rs:36-93: the far root when the near one leaves [zmin, zmax]
      const float a = ld.x * ld.x + ld.y * ld.y;
      if (fabsf(a) < kObjEps) return false;
      float t1, t2;
      if (!quad_roots(a, 2.f * (lo.x * ld.x + lo.y * ld.y), lo.x * lo.x + lo.y * lo.y - q.radius * q.radius, t1, t2))
        return false;
      t = t1 > kObjEps ? t1 : t2;
      if (t <= kObjEps) return false;
      p = lo + ld * t;
      if (p.z < q.zmin || p.z > q.zmax) {
        if (t == t1 && t2 > kObjEps) {
          t = t2;
          p = lo + ld * t;
          if (p.z < q.zmin || p.z > q.zmax) return false;
        } else {
          return false;
        }
      }
      if (!theta_ok(p.x, p.y, q.thetamax)) return false;
      ln = normalize(v3(p.x, p.y, 0.f));
      break;
    }
    case 2: {  // cone.rs:48-101: x^2 + y^2 = k z^2, 0 <= z <= height
      const float a = ld.x * ld.x + ld.y * ld.y - q.k * ld.z * ld.z;
      const float b = 2.f * (lo.x * ld.x + lo.y * ld.y - q.k * lo.z * ld.z);
      const float c = lo.x * lo.x + lo.y * lo.y - q.k * lo.z * lo.z;
      bool hit = false;
      if (fabsf(a) < kObjEps) {  // one root, no theta window (cone.rs:63-75)
        if (fabsf(b) < kObjEps) return false;
        t = -c / b;
        if (t > kObjEps) {
          p = lo + ld * t;
          hit = p.z >= 0.f && p.z <= q.height;
        }
      } else {
        float t1, t2;
        if (!quad_roots(a, b, c, t1, t2)) return false;
        for (int i = 0; i < 2 && !hit; i++) {
          const float tc = i == 0 ? t1 : t2;
          if (!(tc > kObjEps)) continue;
          const V3 pc = lo + ld * tc;
          if (pc.z >= 0.f && pc.z <= q.height && theta_ok(pc.x, pc.y, q.thetamax)) {
            t = tc;
            p = pc;
            hit = true;
          }
        }
      }
      if (!hit) return false;
      const float r = sqrtf(p.x * p.x + p.y * p.y);
      ln = normalize(v3(p.x / r, p.y / r, -q.radius / q.height));  // cone.rs:118-125
      break;
    }
    case 3: {  // extra.rs Disk: the plane z = height within radius
      if (fabsf(ld.z) < kObjEps) return false;
      t = (q.height - lo.z) / ld.z;
      if (t <= kObjEps) return false;
      p = lo + ld * t;
      if (p.x * p.x + p.y * p.y > q.radius * q.radius || !theta_ok(p.x, p.y, q.thetamax)) return false;
      ln = v3(0.f, 0.f, 1.f);
      break;
    }
    case 5: {  // extra.rs Hyperboloid: the swept line x^2 + y^2 = pa z^2 + pb z + pc
      const float zlo = fminf(q.z1, q.z2), zhi = fmaxf(q.z1, q.z2);
      if (fabsf(q.z2 - q.z1) < kObjEps) {  // an annulus in the plane z = z1
        if (fabsf(ld.z) < kObjEps) return false;
        t = (q.z1 - lo.z) / ld.z;
        if (t <= kObjEps) return false;
        p = lo + ld * t;
        const float rr = p.x * p.x + p.y * p.y;
        if (rr < fminf(q.r1, q.r2) || rr > fmaxf(q.r1, q.r2) || !theta_ok(p.x, p.y, q.thetamax)) return false;
        ln = v3(0.f, 0.f, 1.f);
        break;
      }
      const float a = ld.x * ld.x + ld.y * ld.y - q.pa * ld.z * ld.z;
      const float b = 2.f * (lo.x * ld.x + lo.y * ld.y - q.pa * lo.z * ld.z) - q.pb * ld.z;
      const float c = lo.x * lo.x + lo.y * lo.y - q.pa * lo.z * lo.z - q.pb * lo.z - q.pc;
      float t1, t2;
      int nc = 2;
      if (fabsf(a) < kObjEps) {
        if (fabsf(b) < kObjEps) return false;
        t1 = -c / b;
        t2 = 0.f;
        nc = 1;
      } else if (!quad_roots(a, b, c, t1, t2)) {
        return false;
      }
      bool hit = false;
      for (int i = 0; i < nc && !hit; i++) {
        const float tc = i == 0 ? t1 : t2;
        if (tc <= kObjEps) continue;
        const V3 pc = lo + ld * tc;
        if (pc.z < zlo || pc.z > zhi || !theta_ok(pc.x, pc.y, q.thetamax)) continue;
        t = tc;
        p = pc;
        hit = true;
      }
      if (!hit) return false;
      ln = normalize(v3(2.f * p.x, 2.f * p.y, -(2.f * q.pa * p.z + q.pb)));
      break;
    }
    case 6: {  // extra.rs Torus (f64 quartic): the first root in the phi and theta windows
      const double R = q.radius, r = q.height;
      const double ox = lo.x, oy = lo.y, oz = lo.z, dx = ld.x, dy = ld.y, dz = ld.z;
      const double g = dx * dx + dy * dy + dz * dz;
      const double hh = 2.0 * (ox * dx + oy * dy + oz * dz);
      const double ii = ox * ox + oy * oy + oz * oz + R * R - r * r;
      const double j = dx * dx + dy * dy, k = 2.0 * (ox * dx + oy * dy), l = ox * ox + oy * oy;
      const double f4 = 4.0 * R * R;
      double roots[4];
      const int nr = quartic_roots(g * g, 2.0 * g * hh, hh * hh + 2.0 * g * ii - f4 * j, 2.0 * hh * ii - f4 * k,
                                   ii * ii - f4 * l, roots);
      bool hit = false;
      for (int i = 0; i < nr && !hit; i++) {
        if (roots[i] <= 1e-6) continue;
        const double px = ox + dx * roots[i], py = oy + dy * roots[i], pz = oz + dz * roots[i];
        if (q.zmax - q.zmin < 360.f) {  // phi window (extra.rs Torus::phi_ok)
          const double rxy = sqrt(px * px + py * py);
          double phi = atan2(pz, rxy - R) * (180.0 / 3.14159265358979323846);
          while (phi < q.zmin) phi += 360.0;
          while (phi >= q.zmin + 360.0) phi -= 360.0;
          if (phi > q.zmax) continue;
        }
        if (!theta_ok((float)px, (float)py, q.thetamax)) continue;
        const double rxy = fmax(sqrt(px * px + py * py), 1e-12), sc = 1.0 - R / rxy;
        t = (float)roots[i];
        p = v3((float)px, (float)py, (float)pz);
        ln = normalize(v3((float)(px * sc), (float)(py * sc), (float)pz));
        hit = true;
      }
      if (!hit) return false;
      break;
    }
    case 4: {  // extra.rs Paraboloid: x^2 + y^2 = k z, first root in the windows
      if (fabsf(q.zmax) < kObjEps) return false;
      const float a = ld.x * ld.x + ld.y * ld.y;
      const float b = 2.f * (lo.x * ld.x + lo.y * ld.y) - q.k * ld.z;
      const float c = lo.x * lo.x + lo.y * lo.y - q.k * lo.z;
      float t1, t2;
      int nc = 2;
      if (fabsf(a) < kObjEps) {
        if (fabsf(b) < kObjEps) return false;
        t1 = -c / b;
        t2 = 0.f;
        nc = 1;
      } else if (!quad_roots(a, b, c, t1, t2)) {
        return false;
      }
      bool hit = false;
      for (int i = 0; i < nc && !hit; i++) {
        const float tc = i == 0 ? t1 : t2;
        if (tc <= kObjEps) continue;
        const V3 pc = lo + ld * tc;
        if (pc.z < q.zmin || pc.z > q.zmax || !theta_ok(pc.x, pc.y, q.thetamax)) continue;
        t = tc;
        p = pc;
        hit = true;
      }
      if (!hit) return false;
      ln = normalize(v3(2.f * p.x, 2.f * p.y, -q.k));
      break;
    }
    default:
      return false;
  }
  const V3 wp = q_point(q.fwd, p);
  const V3 dw = wp - o;
  const float tw = sqrtf(dot(dw, dw));
  if (!(tw < tmax)) return false;
  h.t = tw;
  h.n = normalize(q_normal(q.inv, ln));
  h.material = q.material;
  h.front = q.kind == 0 ? dot(d, h.n) < 0.f : true;
  return true;
}

// scene/mod.rs:105-146 over the loose objects (strictly closer wins).
// QUAD: the scene holds quadrics (kernels without them keep the quadric
// solvers, and their registers, out of their code).
template <bool QUAD>
__device__ __forceinline__ bool closest_hit(const PtSceneView& S, V3 o, V3 d, Hit& best) {
  bool found = false;
  float tmax = INFINITY;
  for (uint32_t i = 0; i < S.num_tris; i++) {
    Hit h;
    if (hit_triangle(S.tris[i], o, d, tmax, h)) {
      best = h;
      best.obj = (kObjTri << 30) | i;
      tmax = h.t;
      found = true;
    }
  }
  for (uint32_t i = 0; i < S.num_spheres; i++) {
    Hit h;
    if (hit_sphere(S.spheres[i], o, d, tmax, h)) {
      best = h;
      best.obj = (kObjSphere << 30) | i;
      tmax = h.t;
      found = true;
    }
  }
  for (uint32_t i = 0; QUAD && i < S.num_quadrics; i++) {
    Hit h;
    if (hit_quadric(S.quadrics[i], o, d, tmax, h)) {
      best = h;
      best.obj = (kObjQuadric << 30) | i;
      tmax = h.t;
      found = true;
    }
  }
  return found;
}

// Hit point and push of a loose-object hit `h` of ray (o, d) (see Hit):
// triangles as v0 + e1 u + e2 v from the stored vertex and edges the
// intersection test uses (the exact point lies on that plane for any u, v;
// two products and two sums bound the rounding by gamma(3) of the summed
// magnitudes), spheres reprojected onto the radius (gamma(5) of the offset
// from the centre, plus the centre's addition), quadrics along the ray with
// a wide bound covering their f32 root solve.
__device__ __forceinline__ void loose_spawn(const PtSceneView& S, V3 o, V3 d, Hit& h) {
  const uint32_t kind = h.obj >> 30, i = h.obj & kObjIndex;
  if (kind == kObjTri) {
    const PtTriangle& tr = S.tris[i];
    const V3 v0 = v3(tr.v0), e1 = v3(tr.e1), e2 = v3(tr.e2);
    const V3 pv = cross(d, e2);
    const float inv = 1.f / dot(e1, pv);
    const V3 tv = o - v0;
    const float u = dot(tv, pv) * inv;
    const float v = dot(d, cross(tv, e1)) * inv;
    const V3 a = e1 * u, b = e2 * v;
    h.p = v0 + a + b;
    h.e = bound_push(v3(tr.n), (vabs(v0) + vabs(a) + vabs(b)) * err_gamma(3), d);
  } else if (kind == kObjSphere) {
    const PtSphere& sp = S.spheres[i];
    const V3 c = v3(sp.c[0], sp.c[1], sp.c[2]);
    V3 pl = (o - c) + d * h.t;
    pl = pl * (sp.r / fmaxf(len(pl), 1e-30f));
    h.p = c + pl;
    h.e = bound_push(pl / sp.r, vabs(pl) * err_gamma(6) + vabs(c) * err_gamma(1), d);
  } else {
    const V3 td = d * h.t;
    h.p = o + td;
    h.e = bound_push(h.n, (vabs(o) + vabs(td)) * err_gamma(64), d);
  }
}

// Opaque-scene shadow test (pt/mod.rs:1243-1306): blocked when anything is
// hit before `tmax`.
template <bool QUAD>
__device__ __forceinline__ bool occluded(const PtSceneView& S, V3 o, V3 d, float tmax) {
  for (uint32_t i = 0; i < S.num_tris; i++) {
    Hit h;
    if (hit_triangle(S.tris[i], o, d, tmax, h)) return true;
  }
  for (uint32_t i = 0; i < S.num_spheres; i++) {
    float t0, t1;
    if (!sphere_roots(S.spheres[i], o, d, t0, t1)) continue;
    const float t = t0 > kObjEps ? t0 : t1;
    if (t > kObjEps && t < tmax) return true;
  }
  for (uint32_t i = 0; QUAD && i < S.num_quadrics; i++) {
    Hit h;
    if (hit_quadric(S.quadrics[i], o, d, tmax, h)) return true;
  }
  return false;
}

// ---- participating media (medium.rs:28-101, pt/volume.rs:1-189) ---------------------

__device__ __forceinline__ bool heterogeneous(const PtMedium& m) { return m.grid || m.fbm_octaves; }

// Density multiplier at p: render.rs's cloud field clamp((fbm(p) * 0.5 +
// 0.5 - (1 - coverage)) * sharpness, 0, 1) at the world point, evaluated in
// f64 (medium.rs:55-63); trilinear over the vertex grid, 0 outside it; 1
// for homogeneous media.
__device__ __forceinline__ float density_at(const PtMedium& m, V3 p) {
  if (m.fbm_octaves) {
    const double off[3] = {0.0, 0.0, 0.0};
    const double n = fbm3((double)p.x, (double)p.y, (double)p.z, (double)m.fbm_frequency, m.fbm_octaves,
                          (double)m.fbm_gain, (double)m.fbm_lacunarity, off) * 0.5 + 0.5;
    const double r = (n - (1.0 - (double)m.fbm_coverage)) * (double)m.fbm_sharpness;
    return (float)fmin(fmax(r, 0.0), 1.0);
  }
  if (!m.grid) return 1.f;
  const float u[3] = {(p.x - m.lo[0]) * m.scale[0], (p.y - m.lo[1]) * m.scale[1], (p.z - m.lo[2]) * m.scale[2]};
  const uint32_t n[3] = {m.nx, m.ny, m.nz};
  int i[3];
  float f[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (!(u[a] >= 0.f && u[a] <= (float)(n[a] - 1))) return 0.f;
    i[a] = min((int)u[a], (int)n[a] - 2);
    f[a] = u[a] - (float)i[a];
  }
  const float* g = m.grid + ((size_t)i[2] * m.ny + i[1]) * m.nx + i[0];
  const size_t sy = m.nx, sz = (size_t)m.nx * m.ny;
  const float c00 = g[0] + (g[1] - g[0]) * f[0];
  const float c10 = g[sy] + (g[sy + 1] - g[sy]) * f[0];
  const float c01 = g[sz] + (g[sz + 1] - g[sz]) * f[0];
  const float c11 = g[sz + sy] + (g[sz + sy + 1] - g[sz + sy]) * f[0];
  const float c0 = c00 + (c10 - c00) * f[1], c1 = c01 + (c11 - c01) * f[1];
  return c0 + (c1 - c0) * f[2];
}

// Parametric range of the ray inside the medium's grid box (the whole
// line for homogeneous media). Outside the grid the density is zero, so
// tracking can start at the entry and stop at the exit: null collisions
// there change nothing and the exponential steps are memoryless.
__device__ __forceinline__ bool medium_range(const PtMedium& m, V3 o, V3 d, float t_max, float& t0, float& t1) {
  t0 = 0.f;
  t1 = fminf(t_max, m.max_distance);
  if (!m.grid) return t1 > t0;  // homogeneous, or a procedural field over the whole line
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float inv = safe_rcp(dd[a]);
    float ta = (m.lo[a] - oo[a]) * inv, tb = (m.hi[a] - oo[a]) * inv;
    if (ta > tb) {
      const float x = ta;
      ta = tb;
      tb = x;
    }
    t0 = fmaxf(t0, ta);
    t1 = fminf(t1, tb);
  }
  return t1 > t0;
}

struct MediumEvent {
  bool scatter;
  float t;
  V3 weight;
};

__device__ __forceinline__ float avg3(V3 v) { return (v.x + v.y + v.z) * (1.f / 3.f); }
__device__ __forceinline__ V3 exp3(V3 v, float t) { return v3(expf(-v.x * t), expf(-v.y * t), expf(-v.z * t)); }
__device__ __forceinline__ float exp_step(Pcg32& rng, float sigma) {
  return -logf(fmaxf(1.f - rnd(rng), 1e-12f)) / sigma;
}

// volume.rs:29-47 sample_distance: spectral one-sample MIS in homogeneous
// media (volume.rs:60-102), weighted delta tracking against the
// channel-max majorant in gridded ones (volume.rs:104-149).
__device__ __forceinline__ MediumEvent sample_distance(const PtMedium& m, V3 o, V3 d, float t_max, V3 beta,
                                                      Pcg32& rng) {
  const V3 st = v3(m.sigma_t);
  if (!heterogeneous(m)) {
    const float tm = fminf(t_max, m.max_distance);
    const float sa = avg3(st);
    if (sa <= 1e-9f) return {false, 0.f, v3(1.f, 1.f, 1.f)};
    const V3 b = v3(fmaxf(beta.x, 0.f), fmaxf(beta.y, 0.f), fmaxf(beta.z, 0.f));
    const float tot = b.x + b.y + b.z;
    const float w[3] = {tot > 1e-12f ? b.x / tot : 1.f / 3.f, tot > 1e-12f ? b.y / tot : 1.f / 3.f,
                        tot > 1e-12f ? b.z / tot : 1.f / 3.f};
    const float u = rnd(rng);
    const float sc[3] = {st.x, st.y, st.z};
    const int c = u < w[0] ? 0 : (u < w[0] + w[1] ? 1 : 2);
    const float t = exp_step(rng, fmaxf(sc[c], 1e-9f));
    if (t < tm) {
      const V3 tr = exp3(st, t);
      const float pdf = w[0] * st.x * tr.x + w[1] * st.y * tr.y + w[2] * st.z * tr.z;
      if (!(pdf > 0.f)) return {false, 0.f, v3(0.f, 0.f, 0.f)};
      return {true, t, v3(m.sigma_s) * tr / pdf};
    }
    const V3 tr = exp3(st, tm);
    const float pdf = w[0] * tr.x + w[1] * tr.y + w[2] * tr.z;
    if (!(pdf > 0.f)) return {false, 0.f, v3(0.f, 0.f, 0.f)};
    return {false, 0.f, tr / pdf};
  }
  V3 weight = v3(1.f, 1.f, 1.f);
  const float maj = m.majorant;
  float t, t_end;
  if (maj <= 1e-9f || !medium_range(m, o, d, t_max, t, t_end)) return {false, 0.f, weight};
  for (int i = 0; i < 10000; i++) {
    t += exp_step(rng, maj);
    if (t >= t_end) return {false, 0.f, weight};
    const float rho = density_at(m, o + d * t);
    const V3 sh = st * rho;
    const float p_real = fminf(fmaxf(avg3(sh) / maj, 0.f), 1.f);
    if (rnd(rng) < p_real) return {true, t, weight * v3(m.sigma_s) * (rho / (maj * p_real))};
    const float p_null = fmaxf(1.f - p_real, 1e-9f);
    weight = weight * v3((1.f - sh.x / maj) / p_null, (1.f - sh.y / maj) / p_null, (1.f - sh.z / maj) / p_null);
    if (maxc(weight) < 1e-6f) return {false, 0.f, v3(0.f, 0.f, 0.f)};
  }
  return {false, 0.f, weight};
}

// volume.rs:153-189: closed form (homogeneous) or ratio tracking.
__device__ __forceinline__ V3 medium_transmittance(const PtMedium& m, V3 o, V3 d, float dist, Pcg32& rng) {
  const V3 st = v3(m.sigma_t);
  if (!heterogeneous(m)) return exp3(st, fminf(dist, m.max_distance));
  const float maj = m.majorant;
  float t, t_end;
  V3 tr = v3(1.f, 1.f, 1.f);
  if (maj <= 1e-9f || !medium_range(m, o, d, dist, t, t_end)) return tr;
  for (int i = 0; i < 10000; i++) {
    t += exp_step(rng, maj);
    if (t >= t_end) break;
    const float rho = density_at(m, o + d * t);
    tr = tr * v3(1.f - st.x * rho / maj, 1.f - st.y * rho / maj, 1.f - st.z * rho / maj);
    if (maxc(tr) < 1e-5f) return v3(0.f, 0.f, 0.f);
  }
  return tr;
}

// volume.rs:195-232: equiangular distance sampling towards a point light
// (Kulla and Fajardo 2012) over [0, t_max] of a unit-direction ray, and
// its pdf at t.
__device__ __forceinline__ void equiangular_frame(V3 o, V3 d, float t_max, V3 light, float& delta, float& h,
                                                  float& th_a, float& span) {
  delta = dot(light - o, d);
  h = fmaxf(len(light - (o + d * delta)), 1e-6f);
  th_a = atanf(-delta / h);
  span = fmaxf(atanf((t_max - delta) / h) - th_a, 1e-9f);
}

__device__ __forceinline__ float equiangular_sample(V3 o, V3 d, float t_max, V3 light, float u, float& pdf) {
  float delta, h, th_a, span;
  equiangular_frame(o, d, t_max, light, delta, h, th_a, span);
  const float t = fminf(fmaxf(delta + h * tanf(th_a + u * span), 0.f), t_max);
  pdf = h / (span * (h * h + (t - delta) * (t - delta)));
  return t;
}

__device__ __forceinline__ float equiangular_pdf(V3 o, V3 d, float t_max, V3 light, float t) {
  float delta, h, th_a, span;
  equiangular_frame(o, d, t_max, light, delta, h, th_a, span);
  return h / (span * (h * h + (t - delta) * (t - delta)));
}

// volume.rs:234-240: the approximate distance-sampling pdf that weighs
// equiangular sampling (mean extinction; tail mass past t_max).
__device__ __forceinline__ float distance_pdf_approx(const PtMedium& m, float t, float t_max) {
  const float s = fmaxf((m.sigma_t[0] + m.sigma_t[1] + m.sigma_t[2]) * (1.f / 3.f), 1e-9f);
  return t < t_max ? s * expf(-s * t) : expf(-s * t_max);
}

// medium.rs:73-101
__device__ __forceinline__ float hg_phase(float g, float cos_theta) {
  const float denom = 1.f + g * g + 2.f * g * cos_theta;
  return (1.f - g * g) / (4.f * kPi * denom * sqrtf(fmaxf(denom, 1e-9f)));
}

__device__ __forceinline__ V3 hg_sample(float g, V3 wo, float u1, float u2, float& pdf) {
  float cos_t;
  if (fabsf(g) < 1e-3f) {
    cos_t = 1.f - 2.f * u1;
  } else {
    const float sq = (1.f - g * g) / (1.f + g - 2.f * g * u1);
    cos_t = -(1.f + g * g - sq * sq) / (2.f * g);
  }
  const float sin_t = sqrtf(fmaxf(1.f - cos_t * cos_t, 0.f));
  float sp, cp;
  sincosf(2.f * kPi * u2, &sp, &cp);
  const V3 w = -wo;
  const V3 a = fabsf(w.x) > 0.9f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
  const V3 t = normalize(cross(w, a));
  const V3 b = cross(w, t);
  const V3 wi = normalize(t * (sin_t * cp) + b * (sin_t * sp) + w * cos_t);
  pdf = hg_phase(g, -dot(wo, wi));
  return wi;
}

// ---- camera (camera.rs:35-179, pinhole) -------------------------------------------

__device__ __forceinline__ float filter_1d(const PtCamera& c, float u) {
  if (c.filter == 1) {
    const float half = c.filter_width * 0.5f;
    return u < 0.5f ? (sqrtf(2.f * u) - 1.f) * half : (1.f - sqrtf(2.f * (1.f - u))) * half;
  }
  if (c.filter == 2) {
    const float u1 = fmaxf(u, 1e-12f);
    const double x = (double)u * 2654435761.0;
    const float u2 = (float)(x - floor(x));
    const float sigma = c.filter_width * 0.25f;
    const float g = sqrtf(-2.f * logf(u1)) * cosf(2.f * kPi * u2);
    return clampf(g * sigma, -c.filter_width * 0.5f, c.filter_width * 0.5f);
  }
  return (u - 0.5f) * c.filter_width;
}

// pt/mod.rs:340-356 + camera.rs:150-179: filter offsets (u1, u2), then the
// aperture sample (lu, lv) when the lens is open; thin lens through the
// focal plane at forward distance focal_distance, or orthographic.
__device__ __forceinline__ void camera_ray(const PtCamera& c, uint32_t x, uint32_t y, Pcg32& rng, V3& o, V3& d) {
  const float u1 = rnd(rng), u2 = rnd(rng);
  const float px = (float)x + 0.5f + filter_1d(c, u1);
  const float py = (float)y + 0.5f + filter_1d(c, u2);
  float lu = 0.5f, lv = 0.5f;
  if (c.lens_radius > 0.f) {
    lu = rnd(rng);
    lv = rnd(rng);
  }
  const float u = (px / (float)c.width) * 2.f - 1.f;
  const float v = 1.f - (py / (float)c.height) * 2.f;
  if (c.projection == 1) {
    o = v3(c.eye) + v3(c.right) * (u * c.ortho_w) + v3(c.up) * (v * c.ortho_h);
    d = v3(c.forward);
    return;
  }
  const V3 dir = v3(c.forward) + v3(c.right) * (u * c.half_w) + v3(c.up) * (v * c.half_h);
  if (c.lens_radius > 0.f) {
    const float r = c.lens_radius * sqrtf(lu), phi = 2.f * kPi * lv;
    o = v3(c.eye) + v3(c.right) * (r * cosf(phi)) + v3(c.up) * (r * sinf(phi));
    d = normalize(v3(c.eye) + dir * c.focal_distance - o);
  } else {
    o = v3(c.eye);
    d = normalize(dir);
  }
}

}  // namespace pt
}  // namespace rrhip
