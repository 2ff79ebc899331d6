// Hair fiber scattering on the device (render.rs src/raytracer/pt/hair.rs,
// after Marschner et al. 2003 as reformulated by d'Eon et al. 2011 and
// pbrt-v3): R, TT and TRT lobes plus a residual, each the product of an
// energy-conserving longitudinal lobe Mp, a trimmed-logistic azimuthal
// lobe Np and a Fresnel / absorption attenuation Ap. f is per solid angle
// (no cosine on either side, hair.rs:7-11); directions are in the fiber
// frame, x along the fiber, theta measured from the normal plane.
//
// f32 throughout: the longitudinal lobe in the log domain for v <= 0.1,
// where its exponent's terms of size 1 / v ~ 50 cancel to O(1).
#pragma once
#include <hip/hip_runtime.h>

#include "pt_device.h"
#include "pt_types.h"

namespace rrhip {
namespace pt {

// I0 by its power series, ten terms (hair.rs:101-116).
__device__ __forceinline__ float hair_i0(float x) {
  float sum = 0.f, x2i = 1.f, fact = 1.f, four_i = 1.f;
#pragma unroll
  for (int i = 0; i < 10; i++) {
    if (i > 1) fact *= (float)i;
    sum += x2i / (four_i * fact * fact);
    x2i *= x * x;
    four_i *= 4.f;
  }
  return sum;
}

// log I0: the series below 12, pbrt-v3's asymptotic form above
// (hair.rs:118-124).
__device__ __forceinline__ float hair_log_i0(float x) {
  if (x > 12.f) return x + 0.5f * (-logf(2.f * kPi * x) + 1.f / (8.f * x));
  return logf(hair_i0(x));
}

// d'Eon's longitudinal lobe (hair.rs:127-135).
__device__ __forceinline__ float hair_mp(float cos_ti, float cos_to, float sin_ti, float sin_to, float v) {
  const float a = cos_ti * cos_to / v, b = sin_ti * sin_to / v;
  if (v <= 0.1f) return expf(hair_log_i0(a) - b - 1.f / v + 0.6931f + logf(1.f / (2.f * v)));
  return expf(-b) * hair_i0(a) / (sinhf(1.f / v) * 2.f * v);
}

// Unpolarised dielectric Fresnel reflectance (hair.rs:137-153).
__device__ __forceinline__ float hair_fresnel(float cos_i, float eta_i, float eta_t) {
  cos_i = clampf(cos_i, -1.f, 1.f);
  if (!(cos_i > 0.f)) {
    const float tmp = eta_i;
    eta_i = eta_t;
    eta_t = tmp;
    cos_i = -cos_i;
  }
  const float sin_i = sqrtf(fmaxf(1.f - cos_i * cos_i, 0.f));
  const float sin_t = eta_i / eta_t * sin_i;
  if (sin_t >= 1.f) return 1.f;
  const float cos_t = sqrtf(fmaxf(1.f - sin_t * sin_t, 0.f));
  const float par = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t);
  const float perp = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t);
  return 0.5f * (par * par + perp * perp);
}

__device__ __forceinline__ float hair_logistic(float x, float s) {
  const float e = expf(-fabsf(x) / s);
  return e / (s * (1.f + e) * (1.f + e));
}
__device__ __forceinline__ float hair_logistic_cdf(float x, float s) { return 1.f / (1.f + expf(-x / s)); }

// Np of lobe p at azimuth phi (hair.rs:182-205): the trimmed logistic on
// [-pi, pi] about the lobe's net deflection 2 p gamma_t - 2 gamma_o + p pi.
__device__ __forceinline__ float hair_np(float phi, int p, float s, float gamma_o, float gamma_t) {
  float dphi = phi - (2.f * (float)p * gamma_t - 2.f * gamma_o + (float)p * kPi);
  while (dphi > kPi) dphi -= 2.f * kPi;
  while (dphi < -kPi) dphi += 2.f * kPi;
  return hair_logistic(dphi, s) / (hair_logistic_cdf(kPi, s) - hair_logistic_cdf(-kPi, s));
}

// The outgoing direction's angles, the refracted azimuth and one
// crossing's body transmittance (hair.rs:207-240), and the attenuations
// Ap of R, TT, TRT and the residual (hair.rs:156-180).
struct HairGeom {
  float sin_to, cos_to, phi_o, gamma_o, gamma_t;
  V3 a[4];
};

__device__ __forceinline__ HairGeom hair_geom(const PtMaterial& m, V3 wo, float h) {
  HairGeom g;
  g.sin_to = clampf(wo.x, -1.f, 1.f);
  g.cos_to = sqrtf(fmaxf(1.f - g.sin_to * g.sin_to, 0.f));
  g.phi_o = atan2f(wo.z, wo.y);
  g.gamma_o = asinf(clampf(h, -1.f, 1.f));
  const float eta = m.hair_eta;
  const float sin_tt = g.sin_to / eta;
  const float cos_tt = sqrtf(fmaxf(1.f - sin_tt * sin_tt, 0.f));
  const float etap = sqrtf(fmaxf(eta * eta - g.sin_to * g.sin_to, 0.f)) / fmaxf(g.cos_to, 1e-9f);
  const float sin_gt = clampf(h / etap, -1.f, 1.f);
  const float cos_gt = sqrtf(fmaxf(1.f - sin_gt * sin_gt, 0.f));
  g.gamma_t = asinf(sin_gt);
  const float l = 2.f * cos_gt / fmaxf(cos_tt, 1e-9f);
  const V3 T = v3(expf(-m.hair_sigma_a[0] * l), expf(-m.hair_sigma_a[1] * l), expf(-m.hair_sigma_a[2] * l));
  const float cos_go = sqrtf(fmaxf(1.f - h * h, 0.f));
  const float f = hair_fresnel(g.cos_to * cos_go, 1.f, eta);
  g.a[0] = v3(f, f, f);
  g.a[1] = T * ((1.f - f) * (1.f - f));
  g.a[2] = g.a[1] * T * f;
  const V3 tf = T * f;
  g.a[3] = g.a[2] * tf * v3(1.f / fmaxf(1.f - tf.x, 1e-6f), 1.f / fmaxf(1.f - tf.y, 1e-6f),
                            1.f / fmaxf(1.f - tf.z, 1e-6f));
  return g;
}

// Lobe selection weights: Ap's luminances normalised (hair.rs:268-286).
__device__ __forceinline__ void hair_weights(const HairGeom& g, float w[4]) {
  float total = 0.f;
#pragma unroll
  for (int p = 0; p < 4; p++) {
    w[p] = lum(g.a[p]);
    total += w[p];
  }
  if (total > 1e-12f) {
#pragma unroll
    for (int p = 0; p < 4; p++) w[p] /= total;
  } else {
    w[0] = 1.f;
    w[1] = w[2] = w[3] = 0.f;
  }
}

// f and the sampling pdf towards wi (hair.rs:242-309).
__device__ __forceinline__ void hair_eval(const PtMaterial& m, const HairGeom& g, V3 wi, V3& f, float& pdf) {
  const float sin_ti = clampf(wi.x, -1.f, 1.f);
  const float cos_ti = sqrtf(fmaxf(1.f - sin_ti * sin_ti, 0.f));
  const float phi = atan2f(wi.z, wi.y) - g.phi_o;
  float w[4];
  hair_weights(g, w);
  f = v3(0.f, 0.f, 0.f);
  pdf = 0.f;
#pragma unroll
  for (int p = 0; p < 4; p++) {
    const float mp = hair_mp(cos_ti, g.cos_to, sin_ti, g.sin_to, m.hair_v[p]);
    const float np = p < 3 ? hair_np(phi, p, m.hair_s, g.gamma_o, g.gamma_t) : 1.f / (2.f * kPi);
    f = f + g.a[p] * (mp * np);
    pdf += w[p] * mp * np;
  }
  pdf = fmaxf(pdf, 0.f);
}

// hair.rs:311-363 from four uniforms (lobe, two longitudinal, azimuthal):
// the lobe by weight, d'Eon's inverse of Mp's cone rotated about wo's
// angle (sin^2 theta = -w (2 + w) with cos theta = 1 + w, exact near 1),
// the azimuth by the trimmed logistic's inverse CDF about the lobe's
// deflection (uniform for the residual). False where the pdf <= 1e-12.
__device__ __forceinline__ bool hair_sample(const PtMaterial& m, const HairGeom& g, const float u[4], V3& wi, V3& f,
                                            float& pdf) {
  float w[4];
  hair_weights(g, w);
  float ul = u[0];
  int p = 3;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    if (p == 3 && i < 3 && ul < w[i]) p = i;
    if (p == 3) ul -= w[i];
  }
  // the residual is taken when u lands past the first three weights
  // (selects, not an indexed load, keep m in registers)
  const float v = p == 0 ? m.hair_v[0] : p == 1 ? m.hair_v[1] : p == 2 ? m.hair_v[2] : m.hair_v[3];
  const float u1 = fmaxf(u[1], 1e-9f);
  const float wl = v * logf(u1 + (1.f - u1) * expf(-2.f / v));
  const float cos_theta = 1.f + wl;
  const float sin_theta = sqrtf(fmaxf(-wl * (2.f + wl), 0.f));
  const float cos_phi_l = cosf(2.f * kPi * u[2]);
  const float sin_ti = -cos_theta * g.sin_to + sin_theta * cos_phi_l * g.cos_to;
  const float cos_ti = sqrtf(fmaxf(1.f - sin_ti * sin_ti, 0.f));
  float dphi;
  if (p < 3) {
    const float s = m.hair_s;
    const float k = hair_logistic_cdf(kPi, s) - hair_logistic_cdf(-kPi, s);
    const float q = clampf(u[3] * k + hair_logistic_cdf(-kPi, s), 1e-9f, 1.f - 1e-9f);
    const float x = clampf(-s * logf(1.f / q - 1.f), -kPi, kPi);
    dphi = 2.f * (float)p * g.gamma_t - 2.f * g.gamma_o + (float)p * kPi + x;
  } else {
    dphi = 2.f * kPi * u[3];
  }
  const float phi_i = g.phi_o + dphi;
  float sp, cp;
  sincosf(phi_i, &sp, &cp);
  wi = v3(sin_ti, cos_ti * cp, cos_ti * sp);
  hair_eval(m, g, wi, f, pdf);
  return pdf > 1e-12f;
}

// The fiber frame at a curve hit (hair.rs:365-401): x the tangent, z the
// viewer direction off the fiber (the normal's when wo runs along it),
// y = z x x; the azimuthal offset h is the normal's component along y.
struct FiberFrame {
  V3 x, y, z;
  float h;
};
__device__ __forceinline__ FiberFrame fiber_frame(V3 tangent, V3 n, V3 wo) {
  FiberFrame ff;
  ff.x = normalize(tangent);
  V3 z = wo - ff.x * dot(wo, ff.x);
  if (dot(z, z) < 1e-12f) z = n - ff.x * dot(n, ff.x);
  ff.z = normalize(z);
  ff.y = normalize(cross(ff.z, ff.x));
  const V3 np = normalize(n - ff.x * dot(n, ff.x));
  ff.h = clampf(dot(np, ff.y), -1.f, 1.f);
  return ff;
}
__device__ __forceinline__ V3 fiber_local(const FiberFrame& ff, V3 w) { return v3(dot(w, ff.x), dot(w, ff.y), dot(w, ff.z)); }
__device__ __forceinline__ V3 fiber_world(const FiberFrame& ff, V3 w) { return ff.x * w.x + ff.y * w.y + ff.z * w.z; }

}  // namespace pt
}  // namespace rrhip
