// Lobe constants of a material from its PbrParams, shared by the host
// scene build (f64 inputs) and the shade kernels, which re-derive them at
// hits whose material has pattern-bound fields (f32 inputs).
#pragma once
#include <hip/hip_runtime.h>

#include "pt_types.h"

namespace rrhip {

__host__ __device__ inline float pm_sqrt(float x) { return sqrtf(x); }
__host__ __device__ inline double pm_sqrt(double x) { return sqrt(x); }
template <class T>
__host__ __device__ inline T pm_min(T a, T b) {
  return b < a ? b : a;
}
template <class T>
__host__ __device__ inline T pm_max(T a, T b) {
  return a < b ? b : a;
}
template <class T>
__host__ __device__ inline T pm_lum(const T c[3]) {
  return T(0.2126) * c[0] + T(0.7152) * c[1] + T(0.0722) * c[2];
}
template <class T>
__host__ __device__ inline T pm_alpha(T r) {  // bxdf.rs:187-189
  return pm_min(pm_max(r * r, T(2.5e-5)), T(1));
}

// pbr.rs:84-93 (specular F0), 191-202 (lobe weights); bxdf.rs:137-171 (the
// layering scales, Oren-Nayar A / B, GGX alphas); the radiance-cache inputs
// and the albedo AOV (pt/mod.rs:690-696). Writes every lobe field of g;
// identity, media, hull and hair fields are left alone.
template <class T>
__host__ __device__ inline void lobe_constants(const RawPbr<T>& m, PtMaterial& g) {
  constexpr T kPiT = T(3.14159265358979323846);
  T f0[3] = {T(0), T(0), T(0)};
  if (pm_lum(m.specular_face_color) > T(1e-6)) {
    for (int c = 0; c < 3; c++) f0[c] = m.specular_face_color[c];
  } else if (m.specular_ior > T(1)) {
    const T r = (m.specular_ior - T(1)) / (m.specular_ior + T(1));
    for (int c = 0; c < 3; c++) f0[c] = r * r;
  }
  const T wd = m.diffuse_gain * pm_lum(m.diffuse_color);
  const T ws = pm_sqrt(pm_lum(f0));
  const T wc = m.clearcoat_gain * T(0.2);
  const T wf = m.fuzz_gain * pm_lum(m.fuzz_color) * T(0.3);
  const T wg = m.glass_gain;
  const T total = wd + ws + wc + wf + wg;
  g.has_lobes = total > T(1e-9) ? 1u : 0u;
  const T wv[5] = {wd, ws, wc, wf, wg};
  for (int i = 0; i < 5; i++) g.w[i] = g.has_lobes ? (float)(wv[i] / total) : 0.f;
  const T spec_take = pm_lum(f0);
  const T coat_take = T(0.04) * m.clearcoat_gain;
  const T under = pm_min(pm_max((T(1) - spec_take) * (T(1) - coat_take), T(0)), T(1));
  for (int c = 0; c < 3; c++) g.diffuse[c] = (float)(m.diffuse_color[c] * (m.diffuse_gain * under / kPiT));
  const T s2 = m.diffuse_roughness * m.diffuse_roughness;
  g.on_a = (float)(T(1) - s2 / (T(2) * (s2 + T(0.33))));
  g.on_b = (float)(T(0.45) * s2 / (s2 + T(0.09)));
  for (int c = 0; c < 3; c++) {
    g.spec_f0[c] = (float)f0[c];
    g.spec_f90[c] = (float)m.specular_edge_color[c];
  }
  g.spec_alpha = (float)pm_alpha(m.specular_roughness);
  g.coat_f0 = (float)(T(0.04) * m.clearcoat_gain);
  g.coat_f90 = (float)m.clearcoat_gain;
  g.coat_alpha = (float)pm_alpha(m.clearcoat_roughness);
  for (int c = 0; c < 3; c++) g.fuzz[c] = (float)(m.fuzz_color[c] * (m.fuzz_gain / kPiT));
  g.glass_scale = (float)(m.glass_gain * under);
  for (int c = 0; c < 3; c++) g.refraction[c] = (float)m.refraction_color[c];
  g.glass_alpha = (float)pm_alpha(m.glass_roughness);
  g.glass_ior = (float)m.glass_ior;
  for (int c = 0; c < 3; c++) g.emission[c] = (float)(m.emission[c] + m.glow[c]);
  // Radiance-cache inputs: the lobes' summed albedo (clamped to [0, 1])
  // and the roughness of the strongest non-diffuse lobe (1 if none).
  for (int c = 0; c < 3; c++) {
    const T a = m.diffuse_color[c] * (m.diffuse_gain * under) + (ws > T(0) ? f0[c] : T(0)) +
                m.refraction_color[c] * (m.glass_gain * under);
    g.nrc_albedo[c] = (float)pm_min(pm_max(a, T(0)), T(1));
  }
  g.sort_key = 1;
  for (int i = 0; i < 5; i++)
    if (g.w[i] > 0.f) g.sort_key += 1u << i;
  g.nrc_rough = 1.f;
  if (ws > T(0) || wg > T(0)) g.nrc_rough = (float)(wg > ws ? m.glass_roughness : m.specular_roughness);
  // Albedo AOV (pt/mod.rs:690-696): diffuse, the primary specular F0 when
  // that lobe is active, refraction; each channel capped at 1.
  const bool spec_on = pm_lum(f0) > T(1e-6);
  for (int c = 0; c < 3; c++) {
    const T a = m.diffuse_color[c] * m.diffuse_gain + (spec_on ? f0[c] : T(0)) + m.refraction_color[c] * m.glass_gain;
    g.film_albedo[c] = (float)pm_min(a, T(1));
  }
}

}  // namespace rrhip
