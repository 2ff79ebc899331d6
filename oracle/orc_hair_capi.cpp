// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// C entry points over the curve and hair restatement (orc_hair.hpp) for
// the tests' ctypes loader. Hair parameters travel as 6 doubles:
// sigma_a xyz, beta_m, beta_n, eta.
#include <cmath>

#include "orc_hair.hpp"

using orc::HairParams;
using orc::Vec3;

namespace {
HairParams params_of(const double* hp) { return HairParams::make({hp[0], hp[1], hp[2]}, hp[3], hp[4], hp[5]); }
Vec3 v3(const double* a) { return {a[0], a[1], a[2]}; }
void put(double* o, const Vec3& v) {
  o[0] = v.x;
  o[1] = v.y;
  o[2] = v.z;
}
}  // namespace

extern "C" {

// Rays (o xyz, d xyz per ray) against one segment; t = -1 on a miss.
void orc_curve_intersect(const float* p0, const float* p1, float v0, float v1, const double* rays, uint64_t n,
                         double t_max, double* t, double* nrm, double* tan, double* v) {
  orc::CurveSet cs;
  cs.build(p0, p1, &v0, &v1, 1);
  for (uint64_t i = 0; i < n; i++) {
    const orc::Ray r = orc::Ray::raw(v3(rays + 6 * i), v3(rays + 6 * i + 3));
    Vec3 nn, tg;
    double tt, vv;
    if (cs.intersect_segment(r, 0, t_max, &tt, &nn, &tg, &vv)) {
      t[i] = tt;
      put(nrm + 3 * i, nn);
      put(tan + 3 * i, tg);
      v[i] = vv;
    } else {
      t[i] = -1.0;
    }
  }
}

// Dices one curve; returns the segment count (at most `cap` written).
uint64_t orc_dice_curve(const double* ctrl, uint64_t n, const double* basis, uint64_t step, double width_root,
                        double width_tip, uint64_t segs, float* p0, float* p1, float* v0, float* v1, uint64_t cap) {
  std::vector<float> a, b, va, vb;
  orc::dice_curve(ctrl, (size_t)n, basis, (size_t)step, width_root, width_tip, (size_t)segs, &a, &b, &va, &vb);
  const uint64_t m = va.size();
  for (uint64_t i = 0; i < m && i < cap; i++) {
    for (int k = 0; k < 4; k++) {
      p0[4 * i + k] = a[4 * i + k];
      p1[4 * i + k] = b[4 * i + k];
    }
    v0[i] = va[i];
    v1[i] = vb[i];
  }
  return m;
}

// Longitudinal variances v[4] and the azimuthal scale s of HairParams::new.
void orc_hair_params(const double* hp, double* v, double* s) {
  const HairParams h = params_of(hp);
  for (int p = 0; p < 4; p++) v[p] = h.v[p];
  *s = h.s;
}

void orc_hair_sigma_a_from_color(const double* c, double beta_n, double* out) {
  put(out, HairParams::sigma_a_from_color(v3(c), beta_n));
}

// which: 0 i0, 1 log_i0, 2 logistic(x, a), 3 logistic_cdf(x, a),
// 4 trimmed_logistic(x, a, -pi, pi), 5 sample_trimmed_logistic(x, a, -pi, pi)
void orc_hair_scalar(int which, const double* x, uint64_t n, double a, double* out) {
  for (uint64_t i = 0; i < n; i++) {
    switch (which) {
      case 0: out[i] = orc::hair::i0(x[i]); break;
      case 1: out[i] = orc::hair::log_i0(x[i]); break;
      case 2: out[i] = orc::hair::logistic(x[i], a); break;
      case 3: out[i] = orc::hair::logistic_cdf(x[i], a); break;
      case 4: out[i] = orc::hair::trimmed_logistic(x[i], a, -orc::PI, orc::PI); break;
      default: out[i] = orc::hair::sample_trimmed_logistic(x[i], a, -orc::PI, orc::PI); break;
    }
  }
}

// Mp over rows (cos_ti, cos_to, sin_ti, sin_to, v).
void orc_hair_mp(const double* rows, uint64_t n, double* out) {
  for (uint64_t i = 0; i < n; i++) {
    const double* r = rows + 5 * i;
    out[i] = orc::hair::mp(r[0], r[1], r[2], r[3], r[4]);
  }
}

// Ap[4] (rgb each) for cos_to, eta, h and the body transmittance t.
void orc_hair_ap(double cos_to, double eta, double h, const double* t, double* out) {
  Vec3 a[4];
  orc::hair::ap(cos_to, eta, h, v3(t), a);
  for (int p = 0; p < 4; p++) put(out + 3 * p, a[p]);
}

double orc_hair_fresnel(double cos_i, double eta_i, double eta_t) { return orc::hair::fr_dielectric(cos_i, eta_i, eta_t); }

// f (rgb) and pdf of n (wo, wi, h) triples in the fiber frame.
void orc_hair_eval(const double* hp, const double* wo, const double* wi, const double* h, uint64_t n, double* f,
                   double* pdf) {
  const HairParams p = params_of(hp);
  for (uint64_t i = 0; i < n; i++) {
    put(f + 3 * i, orc::hair::f(p, v3(wo + 3 * i), v3(wi + 3 * i), h[i]));
    pdf[i] = orc::hair::pdf(p, v3(wo + 3 * i), v3(wi + 3 * i), h[i]);
  }
}

// hair::sample from explicit uniforms (4 per sample); ok[i] = 0 where it fails.
void orc_hair_sample(const double* hp, const double* wo, const double* h, const double* u, uint64_t n, double* wi,
                     double* f, double* pdf, uint8_t* ok) {
  const HairParams p = params_of(hp);
  for (uint64_t i = 0; i < n; i++) {
    Vec3 w, fv;
    double pv = 0.0;
    ok[i] = orc::hair::sample_u(p, v3(wo + 3 * i), h[i], u + 4 * i, &w, &fv, &pv) ? 1 : 0;
    put(wi + 3 * i, w);
    put(f + 3 * i, fv);
    pdf[i] = pv;
  }
}

// The fiber frame: x, y, z rows and h.
void orc_fiber_frame(const double* tangent, const double* n, const double* wo, double* xyz, double* h) {
  const orc::FiberFrame ff(v3(tangent), v3(n), v3(wo));
  put(xyz, ff.x);
  put(xyz + 3, ff.y);
  put(xyz + 6, ff.z);
  *h = ff.h;
}

}  // extern "C"
