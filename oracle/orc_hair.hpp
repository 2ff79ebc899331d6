// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// f64 restatement of render.rs's curve primitive and hair BSDF:
// - src/geometry/curves.rs: CurveSet (segments as rounded cones: a swept
//   sphere whose radius varies linearly from p0 to p1, spherical caps),
//   the rounded-cone and cap intersectors (curves.rs:72-189) and the
//   dicing of linear / cubic curves into segments (curves.rs:191-258);
// - src/raytracer/pt/hair.rs: the R / TT / TRT + residual fiber BSDF
//   (d'Eon et al. 2011 longitudinal lobes, trimmed-logistic azimuthal
//   lobes, Fresnel + absorption attenuation), its lobe-selection pdf and
//   sampling routine, and the fiber frame at a curve hit.
//
// Parity unpinned: HairParams::sigma_a_from_color (hair.rs:63-77, the
// published fit of Chiang et al. 2016) is checked only for its fixed
// point (white -> no absorption) and its monotonicity; nothing the
// reference holds fixes its coefficients.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "orc_accel.hpp"
#include "orc_math.hpp"
#include "orc_sampling.hpp"

namespace orc {

// curves.rs:15-26: f32 storage (xyz + radius per segment end, the curve
// parameter at each end), one BLAS over the segments' boxes.
struct CurveSet {
  std::vector<std::array<float, 4>> p0, p1;
  std::vector<float> v0, v1;
  Bvh blas;
  Aabb local_bounds;
  void build(const float* p0, const float* p1, const float* v0, const float* v1, size_t n);
  // curves.rs:72-107: closest rounded-cone hit of segment `seg` with
  // 1e-9 < t < t_max (t parametric along the unnormalised direction); the
  // outward normal, the unit tangent (0, 0, 1 for a point) and v.
  bool intersect_segment(const Ray& ray, uint32_t seg, double t_max, double* t, Vec3* n, Vec3* tangent,
                         double* v) const;
};

// curves.rs:112-169 (after Quilez): (t, outward normal) of the rounded cone
// pa (radius ra) -> pb (radius rb); the body accepts only the first root.
bool rounded_cone_intersect(const Ray& ray, const Vec3& pa, double ra, const Vec3& pb, double rb, double t_max,
                            double* t, Vec3* n);
// curves.rs:171-189: the cap sphere (second root when the first is <= 1e-9).
bool curve_sphere_hit(const Ray& ray, const Vec3& center, double radius, double t_max, double* t, Vec3* n);

// curves.rs:191-258: samples along the curve (the control points of a
// linear curve; for a cubic one, segs_per_span points per span of the
// 4x4 basis `basis` (row-major, p(u) = [u^3 u^2 u 1] M G) stepping `step`
// control points, plus the last span's end), widths lerped root -> tip by
// sample index, radii = width / 2, v = index / (samples - 1).
void dice_curve(const double* ctrl, size_t n, const double* basis, size_t step, double width_root, double width_tip,
                size_t segs_per_span, std::vector<float>* p0, std::vector<float>* p1, std::vector<float>* v0,
                std::vector<float>* v1);

// hair.rs:25-53
struct HairParams {
  Vec3 sigma_a;
  double beta_m = 0.3, beta_n = 0.3, eta = 1.55;
  double v[4] = {0, 0, 0, 0};  // longitudinal variance per lobe
  double s = 0;                // azimuthal logistic scale
  static HairParams make(const Vec3& sigma_a, double beta_m, double beta_n, double eta);
  // hair.rs:64-77 (pbrt's SigmaAFromReflectance)
  static Vec3 sigma_a_from_color(const Vec3& c, double beta_n);
};

namespace hair {
double i0(double x);                    // hair.rs:101-116 (10-term series)
double log_i0(double x);                // hair.rs:118-124
double mp(double cos_ti, double cos_to, double sin_ti, double sin_to, double v);  // hair.rs:127-135
double fr_dielectric(double cos_i, double eta_i, double eta_t);                   // hair.rs:137-153
void ap(double cos_to, double eta, double h, const Vec3& t, Vec3 out[4]);          // hair.rs:156-180
double logistic(double x, double s);
double logistic_cdf(double x, double s);
double trimmed_logistic(double x, double s, double a, double b);
double sample_trimmed_logistic(double u, double s, double a, double b);
double phi_fn(int p, double gamma_o, double gamma_t);
double np(double phi, int p, double s, double gamma_o, double gamma_t);
// hair.rs:242-266: per-solid-angle value (no cosine on either side) in the
// fiber frame (x = tangent), azimuthal offset h.
Vec3 f(const HairParams& hp, const Vec3& wo, const Vec3& wi, double h);
// hair.rs:268-309
double pdf(const HairParams& hp, const Vec3& wo, const Vec3& wi, double h);
// hair.rs:311-363 from four uniforms in the order the routine draws them
// (lobe, two longitudinal, azimuthal); false where the pdf is <= 1e-12.
bool sample_u(const HairParams& hp, const Vec3& wo, double h, const double u[4], Vec3* wi, Vec3* fv, double* pv);
bool sample(const HairParams& hp, const Vec3& wo, double h, Pcg32& rng, Vec3* wi, Vec3* fv, double* pv);
}  // namespace hair

// hair.rs:365-401: x = tangent, z = the viewer direction off the fiber, y =
// z x x; h = the normal's component along y.
struct FiberFrame {
  Vec3 x, y, z;
  double h = 0;
  FiberFrame(const Vec3& tangent, const Vec3& n, const Vec3& wo);
  Vec3 to_local(const Vec3& w) const { return {w.dot(x), w.dot(y), w.dot(z)}; }
  Vec3 to_world(const Vec3& w) const { return x * w.x + y * w.y + z * w.z; }
};

}  // namespace orc
