// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// Curves and hair of render.rs (src/geometry/curves.rs,
// src/raytracer/pt/hair.rs) in f64; declarations in orc_hair.hpp.
#include "orc_hair.hpp"

#include <algorithm>
#include <cmath>

namespace orc {

// ---- curves.rs -------------------------------------------------------------

void CurveSet::build(const float* a, const float* b, const float* va, const float* vb, size_t n) {
  p0.resize(n);
  p1.resize(n);
  v0.assign(va, va + n);
  v1.assign(vb, vb + n);
  std::vector<Aabb> boxes(n);
  local_bounds = Aabb{};
  for (size_t i = 0; i < n; i++) {
    for (int k = 0; k < 4; k++) {
      p0[i][k] = a[4 * i + k];
      p1[i][k] = b[4 * i + k];
    }
    // curves.rs:29-58: both end spheres' boxes, from the f32 values in f64
    Aabb bb;
    for (const auto* e : {&p0[i], &p1[i]}) {
      const double r = (double)(*e)[3];
      const Vec3 c{(double)(*e)[0], (double)(*e)[1], (double)(*e)[2]};
      bb.grow_point(c - Vec3{r, r, r});
      bb.grow_point(c + Vec3{r, r, r});
    }
    boxes[i] = bb;
    local_bounds.grow(bb);
  }
  blas.build(boxes, BvhParams{});
}

bool curve_sphere_hit(const Ray& ray, const Vec3& center, double radius, double t_max, double* t, Vec3* n) {
  const Vec3 oc = ray.origin - center;
  const double a = ray.direction.dot(ray.direction);
  const double b = oc.dot(ray.direction);
  const double c = oc.dot(oc) - radius * radius;
  const double disc = b * b - a * c;
  if (disc < 0.0) return false;
  const double s = std::sqrt(disc);
  double r = (-b - s) / a;
  if (r <= 1e-9) r = (-b + s) / a;
  if (r <= 1e-9 || r >= t_max) return false;
  *t = r;
  *n = (ray.at(r) - center).normalize();
  return true;
}

bool rounded_cone_intersect(const Ray& ray, const Vec3& pa, double ra, const Vec3& pb, double rb, double t_max,
                            double* t, Vec3* n) {
  const Vec3 ba = pb - pa, oa = ray.origin - pa, d = ray.direction;
  const double rr = ra - rb;
  const double m0 = ba.dot(ba), m1 = ba.dot(oa), m2 = ba.dot(d), m3 = d.dot(oa), m5 = oa.dot(oa);
  const double dd = d.dot(d);
  if (m0 < 1e-24) return curve_sphere_hit(ray, pa, std::fmax(ra, rb), t_max, t, n);
  // the swept sphere's body: a quadratic in t, valid for 0 < y < d2 with
  // y = ba . (o + t d - pa)
  const double d2 = m0 - rr * rr;
  const double k2 = d2 * dd - m2 * m2;
  const double k1 = d2 * m3 - m1 * m2 + m2 * rr * ra;
  const double k0 = d2 * m5 - m1 * m1 + m1 * rr * ra * 2.0 - m0 * ra * ra;
  if (std::fabs(k2) > 1e-24) {
    const double h = k1 * k1 - k0 * k2;
    if (h >= 0.0) {
      const double r = (-k1 - std::sqrt(h)) / k2;
      const double y = m1 + r * m2;
      if (r > 1e-9 && r < t_max && y > 0.0 && y < d2) {
        *t = r;
        *n = ((oa + d * r) * d2 - ba * y).normalize();
        return true;
      }
    }
  }
  // caps: the start sphere owns y <= 0, the end sphere y >= d2; the end
  // cap is searched below the start cap's hit
  bool found = false;
  double tc;
  Vec3 nc;
  if (curve_sphere_hit(ray, pa, ra, t_max, &tc, &nc) && m1 + tc * m2 <= 0.0) {
    *t = tc;
    *n = nc;
    found = true;
  }
  const double lim = found ? *t : t_max;
  if (curve_sphere_hit(ray, pb, rb, lim, &tc, &nc) && m1 + tc * m2 >= d2) {
    *t = tc;
    *n = nc;
    found = true;
  }
  return found;
}

bool CurveSet::intersect_segment(const Ray& ray, uint32_t seg, double t_max, double* t, Vec3* n, Vec3* tangent,
                                 double* v) const {
  const auto& a = p0[seg];
  const auto& b = p1[seg];
  const Vec3 pa{(double)a[0], (double)a[1], (double)a[2]}, pb{(double)b[0], (double)b[1], (double)b[2]};
  if (!rounded_cone_intersect(ray, pa, (double)a[3], pb, (double)b[3], t_max, t, n)) return false;
  const Vec3 axis = pb - pa;
  const double len2 = axis.length_squared();
  const double along = len2 > 1e-24 ? std::clamp((ray.at(*t) - pa).dot(axis) / len2, 0.0, 1.0) : 0.5;
  *v = (double)v0[seg] + (double)(v1[seg] - v0[seg]) * along;
  *tangent = len2 > 1e-24 ? axis.normalize() : Vec3{0, 0, 1};
  return true;
}

void dice_curve(const double* ctrl, size_t n, const double* basis, size_t step, double width_root, double width_tip,
                size_t segs_per_span, std::vector<float>* p0, std::vector<float>* p1, std::vector<float>* v0,
                std::vector<float>* v1) {
  std::vector<std::array<double, 3>> pts;
  if (!basis) {
    for (size_t i = 0; i < n; i++) pts.push_back({ctrl[3 * i], ctrl[3 * i + 1], ctrl[3 * i + 2]});
  } else {
    if (n < 4) return;
    const size_t spans = (n - 4) / step + 1;
    for (size_t span = 0; span < spans; span++) {
      const double* g = ctrl + 3 * span * step;
      const size_t count = span + 1 == spans ? segs_per_span + 1 : segs_per_span;
      for (size_t i = 0; i < count; i++) {
        const double u = (double)i / (double)segs_per_span;
        const double uv[4] = {u * u * u, u * u, u, 1.0};
        double w[4];
        for (int c = 0; c < 4; c++)
          w[c] = uv[0] * basis[0 * 4 + c] + uv[1] * basis[1 * 4 + c] + uv[2] * basis[2 * 4 + c] +
                 uv[3] * basis[3 * 4 + c];
        std::array<double, 3> p;
        for (int ax = 0; ax < 3; ax++)
          p[ax] = w[0] * g[ax] + w[1] * g[3 + ax] + w[2] * g[6 + ax] + w[3] * g[9 + ax];
        pts.push_back(p);
      }
    }
  }
  if (pts.size() < 2) return;
  const double denom = (double)(pts.size() - 1);
  for (size_t i = 0; i + 1 < pts.size(); i++) {
    const double va = (double)i / denom, vb = (double)(i + 1) / denom;
    const double ra = width_root + (width_tip - width_root) * va;
    const double rb = width_root + (width_tip - width_root) * vb;
    for (int ax = 0; ax < 3; ax++) p0->push_back((float)pts[i][ax]);
    p0->push_back((float)(ra * 0.5));
    for (int ax = 0; ax < 3; ax++) p1->push_back((float)pts[i + 1][ax]);
    p1->push_back((float)(rb * 0.5));
    v0->push_back((float)va);
    v1->push_back((float)vb);
  }
}

// ---- hair.rs -----------------------------------------------------------------

namespace {
constexpr double kSqrtPiOver8 = 0.626657069;  // hair.rs:23
}

HairParams HairParams::make(const Vec3& sigma_a, double beta_m, double beta_n, double eta) {
  HairParams h;
  h.sigma_a = sigma_a;
  h.beta_m = std::clamp(beta_m, 0.01, 1.0);
  h.beta_n = std::clamp(beta_n, 0.01, 1.0);
  h.eta = eta;
  const double bm = h.beta_m, bn = h.beta_n;
  const double root = 0.726 * bm + 0.812 * bm * bm + 3.7 * std::pow(bm, 20);
  const double v0 = root * root;
  h.v[0] = v0;
  h.v[1] = 0.25 * v0;
  h.v[2] = 4.0 * v0;
  h.v[3] = 4.0 * v0;
  h.s = kSqrtPiOver8 * (0.265 * bn + 1.194 * bn * bn + 5.372 * std::pow(bn, 22));
  return h;
}

Vec3 HairParams::sigma_a_from_color(const Vec3& c, double beta_n) {
  const double b = beta_n;
  const double denom = 5.969 - 0.215 * b + 2.532 * b * b - 10.73 * std::pow(b, 3) + 5.574 * std::pow(b, 4) +
                       0.245 * std::pow(b, 5);
  const auto one = [&](double x) {
    const double l = std::log(std::clamp(x, 1e-4, 0.999)) / denom;
    return l * l;
  };
  return {one(c.x), one(c.y), one(c.z)};
}

namespace hair {

double i0(double x) {
  double sum = 0.0, x2i = 1.0, fact = 1.0, four_i = 1.0;
  for (int i = 0; i < 10; i++) {
    if (i > 1) fact *= (double)i;
    sum += x2i / (four_i * fact * fact);
    x2i *= x * x;
    four_i *= 4.0;
  }
  return sum;
}

double log_i0(double x) {
  if (x > 12.0) return x + 0.5 * (-std::log(2.0 * PI * x) + 1.0 / (8.0 * x));
  return std::log(i0(x));
}

double mp(double cos_ti, double cos_to, double sin_ti, double sin_to, double v) {
  const double a = cos_ti * cos_to / v, b = sin_ti * sin_to / v;
  if (v <= 0.1) return std::exp(log_i0(a) - b - 1.0 / v + 0.6931 + std::log(1.0 / (2.0 * v)));
  return std::exp(-b) * i0(a) / (std::sinh(1.0 / v) * 2.0 * v);
}

double fr_dielectric(double cos_i, double eta_i, double eta_t) {
  cos_i = std::clamp(cos_i, -1.0, 1.0);
  if (!(cos_i > 0.0)) {
    std::swap(eta_i, eta_t);
    cos_i = -cos_i;
  }
  const double sin_i = std::sqrt(std::fmax(1.0 - cos_i * cos_i, 0.0));
  const double sin_t = eta_i / eta_t * sin_i;
  if (sin_t >= 1.0) return 1.0;
  const double cos_t = std::sqrt(std::fmax(1.0 - sin_t * sin_t, 0.0));
  const double par = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t);
  const double perp = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t);
  return (par * par + perp * perp) / 2.0;
}

void ap(double cos_to, double eta, double h, const Vec3& t, Vec3 out[4]) {
  const double cos_gamma_o = std::sqrt(std::fmax(1.0 - h * h, 0.0));
  const double fr = fr_dielectric(cos_to * cos_gamma_o, 1.0, eta);
  out[0] = Vec3{fr, fr, fr};
  out[1] = t * ((1.0 - fr) * (1.0 - fr));
  out[2] = Vec3{out[1].x * t.x * fr, out[1].y * t.y * fr, out[1].z * t.z * fr};
  // the residual: the geometric series of the bounces after TRT
  const Vec3& l = out[2];
  out[3] = Vec3{l.x * fr * t.x / std::fmax(1.0 - t.x * fr, 1e-6), l.y * fr * t.y / std::fmax(1.0 - t.y * fr, 1e-6),
                l.z * fr * t.z / std::fmax(1.0 - t.z * fr, 1e-6)};
}

double logistic(double x, double s) {
  const double e = std::exp(-std::fabs(x) / s);
  return e / (s * (1.0 + e) * (1.0 + e));
}

double logistic_cdf(double x, double s) { return 1.0 / (1.0 + std::exp(-x / s)); }

double trimmed_logistic(double x, double s, double a, double b) {
  return logistic(x, s) / (logistic_cdf(b, s) - logistic_cdf(a, s));
}

double sample_trimmed_logistic(double u, double s, double a, double b) {
  const double k = logistic_cdf(b, s) - logistic_cdf(a, s);
  const double q = std::clamp(u * k + logistic_cdf(a, s), 1e-9, 1.0 - 1e-9);
  return std::clamp(-s * std::log(1.0 / q - 1.0), a, b);
}

double phi_fn(int p, double gamma_o, double gamma_t) {
  return 2.0 * (double)p * gamma_t - 2.0 * gamma_o + (double)p * PI;
}

double np(double phi, int p, double s, double gamma_o, double gamma_t) {
  double dphi = phi - phi_fn(p, gamma_o, gamma_t);
  while (dphi > PI) dphi -= 2.0 * PI;
  while (dphi < -PI) dphi += 2.0 * PI;
  return trimmed_logistic(dphi, s, -PI, PI);
}

namespace {

// hair.rs:207-240: the outgoing direction's angles, the refracted
// azimuth inside the fiber and the body transmittance of one crossing.
struct Geom {
  double sin_to, cos_to, phi_o, gamma_o, gamma_t;
  Vec3 t;
};

Geom geom(const HairParams& hp, const Vec3& wo, double h) {
  Geom g;
  g.sin_to = std::clamp(wo.x, -1.0, 1.0);
  g.cos_to = std::sqrt(std::fmax(1.0 - g.sin_to * g.sin_to, 0.0));
  g.phi_o = std::atan2(wo.z, wo.y);
  g.gamma_o = std::asin(std::clamp(h, -1.0, 1.0));
  const double sin_tt = g.sin_to / hp.eta;
  const double cos_tt = std::sqrt(std::fmax(1.0 - sin_tt * sin_tt, 0.0));
  const double etap = std::sqrt(std::fmax(hp.eta * hp.eta - g.sin_to * g.sin_to, 0.0)) / std::fmax(g.cos_to, 1e-9);
  const double sin_gt = std::clamp(h / etap, -1.0, 1.0);
  const double cos_gt = std::sqrt(std::fmax(1.0 - sin_gt * sin_gt, 0.0));
  g.gamma_t = std::asin(sin_gt);
  const double len = 2.0 * cos_gt / std::fmax(cos_tt, 1e-9);
  g.t = Vec3{std::exp(-hp.sigma_a.x * len), std::exp(-hp.sigma_a.y * len), std::exp(-hp.sigma_a.z * len)};
  return g;
}

// hair.rs:268-286: lobe selection weights, the luminance of each Ap
// normalised (all on R when they vanish).
void lobe_weights(const HairParams& hp, const Geom& g, double h, double w[4]) {
  Vec3 a[4];
  ap(g.cos_to, hp.eta, h, g.t, a);
  double total = 0.0;
  for (int p = 0; p < 4; p++) {
    w[p] = luminance(a[p]);
    total += w[p];
  }
  if (total > 1e-12) {
    for (int p = 0; p < 4; p++) w[p] /= total;
  } else {
    w[0] = 1.0;
  }
}

}  // namespace

Vec3 f(const HairParams& hp, const Vec3& wo, const Vec3& wi, double h) {
  const Geom g = geom(hp, wo, h);
  const double sin_ti = std::clamp(wi.x, -1.0, 1.0);
  const double cos_ti = std::sqrt(std::fmax(1.0 - sin_ti * sin_ti, 0.0));
  const double phi = std::atan2(wi.z, wi.y) - g.phi_o;
  Vec3 a[4];
  ap(g.cos_to, hp.eta, h, g.t, a);
  Vec3 sum = Vec3::zero();
  for (int p = 0; p < 4; p++) {
    const double m = mp(cos_ti, g.cos_to, sin_ti, g.sin_to, hp.v[p]);
    const double nn = p < 3 ? np(phi, p, hp.s, g.gamma_o, g.gamma_t) : 1.0 / (2.0 * PI);
    sum = sum + a[p] * (m * nn);
  }
  return sum;
}

double pdf(const HairParams& hp, const Vec3& wo, const Vec3& wi, double h) {
  const Geom g = geom(hp, wo, h);
  const double sin_ti = std::clamp(wi.x, -1.0, 1.0);
  const double cos_ti = std::sqrt(std::fmax(1.0 - sin_ti * sin_ti, 0.0));
  const double phi = std::atan2(wi.z, wi.y) - g.phi_o;
  double w[4];
  lobe_weights(hp, g, h, w);
  double sum = 0.0;
  for (int p = 0; p < 4; p++) {
    const double m = mp(cos_ti, g.cos_to, sin_ti, g.sin_to, hp.v[p]);
    const double nn = p < 3 ? np(phi, p, hp.s, g.gamma_o, g.gamma_t) : 1.0 / (2.0 * PI);
    sum += w[p] * m * nn;
  }
  return std::fmax(sum, 0.0);
}

bool sample_u(const HairParams& hp, const Vec3& wo, double h, const double u[4], Vec3* wi, Vec3* fv, double* pv) {
  const Geom g = geom(hp, wo, h);
  double w[4];
  lobe_weights(hp, g, h, w);
  // the lobe: the first whose weight exceeds what remains of u
  double ul = u[0];
  int p = 3;
  for (int i = 0; i < 4; i++) {
    if (ul < w[i]) {
      p = i;
      break;
    }
    ul -= w[i];
  }
  // longitudinal: d'Eon's inverse of Mp's cone, rotated about wo's angle
  const double v = hp.v[p];
  const double u1 = std::fmax(u[1], 1e-9);
  const double cos_theta = 1.0 + v * std::log(u1 + (1.0 - u1) * std::exp(-2.0 / v));
  const double sin_theta = std::sqrt(std::fmax(1.0 - cos_theta * cos_theta, 0.0));
  const double cos_phi_l = std::cos(2.0 * PI * u[2]);
  const double sin_ti = -cos_theta * g.sin_to + sin_theta * cos_phi_l * g.cos_to;
  const double cos_ti = std::sqrt(std::fmax(1.0 - sin_ti * sin_ti, 0.0));
  // azimuthal
  const double dphi = p < 3 ? phi_fn(p, g.gamma_o, g.gamma_t) + sample_trimmed_logistic(u[3], hp.s, -PI, PI)
                            : 2.0 * PI * u[3];
  const double phi_i = g.phi_o + dphi;
  const Vec3 d{sin_ti, cos_ti * std::cos(phi_i), cos_ti * std::sin(phi_i)};
  const double pd = pdf(hp, wo, d, h);
  if (pd <= 1e-12) return false;
  *wi = d;
  *pv = pd;
  *fv = f(hp, wo, d, h);
  return true;
}

bool sample(const HairParams& hp, const Vec3& wo, double h, Pcg32& rng, Vec3* wi, Vec3* fv, double* pv) {
  double u[4];
  for (double& x : u) x = rng.next_f64();
  return sample_u(hp, wo, h, u, wi, fv, pv);
}

}  // namespace hair

FiberFrame::FiberFrame(const Vec3& tangent, const Vec3& n, const Vec3& wo) {
  x = tangent.normalize();
  Vec3 zz = wo - x * wo.dot(x);
  if (zz.length_squared() < 1e-12) zz = n - x * n.dot(x);
  z = zz.normalize();
  y = z.cross(x).normalize();
  const Vec3 n_perp = (n - x * n.dot(x)).normalize();
  h = std::clamp(n_perp.dot(y), -1.0, 1.0);
}

}  // namespace orc
