// ORACLE — test infrastructure only (see orc_math.hpp header).
#include "orc_pt.hpp"

#include <algorithm>
#include <cmath>

namespace orc {

namespace {

double lum(const Vec3& c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }

Vec3 v3(const double a[3]) { return {a[0], a[1], a[2]}; }

double clampd(double x, double lo, double hi) { return std::fmin(std::fmax(x, lo), hi); }

double powi5(double x) { return x * x * x * x * x; }

double powi4(double x) { return x * x * x * x; }

// bxdf.rs:57-60
Vec3 schlick(const Vec3& f0, const Vec3& f90, double cos) {
  double m = powi5(clampd(1.0 - cos, 0.0, 1.0));
  return f0 + (f90 - f0) * m;
}

// bxdf.rs:63-73
double fresnel_dielectric(double cos_i, double eta) {
  cos_i = clampd(cos_i, 0.0, 1.0);
  double sin2_t = (1.0 - cos_i * cos_i) / (eta * eta);
  if (sin2_t >= 1.0) return 1.0;
  double cos_t = std::sqrt(1.0 - sin2_t);
  double r_par = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  double r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  return 0.5 * (r_par * r_par + r_perp * r_perp);
}

// bxdf.rs:77-101
double ggx_d(const Vec3& h, double alpha) {
  if (h.z <= 0.0) return 0.0;
  double a2 = alpha * alpha;
  double d = h.z * h.z * (a2 - 1.0) + 1.0;
  return a2 / (PI * d * d);
}

double ggx_lambda(const Vec3& w, double alpha) {
  double cos2 = w.z * w.z;
  if (cos2 <= 0.0) return 0.0;
  double tan2 = std::fmax(1.0 - cos2, 0.0) / cos2;
  return (std::sqrt(1.0 + alpha * alpha * tan2) - 1.0) * 0.5;
}

double ggx_g1(const Vec3& w, double alpha) { return 1.0 / (1.0 + ggx_lambda(w, alpha)); }

double ggx_g2(const Vec3& wo, const Vec3& wi, double alpha) {
  return 1.0 / (1.0 + ggx_lambda(wo, alpha) + ggx_lambda(wi, alpha));
}

// bxdf.rs:104-121 (Heitz 2018 VNDF)
Vec3 ggx_sample_vndf(const Vec3& wo, double alpha, double u1, double u2) {
  Vec3 v = Vec3{alpha * wo.x, alpha * wo.y, wo.z}.normalize();
  double lensq = v.x * v.x + v.y * v.y;
  Vec3 t1 = lensq > 1e-12 ? Vec3{-v.y, v.x, 0.0} / std::sqrt(lensq) : Vec3{1.0, 0.0, 0.0};
  Vec3 t2 = v.cross(t1);
  double r = std::sqrt(u1);
  double phi = 2.0 * PI * u2;
  double p1 = r * std::cos(phi);
  double p2 = r * std::sin(phi);
  double s = 0.5 * (1.0 + v.z);
  p2 = (1.0 - s) * std::sqrt(std::fmax(1.0 - p1 * p1, 0.0)) + s * p2;
  Vec3 nh = t1 * p1 + t2 * p2 + v * std::sqrt(std::fmax(1.0 - p1 * p1 - p2 * p2, 0.0));
  return Vec3{alpha * nh.x, alpha * nh.y, std::fmax(nh.z, 1e-9)}.normalize();
}

// bxdf.rs:124-130
double ggx_pdf_h(const Vec3& wo, const Vec3& h, double alpha) {
  double odh = wo.dot(h);
  if (odh <= 0.0 || wo.z <= 0.0) return 0.0;
  return ggx_g1(wo, alpha) * odh * ggx_d(h, alpha) / wo.z;
}

// bxdf.rs:137-143
double under_layer_scale(const PbrParams& p) {
  Vec3 f0 = p.specular_f0();
  double spec_take = 0.2126 * f0.x + 0.7152 * f0.y + 0.0722 * f0.z;
  double coat_take = 0.04 * p.clearcoat_gain;
  return clampd((1.0 - spec_take) * (1.0 - coat_take), 0.0, 1.0);
}

// bxdf.rs:145-165 (Oren-Nayar)
Vec3 eval_diffuse(const PbrParams& p, const Vec3& wo, const Vec3& wi) {
  double sigma2 = p.diffuse_roughness * p.diffuse_roughness;
  double a = 1.0 - sigma2 / (2.0 * (sigma2 + 0.33));
  double b = 0.45 * sigma2 / (sigma2 + 0.09);
  double sin_o = std::sqrt(std::fmax(1.0 - wo.z * wo.z, 0.0));
  double sin_i = std::sqrt(std::fmax(1.0 - wi.z * wi.z, 0.0));
  double cos_dphi = (sin_o > 1e-6 && sin_i > 1e-6)
                        ? clampd((wo.x * wi.x + wo.y * wi.y) / (sin_o * sin_i), -1.0, 1.0)
                        : 0.0;
  double sin_a, tan_b;
  if (wo.z < wi.z) {
    sin_a = sin_o;
    tan_b = sin_i / std::fmax(wi.z, 1e-6);
  } else {
    sin_a = sin_i;
    tan_b = sin_o / std::fmax(wo.z, 1e-6);
  }
  return p.diffuse_color * (p.diffuse_gain * under_layer_scale(p) / PI) *
         (a + b * std::fmax(cos_dphi, 0.0) * sin_a * tan_b);
}

// bxdf.rs:167-171
Vec3 eval_fuzz(const PbrParams& p, const Vec3& wo, const Vec3& wi) {
  double rim = 0.5 * (powi4(1.0 - wi.z) + powi4(1.0 - wo.z));
  return p.fuzz_color * (p.fuzz_gain / PI) * rim;
}

// bxdf.rs:173-179
Vec3 eval_specular_lobe(const Vec3& f0, const Vec3& f90, double alpha, const Vec3& wo,
                        const Vec3& wi) {
  Vec3 h = (wo + wi).normalize();
  double d = ggx_d(h, alpha);
  double g = ggx_g2(wo, wi, alpha);
  Vec3 f = schlick(f0, f90, std::fmax(wo.dot(h), 0.0));
  return f * (d * g / std::fmax(4.0 * wo.z * wi.z, 1e-9));
}

// bxdf.rs:181-185
double spec_pdf(const Vec3& wo, const Vec3& wi, double alpha) {
  Vec3 h = (wo + wi).normalize();
  double odh = std::fmax(wo.dot(h), 1e-9);
  return ggx_pdf_h(wo, h, alpha) / (4.0 * odh);
}

// bxdf.rs:187-189
double alpha_of(double roughness) { return clampd(roughness * roughness, 2.5e-5, 1.0); }

// bxdf.rs:193-232
void glass_transmit(const PbrParams& p, const Vec3& wo, const Vec3& wi, double eta, Vec3* f_out,
                    double* pdf_out) {
  *f_out = Vec3::zero();
  *pdf_out = 0.0;
  double alpha = alpha_of(p.glass_roughness);
  Vec3 h = (wo + wi * eta).normalize();
  if (h.z < 0.0) h = -h;
  double odh = wo.dot(h), idh = wi.dot(h);
  if (odh <= 0.0 || idh >= 0.0) return;
  double fres = fresnel_dielectric(odh, eta);
  double d = ggx_d(h, alpha);
  double g = ggx_g2(wo, Vec3{wi.x, wi.y, -wi.z}, alpha);
  double sqrt_denom = odh + eta * idh;
  if (std::fabs(sqrt_denom) < 1e-9) return;
  double factor = 1.0 / (eta * eta);
  *f_out = p.refraction_color * (p.glass_gain * under_layer_scale(p) * (1.0 - fres) * d * g * factor *
                                 std::fabs(idh * odh / (wi.z * wo.z)) * (eta * eta) /
                                 (sqrt_denom * sqrt_denom));
  double dwh_dwi = std::fabs(eta * eta * idh) / (sqrt_denom * sqrt_denom);
  *pdf_out = ggx_pdf_h(wo, h, alpha) * dwh_dwi;
}

// bxdf.rs:236-238
double glass_reflect_prob(const Vec3& wo, const Vec3& h, double eta) {
  return fresnel_dielectric(std::fmax(wo.dot(h), 0.0), eta);
}

// bxdf.rs:301-306
Vec3 cosine_sample_local(Pcg32& rng) {
  double u = rng.next_f64(), v = rng.next_f64();
  double r = std::sqrt(u);
  double phi = 2.0 * PI * v;
  return {r * std::cos(phi), r * std::sin(phi), std::sqrt(std::fmax(1.0 - u, 0.0))};
}

double power_heuristic(double a, double b) {  // pt/mod.rs:992-996
  double a2 = a * a, b2 = b * b;
  return (a2 + b2 <= 0.0) ? 0.0 : a2 / (a2 + b2);
}

}  // namespace

// pbr.rs:84-93
Vec3 PbrParams::specular_f0() const {
  if (lum(specular_face_color) > 1e-6) return specular_face_color;
  if (specular_ior > 1.0) {
    double r = (specular_ior - 1.0) / (specular_ior + 1.0);
    return Vec3::one() * (r * r);
  }
  return Vec3::zero();
}

bool PbrParams::has_specular() const { return lum(specular_f0()) > 1e-6; }

// pbr.rs:191-202
bool PbrParams::lobe_weights(double w[5]) const {
  double wd = diffuse_gain * lum(diffuse_color);
  double ws = std::sqrt(lum(specular_f0()));
  double wc = clearcoat_gain * 0.2;
  double wf = fuzz_gain * lum(fuzz_color) * 0.3;
  double wg = glass_gain;
  double total = wd + ws + wc + wf + wg;
  if (total <= 1e-9) return false;
  w[0] = wd / total;
  w[1] = ws / total;
  w[2] = wc / total;
  w[3] = wf / total;
  w[4] = wg / total;
  return true;
}

// bxdf.rs:23-32
Frame::Frame(const Vec3& n_) : n(n_) {
  Vec3 t0 = std::fabs(n.x) > 0.9 ? Vec3{0.0, 1.0, 0.0} : Vec3{1.0, 0.0, 0.0};
  t = n.cross(t0).normalize();
  b = n.cross(t);
}

// bxdf.rs:241-299
void bsdf_eval_pdf(const PbrParams& p, const Vec3& wo, const Vec3& wi, double eta_rel, Vec3* f_out,
                   double* pdf_out) {
  double w[5];
  *f_out = Vec3::zero();
  *pdf_out = 0.0;
  if (!p.lobe_weights(w)) return;
  const double wd = w[0], ws = w[1], wc = w[2], wf = w[3], wg = w[4];
  if (wi.z > 0.0) {
    Vec3 f = Vec3::zero();
    double pdf = 0.0;
    double cos_pdf = wi.z / PI;
    if (wd > 0.0) {
      f = f + eval_diffuse(p, wo, wi);
      pdf += wd * cos_pdf;
    }
    if (wf > 0.0) {
      f = f + eval_fuzz(p, wo, wi);
      pdf += wf * cos_pdf;
    }
    if (ws > 0.0) {
      double alpha = alpha_of(p.specular_roughness);
      f = f + eval_specular_lobe(p.specular_f0(), p.specular_edge_color, alpha, wo, wi);
      pdf += ws * spec_pdf(wo, wi, alpha);
    }
    if (wc > 0.0) {
      double alpha = alpha_of(p.clearcoat_roughness);
      Vec3 f0 = Vec3{0.04, 0.04, 0.04} * p.clearcoat_gain;
      f = f + eval_specular_lobe(f0, Vec3::one() * p.clearcoat_gain, alpha, wo, wi);
      pdf += wc * spec_pdf(wo, wi, alpha);
    }
    if (wg > 0.0) {
      double alpha = alpha_of(p.glass_roughness);
      Vec3 h = (wo + wi).normalize();
      double fres = glass_reflect_prob(wo, h, eta_rel);
      double d = ggx_d(h, alpha);
      double g = ggx_g2(wo, wi, alpha);
      f = f + Vec3::one() * (p.glass_gain * under_layer_scale(p) * fres * d * g /
                             std::fmax(4.0 * wo.z * wi.z, 1e-9));
      pdf += wg * fres * spec_pdf(wo, wi, alpha);
    }
    *f_out = f;
    *pdf_out = pdf;
  } else {
    if (wg <= 0.0) return;
    Vec3 f;
    double pdf_lobe;
    glass_transmit(p, wo, wi, eta_rel, &f, &pdf_lobe);
    Vec3 h = (wo + wi * eta_rel).normalize();
    if (h.z < 0.0) h = -h;
    double fres = glass_reflect_prob(wo, h, eta_rel);
    *f_out = f;
    *pdf_out = wg * (1.0 - fres) * pdf_lobe;
  }
}

// bxdf.rs:309-374
bool bsdf_sample(const PbrParams& p, const Vec3& wo, double eta_rel, Pcg32& rng, BsdfSample* s) {
  double w[5];
  if (!p.lobe_weights(w)) return false;
  const double wd = w[0], ws = w[1], wc = w[2], wf = w[3], wg = w[4];
  double pick = rng.next_f64();
  bool transmitted = false;
  bool specular_lobe = pick >= wd + wf;
  Vec3 wi;
  if (pick < wd + wf) {
    wi = cosine_sample_local(rng);
  } else if (pick < wd + wf + ws) {
    double alpha = alpha_of(p.specular_roughness);
    double u1 = rng.next_f64(), u2 = rng.next_f64();
    Vec3 h = ggx_sample_vndf(wo, alpha, u1, u2);
    wi = (-wo).reflect(h);
    if (wi.z <= 0.0) return false;
  } else if (pick < wd + wf + ws + wc) {
    double alpha = alpha_of(p.clearcoat_roughness);
    double u1 = rng.next_f64(), u2 = rng.next_f64();
    Vec3 h = ggx_sample_vndf(wo, alpha, u1, u2);
    wi = (-wo).reflect(h);
    if (wi.z <= 0.0) return false;
  } else if (wg > 0.0) {
    double alpha = alpha_of(p.glass_roughness);
    double u1 = rng.next_f64(), u2 = rng.next_f64();
    Vec3 h = ggx_sample_vndf(wo, alpha, u1, u2);
    double fres = glass_reflect_prob(wo, h, eta_rel);
    if (rng.next_f64() < fres) {
      wi = (-wo).reflect(h);
      if (wi.z <= 0.0) return false;
    } else {
      double cos_i = wo.dot(h);
      double sin2_t = (1.0 - cos_i * cos_i) / (eta_rel * eta_rel);
      if (sin2_t >= 1.0) return false;
      double cos_t = std::sqrt(1.0 - sin2_t);
      wi = ((-wo) / eta_rel + h * (cos_i / eta_rel - cos_t)).normalize();
      if (wi.z >= 0.0) return false;
      transmitted = true;
    }
  } else {
    return false;
  }
  Vec3 f;
  double pdf;
  bsdf_eval_pdf(p, wo, wi, eta_rel, &f, &pdf);
  if (pdf <= 0.0) return false;
  s->wi = wi;
  s->f = f;
  s->pdf = pdf;
  s->transmitted = transmitted;
  s->specular_lobe = specular_lobe;
  return true;
}

// ---- light BVH (light_sampler.rs) -------------------------------------------

namespace {

double lum_clamped(const Vec3& c) { return std::fmax(0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z, 1e-9); }

// light_sampler.rs:51-70
bool finite_light_info(const Light& l, Vec3* center, double* radius, double* power) {
  double pw = lum_clamped(l.radiance());
  switch (l.kind) {
    case RRHIP_LIGHT_POINT:
      *center = l.position;
      *radius = 0.0;
      *power = pw * 4.0 * PI;
      return true;
    case RRHIP_LIGHT_RECT:
      *center = l.corner + l.edge1 * 0.5 + l.edge2 * 0.5;
      *radius = (l.edge1.length() + l.edge2.length()) * 0.5;
      *power = pw * l.area * PI;
      return true;
    case RRHIP_LIGHT_SPHERE:
      *center = l.center;
      *radius = l.radius;
      *power = pw * 4.0 * PI * l.radius * l.radius * PI;
      return true;
    case RRHIP_LIGHT_DISK:
      *center = l.center;
      *radius = l.e1.length();
      *power = pw * l.area * PI;
      return true;
    default:
      return false;
  }
}

// light_sampler.rs:74-87
double node_importance(const LightBvhNode& n, const Vec3& p) {
  double cx = ((double)n.min[0] + (double)n.max[0]) * 0.5;
  double cy = ((double)n.min[1] + (double)n.max[1]) * 0.5;
  double cz = ((double)n.min[2] + (double)n.max[2]) * 0.5;
  double dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
  double d2 = dx * dx + dy * dy + dz * dz;
  double rx = (double)(n.max[0] - n.min[0]) * 0.5;
  double ry = (double)(n.max[1] - n.min[1]) * 0.5;
  double rz = (double)(n.max[2] - n.min[2]) * 0.5;
  double r2 = rx * rx + ry * ry + rz * rz;
  return (double)n.power / std::fmax(std::fmax(d2, r2), 1e-6);
}

struct BuildEntry {
  uint32_t light;
  Vec3 center;
  double radius, power;
};

// light_sampler.rs:242-312 (median split on the widest axis)
uint32_t build_recursive(const std::vector<BuildEntry>& e, std::vector<size_t>& order, size_t start,
                         size_t end, uint32_t parent, std::vector<LightBvhNode>& nodes,
                         std::vector<uint32_t>& light_leaf) {
  uint32_t index = (uint32_t)nodes.size();
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  double power = 0.0;
  for (size_t k = start; k < end; k++) {
    const BuildEntry& b = e[order[k]];
    double c[3] = {b.center.x, b.center.y, b.center.z};
    for (int a = 0; a < 3; a++) {
      mn[a] = std::fmin(mn[a], (float)(c[a] - b.radius));
      mx[a] = std::fmax(mx[a], (float)(c[a] + b.radius));
    }
    power += b.power;
  }
  LightBvhNode node{};
  for (int a = 0; a < 3; a++) {
    node.min[a] = mn[a];
    node.max[a] = mx[a];
  }
  node.power = (float)power;
  node.parent = parent;
  node.a = 0;
  node.b = 0xFFFFFFFFu;
  nodes.push_back(node);
  if (end - start == 1) {
    uint32_t light = e[order[start]].light;
    nodes[index].a = light;
    light_leaf[light] = index;
    return index;
  }
  int axis = 0;
  float widest = -1.0f;
  for (int a = 0; a < 3; a++) {
    float w = mx[a] - mn[a];
    if (w > widest) {
      widest = w;
      axis = a;
    }
  }
  size_t mid = (start + end) / 2;
  std::stable_sort(order.begin() + start, order.begin() + end, [&](size_t x, size_t y) {
    return e[x].center[axis] < e[y].center[axis];
  });
  uint32_t a = build_recursive(e, order, start, mid, index, nodes, light_leaf);
  uint32_t b = build_recursive(e, order, mid, end, index, nodes, light_leaf);
  nodes[index].a = a;
  nodes[index].b = b;
  return index;
}

}  // namespace

// ---- EnvMap (envmap.rs) -----------------------------------------------------

static double texel_lum(const float* c) { return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]; }

void EnvMap::from_data(uint32_t w, uint32_t h, const float* rgb) {  // envmap.rs:36-54
  width = w;
  height = h;
  data.assign(rgb, rgb + 3ull * w * h);
  marginal.assign(h + 1, 0.0);
  conditional.assign((size_t)(w + 1) * h, 0.0);
  for (uint32_t y = 0; y < h; y++) {
    const double sin_theta = std::sin(PI * (y + 0.5) / h);
    const size_t base = (size_t)y * (w + 1);
    double acc = 0.0;
    for (uint32_t x = 0; x < w; x++) {
      acc += texel_lum(&data[3 * ((size_t)y * w + x)]) * sin_theta;
      conditional[base + x + 1] = acc;
    }
    marginal[y + 1] = marginal[y] + acc;
  }
  total = std::fmax(marginal[h], 1e-12);
}

double EnvMap::mean_luminance() const {  // envmap.rs:61-66
  const double weight_sum = (double)width * height * (2.0 / PI);
  return total / std::fmax(weight_sum, 1e-12);
}

void EnvMap::texel_of(const Vec3& dir, uint32_t* x, uint32_t* y) const {  // envmap.rs:84-91
  const Vec3 d = dir.normalize();
  double u = std::atan2(d.x, -d.z) / (2.0 * PI);
  u -= std::floor(u);
  const double v = std::fmin(std::fmax(std::acos(std::fmin(std::fmax(d.y, -1.0), 1.0)) / PI, 0.0), 1.0);
  *x = std::min((uint32_t)(u * width), width - 1);
  *y = std::min((uint32_t)(v * height), height - 1);
}

Vec3 EnvMap::texel(uint32_t x, uint32_t y) const {
  const float* c = &data[3 * ((size_t)std::min(y, height - 1) * width + std::min(x, width - 1))];
  return Vec3{c[0], c[1], c[2]};
}

Vec3 EnvMap::eval(const Vec3& dir) const {
  uint32_t x, y;
  texel_of(dir, &x, &y);
  return texel(x, y);
}

double EnvMap::pdf(const Vec3& dir) const {  // envmap.rs:134-148
  uint32_t x, y;
  texel_of(dir, &x, &y);
  const double sin_theta = std::sin(PI * (y + 0.5) / height);
  if (sin_theta < 1e-9) return 0.0;
  const double texel_weight = texel_lum(&data[3 * ((size_t)y * width + x)]) * sin_theta;
  return texel_weight * width * height / (total * 2.0 * PI * PI * sin_theta);
}

// The index the reference's binary search lands on: the last entry <= t.
static uint32_t last_not_above(const double* c, size_t n, double t) {
  const size_t i = (size_t)(std::upper_bound(c, c + n, t) - c);
  return (uint32_t)(i == 0 ? 0 : i - 1);
}

void EnvMap::sample(double u1, double u2, Vec3* dir, Vec3* radiance, double* pdf_out) const {  // envmap.rs:99-131
  const uint32_t y = std::min(last_not_above(marginal.data(), height + 1, u1 * total), height - 1);
  const double row_total = std::fmax(marginal[y + 1] - marginal[y], 1e-12);
  const double* cond = &conditional[(size_t)y * (width + 1)];
  const uint32_t x = std::min(last_not_above(cond, width + 1, u2 * row_total), width - 1);
  const double u = (x + 0.5) / width, v = (y + 0.5) / height;
  const double phi = u * 2.0 * PI, theta = v * PI;
  const double sin_t = std::sin(theta);
  *dir = Vec3{sin_t * std::sin(phi), std::cos(theta), -sin_t * std::cos(phi)};
  *radiance = texel(x, y);
  *pdf_out = pdf(*dir);
}

// light_sampler.rs:97-142
void LightSampler::build(const std::vector<Light>& lights) {
  nodes.clear();
  infinite.clear();
  infinite_power.clear();
  std::vector<BuildEntry> finite;
  for (size_t i = 0; i < lights.size(); i++) {
    Vec3 c;
    double r, pw;
    if (finite_light_info(lights[i], &c, &r, &pw)) {
      finite.push_back({(uint32_t)i, c, r, pw});
    } else {
      infinite.push_back((uint32_t)i);
      // a dome with a map: its mean luminance times the intensity
      // (light_sampler.rs:111-118)
      const double base = lights[i].env ? std::fmax(lights[i].env->mean_luminance() * lights[i].intensity, 1e-9)
                                        : lum_clamped(lights[i].radiance());
      infinite_power.push_back(base * 4.0 * PI);
    }
  }
  infinite_total = 0;
  for (double w : infinite_power) infinite_total += w;
  light_leaf.assign(lights.size(), 0xFFFFFFFFu);
  if (!finite.empty()) {
    std::vector<size_t> order(finite.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = k;
    build_recursive(finite, order, 0, finite.size(), 0xFFFFFFFFu, nodes, light_leaf);
  }
  double finite_total = 0;
  for (auto& f : finite) finite_total += f.power;
  if (infinite.empty())
    p_infinite = 0.0;
  else if (finite.empty())
    p_infinite = 1.0;
  else
    p_infinite = clampd(infinite_total / (infinite_total + finite_total), 0.1, 0.9);
}

// light_sampler.rs:150-208
bool LightSampler::sample(const Vec3& p, double u, size_t* light, double* pmf_out) const {
  bool has_finite = !nodes.empty(), has_infinite = !infinite.empty();
  if (!has_finite && !has_infinite) return false;
  double pmf = 1.0;
  bool inf;
  if (has_finite && has_infinite) {
    if (u < p_infinite) {
      u /= p_infinite;
      pmf *= p_infinite;
      inf = true;
    } else {
      u = (u - p_infinite) / (1.0 - p_infinite);
      pmf *= 1.0 - p_infinite;
      inf = false;
    }
  } else {
    inf = has_infinite;
  }
  if (inf) {
    double acc = 0.0, target = u * infinite_total;
    for (size_t k = 0; k < infinite_power.size(); k++) {
      acc += infinite_power[k];
      if (target <= acc || k == infinite.size() - 1) {
        *light = infinite[k];
        *pmf_out = pmf * infinite_power[k] / std::fmax(infinite_total, 1e-12);
        return true;
      }
    }
    return false;
  }
  size_t node = 0;
  for (;;) {
    const LightBvhNode& nd = nodes[node];
    if (nd.b == 0xFFFFFFFFu) {
      *light = nd.a;
      *pmf_out = pmf;
      return true;
    }
    double ia = node_importance(nodes[nd.a], p), ib = node_importance(nodes[nd.b], p);
    double total = ia + ib;
    double pa = total > 1e-30 ? ia / total : 0.5;
    if (u < pa) {
      u = std::fmin(u / pa, 1.0 - 1e-12);
      pmf *= pa;
      node = nd.a;
    } else {
      u = std::fmin((u - pa) / (1.0 - pa), 1.0 - 1e-12);
      pmf *= 1.0 - pa;
      node = nd.b;
    }
  }
}

// light_sampler.rs:212-238
double LightSampler::pmf(const Vec3& p, size_t light) const {
  bool has_finite = !nodes.empty(), has_infinite = !infinite.empty();
  for (size_t k = 0; k < infinite.size(); k++) {
    if (infinite[k] == light) {
      double group = has_finite ? p_infinite : 1.0;
      return group * infinite_power[k] / std::fmax(infinite_total, 1e-12);
    }
  }
  uint32_t leaf = light_leaf[light];
  if (leaf == 0xFFFFFFFFu) return 0.0;
  double pmf = has_infinite ? 1.0 - p_infinite : 1.0;
  uint32_t node = leaf;
  while (nodes[node].parent != 0xFFFFFFFFu) {
    uint32_t parent = nodes[node].parent;
    const LightBvhNode& pn = nodes[parent];
    double ia = node_importance(nodes[pn.a], p), ib = node_importance(nodes[pn.b], p);
    double total = ia + ib;
    double mine = node == pn.a ? ia : ib;
    pmf *= total > 1e-30 ? mine / total : 0.5;
    node = parent;
  }
  return pmf;
}

// ---- camera (camera.rs) -------------------------------------------------------

// camera.rs:36-61
double Camera::filter_1d(double u) const {
  switch (filter) {
    case 1: {
      double half = filter_width * 0.5;
      return u < 0.5 ? (std::sqrt(2.0 * u) - 1.0) * half : (1.0 - std::sqrt(2.0 * (1.0 - u))) * half;
    }
    case 2: {
      double u1 = std::fmax(u, 1e-12);
      double x = u * 2654435761.0;
      double u2 = x - std::floor(x);
      double sigma = filter_width / 4.0;
      double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
      return clampd(g * sigma, -filter_width * 0.5, filter_width * 0.5);
    }
    default:
      return (u - 0.5) * filter_width;
  }
}

// camera.rs:150-179 (pinhole)
// camera.rs:134-146
Ray Camera::apply_motion(const Ray& r, double t) const {
  if (!has_motion_inv || t <= 0.0) return r;
  Mat4 m;
  const Mat4 id = Mat4::identity();
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) m.m[i][j] = id.m[i][j] + (motion_inv.m[i][j] - id.m[i][j]) * t;
  Ray out = Ray::make(m.transform_point(r.origin), m.transform_vec(r.direction));
  out.time = r.time;
  return out;
}

// camera.rs:150-179
Ray Camera::generate_ray_lens(double px, double py, double lu, double lv) const {
  double u = (px / (double)width) * 2.0 - 1.0;
  double v = 1.0 - (py / (double)height) * 2.0;
  if (projection == 1) return Ray::make(eye + right * (u * ortho_half[0]) + up * (v * ortho_half[1]), forward);
  double aspect = (double)width / (double)height;
  double half_height = std::tan(fov * PI / 180.0 / 2.0);
  double half_width = aspect * half_height;
  Vec3 dir = forward + right * (u * half_width) + up * (v * half_height);
  if (lens_radius <= 0.0) return Ray::make(eye, dir);
  const double r = lens_radius * std::sqrt(lu), phi = 2.0 * PI * lv;
  const Vec3 origin = eye + right * (r * std::cos(phi)) + up * (r * std::sin(phi));
  const Vec3 focus = eye + dir * focal_distance;
  return Ray::make(origin, focus - origin);
}

// ---- geometry -----------------------------------------------------------------

// extra.rs:673-720 with an identity transform (world-space vertices)
bool TriangleObj::intersect(const Ray& ray, Intersection* out) const {
  Vec3 lo = ray.origin, ld = ray.direction.normalize();
  Vec3 e1 = v1 - v0, e2 = v2 - v0;
  Vec3 pvec = ld.cross(e2);
  double det = e1.dot(pvec);
  if (std::fabs(det) < 1e-12) return false;
  double inv_det = 1.0 / det;
  Vec3 tvec = lo - v0;
  double u = tvec.dot(pvec) * inv_det;
  if (!(u >= 0.0 && u <= 1.0)) return false;
  Vec3 qvec = tvec.cross(e1);
  double v = ld.dot(qvec) * inv_det;
  if (v < 0.0 || u + v > 1.0) return false;
  double t = e2.dot(qvec) * inv_det;
  if (t <= EPSILON) return false;
  Vec3 p = lo + ld * t;
  Vec3 n = e1.cross(e2).normalize();
  bool front = n.dot(ld) < 0.0;
  if (!front) n = -n;
  out->t = (p - ray.origin).length();
  out->point = p;
  out->normal = n.normalize();
  out->front_face = front;
  out->material = material;
  return true;
}

// sphere.rs:38-109 (full sphere, translation transform)
bool SphereObj::intersect(const Ray& ray, Intersection* out) const {
  Vec3 lo = ray.origin - center;
  Vec3 ld = ray.direction.normalize();
  double a = ld.dot(ld);
  double b = 2.0 * lo.dot(ld);
  double c = lo.dot(lo) - radius * radius;
  double disc = b * b - 4.0 * a * c;
  if (disc < 0.0) return false;
  double sq = std::sqrt(disc);
  double t1 = (-b - sq) / (2.0 * a), t2 = (-b + sq) / (2.0 * a);
  double t;
  if (t1 > EPSILON)
    t = t1;
  else if (t2 > EPSILON)
    t = t2;
  else
    return false;
  Vec3 lp = lo + ld * t;
  Vec3 wp = lp + center;
  out->t = (wp - ray.origin).length();
  out->point = wp;
  out->normal = lp.normalize().normalize();
  out->front_face = ray.direction.dot(out->normal) < 0.0;
  out->material = material;
  // sphere.rs:86-100 with thetamax 360, z in [-r, r], unit scale
  double phi = std::atan2(lp.y, lp.x) * 180.0 / PI;
  if (phi < 0.0) phi += 360.0;
  const double zr = std::fmax(2.0 * radius, 1e-12);
  const double ring = std::fmax(std::sqrt(lp.x * lp.x + lp.y * lp.y), 1e-6 * std::fmax(radius, 1e-12));
  out->st[0] = std::clamp(phi / 360.0, 0.0, 1.0);
  out->st[1] = std::clamp((lp.z + radius) / zr, 0.0, 1.0);
  out->st_density = 1.0 / std::sqrt(std::fmax(ring * 2.0 * PI * (zr * radius / ring), 1e-24));
  return true;
}

Mat4 InstanceObj::inverse_at(double time) const {
  if (!moving || !(time > 0.0)) return inverse;
  Mat4 fwd, inv;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) fwd.m[r][c] = transform.m[r][c] + (transform1.m[r][c] - transform.m[r][c]) * time;
  if (!fwd.inverse(inv)) return Mat4::identity();
  return inv;
}

// mesh.rs:308-381 -> the parametric hit and world normal of one instance
bool Scene::instance_hit(const InstanceObj& inst, const Ray& ray, double t_max, double* t_param,
                         Vec3* world_normal, Vec3* world_tangent, Intersection* st_out) const {
  const Mat4 inverse = inst.inverse_at(ray.time);
  Ray local = Ray::raw(inverse.transform_point(ray.origin), inverse.transform_vec(ray.direction));
  local.time = ray.time;  // mesh.rs:319-323: the local ray keeps the shutter time
  if (world_tangent) *world_tangent = Vec3::zero();
  if (inst.curves) {  // mesh.rs:324-343: the closest segment, its normal and tangent
    const CurveSet& cs = curve_sets[inst.mesh];
    Vec3 bn, bt;
    auto isect = [&](uint32_t seg, double tcur, double* tout) {
      double tt, vv;
      Vec3 nn, tg;
      if (!cs.intersect_segment(local, seg, tcur, &tt, &nn, &tg, &vv)) return false;
      *tout = tt;
      return true;
    };
    uint32_t seg;
    double t;
    if (!cs.blas.traverse(local, t_max, isect, &seg, &t)) return false;
    double tt, vv;
    cs.intersect_segment(local, seg, INF, &tt, &bn, &bt, &vv);
    *t_param = t;
    *world_normal = inverse.transform_normal(bn).normalize();
    // the tangent goes through the shutter-open transform (mesh.rs:334)
    if (world_tangent) *world_tangent = inst.transform.transform_vec(bt).normalize();
    if (st_out) {  // mesh.rs:339-342: st = [0.5, v], no density
      st_out->st[0] = 0.5;
      st_out->st[1] = vv;
      st_out->st_density = 0.0;
      st_out->st_grad[0] = st_out->st_grad[1] = Vec3::zero();
    }
    return true;
  }
  const MeshObj& mo = meshes[inst.mesh];
  const Mesh& m = mo.mesh;
  double hu = 0, hv = 0;
  auto isect = [&](uint32_t tri, double tcur, double* tout) {
    double tt, uu, vv;
    if (!m.intersect_triangle(local, tri, 1e-9, tcur, &tt, &uu, &vv) || !(tt < tcur)) return false;
    hu = uu;
    hv = vv;
    *tout = tt;
    return true;
  };
  uint32_t tri;
  double t;
  if (!m.blas.traverse(local, t_max, isect, &tri, &t)) return false;
  // The payload of the winning triangle (the traversal keeps the closest).
  double tt;
  m.intersect_triangle(local, tri, 1e-9, INF, &tt, &hu, &hv);
  uint32_t i0 = m.indices[3 * tri], i1 = m.indices[3 * tri + 1], i2 = m.indices[3 * tri + 2];
  // mesh.rs:199-220: vertex normals for the rest pose; a deforming mesh
  // takes the geometric normal at the shutter time
  Vec3 ln;
  if (!mo.normals.empty() && m.positions1.empty()) {
    auto nv = [&](uint32_t i) {
      return Vec3{(double)mo.normals[3 * i], (double)mo.normals[3 * i + 1], (double)mo.normals[3 * i + 2]};
    };
    double w = 1.0 - hu - hv;
    ln = (nv(i0) * w + nv(i1) * hu + nv(i2) * hv).normalize();
  } else {
    Vec3 v0 = m.vertex_at(i0, ray.time);
    ln = (m.vertex_at(i1, ray.time) - v0).cross(m.vertex_at(i2, ray.time) - v0).normalize();
  }
  *t_param = t;
  *world_normal = inverse.transform_normal(ln).normalize();
  if (st_out) {
    // mesh.rs:145-197, 357-370: st interpolated at the hit, density the
    // square root of st area over geometric area over the instance scale,
    // world gradients of s and t through the (static) inverse transform
    st_out->st[0] = st_out->st[1] = st_out->st_density = 0.0;
    st_out->st_grad[0] = st_out->st_grad[1] = Vec3::zero();
    if (!mo.st.empty()) {
      auto stv = [&](uint32_t i, int c) { return (double)mo.st[2 * i + c]; };
      const double w = 1.0 - hu - hv;
      st_out->st[0] = stv(i0, 0) * w + stv(i1, 0) * hu + stv(i2, 0) * hv;
      st_out->st[1] = stv(i0, 1) * w + stv(i1, 1) * hu + stv(i2, 1) * hv;
      const double st_area = 0.5 * std::fabs((double)(mo.st[2 * i1] - mo.st[2 * i0]) * (double)(mo.st[2 * i2 + 1] - mo.st[2 * i0 + 1]) -
                                             (double)(mo.st[2 * i2] - mo.st[2 * i0]) * (double)(mo.st[2 * i1 + 1] - mo.st[2 * i0 + 1]));
      const Vec3 p0 = m.vertex(i0), e1 = m.vertex(i1) - p0, e2 = m.vertex(i2) - p0;
      const double geo_area = 0.5 * e1.cross(e2).length();
      st_out->st_density = (geo_area > 1e-18 ? std::sqrt(st_area / geo_area) : 0.0) / inst.scale;
      const Vec3 n = e1.cross(e2);
      const double n2 = n.length_squared();
      if (n2 >= 1e-24) {
        const Vec3 dual1 = e2.cross(n) / n2, dual2 = e1.cross(n) / -n2;
        const double ds1 = (double)(mo.st[2 * i1] - mo.st[2 * i0]), dt1 = (double)(mo.st[2 * i1 + 1] - mo.st[2 * i0 + 1]);
        const double ds2 = (double)(mo.st[2 * i2] - mo.st[2 * i0]), dt2 = (double)(mo.st[2 * i2 + 1] - mo.st[2 * i0 + 1]);
        st_out->st_grad[0] = inst.inverse.transform_normal(dual1 * ds1 + dual2 * ds2);
        st_out->st_grad[1] = inst.inverse.transform_normal(dual1 * dt1 + dual2 * dt2);
      }
    }
  }
  return true;
}

// scene/mod.rs:105-146
// extra.rs:44-55
static bool theta_ok(double x, double y, double thetamax) {
  if (thetamax >= 360.0) return true;
  double phi = std::atan2(y, x) * 180.0 / PI;
  if (phi < 0.0) phi += 360.0;
  return phi <= thetamax;
}

// Real roots of a cubic / quartic, ascending (extra.rs solve_cubic /
// solve_quartic: Cardano or the trigonometric form; Ferrari through the
// resolvent cubic, Newton-polished, residual-filtered, near-duplicates
// merged).
static std::vector<double> cubic_real_roots(double a3, double a2, double a1, double a0) {
  std::vector<double> v;
  if (std::fabs(a3) < 1e-12) {
    if (std::fabs(a2) < 1e-12) {
      if (std::fabs(a1) >= 1e-12) v.push_back(-a0 / a1);
      return v;
    }
    const double disc = a1 * a1 - 4.0 * a2 * a0;
    if (disc < 0.0) return v;
    const double s = std::sqrt(disc);
    v = {(-a1 - s) / (2.0 * a2), (-a1 + s) / (2.0 * a2)};
    std::sort(v.begin(), v.end());
    return v;
  }
  const double a = a2 / a3, b = a1 / a3, c = a0 / a3;
  const double p = b - a * a / 3.0, q = 2.0 * a * a * a / 27.0 - a * b / 3.0 + c, shift = -a / 3.0;
  const double disc = q * q / 4.0 + p * p * p / 27.0;
  if (disc > 0.0) {
    const double s = std::sqrt(disc);
    v.push_back(std::cbrt(-q / 2.0 + s) + std::cbrt(-q / 2.0 - s) + shift);
  } else {
    const double m = 2.0 * std::sqrt(std::fmax(-p / 3.0, 0.0));
    if (m < 1e-30) {
      v.push_back(shift);
    } else {
      const double th = std::acos(std::fmin(std::fmax(3.0 * q / (p * m), -1.0), 1.0)) / 3.0;
      for (int k = 0; k < 3; k++) v.push_back(m * std::cos(th - 2.0 * PI * k / 3.0) + shift);
    }
  }
  std::sort(v.begin(), v.end());
  return v;
}

static std::vector<double> quartic_real_roots(double a4, double a3, double a2, double a1, double a0) {
  if (std::fabs(a4) < 1e-12) return cubic_real_roots(a3, a2, a1, a0);
  const double b = a3 / a4, c = a2 / a4, d = a1 / a4, e = a0 / a4;
  const double p = c - 3.0 * b * b / 8.0;
  const double q = d - b * c / 2.0 + b * b * b / 8.0;
  const double r = e - b * d / 4.0 + b * b * c / 16.0 - 3.0 * b * b * b * b / 256.0;
  const double shift = -b / 4.0;
  const auto poly = [&](double t) { return (((a4 * t + a3) * t + a2) * t + a1) * t + a0; };
  const auto dpoly = [&](double t) { return ((4.0 * a4 * t + 3.0 * a3) * t + 2.0 * a2) * t + a1; };
  const auto polish = [&](double t) {
    for (int it = 0; it < 8; it++) {
      const double df = dpoly(t);
      if (std::fabs(df) < 1e-30) break;
      const double step = poly(t) / df;
      t -= step;
      if (std::fabs(step) < 1e-12 * std::fmax(std::fabs(t), 1.0)) break;
    }
    return t;
  };
  std::vector<double> roots;
  if (std::fabs(q) < 1e-12) {
    const double disc = p * p - 4.0 * r;
    if (disc >= 0.0) {
      const double s = std::sqrt(disc);
      for (double y2 : {(-p - s) / 2.0, (-p + s) / 2.0})
        if (y2 >= 0.0) {
          const double y = std::sqrt(y2);
          roots.push_back(polish(y + shift));
          roots.push_back(polish(-y + shift));
        }
    }
  } else {
    double z = -1.0;
    for (double zz : cubic_real_roots(1.0, 2.0 * p, p * p - 4.0 * r, -q * q))
      if (zz > 1e-12 && zz > z) z = zz;
    if (z > 0.0) {
      const double m = std::sqrt(z), n = q / (2.0 * m);
      for (int sg = 0; sg < 2; sg++) {
        const double bq = sg == 0 ? m : -m, cq = (p + z) / 2.0 + (sg == 0 ? -n : n);
        const double disc = bq * bq - 4.0 * cq;
        if (disc >= 0.0) {
          const double s = std::sqrt(disc);
          roots.push_back(polish((-bq - s) / 2.0 + shift));
          roots.push_back(polish((-bq + s) / 2.0 + shift));
        }
      }
    }
  }
  std::vector<double> kept;
  for (double t : roots)
    if (std::isfinite(t) && std::fabs(poly(t)) < 1e-6 * (1.0 + std::pow(std::fabs(t), 4) * std::fabs(a4)))
      kept.push_back(t);
  std::sort(kept.begin(), kept.end());
  std::vector<double> out;
  for (double t : kept)
    if (out.empty() || std::fabs(t - out.back()) >= 1e-9) out.push_back(t);
  return out;
}

bool QuadricObj::intersect(const Ray& ray, Intersection* out) const {
  const Vec3 lo = inverse.transform_point(ray.origin);
  const Vec3 ld = inverse.transform_vec(ray.direction).normalize();
  double t = 0.0;
  Vec3 p, ln;
  switch (kind) {
    case RRHIP_QUADRIC_SPHERE: {  // sphere.rs:38-75
      const double a = ld.dot(ld), b = 2.0 * lo.dot(ld), c = lo.dot(lo) - radius * radius;
      const double disc = b * b - 4.0 * a * c;
      if (disc < 0.0) return false;
      const double s = std::sqrt(disc);
      const double t1 = (-b - s) / (2.0 * a), t2 = (-b + s) / (2.0 * a);
      if (t1 > EPSILON)
        t = t1;
      else if (t2 > EPSILON)
        t = t2;
      else
        return false;
      p = lo + ld * t;
      if (p.z < zmin || p.z > zmax || !theta_ok(p.x, p.y, thetamax)) return false;
      ln = p.normalize();
      break;
    }
    case RRHIP_QUADRIC_CYLINDER: {  // cylinder.rs:36-93
      const double a = ld.x * ld.x + ld.y * ld.y;
      const double b = 2.0 * (lo.x * ld.x + lo.y * ld.y);
      const double c = lo.x * lo.x + lo.y * lo.y - radius * radius;
      if (std::fabs(a) < EPSILON) return false;
      const double disc = b * b - 4.0 * a * c;
      if (disc < 0.0) return false;
      const double s = std::sqrt(disc);
      const double t1 = (-b - s) / (2.0 * a), t2 = (-b + s) / (2.0 * a);
      t = t1 > EPSILON ? t1 : t2;
      if (t <= EPSILON) return false;
      p = lo + ld * t;
      if (p.z < zmin || p.z > zmax) {
        if (t == t1 && t2 > EPSILON) {
          t = t2;
          p = lo + ld * t;
          if (p.z < zmin || p.z > zmax) return false;
        } else {
          return false;
        }
      }
      if (!theta_ok(p.x, p.y, thetamax)) return false;
      ln = Vec3{p.x, p.y, 0.0}.normalize();
      break;
    }
    case RRHIP_QUADRIC_CONE: {  // cone.rs:48-101, 110-125
      const double k = (radius / height) * (radius / height);
      const double a = ld.x * ld.x + ld.y * ld.y - k * ld.z * ld.z;
      const double b = 2.0 * (lo.x * ld.x + lo.y * ld.y - k * lo.z * ld.z);
      const double c = lo.x * lo.x + lo.y * lo.y - k * lo.z * lo.z;
      bool hit = false;
      if (std::fabs(a) < EPSILON) {
        if (std::fabs(b) < EPSILON) return false;
        t = -c / b;
        if (t > EPSILON) {
          p = lo + ld * t;
          hit = p.z >= 0.0 && p.z <= height;
        }
      } else {
        const double disc = b * b - 4.0 * a * c;
        if (disc < 0.0) return false;
        const double s = std::sqrt(disc);
        const double ts[2] = {(-b - s) / (2.0 * a), (-b + s) / (2.0 * a)};
        for (double tc : ts) {
          if (!(tc > EPSILON)) continue;
          const Vec3 pc = lo + ld * tc;
          if (pc.z >= 0.0 && pc.z <= height) {
            if (!theta_ok(pc.x, pc.y, thetamax)) continue;
            t = tc;
            p = pc;
            hit = true;
            break;
          }
        }
      }
      if (!hit) return false;
      const double r = std::sqrt(p.x * p.x + p.y * p.y);
      ln = Vec3{p.x / r, p.y / r, -(radius / height)}.normalize();
      break;
    }
    case RRHIP_QUADRIC_DISK: {  // extra.rs Disk::intersect
      if (std::fabs(ld.z) < EPSILON) return false;
      t = (height - lo.z) / ld.z;
      if (t <= EPSILON) return false;
      p = lo + ld * t;
      if (p.x * p.x + p.y * p.y > radius * radius || !theta_ok(p.x, p.y, thetamax)) return false;
      ln = Vec3{0.0, 0.0, 1.0};
      break;
    }
    case RRHIP_QUADRIC_HYPERBOLOID: {  // extra.rs Hyperboloid::intersect
      const double zlo = std::fmin(p1[2], p2[2]), zhi = std::fmax(p1[2], p2[2]);
      if (std::fabs(p2[2] - p1[2]) < EPSILON) {
        if (std::fabs(ld.z) < EPSILON) return false;
        t = (p1[2] - lo.z) / ld.z;
        if (t <= EPSILON) return false;
        p = lo + ld * t;
        const double rr = p.x * p.x + p.y * p.y;
        const double ra = p1[0] * p1[0] + p1[1] * p1[1], rb = p2[0] * p2[0] + p2[1] * p2[1];
        if (rr < std::fmin(ra, rb) || rr > std::fmax(ra, rb) || !theta_ok(p.x, p.y, thetamax)) return false;
        ln = Vec3{0.0, 0.0, 1.0};
        break;
      }
      const double dz = p2[2] - p1[2];
      const double sx = (p2[0] - p1[0]) / dz, sy = (p2[1] - p1[1]) / dz;
      const double cx = p1[0] - p1[2] * sx, cy = p1[1] - p1[2] * sy;
      const double pa = sx * sx + sy * sy, pb = 2.0 * (cx * sx + cy * sy), pc = cx * cx + cy * cy;
      const double a = ld.x * ld.x + ld.y * ld.y - pa * ld.z * ld.z;
      const double b = 2.0 * (lo.x * ld.x + lo.y * ld.y - pa * lo.z * ld.z) - pb * ld.z;
      const double c = lo.x * lo.x + lo.y * lo.y - pa * lo.z * lo.z - pb * lo.z - pc;
      double cands[2];
      int nc = 2;
      if (std::fabs(a) < EPSILON) {
        if (std::fabs(b) < EPSILON) return false;
        cands[0] = -c / b;
        nc = 1;
      } else {
        const double disc = b * b - 4.0 * a * c;
        if (disc < 0.0) return false;
        const double s = std::sqrt(disc);
        cands[0] = (-b - s) / (2.0 * a);
        cands[1] = (-b + s) / (2.0 * a);
      }
      bool hit = false;
      for (int i = 0; i < nc && !hit; i++) {
        const double tc = cands[i];
        if (tc <= EPSILON) continue;
        const Vec3 pcand = lo + ld * tc;
        if (pcand.z < zlo || pcand.z > zhi || !theta_ok(pcand.x, pcand.y, thetamax)) continue;
        t = tc;
        p = pcand;
        hit = true;
      }
      if (!hit) return false;
      ln = Vec3{2.0 * p.x, 2.0 * p.y, -(2.0 * pa * p.z + pb)}.normalize();
      break;
    }
    case RRHIP_QUADRIC_TORUS: {  // extra.rs Torus::intersect / phi_ok
      const double R = radius, r = height;
      const double g = ld.dot(ld), h = 2.0 * lo.dot(ld), i = lo.dot(lo) + R * R - r * r;
      const double j = ld.x * ld.x + ld.y * ld.y, k = 2.0 * (lo.x * ld.x + lo.y * ld.y), l = lo.x * lo.x + lo.y * lo.y;
      const double f4 = 4.0 * R * R;
      bool hit = false;
      for (double tc : quartic_real_roots(g * g, 2.0 * g * h, h * h + 2.0 * g * i - f4 * j, 2.0 * h * i - f4 * k,
                                          i * i - f4 * l)) {
        if (tc <= EPSILON) continue;
        const Vec3 pcand = lo + ld * tc;
        if (zmax - zmin < 360.0) {
          const double rxy = std::sqrt(pcand.x * pcand.x + pcand.y * pcand.y);
          double phi = std::atan2(pcand.z, rxy - R) * 180.0 / PI;
          while (phi < zmin) phi += 360.0;
          while (phi >= zmin + 360.0) phi -= 360.0;
          if (phi > zmax) continue;
        }
        if (!theta_ok(pcand.x, pcand.y, thetamax)) continue;
        const double rxy = std::fmax(std::sqrt(pcand.x * pcand.x + pcand.y * pcand.y), 1e-12), sc = 1.0 - R / rxy;
        t = tc;
        p = pcand;
        ln = Vec3{p.x * sc, p.y * sc, p.z}.normalize();
        hit = true;
        break;
      }
      if (!hit) return false;
      break;
    }
    case RRHIP_QUADRIC_PARABOLOID: {  // extra.rs Paraboloid::intersect
      if (std::fabs(zmax) < EPSILON) return false;
      const double k = radius * radius / zmax;
      const double a = ld.x * ld.x + ld.y * ld.y;
      const double b = 2.0 * (lo.x * ld.x + lo.y * ld.y) - k * ld.z;
      const double c = lo.x * lo.x + lo.y * lo.y - k * lo.z;
      double cands[2];
      int nc = 2;
      if (std::fabs(a) < EPSILON) {
        if (std::fabs(b) < EPSILON) return false;
        cands[0] = -c / b;
        nc = 1;
      } else {
        const double disc = b * b - 4.0 * a * c;
        if (disc < 0.0) return false;
        const double s = std::sqrt(disc);
        cands[0] = (-b - s) / (2.0 * a);
        cands[1] = (-b + s) / (2.0 * a);
      }
      bool hit = false;
      for (int i = 0; i < nc && !hit; i++) {
        const double tc = cands[i];
        if (tc <= EPSILON) continue;
        const Vec3 pc = lo + ld * tc;
        if (pc.z < zmin || pc.z > zmax || !theta_ok(pc.x, pc.y, thetamax)) continue;
        t = tc;
        p = pc;
        hit = true;
      }
      if (!hit) return false;
      ln = Vec3{2.0 * p.x, 2.0 * p.y, -k}.normalize();
      break;
    }
    default:
      return false;
  }
  const Vec3 wp = transform.transform_point(p);
  out->t = (wp - ray.origin).length();
  out->point = wp;
  out->normal = inverse.transform_normal(ln).normalize();
  out->front_face = kind == RRHIP_QUADRIC_SPHERE ? ray.direction.dot(out->normal) < 0.0 : true;
  out->material = material;
  quadric_st(p, out);
  return true;
}

// The RiSpec parameterisation of each quadric and its st density: u along
// theta, v along the profile, density 1 / sqrt(|dp/du| |dp/dv|) over the
// transform's scale (sphere.rs:86-100, cylinder.rs:100-109, extra.rs Torus
// 272-289, Disk 364-369, Paraboloid 461-470, Hyperboloid 603-618). Cones
// carry none (cone.rs).
static double phi_u(double x, double y, double thetamax) {  // extra.rs:35-41
  double phi = std::atan2(y, x) * 180.0 / PI;
  if (phi < 0.0) phi += 360.0;
  return std::clamp(phi / std::fmax(thetamax, 1e-9), 0.0, 1.0);
}

void QuadricObj::quadric_st(const Vec3& p, Intersection* out) const {
  const double rad_theta = thetamax * PI / 180.0;
  double u = 0, v = 0, d = 0;
  switch (kind) {
    case RRHIP_QUADRIC_SPHERE: {
      u = phi_u(p.x, p.y, thetamax);  // sphere.rs clamps phi / thetamax the same way
      const double zr = std::fmax(zmax - zmin, 1e-12);
      v = std::clamp((p.z - zmin) / zr, 0.0, 1.0);
      const double ring = std::fmax(std::sqrt(p.x * p.x + p.y * p.y), 1e-6 * std::fmax(radius, 1e-12));
      d = 1.0 / std::sqrt(std::fmax(ring * rad_theta * (zr * radius / ring), 1e-24)) / scale;
      break;
    }
    case RRHIP_QUADRIC_CYLINDER: {
      u = phi_u(p.x, p.y, thetamax);
      const double zr = std::fmax(zmax - zmin, 1e-12);
      v = std::clamp((p.z - zmin) / zr, 0.0, 1.0);
      d = 1.0 / std::sqrt(std::fmax(std::fmax(radius, 1e-12) * rad_theta * zr, 1e-24)) / scale;
      break;
    }
    case RRHIP_QUADRIC_DISK: {
      const double rh = std::sqrt(p.x * p.x + p.y * p.y);
      u = phi_u(p.x, p.y, thetamax);
      v = std::clamp((radius - rh) / std::fmax(radius, 1e-12), 0.0, 1.0);
      const double dpdu = std::fmax(rh, 1e-6 * std::fmax(radius, 1e-12)) * rad_theta;
      d = 1.0 / std::sqrt(std::fmax(dpdu * radius, 1e-24)) / scale;
      break;
    }
    case RRHIP_QUADRIC_PARABOLOID: {
      u = phi_u(p.x, p.y, thetamax);
      const double zr = std::fmax(zmax - zmin, 1e-12);
      v = std::clamp((p.z - zmin) / zr, 0.0, 1.0);
      const double ring = std::fmax(std::sqrt(p.x * p.x + p.y * p.y), 1e-6 * std::fmax(radius, 1e-12));
      d = 1.0 / std::sqrt(std::fmax(ring * rad_theta * zr, 1e-24)) / scale;
      break;
    }
    case RRHIP_QUADRIC_HYPERBOLOID: {
      if (std::fabs(p2[2] - p1[2]) < EPSILON) break;  // the annulus case carries none here
      u = phi_u(p.x, p.y, thetamax);
      v = std::clamp((p.z - p1[2]) / (p2[2] - p1[2]), 0.0, 1.0);
      const double ring = std::fmax(std::sqrt(p.x * p.x + p.y * p.y), 1e-12);
      const double dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
      const double sweep = std::fmax(std::sqrt(dx * dx + dy * dy + dz * dz), 1e-12);
      d = 1.0 / std::sqrt(std::fmax(ring * rad_theta * sweep, 1e-24)) / scale;
      break;
    }
    case RRHIP_QUADRIC_TORUS: {
      const double R = radius, r = height;
      const double rxy = std::fmax(std::sqrt(p.x * p.x + p.y * p.y), 1e-12);
      u = phi_u(p.x, p.y, thetamax);
      double vphi = std::atan2(p.z, rxy - R) * 180.0 / PI;
      while (vphi < zmin) vphi += 360.0;
      while (vphi >= zmin + 360.0) vphi -= 360.0;
      const double vr = std::fmax(zmax - zmin, 1e-9);
      v = std::clamp((vphi - zmin) / vr, 0.0, 1.0);
      d = 1.0 / std::sqrt(std::fmax(rxy * rad_theta * (std::fmax(r, 1e-12) * vr * PI / 180.0), 1e-24)) / scale;
      break;
    }
    default:
      break;
  }
  out->st[0] = u;
  out->st[1] = v;
  out->st_density = d;
}

bool Scene::intersect(const Ray& ray, Intersection* out) const {
  bool found = false;
  double closest_t = INF;
  Intersection h;
  for (const auto& tri : triangles) {
    if (tri.intersect(ray, &h) && h.t < closest_t) {
      closest_t = h.t;
      *out = h;
      found = true;
    }
  }
  for (const auto& sp : spheres) {
    if (sp.intersect(ray, &h) && h.t < closest_t) {
      closest_t = h.t;
      *out = h;
      found = true;
    }
  }
  for (const auto& q : quadrics) {
    if (q.intersect(ray, &h) && h.t < closest_t) {
      closest_t = h.t;
      *out = h;
      found = true;
    }
  }
  if (!instances.empty()) {
    double dir_len = std::fmax(ray.direction.length(), 1e-300);
    double budget = closest_t / dir_len;
    uint32_t best_inst = 0;
    double best_t = 0;
    Vec3 best_n;
    auto isect = [&](uint32_t id, double tcur, double* tout) {
      double tp;
      Vec3 wn;
      if (!instance_hit(instances[id], ray, tcur, &tp, &wn)) return false;
      if (!(tp < tcur)) return false;
      *tout = tp;
      return true;
    };
    uint32_t win;
    double wt;
    if (tlas.traverse(ray, budget, isect, &win, &wt)) {
      best_inst = win;
      best_t = wt;
      Vec3 best_tan;
      Intersection best_st;
      instance_hit(instances[best_inst], ray, INF, &best_t, &best_n, &best_tan, &best_st);
      // mesh.rs:427-439 (to_intersection)
      Vec3 p = ray.at(best_t);
      Vec3 n = best_n;
      bool front = n.dot(ray.direction) < 0.0;
      if (!front) n = -n;
      double t = (p - ray.origin).length();
      if (t < closest_t) {
        out->t = t;
        out->point = p;
        out->normal = n.normalize();
        out->front_face = front;
        out->material = instances[best_inst].material;
        out->tangent = best_tan;
        out->st[0] = best_st.st[0];
        out->st[1] = best_st.st[1];
        out->st_density = best_st.st_density;
        out->st_grad[0] = best_st.st_grad[0];
        out->st_grad[1] = best_st.st_grad[1];
        found = true;
      }
    }
  }
  return found;
}

// ---- integrator (pt/mod.rs) -------------------------------------------------------

// pt/mod.rs:340-356 (no lens, no motion)
Ray Scene::camera_ray(uint32_t x, uint32_t y, Pcg32& rng) const {
  // pt/mod.rs:340-356
  double u1 = rng.next_f64(), u2 = rng.next_f64();
  double dx = camera.filter_1d(u1), dy = camera.filter_1d(u2);
  double lu = 0.5, lv = 0.5;
  if (camera.lens_radius > 0.0) {
    lu = rng.next_f64();
    lv = rng.next_f64();
  }
  Ray ray = camera.generate_ray_lens((double)x + 0.5 + dx, (double)y + 0.5 + dy, lu, lv);
  if (has_motion) {
    const double t = rng.next_f64();
    ray.time = t;
    ray = camera.apply_motion(ray, t);
  }
  return ray;
}

// ---- participating media (medium.rs:28-101, pt/volume.rs:29-189) -------------------

double Medium::majorant() const {
  const Vec3 st = sigma_t();
  return std::fmax(std::fmax(st.x, st.y), st.z);
}

// medium.rs:55-63 for the fBm cloud field (offset 0, at the world point);
// trilinear over the vertex grid, zero outside; 1 when homogeneous.
double Medium::density_at(const Vec3& p) const {
  if (fbm_octaves > 0) {
    FbmParams fp;
    fp.frequency = fbm_frequency;
    fp.octaves = fbm_octaves;
    fp.gain = fbm_gain;
    fp.lacunarity = fbm_lacunarity;
    const double n = fbm(p.x, p.y, p.z, fp) * 0.5 + 0.5;
    return std::clamp((n - (1.0 - fbm_coverage)) * fbm_sharpness, 0.0, 1.0);
  }
  if (nx == 0) return 1.0;
  const double pc[3] = {p.x, p.y, p.z};
  const uint32_t n[3] = {nx, ny, nz};
  int i[3];
  double f[3];
  for (int a = 0; a < 3; a++) {
    const double u = (pc[a] - lo[a]) * ((double)(n[a] - 1) / (hi[a] - lo[a]));
    if (!(u >= 0.0 && u <= (double)(n[a] - 1))) return 0.0;
    i[a] = std::min((int)u, (int)n[a] - 2);
    f[a] = u - (double)i[a];
  }
  auto at = [&](int dx, int dy, int dz) {
    return (double)grid[((size_t)(i[2] + dz) * ny + (i[1] + dy)) * nx + (i[0] + dx)];
  };
  const double c00 = at(0, 0, 0) + (at(1, 0, 0) - at(0, 0, 0)) * f[0];
  const double c10 = at(0, 1, 0) + (at(1, 1, 0) - at(0, 1, 0)) * f[0];
  const double c01 = at(0, 0, 1) + (at(1, 0, 1) - at(0, 0, 1)) * f[0];
  const double c11 = at(0, 1, 1) + (at(1, 1, 1) - at(0, 1, 1)) * f[0];
  const double c0 = c00 + (c10 - c00) * f[1], c1 = c01 + (c11 - c01) * f[1];
  return c0 + (c1 - c0) * f[2];
}

namespace {

// Part of the ray inside the grid box (density is zero outside, so delta
// and ratio tracking start at the entry and stop at the exit).
bool medium_range(const Medium& m, const Vec3& o, const Vec3& d, double t_max, double* t0, double* t1) {
  *t0 = 0.0;
  *t1 = std::fmin(t_max, m.max_distance);
  if (m.nx == 0) return *t1 > *t0;
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  for (int a = 0; a < 3; a++) {
    const double inv = 1.0 / dd[a];
    double ta = (m.lo[a] - oo[a]) * inv, tb = (m.hi[a] - oo[a]) * inv;
    if (ta > tb) std::swap(ta, tb);
    if (std::isnan(ta)) ta = -INFINITY;  // 0 * inf when the origin lies on a slab plane
    if (std::isnan(tb)) tb = INFINITY;
    *t0 = std::fmax(*t0, ta);
    *t1 = std::fmin(*t1, tb);
  }
  return *t1 > *t0;
}

double avg3(const Vec3& v) { return (v.x + v.y + v.z) / 3.0; }
Vec3 exp3(const Vec3& v, double t) { return {std::exp(-v.x * t), std::exp(-v.y * t), std::exp(-v.z * t)}; }
double exp_step(Pcg32& rng, double sigma) { return -std::log(std::fmax(1.0 - rng.next_f64(), 1e-12)) / sigma; }

}  // namespace

MediumEvent sample_distance(const Medium& m, const Vec3& o, const Vec3& d, double t_max, const Vec3& beta,
                            Pcg32& rng) {
  const Vec3 st = m.sigma_t();
  MediumEvent ev;
  if (!m.heterogeneous()) {  // volume.rs:60-102
    const double tm = std::fmin(t_max, m.max_distance);
    if (avg3(st) <= 1e-9) return ev;
    const Vec3 b{std::fmax(beta.x, 0.0), std::fmax(beta.y, 0.0), std::fmax(beta.z, 0.0)};
    const double tot = b.x + b.y + b.z;
    const double w[3] = {tot > 1e-12 ? b.x / tot : 1.0 / 3.0, tot > 1e-12 ? b.y / tot : 1.0 / 3.0,
                         tot > 1e-12 ? b.z / tot : 1.0 / 3.0};
    const double u = rng.next_f64();
    const double sc[3] = {st.x, st.y, st.z};
    const int c = u < w[0] ? 0 : (u < w[0] + w[1] ? 1 : 2);
    const double t = exp_step(rng, std::fmax(sc[c], 1e-9));
    if (t < tm) {
      const Vec3 tr = exp3(st, t);
      const double pdf = w[0] * st.x * tr.x + w[1] * st.y * tr.y + w[2] * st.z * tr.z;
      if (pdf <= 1e-300) return MediumEvent{false, 0.0, Vec3::zero()};
      return MediumEvent{true, t, m.sigma_s * tr / pdf};
    }
    const Vec3 tr = exp3(st, tm);
    const double pdf = w[0] * tr.x + w[1] * tr.y + w[2] * tr.z;
    if (pdf <= 1e-300) return MediumEvent{false, 0.0, Vec3::zero()};
    return MediumEvent{false, 0.0, tr / pdf};
  }
  // volume.rs:104-149 weighted delta tracking
  const double maj = m.majorant();
  double t, t_end;
  if (maj <= 1e-9 || !medium_range(m, o, d, t_max, &t, &t_end)) return ev;
  Vec3 weight{1, 1, 1};
  for (int i = 0; i < 10000; i++) {
    t += exp_step(rng, maj);
    if (t >= t_end) return MediumEvent{false, 0.0, weight};
    const double rho = m.density_at(o + d * t);
    const Vec3 sh = st * rho;
    const double p_real = std::fmin(std::fmax(avg3(sh) / maj, 0.0), 1.0);
    if (rng.next_f64() < p_real) return MediumEvent{true, t, weight * m.sigma_s * (rho / (maj * p_real))};
    const double p_null = std::fmax(1.0 - p_real, 1e-9);
    weight = weight * Vec3{(1.0 - sh.x / maj) / p_null, (1.0 - sh.y / maj) / p_null, (1.0 - sh.z / maj) / p_null};
    if (max_component(weight) < 1e-6) return MediumEvent{false, 0.0, Vec3::zero()};
  }
  return MediumEvent{false, 0.0, weight};
}

// volume.rs:153-189
Vec3 medium_transmittance(const Medium& m, const Vec3& o, const Vec3& d, double dist, Pcg32& rng) {
  const Vec3 st = m.sigma_t();
  if (!m.heterogeneous()) return exp3(st, std::fmin(dist, m.max_distance));
  const double maj = m.majorant();
  double t, t_end;
  Vec3 tr{1, 1, 1};
  if (maj <= 1e-9 || !medium_range(m, o, d, dist, &t, &t_end)) return tr;
  for (int i = 0; i < 10000; i++) {
    t += exp_step(rng, maj);
    if (t >= t_end) break;
    const double rho = m.density_at(o + d * t);
    tr = tr * Vec3{1.0 - st.x * rho / maj, 1.0 - st.y * rho / maj, 1.0 - st.z * rho / maj};
    if (max_component(tr) < 1e-5) return Vec3::zero();
  }
  return tr;
}

// medium.rs:73-101
double hg_phase(double g, double cos_theta) {
  const double denom = 1.0 + g * g + 2.0 * g * cos_theta;
  return (1.0 - g * g) / (4.0 * PI * denom * std::sqrt(std::fmax(denom, 1e-9)));
}

Vec3 hg_sample(double g, const Vec3& wo, double u1, double u2, double* pdf) {
  double cos_t;
  if (std::fabs(g) < 1e-3) {
    cos_t = 1.0 - 2.0 * u1;
  } else {
    const double sq = (1.0 - g * g) / (1.0 + g - 2.0 * g * u1);
    cos_t = -(1.0 + g * g - sq * sq) / (2.0 * g);
  }
  const double sin_t = std::sqrt(std::fmax(1.0 - cos_t * cos_t, 0.0));
  const double phi = 2.0 * PI * u2;
  const Vec3 w = -wo;
  const Vec3 a = std::fabs(w.x) > 0.9 ? Vec3{0, 1, 0} : Vec3{1, 0, 0};
  const Vec3 t = w.cross(a).normalize();
  const Vec3 b = w.cross(t);
  const Vec3 wi = (t * (sin_t * std::cos(phi)) + b * (sin_t * std::sin(phi)) + w * cos_t).normalize();
  *pdf = hg_phase(g, -wo.dot(wi));
  return wi;
}

// volume.rs:195-216: t with density proportional to 1/r^2 towards the
// light over [0, t_max], h the light's distance from the ray's line.
double equiangular_sample(const Vec3& o, const Vec3& d, double t_max, const Vec3& light, double u, double* pdf) {
  const double delta = (light - o).dot(d);
  const double h = std::fmax((light - (o + d * delta)).length(), 1e-6);
  const double th_a = std::atan(-delta / h);
  const double th_b = std::atan((t_max - delta) / h);
  const double span = std::fmax(th_b - th_a, 1e-9);
  double t = delta + h * std::tan(th_a + u * span);
  t = std::fmin(std::fmax(t, 0.0), t_max);
  *pdf = h / (span * (h * h + (t - delta) * (t - delta)));
  return t;
}

// volume.rs:219-232
double equiangular_pdf(const Vec3& o, const Vec3& d, double t_max, const Vec3& light, double t) {
  const double delta = (light - o).dot(d);
  const double h = std::fmax((light - (o + d * delta)).length(), 1e-6);
  const double th_a = std::atan(-delta / h);
  const double th_b = std::atan((t_max - delta) / h);
  const double span = std::fmax(th_b - th_a, 1e-9);
  return h / (span * (h * h + (t - delta) * (t - delta)));
}

// volume.rs:234-240: an exponential in the mean extinction (a collision
// before t_max) or its tail mass (no collision).
double distance_pdf_approx(const Medium& m, double t, double t_max) {
  const Vec3 st = m.sigma_t();
  const double s = std::fmax((st.x + st.y + st.z) / 3.0, 1e-9);
  return t < t_max ? s * std::exp(-s * t) : std::exp(-s * t_max);
}

bool Material::hull() const {
  double w[5];
  return interior >= 0 && !pbr.lobe_weights(w) && max_component(emission + pbr.glow) <= 0.0;
}

// pt/mod.rs:1238-1306. The stack starts from the atmosphere with the
// shading point's medium on top (when it is another one).
Vec3 Scene::visibility(const Vec3& origin0, const Vec3& dir, double remaining, Pcg32& rng, uint64_t* rays,
                       int medium, double time) const {
  MediumStack ms;
  ms.atmosphere = atmosphere;
  if (medium >= 0 && medium != atmosphere) ms.push(medium);
  Vec3 vis{1, 1, 1};
  Vec3 origin = origin0;
  for (int i = 0; i < 16; i++) {
    if (rays) ++*rays;
    Intersection h;
    Ray r = Ray::make(origin, dir);
    r.time = time;
    const bool hit = intersect(r, &h);
    const double seg = std::fmin(hit ? h.t : remaining, remaining);
    const int cur = ms.current();
    if (cur >= 0) {
      vis = vis * medium_transmittance(media[(size_t)cur], origin, dir, seg, rng);
      if (max_component(vis) < 1e-4) return Vec3::zero();
    }
    if (!hit || h.t >= remaining) return vis;
    const Material& m = materials[h.material];
    if (m.hull()) {
      if (h.front_face)
        ms.push(m.interior);
      else
        ms.pop();
      origin = h.point + dir * ray_offset;
      remaining -= h.t + ray_offset;
      continue;
    }
    double presence = 1.0;
    shade_params(m, h, dir, 0.0, &presence);
    if (presence >= 1.0) return Vec3::zero();
    vis = vis * (1.0 - presence);
    if (max_component(vis) < 1e-4) return Vec3::zero();
    origin = h.point + dir * ray_offset;
    remaining -= h.t + ray_offset;
  }
  return vis;
}

// pt/mod.rs:1177-1209
Vec3 Scene::area_contribution(const Vec3& p, const Vec3& n, const Vec3& sample_point,
                              const Vec3& light_normal, double area, const Light& light, const EvalFn& eval,
                              double pick_pmf, Pcg32& rng, uint64_t* rays, int medium, double time) const {
  Vec3 to_light = sample_point - p;
  double dist2 = std::fmax(to_light.length_squared(), 1e-12);
  Vec3 wi = to_light.normalize();
  double cos_light = std::fabs(wi.dot(light_normal));
  if (cos_light < 1e-9 || area <= 0.0) return Vec3::zero();
  Vec3 f;
  double bsdf_pdf;
  eval(wi, &f, &bsdf_pdf);
  if (max_component(f) <= 0.0) return Vec3::zero();
  // visibility_to (pt/mod.rs:1212-1224)
  Vec3 dir = sample_point - p;
  double dist = dir.length();
  Vec3 vis = visibility(p + n * ray_offset, dir.normalize(), dist - 1e-3, rng, rays, medium, time);
  if (max_component(vis) <= 0.0) return Vec3::zero();
  double pdf_sa = pick_pmf * dist2 / (cos_light * area);
  double weight = power_heuristic(pdf_sa, bsdf_pdf);
  return f * light.radiance() * vis * (weight / pdf_sa);
}

// pt/mod.rs:1032-1174
Vec3 Scene::sample_light(const Light& light, const Vec3& p, const Vec3& n, const EvalFn& eval, double pick_pmf,
                         Pcg32& rng, uint64_t* rays, int medium, double time) const {
  switch (light.kind) {
    case RRHIP_LIGHT_POINT: {
      Vec3 to_light = light.position - p;
      double dist2 = std::fmax(to_light.length_squared(), 1e-12);
      Vec3 wi = to_light.normalize();
      Vec3 f;
      double pdf;
      eval(wi, &f, &pdf);
      if (max_component(f) <= 0.0) return Vec3::zero();
      Vec3 dir = light.position - p;
      Vec3 vis = visibility(p + n * ray_offset, dir.normalize(), dir.length() - 1e-3, rng, rays, medium, time);
      if (max_component(vis) <= 0.0) return Vec3::zero();
      return f * light.radiance() * vis / (dist2 * pick_pmf);
    }
    case RRHIP_LIGHT_DISTANT: {
      Vec3 base = -light.direction;
      Vec3 wi = base;
      if (light.angular_radius > 1e-5) {
        double cos_max = std::cos(light.angular_radius);
        double u = rng.next_f64(), v = rng.next_f64();
        double cos_t = 1.0 - u * (1.0 - cos_max);
        double sin_t = std::sqrt(std::fmax(1.0 - cos_t * cos_t, 0.0));
        double phi = 2.0 * PI * v;
        Frame lf(base);
        wi = lf.to_world(Vec3{sin_t * std::cos(phi), sin_t * std::sin(phi), cos_t});
      }
      Vec3 f;
      double pdf;
      eval(wi, &f, &pdf);
      if (max_component(f) <= 0.0) return Vec3::zero();
      Vec3 vis = visibility(p + n * ray_offset, wi, 1e8, rng, rays, medium, time);
      if (max_component(vis) <= 0.0) return Vec3::zero();
      return f * light.radiance() * vis / pick_pmf;
    }
    case RRHIP_LIGHT_RECT: {
      double u = rng.next_f64(), v = rng.next_f64();
      Vec3 sp = light.corner + light.edge1 * u + light.edge2 * v;
      return area_contribution(p, n, sp, light.normal, light.area, light, eval, pick_pmf, rng, rays, medium, time);
    }
    case RRHIP_LIGHT_DISK: {
      double u = rng.next_f64(), v = rng.next_f64();
      double r = std::sqrt(u), phi = 2.0 * PI * v;
      Vec3 sp = light.center + light.e1 * (r * std::cos(phi)) + light.e2 * (r * std::sin(phi));
      return area_contribution(p, n, sp, light.normal, light.area, light, eval, pick_pmf, rng, rays, medium, time);
    }
    case RRHIP_LIGHT_SPHERE: {
      Vec3 to_center = light.center - p;
      double dist2 = to_center.length_squared();
      double r = light.radius;
      if (dist2 <= r * r * 1.0001) return Vec3::zero();
      double sin2 = std::fmin(r * r / dist2, 1.0);
      double cos_max = std::sqrt(std::fmax(1.0 - sin2, 0.0));
      double u = rng.next_f64(), v = rng.next_f64();
      double cos_t = 1.0 - u * (1.0 - cos_max);
      double sin_t = std::sqrt(std::fmax(1.0 - cos_t * cos_t, 0.0));
      double phi = 2.0 * PI * v;
      Frame lf(to_center.normalize());
      Vec3 wi = lf.to_world(Vec3{sin_t * std::cos(phi), sin_t * std::sin(phi), cos_t});
      Vec3 f;
      double bsdf_pdf;
      eval(wi, &f, &bsdf_pdf);
      if (max_component(f) <= 0.0) return Vec3::zero();
      double dist = std::sqrt(dist2) * cos_t - std::sqrt(std::fmax(r * r - dist2 * sin_t * sin_t, 0.0));
      Vec3 target = p + wi * dist;
      Vec3 dir = target - p;
      Vec3 vis = visibility(p + n * ray_offset, dir.normalize(), dir.length() - 1e-3, rng, rays, medium, time);
      if (max_component(vis) <= 0.0) return Vec3::zero();
      double pdf_sa = pick_pmf / std::fmax(2.0 * PI * (1.0 - cos_max), 1e-12);
      double weight = power_heuristic(pdf_sa, bsdf_pdf);
      return f * light.radiance() * vis * (weight / pdf_sa);
    }
    case RRHIP_LIGHT_DOME: {  // pt/mod.rs:1138-1172
      double u = rng.next_f64(), v = rng.next_f64();
      Vec3 wi, rad = light.radiance();
      double pdf_d = 1.0 / (4.0 * PI);
      if (light.env) {
        Vec3 texel;
        light.env->sample(u, v, &wi, &texel, &pdf_d);
        if (pdf_d <= 0.0) return Vec3::zero();
        rad = texel * light.radiance();
      } else {
        double z = 1.0 - 2.0 * u;
        double rr = std::sqrt(std::fmax(1.0 - z * z, 0.0));
        double phi = 2.0 * PI * v;
        wi = Vec3{rr * std::cos(phi), rr * std::sin(phi), z};
      }
      double pdf_sa = pdf_d * pick_pmf;
      Vec3 f;
      double bsdf_pdf;
      eval(wi, &f, &bsdf_pdf);
      if (max_component(f) <= 0.0) return Vec3::zero();
      Vec3 vis = visibility(p + n * ray_offset, wi, 1e8, rng, rays, medium, time);
      if (max_component(vis) <= 0.0) return Vec3::zero();
      double weight = power_heuristic(pdf_sa, bsdf_pdf);
      return f * rad * vis * (weight / pdf_sa);
    }
    default:
      return Vec3::zero();
  }
}

// pt/mod.rs:1000-1023
static double light_pdf_solid_angle(const Light& light, const Vec3& origin, const Intersection& hit) {
  if (light.kind == RRHIP_LIGHT_RECT || light.kind == RRHIP_LIGHT_DISK) {
    Vec3 d = hit.point - origin;
    double dist2 = d.length_squared();
    double cos_light = std::fabs(d.normalize().dot(light.normal));
    if (cos_light < 1e-9 || light.area <= 0.0) return 0.0;
    return dist2 / (cos_light * light.area);
  }
  if (light.kind == RRHIP_LIGHT_SPHERE) {
    double dist2 = (light.center - origin).length_squared();
    double sin2 = std::fmin(light.radius * light.radius / dist2, 1.0);
    double cos_max = std::sqrt(std::fmax(1.0 - sin2, 0.0));
    double sa = 2.0 * PI * (1.0 - cos_max);
    if (sa < 1e-12) return 0.0;
    return 1.0 / sa;
  }
  return 0.0;
}

PatternNode pattern_from_desc(const rrhip_pattern_desc& pd) {
  PatternNode n;
  n.kind = (int)pd.kind;
  n.texture = pd.texture;
  if (pd.udim) n.udim.assign(pd.udim, pd.udim + 100);
  n.wrap = (Wrap)pd.wrap;
  n.scale[0] = pd.scale[0];
  n.scale[1] = pd.scale[1];
  n.color_a = Vec3{pd.color_a[0], pd.color_a[1], pd.color_a[2]};
  n.color_b = Vec3{pd.color_b[0], pd.color_b[1], pd.color_b[2]};
  n.frequency = pd.frequency;
  n.octaves = pd.octaves;
  n.gain = pd.gain;
  n.lacunarity = pd.lacunarity;
  for (int k = 0; k < 3; k++) {
    n.input[k] = pd.input[k];
    n.constant[k] = Vec3{pd.constant[k][0], pd.constant[k][1], pd.constant[k][2]};
  }
  n.cc_gain = Vec3{pd.cc_gain[0], pd.cc_gain[1], pd.cc_gain[2]};
  n.cc_offset = Vec3{pd.cc_offset[0], pd.cc_offset[1], pd.cc_offset[2]};
  n.gamma = pd.gamma;
  n.saturation = pd.saturation;
  n.use_t = pd.use_t != 0;
  for (uint32_t k = 0; k < pd.num_stops; k++) {
    n.positions.push_back(pd.positions[k]);
    n.colors.push_back(Vec3{pd.colors[3 * k], pd.colors[3 * k + 1], pd.colors[3 * k + 2]});
  }
  return n;
}

PbrParams Scene::shade_params(const Material& m, const Intersection& hit, const Vec3& ray_dir, double cone_width,
                              double* presence) const {
  *presence = std::fmin(std::fmax(m.presence, 0.0), 1.0);
  if (m.bindings.empty()) return m.pbr;
  // the ray cone's disk projected on the surface: minor axis the cone
  // width, major axis stretched by 1 / |cos| along the view direction; the
  // st Jacobian turns both into texture-space axes (EWA)
  double dmaj[2] = {0, 0}, dmin[2] = {0, 0};
  const Vec3 gs = hit.st_grad[0], gt = hit.st_grad[1];
  if (gs.length_squared() + gt.length_squared() > 1e-18 && cone_width > 0.0) {
    const Vec3 n = hit.normal, d = ray_dir.normalize();
    const double cs = std::fmax(std::fabs(d.dot(n)), 1.0 / 16.0);
    Vec3 t1 = d - n * d.dot(n);
    if (t1.length_squared() < 1e-12) t1 = std::fabs(n.x) < 0.9 ? Vec3{1, 0, 0}.cross(n) : Vec3{0, 1, 0}.cross(n);
    t1 = t1.normalize() * (cone_width / cs);
    const Vec3 t2 = n.cross(t1).normalize() * cone_width;
    dmaj[0] = gs.dot(t1);
    dmaj[1] = gt.dot(t1);
    dmin[0] = gs.dot(t2);
    dmin[1] = gt.dot(t2);
  }
  ShadeCtx ctx;
  ctx.st[0] = hit.st[0];
  ctx.st[1] = hit.st[1];
  ctx.p = hit.point;
  ctx.n = hit.normal;
  ctx.footprint = cone_width * hit.st_density;
  for (int k = 0; k < 2; k++) {
    ctx.dst_major[k] = dmaj[k];
    ctx.dst_minor[k] = dmin[k];
  }
  PbrParams p = m.pbr;
  double pres = m.presence;
  for (const auto& b : m.bindings) {  // BoundField::apply (pattern.rs:256-272)
    const Vec3 v = pattern_eval(patterns, textures, b.second, ctx);
    switch (b.first) {
      case BIND_DIFFUSE_COLOR: p.diffuse_color = v; break;
      case BIND_DIFFUSE_GAIN: p.diffuse_gain = v.x; break;
      case BIND_DIFFUSE_ROUGHNESS: p.diffuse_roughness = v.x; break;
      case BIND_SPECULAR_FACE_COLOR: p.specular_face_color = v; break;
      case BIND_SPECULAR_EDGE_COLOR: p.specular_edge_color = v; break;
      case BIND_SPECULAR_ROUGHNESS: p.specular_roughness = std::clamp(v.x, 0.005, 1.0); break;
      case BIND_CLEARCOAT_GAIN: p.clearcoat_gain = v.x; break;
      case BIND_FUZZ_GAIN: p.fuzz_gain = v.x; break;
      case BIND_FUZZ_COLOR: p.fuzz_color = v; break;
      case BIND_GLASS_GAIN: p.glass_gain = v.x; break;
      case BIND_GLASS_ROUGHNESS: p.glass_roughness = std::clamp(v.x, 0.005, 1.0); break;
      case BIND_REFRACTION_COLOR: p.refraction_color = v; break;
      case BIND_GLOW: p.glow = v; break;
      case BIND_PRESENCE: pres = std::clamp(v.x, 0.0, 1.0); break;
      default: break;
    }
  }
  *presence = std::fmin(std::fmax(pres, 0.0), 1.0);
  return p;
}

// pt/mod.rs:449-990, surface subset
Vec3 Scene::trace(Ray ray, Pcg32& rng, const RenderParams& rp, uint64_t* rays, AuxSample* aux) const {
  const double time = ray.time;  // the path's shutter time
  Vec3 l = Vec3::zero(), beta = Vec3::one();
  // AOVs (pt/mod.rs:457-470): which lobe the first continuation used
  // (-1 none yet, 0 diffuse-like, 1 specular-like); contributions go to
  // the diffuse AOV unless it was specular-like.
  int first_specular = -1;
  bool aux_set = aux == nullptr;
  double travel = 0.0;
  const auto emit = [&](const Vec3& c) {
    l = l + c;
    if (aux && first_specular != 1) aux->diffuse = aux->diffuse + c;
  };
  double prev_pdf = 0.0;
  Vec3 prev_origin = ray.origin;
  bool from_camera = true;
  uint32_t depth = 0;
  MediumStack ms;  // pt/mod.rs:482
  ms.atmosphere = atmosphere;
  uint32_t skips = 0;  // hull crossings and cutout pass-throughs (pt/mod.rs:476)
  // The ray cone for texture footprints (pt/mod.rs:477-480, 687): its width
  // grows by t times the spread at every surface hit; the spread starts at
  // one pixel's angle (pt/mod.rs:76-77) and widens after rough bounces.
  double cone_width = 0.0;
  double cone_spread = std::tan(camera.fov * PI / 180.0 / 2.0) * 2.0 / (double)camera.height;
  while (depth < rp.max_depth) {
    if (rays) ++*rays;
    Intersection hit;
    const bool hit_found = intersect(ray, &hit);
    // The medium this segment runs through (pt/mod.rs:486-633).
    const int medium = ms.current();
    if (medium >= 0) {
      const Medium& md = media[(size_t)medium];
      const double t_limit = hit_found ? hit.t : 1e8;
      const double t_seg = std::fmin(t_limit, md.max_distance);
      size_t seg_li = 0;
      double seg_pick = 0.0;
      bool has_pick = false;
      if (!lights.empty()) {
        const Vec3 mid = ray.at(std::fmin(t_seg * 0.5, 50.0));
        has_pick = light_sampler.sample(mid, rng.next_f64(), &seg_li, &seg_pick);
      }
      const bool point_pick = has_pick && lights[seg_li].kind == RRHIP_LIGHT_POINT;
      // Equiangular single-scatter connection to a picked point light
      // (pt/mod.rs:501-551), MIS-weighted against distance sampling.
      if (point_pick && t_seg > 1e-6) {
        const Light& lt = lights[seg_li];
        double pdf_eq;
        const double t_eq = equiangular_sample(ray.origin, ray.direction, t_seg, lt.position, rng.next_f64(), &pdf_eq);
        if (pdf_eq > 1e-12) {
          const Vec3 p = ray.at(t_eq);
          const double rho = md.density_at(p);
          if (rho > 0.0) {
            const Vec3 tr = medium_transmittance(md, ray.origin, ray.direction, t_eq, rng);
            const Vec3 sigma_s = md.sigma_s * rho;
            const Vec3 to_l = lt.position - p;
            const double dist2 = std::fmax(to_l.length_squared(), 1e-12);
            const Vec3 wi = to_l.normalize();
            const Vec3 wo = -ray.direction.normalize();
            const double ph = hg_phase(md.g, -wo.dot(wi));
            const Vec3 vis = visibility(p, wi, std::sqrt(dist2) - 1e-3, rng, rays, medium, time);
            if (max_component(vis) > 0.0) {
              const double pdf_d = distance_pdf_approx(md, t_eq, t_seg);
              const double w_eq = pdf_eq / (pdf_eq + pdf_d);
              const Vec3 c = sigma_s * tr * vis * lt.radiance() * (ph * w_eq / (dist2 * pdf_eq * seg_pick));
              emit(beta * c);
            }
          }
        }
      }
      const MediumEvent ev = sample_distance(md, ray.origin, ray.direction, t_limit, beta, rng);
      beta = beta * ev.weight;
      if (ev.scatter) {
        const Vec3 p = ray.at(ev.t);
        if (!aux_set) {  // pt/mod.rs:559-570
          aux_set = true;
          const Vec3 st = md.sigma_t();
          aux->albedo = Vec3{md.sigma_s.x / std::fmax(st.x, 1e-9), md.sigma_s.y / std::fmax(st.y, 1e-9),
                             md.sigma_s.z / std::fmax(st.z, 1e-9)};
          aux->normal = -ray.direction.normalize();
          aux->depth = travel + ev.t;
          aux->id = 0;
        }
        if (first_specular < 0) first_specular = 0;
        if (max_component(md.emission) > 0.0) emit(beta * md.emission);
        const Vec3 wo = -ray.direction.normalize();
        if (has_pick) {
          const double g = md.g;
          const EvalFn eval = [&](const Vec3& wi, Vec3* f, double* pdf) {
            const double ph = hg_phase(g, -wo.dot(wi));
            *f = Vec3{ph, ph, ph};
            *pdf = ph;
          };
          // collision NEE; for a point light paired with equiangular
          // sampling by the MIS weight of pt/mod.rs:583-598
          double w_dist = 1.0;
          if (point_pick) {
            const double pdf_d = distance_pdf_approx(md, ev.t, t_seg);
            const double pdf_e = equiangular_pdf(ray.origin, ray.direction, t_seg, lights[seg_li].position, ev.t);
            w_dist = pdf_d / (pdf_d + pdf_e);
          }
          emit(beta * sample_light(lights[seg_li], p, Vec3::zero(), eval, seg_pick, rng, rays, medium, time) * w_dist);
        }
        const double u1 = rng.next_f64(), u2 = rng.next_f64();
        double pdf;
        const Vec3 wi = hg_sample(md.g, wo, u1, u2, &pdf);
        prev_pdf = pdf;
        cone_spread += 0.4;
        prev_origin = p;
        from_camera = false;
        ray = Ray::make(p, wi);
        ray.time = time;
        depth++;
        if (depth >= rp.rr_start) {
          double q = std::fmin(max_component(beta), 0.95);
          if (q <= 0.0 || rng.next_f64() > q) break;
          beta = beta / q;
        }
        continue;
      }
      if (!(max_component(beta) > 0.0)) break;
    }
    if (!hit_found) {
      if (dome >= 0) {  // pt/mod.rs:645-654 with dome_radiance / dome_pdf (375-387)
        const Light& d = lights[dome];
        const Vec3 dir = ray.direction.normalize();
        double weight = 1.0;
        if (!from_camera) {
          double pick = light_sampler.pmf(prev_origin, (size_t)dome);
          const double pdf_dome = d.env ? d.env->pdf(dir) : 1.0 / (4.0 * PI);
          weight = power_heuristic(prev_pdf, pdf_dome * pick);
        }
        emit(beta * (d.env ? d.env->eval(dir) * d.radiance() : d.radiance()) * weight);
      } else {
        emit(beta * background);
      }
      break;
    }
    const Material& mat = materials[hit.material];
    // Invisible volume hull: the crossing toggles the medium
    // (pt/mod.rs:662-683).
    if (mat.hull() && skips < 16) {
      skips++;
      travel += hit.t;
      if (hit.front_face)
        ms.push(mat.interior);
      else
        ms.pop();
      ray = Ray::make(hit.point + ray.direction.normalize() * ray_offset, ray.direction);
      ray.time = time;
      continue;
    }
    // The first real hit's AOVs, cutouts included (pt/mod.rs:688-700),
    // from the parameters the material's patterns give here.
    travel += hit.t;
    cone_width += hit.t * cone_spread;
    double presence = 1.0;
    const PbrParams pbr = shade_params(mat, hit, ray.direction, cone_width, &presence);
    if (!aux_set) {
      aux_set = true;
      const PbrParams& pb = pbr;
      const Vec3 a = pb.diffuse_color * pb.diffuse_gain + pb.specular_f0() * (pb.has_specular() ? 1.0 : 0.0) +
                     pb.refraction_color * pb.glass_gain;
      aux->albedo = Vec3{std::fmin(a.x, 1.0), std::fmin(a.y, 1.0), std::fmin(a.z, 1.0)};
      const Vec3 wo_a = -ray.direction.normalize();
      aux->normal = hit.normal.dot(wo_a) >= 0.0 ? hit.normal : -hit.normal;
      aux->depth = travel;
      aux->id = mat.id;
    }
    // Presence cutout: pass through with probability 1 - presence
    // (pt/mod.rs:702-710).
    if (presence < 1.0 && rng.next_f64() >= presence && skips < 16) {
      skips++;
      ray = Ray::make(hit.point + ray.direction.normalize() * ray_offset, ray.direction);
      ray.time = time;
      continue;
    }
    Vec3 wo = -ray.direction.normalize();
    bool entering = hit.front_face;
    Vec3 n = hit.normal.dot(wo) >= 0.0 ? hit.normal : -hit.normal;
    double eta_rel = entering ? pbr.glass_ior : 1.0 / pbr.glass_ior;
    Vec3 emission = mat.emission + pbr.glow;
    if (max_component(emission) > 0.0) {
      double weight = 1.0;
      if (!from_camera && mat.area_light >= 0) {
        const Light& lt = lights[mat.area_light];
        double pick = light_sampler.pmf(prev_origin, (size_t)mat.area_light);
        double pdf_light = light_pdf_solid_angle(lt, prev_origin, hit) * pick;
        weight = power_heuristic(prev_pdf, pdf_light);
      }
      emit(beta * emission * weight);
    }
    // Hair fibers scatter over the whole sphere with the fiber BSDF, in the
    // fiber frame of the curve hit (pt/mod.rs:738-800): NEE with f and its
    // pdf (no cosine), then a sampled continuation that leaves from the
    // side of the normal it points to.
    if (mat.has_hair && hit.tangent.length_squared() > 0.5) {
      const FiberFrame ff(hit.tangent, hit.normal, wo);
      const Vec3 wo_f = ff.to_local(wo);
      const HairParams& hp = mat.hair;
      if (!lights.empty()) {
        const EvalFn eval = [&](const Vec3& wi_world, Vec3* f, double* pdf) {
          const Vec3 wi_f = ff.to_local(wi_world);
          *f = hair::f(hp, wo_f, wi_f, ff.h);
          *pdf = hair::pdf(hp, wo_f, wi_f, ff.h);
        };
        size_t li;
        double pmf;
        if (light_sampler.sample(hit.point, rng.next_f64(), &li, &pmf))
          emit(beta * sample_light(lights[li], hit.point, n, eval, pmf, rng, rays, medium, time));
      }
      Vec3 wi_f, fv;
      double pv;
      if (!hair::sample(hp, wo_f, ff.h, rng, &wi_f, &fv, &pv)) break;
      const Vec3 wi_world = ff.to_world(wi_f);
      beta = beta * fv / pv;
      if (first_specular < 0) first_specular = 0;
      prev_pdf = pv;
      prev_origin = hit.point;
      from_camera = false;
      cone_spread += 0.4;
      const Vec3 side = wi_world.dot(hit.normal) >= 0.0 ? hit.normal : -hit.normal;
      ray = Ray::make(hit.point + side * ray_offset, wi_world);
      ray.time = time;
      depth++;
      if (depth >= rp.rr_start) {
        const double q = std::fmin(max_component(beta), 0.95);
        if (q <= 0.0 || rng.next_f64() > q) break;
        beta = beta / q;
      }
      continue;
    }
    Frame frame(n);
    Vec3 wo_l = frame.to_local(wo);
    double w5[5];
    if (!lights.empty() && pbr.lobe_weights(w5)) {
      size_t li;
      double pmf;
      if (light_sampler.sample(hit.point, rng.next_f64(), &li, &pmf)) {
        const EvalFn eval = [&](const Vec3& wi_world, Vec3* f, double* pdf) {
          Vec3 wi_l = frame.to_local(wi_world);
          if (wi_l.z <= 0.0) {
            *f = Vec3::zero();
            *pdf = 0.0;
            return;
          }
          bsdf_eval_pdf(pbr, wo_l, wi_l, eta_rel, f, pdf);
          *f = *f * wi_l.z;
        };
        Vec3 c = sample_light(lights[li], hit.point, n, eval, pmf, rng, rays, medium, time);
        emit(beta * c);
      }
    }
    BsdfSample s;
    if (!bsdf_sample(pbr, wo_l, eta_rel, rng, &s)) break;
    Vec3 wi_world = frame.to_world(s.wi);
    beta = beta * s.f * (std::fabs(s.wi.z) / s.pdf);
    if (first_specular < 0) first_specular = s.specular_lobe ? 1 : 0;
    prev_pdf = s.pdf;
    prev_origin = hit.point;
    from_camera = false;
    cone_spread += std::fmin(1.0 / (1.0 + s.pdf), 0.4);
    if (s.transmitted && mat.interior >= 0) {  // pt/mod.rs:969-975
      if (entering)
        ms.push(mat.interior);
      else
        ms.pop();
    }
    double offset = s.transmitted ? -ray_offset : ray_offset;
    ray = Ray::make(hit.point + n * offset, wi_world);
    ray.time = time;
    depth++;
    if (depth >= rp.rr_start) {
      double q = std::fmin(max_component(beta), 0.95);
      if (q <= 0.0 || rng.next_f64() > q) break;
      beta = beta / q;
    }
  }
  return l;
}

// ---- scene loading ------------------------------------------------------------------

static bool mat_inverse(const Mat4& m, Mat4* out) {
  // Gauss-Jordan with partial pivoting (f64).
  double a[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      a[i][j] = m.m[i][j];
      a[i][j + 4] = i == j ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; c++) {
    int piv = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    if (std::fabs(a[piv][c]) < 1e-300) return false;
    if (piv != c)
      for (int j = 0; j < 8; j++) std::swap(a[c][j], a[piv][j]);
    double inv = 1.0 / a[c][c];
    for (int j = 0; j < 8; j++) a[c][j] *= inv;
    for (int r = 0; r < 4; r++) {
      if (r == c) continue;
      double f = a[r][c];
      if (f != 0.0)
        for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) out->m[i][j] = a[i][j + 4];
  return true;
}

bool Mat4::inverse(Mat4& out) const { return mat_inverse(*this, &out); }

void Scene::load(const rrhip_scene_desc& d) {
  const rrhip_camera_desc& c = d.camera;
  camera.width = c.width;
  camera.height = c.height;
  camera.fov = c.fov_deg;
  camera.eye = v3(c.eye);
  camera.filter = (int)c.filter;
  camera.filter_width = c.filter_width > 0 ? c.filter_width : 1.0;
  camera.projection = (int)c.projection;
  camera.lens_radius = c.lens_radius;
  camera.focal_distance = c.focal_distance;
  camera.ortho_half[0] = c.ortho_half[0];
  camera.ortho_half[1] = c.ortho_half[1];
  camera.has_motion_inv = c.has_motion_inv != 0;
  camera.motion_inv = Mat4::identity();
  if (camera.has_motion_inv)
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 4; cc++) camera.motion_inv.m[r][cc] = c.motion_inv[4 * r + cc];
  // RenderMan frame of Camera::new (camera.rs:96-119) aimed along look_at
  camera.forward = (v3(c.look_at) - camera.eye).normalize();
  camera.right = v3(c.up).cross(camera.forward).normalize();
  camera.up = camera.forward.cross(camera.right).normalize();
  background = v3(d.background);
  media.clear();
  for (uint32_t i = 0; i < d.num_media; i++) {
    const rrhip_medium_desc& md = d.media[i];
    Medium m;
    m.sigma_a = v3(md.sigma_a);
    m.sigma_s = v3(md.sigma_s);
    m.emission = v3(md.emission);
    m.g = md.g;
    m.max_distance = md.max_distance > 0.0 ? md.max_distance : 1e300;
    if (md.procedural == RRHIP_MEDIUM_FBM) {  // the device keeps the parameters in f32
      m.fbm_octaves = std::max(md.octaves, 1u);
      m.fbm_frequency = (float)md.frequency;
      m.fbm_gain = (float)md.gain;
      m.fbm_lacunarity = (float)md.lacunarity;
      m.fbm_coverage = (float)md.coverage;
      m.fbm_sharpness = (float)md.sharpness;
    }
    if (md.nx > 0) {
      m.nx = md.nx;
      m.ny = md.ny;
      m.nz = md.nz;
      m.grid.assign(md.density, md.density + (size_t)md.nx * md.ny * md.nz);
      for (float& v : m.grid) v = std::min(std::max(v, 0.f), 1.f);
      for (int a = 0; a < 3; a++) {
        m.lo[a] = md.grid_lo[a];
        m.hi[a] = md.grid_hi[a];
      }
    }
    media.push_back(m);
  }
  atmosphere = d.num_media > 0 ? d.atmosphere : -1;

  materials.clear();
  for (uint32_t i = 0; i < d.num_materials; i++) {
    const rrhip_material_desc& m = d.materials[i];
    Material mt;
    PbrParams& p = mt.pbr;
    p.diffuse_gain = m.diffuse_gain;
    p.diffuse_color = v3(m.diffuse_color);
    p.diffuse_roughness = m.diffuse_roughness;
    p.specular_face_color = v3(m.specular_face_color);
    p.specular_edge_color = v3(m.specular_edge_color);
    p.specular_roughness = m.specular_roughness;
    p.specular_ior = m.specular_ior;
    p.clearcoat_gain = m.clearcoat_gain;
    p.clearcoat_roughness = m.clearcoat_roughness;
    p.fuzz_gain = m.fuzz_gain;
    p.fuzz_color = v3(m.fuzz_color);
    p.glass_gain = m.glass_gain;
    p.glass_ior = m.glass_ior;
    p.glass_roughness = m.glass_roughness;
    p.refraction_color = v3(m.refraction_color);
    p.glow = v3(m.glow);
    mt.emission = v3(m.emission);
    mt.area_light = m.area_light;
    mt.presence = 1.0 - m.cutout;
    mt.interior = (int)m.interior - 1;
    mt.id = m.id;
    if (m.hair) {  // hair.rs:40-53
      mt.has_hair = true;
      mt.hair = HairParams::make(v3(m.hair_sigma_a), m.hair_beta_m, m.hair_beta_n, m.hair_eta);
    }
    materials.push_back(mt);
  }

  lights.clear();
  dome = -1;
  for (uint32_t i = 0; i < d.num_lights; i++) {
    const rrhip_light_desc& ld = d.lights[i];
    Light l;
    l.kind = (int)ld.kind;
    l.intensity = ld.intensity;
    l.color = v3(ld.color);
    switch (l.kind) {
      case RRHIP_LIGHT_POINT:
        l.position = v3(ld.p);
        break;
      case RRHIP_LIGHT_DISTANT:
        l.direction = v3(ld.p).normalize();
        l.angular_radius = ld.radius;
        break;
      case RRHIP_LIGHT_RECT: {  // light.rs:99-116
        l.corner = v3(ld.p);
        l.edge1 = v3(ld.e1);
        l.edge2 = v3(ld.e2);
        Vec3 cr = l.edge1.cross(l.edge2);
        l.area = cr.length();
        l.normal = cr.normalize();
        break;
      }
      case RRHIP_LIGHT_SPHERE:
        l.center = v3(ld.p);
        l.radius = ld.radius;
        break;
      case RRHIP_LIGHT_DISK:  // light.rs:129-140
        l.center = v3(ld.p);
        l.e1 = v3(ld.e1);
        l.e2 = v3(ld.e2);
        l.normal = l.e1.cross(l.e2).normalize();
        l.area = PI * l.e1.length() * l.e2.length();
        break;
      case RRHIP_LIGHT_DOME:
        if (dome < 0) dome = (int)i;
        break;
    }
    lights.push_back(l);
  }
  if (d.env_rgb && dome >= 0) {
    auto env = std::make_shared<EnvMap>();
    env->from_data(d.env_width, d.env_height, d.env_rgb);
    lights[(size_t)dome].env = env;
  }
  light_sampler.build(lights);

  triangles.clear();
  for (uint32_t i = 0; i < d.num_triangles; i++) {
    const rrhip_triangle_desc& t = d.triangles[i];
    triangles.push_back({v3(t.v0), v3(t.v1), v3(t.v2), t.material});
  }
  spheres.clear();
  for (uint32_t i = 0; i < d.num_spheres; i++) {
    const rrhip_sphere_desc& s = d.spheres[i];
    spheres.push_back({v3(s.center), s.radius, s.material});
  }
  quadrics.clear();
  for (uint32_t i = 0; i < d.num_quadrics; i++) {
    const rrhip_quadric_desc& q = d.quadrics[i];
    QuadricObj o;
    o.kind = q.kind;
    o.material = q.material;
    o.radius = q.radius;
    o.zmin = q.zmin;
    o.zmax = q.zmax;
    o.height = q.height;
    o.thetamax = q.thetamax;
    for (int c = 0; c < 3; c++) {
      o.p1[c] = q.p1[c];
      o.p2[c] = q.p2[c];
    }
    o.transform = Mat4::identity();
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 4; c++) o.transform.m[r][c] = q.object_to_world[4 * r + c];
    if (!o.transform.inverse(o.inverse)) o.inverse = Mat4::identity();
    o.scale = o.transform.approx_scale();
    quadrics.push_back(o);
  }
  meshes.clear();
  meshes.resize(d.num_meshes);
  BvhParams bp;
  for (uint32_t i = 0; i < d.num_meshes; i++) {
    const rrhip_mesh_desc& md = d.meshes[i];
    meshes[i].mesh.build(md.positions, md.num_vertices, md.indices, md.num_triangles, bp, md.positions1);
    if (md.normals) meshes[i].normals.assign(md.normals, md.normals + 3 * md.num_vertices);
    if (md.st) meshes[i].st.assign(md.st, md.st + 2 * md.num_vertices);
  }
  curve_sets.clear();
  curve_sets.resize(d.num_curve_sets);
  for (uint32_t i = 0; i < d.num_curve_sets; i++) {
    const rrhip_curves_desc& cd = d.curve_sets[i];
    curve_sets[i].build(cd.p0, cd.p1, cd.v0, cd.v1, (size_t)cd.num_segments);
  }
  instances.clear();
  std::vector<Aabb> ib;
  for (uint32_t i = 0; i < d.num_instances; i++) {
    const rrhip_instance_desc& id = d.instances[i];
    InstanceObj in;
    in.mesh = id.mesh;
    in.material = id.material;
    in.curves = id.kind == RRHIP_INSTANCE_CURVES;
    in.transform = Mat4::identity();
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 4; cc++) in.transform.m[r][cc] = id.object_to_world[4 * r + cc];
    if (!in.transform.inverse(in.inverse)) in.inverse = Mat4::identity();
    in.moving = id.moving != 0;
    in.scale = in.transform.approx_scale();
    in.transform1 = Mat4::identity();
    if (in.moving)
      for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 4; cc++) in.transform1.m[r][cc] = id.object_to_world1[4 * r + cc];
    // mesh.rs:266-292 world bounds from the 8 transformed corners, at both
    // shutter ends when moving
    const Aabb& lb = in.curves ? curve_sets[in.mesh].local_bounds : meshes[in.mesh].mesh.local_bounds;
    Aabb wb;
    for (int e = 0; e < (in.moving ? 2 : 1); e++)
      for (int k = 0; k < 8; k++) {
        Vec3 corner{(k & 1) ? lb.max.x : lb.min.x, (k & 2) ? lb.max.y : lb.min.y,
                    (k & 4) ? lb.max.z : lb.min.z};
        wb.grow_point((e ? in.transform1 : in.transform).transform_point(corner));
      }
    in.world_bounds = wb;
    ib.push_back(wb);
    instances.push_back(in);
  }
  tlas.build(ib, bp);
  // textures (the caller's level 0; the chain rebuilt as tex.rs writes it),
  // pattern nodes and the materials' bindings
  textures.assign(d.num_textures, Texture{});
  for (uint32_t i = 0; i < d.num_textures; i++)
    textures[i].build_chain(d.textures[i].width, d.textures[i].height, d.textures[i].texels);
  patterns.assign(d.num_patterns, PatternNode{});
  for (uint32_t i = 0; i < d.num_patterns; i++) patterns[i] = pattern_from_desc(d.patterns[i]);
  for (uint32_t i = 0; i < d.num_bindings; i++)
    materials[d.bindings[i].material].bindings.push_back({(int)d.bindings[i].field, (int)d.bindings[i].pattern});
  has_motion = camera.has_motion_inv;  // scene_builder.rs:239-241
  for (const InstanceObj& in : instances) has_motion = has_motion || in.moving;
  for (const MeshObj& mo : meshes) has_motion = has_motion || !mo.mesh.positions1.empty();
}

}  // namespace orc
