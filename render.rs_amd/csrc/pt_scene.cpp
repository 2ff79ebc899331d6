// Host build of path-tracing scenes (see pt_scene.h).
#include "pt_scene.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace rrhip {
namespace {

constexpr double kPi = 3.14159265358979323846;

struct D3 {
  double x, y, z;
};
D3 d3(const double a[3]) { return {a[0], a[1], a[2]}; }
D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double length(D3 a) { return std::sqrt(dot(a, a)); }
D3 normalize(D3 a) {  // vector.rs:31-38
  double l = length(a);
  return l > 1e-6 ? mul(a, 1.0 / l) : a;
}
double lum(D3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
void put(float* dst, D3 v) {
  dst[0] = (float)v.x;
  dst[1] = (float)v.y;
  dst[2] = (float)v.z;
}
bool finite3(const double a[3]) { return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]); }

double alpha_of(double r) { return std::min(std::max(r * r, 2.5e-5), 1.0); }  // bxdf.rs:187-189

// pbr.rs:84-93, 191-202; bxdf.rs:137-143, 145-171 (lobe constants)
PtMaterial make_material(const rrhip_material_desc& m) {
  PtMaterial g{};
  D3 face = d3(m.specular_face_color);
  D3 f0{0, 0, 0};
  if (lum(face) > 1e-6) {
    f0 = face;
  } else if (m.specular_ior > 1.0) {
    double r = (m.specular_ior - 1.0) / (m.specular_ior + 1.0);
    f0 = {r * r, r * r, r * r};
  }
  double wd = m.diffuse_gain * lum(d3(m.diffuse_color));
  double ws = std::sqrt(lum(f0));
  double wc = m.clearcoat_gain * 0.2;
  double wf = m.fuzz_gain * lum(d3(m.fuzz_color)) * 0.3;
  double wg = m.glass_gain;
  double total = wd + ws + wc + wf + wg;
  g.has_lobes = total > 1e-9 ? 1u : 0u;
  if (g.has_lobes) {
    g.w[0] = (float)(wd / total);
    g.w[1] = (float)(ws / total);
    g.w[2] = (float)(wc / total);
    g.w[3] = (float)(wf / total);
    g.w[4] = (float)(wg / total);
  }
  double spec_take = lum(f0);
  double coat_take = 0.04 * m.clearcoat_gain;
  double under = std::min(std::max((1.0 - spec_take) * (1.0 - coat_take), 0.0), 1.0);
  put(g.diffuse, mul(d3(m.diffuse_color), m.diffuse_gain * under / kPi));
  double s2 = m.diffuse_roughness * m.diffuse_roughness;
  g.on_a = (float)(1.0 - s2 / (2.0 * (s2 + 0.33)));
  g.on_b = (float)(0.45 * s2 / (s2 + 0.09));
  put(g.spec_f0, f0);
  put(g.spec_f90, d3(m.specular_edge_color));
  g.spec_alpha = (float)alpha_of(m.specular_roughness);
  g.coat_f0 = (float)(0.04 * m.clearcoat_gain);
  g.coat_f90 = (float)m.clearcoat_gain;
  g.coat_alpha = (float)alpha_of(m.clearcoat_roughness);
  put(g.fuzz, mul(d3(m.fuzz_color), m.fuzz_gain / kPi));
  g.glass_scale = (float)(m.glass_gain * under);
  put(g.refraction, d3(m.refraction_color));
  g.glass_alpha = (float)alpha_of(m.glass_roughness);
  g.glass_ior = (float)m.glass_ior;
  put(g.emission, add(d3(m.emission), d3(m.glow)));
  g.area_light = m.area_light;
  return g;
}

// light_sampler.rs:46-70 (power of a finite light; false for infinite)
bool finite_info(const rrhip_light_desc& l, D3* center, double* radius, double* power) {
  double pw = std::max(lum(mul(d3(l.color), l.intensity)), 1e-9);
  switch (l.kind) {
    case RRHIP_LIGHT_POINT:
      *center = d3(l.p);
      *radius = 0.0;
      *power = pw * 4.0 * kPi;
      return true;
    case RRHIP_LIGHT_RECT: {
      D3 e1 = d3(l.e1), e2 = d3(l.e2);
      *center = add(add(d3(l.p), mul(e1, 0.5)), mul(e2, 0.5));
      *radius = (length(e1) + length(e2)) * 0.5;
      *power = pw * length(cross(e1, e2)) * kPi;
      return true;
    }
    case RRHIP_LIGHT_SPHERE:
      *center = d3(l.p);
      *radius = l.radius;
      *power = pw * 4.0 * kPi * l.radius * l.radius * kPi;
      return true;
    case RRHIP_LIGHT_DISK: {
      D3 e1 = d3(l.e1), e2 = d3(l.e2);
      *center = d3(l.p);
      *radius = length(e1);
      *power = pw * kPi * length(e1) * length(e2) * kPi;
      return true;
    }
    default:
      return false;
  }
}

struct Entry {
  uint32_t light;
  double c[3];
  double radius, power;
};

uint32_t light_build(const std::vector<Entry>& e, std::vector<size_t>& order, size_t start, size_t end,
                     uint32_t parent, PtSceneHost* out) {
  const uint32_t index = (uint32_t)out->lnodes.size();
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  double power = 0.0;  // f64 power sums
  for (size_t k = start; k < end; k++) {
    const Entry& b = e[order[k]];
    for (int a = 0; a < 3; a++) {
      mn[a] = std::min(mn[a], (float)(b.c[a] - b.radius));
      mx[a] = std::max(mx[a], (float)(b.c[a] + b.radius));
    }
    power += b.power;
  }
  LightNode n{};
  for (int a = 0; a < 3; a++) {
    n.min[a] = mn[a];
    n.max[a] = mx[a];
  }
  n.power = (float)power;
  n.parent = parent;
  n.a = 0;
  n.b = 0xFFFFFFFFu;
  out->lnodes.push_back(n);
  if (end - start == 1) {
    const uint32_t light = e[order[start]].light;
    out->lnodes[index].a = light;
    out->light_leaf[light] = index;
    return index;
  }
  int axis = 0;
  float widest = -1.0f;
  for (int a = 0; a < 3; a++) {
    if (mx[a] - mn[a] > widest) {
      widest = mx[a] - mn[a];
      axis = a;
    }
  }
  const size_t mid = (start + end) / 2;
  std::stable_sort(order.begin() + start, order.begin() + end,
                   [&](size_t x, size_t y) { return e[x].c[axis] < e[y].c[axis]; });
  const uint32_t a = light_build(e, order, start, mid, index, out);
  const uint32_t b = light_build(e, order, mid, end, index, out);
  out->lnodes[index].a = a;
  out->lnodes[index].b = b;
  return index;
}

}  // namespace

bool build_pt_scene(const rrhip_scene_desc& d, PtSceneHost* out, std::string* err) {
  *out = PtSceneHost{};
  if (d.num_materials && !d.materials) return *err = "null materials", false;
  if (d.num_lights && !d.lights) return *err = "null lights", false;
  if (d.num_triangles && !d.triangles) return *err = "null triangles", false;
  if (d.num_spheres && !d.spheres) return *err = "null spheres", false;
  for (uint32_t i = 0; i < d.num_materials; i++) {
    const rrhip_material_desc& m = d.materials[i];
    if (m.area_light >= (int32_t)d.num_lights) return *err = "material area_light out of range", false;
    out->materials.push_back(make_material(m));
  }
  // Lights (light.rs:59-151)
  for (uint32_t i = 0; i < d.num_lights; i++) {
    const rrhip_light_desc& l = d.lights[i];
    if (l.kind > RRHIP_LIGHT_DOME) return *err = "unknown light kind", false;
    if (!finite3(l.p) || !finite3(l.e1) || !finite3(l.e2) || !finite3(l.color))
      return *err = "light parameters must be finite", false;
    PtLight g{};
    g.kind = l.kind;
    put(g.radiance, mul(d3(l.color), l.intensity));
    switch (l.kind) {
      case RRHIP_LIGHT_DISTANT:
        put(g.p, normalize(d3(l.p)));
        g.radius = (float)l.radius;
        g.cos_max = (float)std::cos(l.radius);
        break;
      case RRHIP_LIGHT_RECT: {
        D3 c = cross(d3(l.e1), d3(l.e2));
        put(g.p, d3(l.p));
        put(g.e1, d3(l.e1));
        put(g.e2, d3(l.e2));
        put(g.n, normalize(c));
        g.area = (float)length(c);
        break;
      }
      case RRHIP_LIGHT_DISK: {
        put(g.p, d3(l.p));
        put(g.e1, d3(l.e1));
        put(g.e2, d3(l.e2));
        put(g.n, normalize(cross(d3(l.e1), d3(l.e2))));
        g.area = (float)(kPi * length(d3(l.e1)) * length(d3(l.e2)));
        break;
      }
      case RRHIP_LIGHT_SPHERE:
        put(g.p, d3(l.p));
        g.radius = (float)l.radius;
        break;
      case RRHIP_LIGHT_DOME:
        if (out->dome < 0) out->dome = (int32_t)i;
        break;
      default:
        put(g.p, d3(l.p));
        break;
    }
    out->lights.push_back(g);
  }
  // Light sampler (light_sampler.rs:97-142)
  std::vector<Entry> finite;
  std::vector<double> inf_power;
  for (uint32_t i = 0; i < d.num_lights; i++) {
    D3 c;
    double r, pw;
    if (finite_info(d.lights[i], &c, &r, &pw)) {
      finite.push_back({i, {c.x, c.y, c.z}, r, pw});
    } else {
      out->inf_lights.push_back(i);
      inf_power.push_back(std::max(lum(mul(d3(d.lights[i].color), d.lights[i].intensity)), 1e-9) * 4.0 * kPi);
    }
  }
  double acc = 0.0;
  for (double p : inf_power) {
    acc += p;
    out->inf_cdf.push_back(acc);
  }
  out->inf_total = acc;
  out->light_leaf.assign(d.num_lights, 0xFFFFFFFFu);
  if (!finite.empty()) {
    std::vector<size_t> order(finite.size());
    std::iota(order.begin(), order.end(), 0);
    light_build(finite, order, 0, finite.size(), 0xFFFFFFFFu, out);
  }
  double finite_total = 0.0;
  for (const Entry& e : finite) finite_total += e.power;
  if (out->inf_lights.empty())
    out->p_infinite = 0.0;
  else if (finite.empty())
    out->p_infinite = 1.0;
  else
    out->p_infinite = std::min(std::max(out->inf_total / (out->inf_total + finite_total), 0.1), 0.9);

  // Loose objects
  for (uint32_t i = 0; i < d.num_triangles; i++) {
    const rrhip_triangle_desc& t = d.triangles[i];
    if (t.material >= d.num_materials) return *err = "triangle material out of range", false;
    if (!finite3(t.v0) || !finite3(t.v1) || !finite3(t.v2)) return *err = "triangle vertices must be finite", false;
    PtTriangle g{};
    D3 v0 = d3(t.v0), e1 = sub(d3(t.v1), v0), e2 = sub(d3(t.v2), v0);
    put(g.v0, v0);
    put(g.e1, e1);
    put(g.e2, e2);
    put(g.n, normalize(cross(e1, e2)));
    g.material = t.material;
    out->tris.push_back(g);
  }
  for (uint32_t i = 0; i < d.num_spheres; i++) {
    const rrhip_sphere_desc& s = d.spheres[i];
    if (s.material >= d.num_materials) return *err = "sphere material out of range", false;
    if (!finite3(s.center) || !(s.radius > 0.0)) return *err = "sphere needs a finite centre and radius > 0", false;
    PtSphere g{};
    put(g.c, d3(s.center));
    g.r = (float)s.radius;
    g.material = s.material;
    out->spheres.push_back(g);
  }
  put(out->background, d3(d.background));

  // Camera: RenderMan frame of Camera::new (camera.rs:96-119), aimed.
  const rrhip_camera_desc& c = d.camera;
  if (c.width == 0 || c.height == 0) return *err = "camera needs width and height", false;
  if (!(c.fov_deg > 0.0 && c.fov_deg < 180.0)) return *err = "camera fov must be in (0, 180)", false;
  PtCamera& cam = out->cam;
  cam.width = c.width;
  cam.height = c.height;
  cam.filter = c.filter <= 2 ? c.filter : 0;
  cam.filter_width = (float)(c.filter_width > 0 ? c.filter_width : 1.0);
  D3 eye = d3(c.eye);
  D3 fwd = normalize(sub(d3(c.look_at), eye));
  D3 right = normalize(cross(d3(c.up), fwd));
  D3 up = normalize(cross(fwd, right));
  put(cam.eye, eye);
  put(cam.forward, fwd);
  put(cam.right, right);
  put(cam.up, up);
  double hh = std::tan(c.fov_deg * kPi / 180.0 / 2.0);
  cam.half_h = (float)hh;
  cam.half_w = (float)(hh * (double)c.width / (double)c.height);
  return true;
}

}  // namespace rrhip
