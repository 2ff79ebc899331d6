// Host build of path-tracing scenes (see pt_scene.h).
#include "pt_scene.h"

#include "pt_material.h"

#include "bvh_build.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
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

// Matrix4::approx_scale: the cube root of the product of the linear part's
// column lengths (the uniform scale an st density divides by).
double approx_scale(const double* m) {
  const auto col = [&](int c) { return std::sqrt(m[c] * m[c] + m[4 + c] * m[4 + c] + m[8 + c] * m[8 + c]); };
  return std::max(std::cbrt(col(0) * col(1) * col(2)), 1e-12);
}

// The PbrParams of a material description (pbr.rs:11-45).
template <class T>
RawPbr<T> raw_pbr(const rrhip_material_desc& m) {
  RawPbr<T> r{};
  const auto cp = [](T* dst, const double* src) {
    for (int c = 0; c < 3; c++) dst[c] = (T)src[c];
  };
  r.diffuse_gain = (T)m.diffuse_gain;
  cp(r.diffuse_color, m.diffuse_color);
  r.diffuse_roughness = (T)m.diffuse_roughness;
  cp(r.specular_face_color, m.specular_face_color);
  cp(r.specular_edge_color, m.specular_edge_color);
  r.specular_roughness = (T)m.specular_roughness;
  r.specular_ior = (T)m.specular_ior;
  r.clearcoat_gain = (T)m.clearcoat_gain;
  r.clearcoat_roughness = (T)m.clearcoat_roughness;
  r.fuzz_gain = (T)m.fuzz_gain;
  cp(r.fuzz_color, m.fuzz_color);
  r.glass_gain = (T)m.glass_gain;
  r.glass_ior = (T)m.glass_ior;
  r.glass_roughness = (T)m.glass_roughness;
  cp(r.refraction_color, m.refraction_color);
  cp(r.glow, m.glow);
  cp(r.emission, m.emission);
  return r;
}

// Lobe constants (pt_material.h) plus the material's identity, medium,
// cutout and hair fiber parameters.
PtMaterial make_material(const rrhip_material_desc& m) {
  PtMaterial g{};
  lobe_constants(raw_pbr<double>(m), g);
  g.area_light = m.area_light;
  g.id = m.id;
  g.interior = (int32_t)m.interior - 1;
  g.presence = (float)std::min(std::max(1.0 - m.cutout, 0.0), 1.0);
  g.hull = !m.hair && g.interior >= 0 && !g.has_lobes &&
           std::max(g.emission[0], std::max(g.emission[1], g.emission[2])) <= 0.f;
  // Hair fiber (HairParams::new, hair.rs:40-53): roughness clamped to
  // [0.01, 1], longitudinal variances per lobe, azimuthal logistic scale
  // (sqrt(pi / 8) times a fit in beta_n).
  g.hair = m.hair ? 1u : 0u;
  if (m.hair) {
    const double bm = std::min(std::max(m.hair_beta_m, 0.01), 1.0);
    const double bn = std::min(std::max(m.hair_beta_n, 0.01), 1.0);
    const double root = 0.726 * bm + 0.812 * bm * bm + 3.7 * std::pow(bm, 20);
    const double v0 = root * root;
    g.hair_v[0] = (float)v0;
    g.hair_v[1] = (float)(0.25 * v0);
    g.hair_v[2] = (float)(4.0 * v0);
    g.hair_v[3] = (float)(4.0 * v0);
    g.hair_s = (float)(0.626657069 * (0.265 * bn + 1.194 * bn * bn + 5.372 * std::pow(bn, 22)));
    g.hair_eta = (float)m.hair_eta;
    put(g.hair_sigma_a, d3(m.hair_sigma_a));
  }
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

// Inverse of a 3x4 affine map (row-major), in f64.
bool affine_inverse(const double m[12], double inv[12]) {
  const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
  const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  if (!(std::fabs(det) > 1e-300) || !std::isfinite(det)) return false;
  const double r[9] = {A / det, -(b * i - c * h) / det, (b * f - c * e) / det,
                       B / det, (a * i - c * g) / det, -(a * f - c * d) / det,
                       C / det, -(a * h - b * g) / det, (a * e - b * d) / det};
  for (int row = 0; row < 3; row++) {
    for (int col = 0; col < 3; col++) inv[4 * row + col] = r[3 * row + col];
    inv[4 * row + 3] = -(r[3 * row] * m[3] + r[3 * row + 1] * m[7] + r[3 * row + 2] * m[11]);
  }
  return true;
}

// Per-mesh BLAS, instance records, TLAS over instance world bounds
// (render.rs: Mesh::new -> Bvh::build, Instance::build's 8-corner world
// bounds, Scene::build_tlas; src/geometry/mesh.rs:48-84, 266-292,
// src/scene/mod.rs:82-86). Node references are rebased into one array.
// BLAS over a curve set's segments (curves.rs:28-62). A thin segment
// running diagonally fills a small part of its box, so a segment longer
// than kPieceLen of its radius is entered as up to kMaxPieces pieces,
// each a box over a stretch of the axis, all referring to a copy of the
// same record (a ray may test a segment once per piece it enters; the
// test is the whole segment's, so hits are unchanged). A piece's box holds
// the spheres at its ends, their radii divided by cos(alpha) (sin(alpha) =
// |r0 - r1| / length: the cone's widest cross-section there) and the
// stretch widened by r sin(alpha) / length on both sides (where the cone
// touches a sphere), taken in f64 and rounded outward to f32.
constexpr double kPieceLen = 4.0;
constexpr int kMaxPieces = 8;

bool build_curve_blas(const rrhip_curves_desc& c, const BuildParams& bp, BuiltBvh* out,
                      std::vector<CurveRec>* recs, std::string* err) {
  const uint64_t n = c.num_segments;
  if (n && (!c.p0 || !c.p1 || !c.v0 || !c.v1)) return *err = "null curve arrays", false;
  std::vector<float> boxes;
  std::vector<uint32_t> seg_of;  // the segment of each piece
  boxes.reserve(6 * n);
  seg_of.reserve(n);
  for (uint64_t i = 0; i < n; i++) {
    const float* e0 = c.p0 + 4 * i;
    const float* e1 = c.p1 + 4 * i;
    const double r0 = e0[3], r1 = e1[3];
    if (!(r0 >= 0.0) || !(r1 >= 0.0) || !std::isfinite(r0) || !std::isfinite(r1))
      return *err = "curve radii must be finite and >= 0", false;
    double ax[3], len2 = 0.0;
    for (int a = 0; a < 3; a++) {
      ax[a] = (double)e1[a] - (double)e0[a];
      len2 += ax[a] * ax[a];
    }
    const double len = std::sqrt(len2), rmax = std::max(r0, r1);
    const double sin_a = len > 0.0 ? std::min(std::fabs(r0 - r1) / len, 1.0) : 0.0;
    const double cos_a = std::sqrt(std::max(1.0 - sin_a * sin_a, 0.0));
    int pieces = 1;
    if (rmax > 0.0 && cos_a > 0.05 && len > kPieceLen * rmax)
      pieces = (int)std::min<double>(kMaxPieces, std::ceil(len / (kPieceLen * rmax)));
    const double grow = len > 0.0 ? rmax * sin_a / len : 0.0;
    for (int j = 0; j < pieces; j++) {
      const double fa = pieces == 1 ? 0.0 : std::max(0.0, (double)j / pieces - grow);
      const double fb = pieces == 1 ? 1.0 : std::min(1.0, (double)(j + 1) / pieces + grow);
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (const double f : {fa, fb}) {
        // one piece: the end spheres as they are (curves.rs:29-58)
        const double r = pieces == 1 ? (f == 0.0 ? r0 : r1) : (r0 + (r1 - r0) * f) / cos_a;
        for (int a = 0; a < 3; a++) {
          const double p = f == 1.0 ? (double)e1[a] : (double)e0[a] + ax[a] * f;
          lo[a] = std::min(lo[a], p - r);
          hi[a] = std::max(hi[a], p + r);
        }
      }
      for (int a = 0; a < 3; a++) boxes.push_back(std::nextafter((float)lo[a], -INFINITY));
      for (int a = 0; a < 3; a++) boxes.push_back(std::nextafter((float)hi[a], INFINITY));
      seg_of.push_back((uint32_t)i);
    }
  }
  std::vector<uint32_t> order;
  if (!build_box_bvh(boxes.data(), seg_of.size(), bp, out, &order, err)) return false;
  recs->resize(order.size());
  for (size_t k = 0; k < order.size(); k++) {
    const uint64_t i = seg_of[order[k]];
    CurveRec& r = (*recs)[k];
    for (int a = 0; a < 3; a++) {
      r.p0[a] = c.p0[4 * i + a];
      r.p1[a] = c.p1[4 * i + a];
    }
    r.r0 = c.p0[4 * i + 3];
    r.r1 = c.p1[4 * i + 3];
    r.v0 = c.v0[i];
    r.v1 = c.v1[i];
    r.prim = (uint32_t)i;
    r.pad = 0;
  }
  return true;
}

bool build_instances(const rrhip_scene_desc& d, PtSceneHost* out, std::string* err) {
  if (d.num_instances == 0) return true;
  if (!d.instances || (d.num_meshes && !d.meshes)) return *err = "null instances or meshes", false;
  if (d.num_curve_sets && !d.curve_sets) return *err = "null curve sets", false;
  if (d.num_instances >= (1u << 28)) return *err = "too many instances (limit 2^28)", false;
  BuildParams bp;
  bp.bins = d.build.bins ? (int)d.build.bins : 32;
  bp.max_leaf = d.build.max_leaf ? (int)d.build.max_leaf : 4;
  // BLAS per mesh, then per curve set (geometry index num_meshes + set).
  const uint32_t ngeo = d.num_meshes + d.num_curve_sets;
  std::vector<BuiltBvh> blas(ngeo);
  std::vector<std::vector<CurveRec>> crecs(d.num_curve_sets);
  std::vector<std::array<float, 6>> local(ngeo);
  std::vector<bool> geo_empty(ngeo);
  int blas_levels = 0;
  for (uint32_t k = 0; k < ngeo; k++) {
    if (k < d.num_meshes) {
      const rrhip_mesh_desc& m = d.meshes[k];
      if (m.num_triangles && (!m.positions || !m.indices)) return *err = "null mesh arrays", false;
      if (!build_triangle_bvh(m.positions, m.num_vertices, m.indices, m.num_triangles, bp, &blas[k], err,
                              m.positions1))
        return false;
      geo_empty[k] = m.num_triangles == 0;
    } else {
      const rrhip_curves_desc& c = d.curve_sets[k - d.num_meshes];
      if (!build_curve_blas(c, bp, &blas[k], &crecs[k - d.num_meshes], err)) return false;
      geo_empty[k] = c.num_segments == 0;
    }
    for (int a = 0; a < 3; a++) {
      local[k][a] = blas[k].root_lo[a];
      local[k][3 + a] = blas[k].root_hi[a];
    }
    blas_levels = std::max(blas_levels, blas[k].nodes.empty() ? 0 : blas[k].depth + 1);
  }
  // Instances and their world bounds.
  std::vector<float> boxes(6 * (size_t)d.num_instances);
  std::vector<PtInstance> raw(d.num_instances);
  std::vector<float> raw_fwd(12 * (size_t)d.num_instances);
  for (uint32_t i = 0; i < d.num_instances; i++) {
    const rrhip_instance_desc& in = d.instances[i];
    if (in.kind != RRHIP_INSTANCE_MESH && in.kind != RRHIP_INSTANCE_CURVES)
      return *err = "instance kind must be RRHIP_INSTANCE_MESH or RRHIP_INSTANCE_CURVES", false;
    const bool curves = in.kind == RRHIP_INSTANCE_CURVES;
    if (in.mesh >= (curves ? d.num_curve_sets : d.num_meshes))
      return *err = curves ? "instance curve set out of range" : "instance mesh out of range", false;
    if (in.material >= d.num_materials) return *err = "instance material out of range", false;
    const uint32_t geo = curves ? d.num_meshes + in.mesh : in.mesh;
    if (curves) out->curves = 1;
    double inv[12];
    if (!affine_inverse(in.object_to_world, inv)) return *err = "instance transform is singular", false;
    PtInstance& g = raw[i];
    for (int k = 0; k < 12; k++) g.inv[k] = (float)inv[k];
    for (int k = 0; k < 12; k++) raw_fwd[12 * (size_t)i + k] = (float)in.object_to_world[k];
    g.mesh = geo;
    g.material = in.material;
    g.motion = 0xFFFFFFFFu;
    g.deform = (!curves && d.meshes[in.mesh].positions1) ? 1u : 0u;
    if (in.moving) {
      for (int k = 0; k < 12; k++)
        if (!std::isfinite(in.object_to_world1[k])) return *err = "instance transform1 must be finite", false;
      PtInstMotion mo;
      for (int k = 0; k < 12; k++) {
        mo.fwd0[k] = (float)in.object_to_world[k];
        mo.fwd1[k] = (float)in.object_to_world1[k];
      }
      g.motion = (uint32_t)out->inst_motion.size();
      out->inst_motion.push_back(mo);
      out->motion = 1;
    }
    // mesh.rs:274-288: world bounds from the 8 transformed corners, at both
    // shutter ends when moving (corners move linearly in between)
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const auto& lb = local[geo];
    const bool empty = geo_empty[geo];
    for (int e = 0; e < (in.moving ? 2 : 1); e++)
      for (int c = 0; c < 8 && !empty; c++) {
        const double p[3] = {(c & 1) ? lb[3] : lb[0], (c & 2) ? lb[4] : lb[1], (c & 4) ? lb[5] : lb[2]};
        for (int r = 0; r < 3; r++) {
          const double* M = (e ? in.object_to_world1 : in.object_to_world) + 4 * r;
          const double w = M[0] * p[0] + M[1] * p[1] + M[2] * p[2] + M[3];
          lo[r] = std::min(lo[r], w);
          hi[r] = std::max(hi[r], w);
        }
      }
    for (int r = 0; r < 3; r++) {
      if (empty) lo[r] = hi[r] = 0.0;
      // Round outward so f32 boxes still contain the f64 corners.
      boxes[6 * (size_t)i + r] = std::nextafter((float)lo[r], -INFINITY);
      boxes[6 * (size_t)i + 3 + r] = std::nextafter((float)hi[r], INFINITY);
    }
  }
  BuiltBvh tlas;
  std::vector<uint32_t> order;
  if (!build_box_bvh(boxes.data(), d.num_instances, bp, &tlas, &order, err)) return false;
  out->nodes = tlas.nodes;
  if (!tlas.nodes.empty()) {
    for (int a = 0; a < 3; a++) {
      out->lo[a] = std::min(out->lo[a], tlas.root_lo[a]);
      out->hi[a] = std::max(out->hi[a], tlas.root_hi[a]);
    }
  }
  for (uint32_t k : order) {
    out->inst_scale.push_back((float)approx_scale(d.instances[k].object_to_world));
    out->insts.push_back(raw[k]);
    out->inst_fwd.insert(out->inst_fwd.end(), raw_fwd.begin() + 12 * (size_t)k, raw_fwd.begin() + 12 * (size_t)k + 12);
  }
  const int tlas_levels = tlas.nodes.empty() ? 0 : tlas.depth + 1;
  // Concatenate the BLASes, rebasing child references (leaves of meshes
  // into the triangle records, of curve sets into the segment records).
  out->meshes.resize(ngeo);
  for (uint32_t k = 0; k < ngeo; k++) {
    const bool curves = k >= d.num_meshes;
    const uint32_t node_base = (uint32_t)out->nodes.size();
    const uint32_t rec_base = (uint32_t)(curves ? out->csegs.size() : out->btris.size());
    for (Node2 n : blas[k].nodes) {
      n.ref0 += n.cnt0 ? rec_base : node_base;
      n.ref1 += n.cnt1 ? rec_base : node_base;
      out->nodes.push_back(n);
    }
    PtMeshInfo& mi = out->meshes[k];
    mi.root = blas[k].nodes.empty() ? 0xFFFFFFFFu : node_base;
    mi.curves = curves ? 1u : 0u;
    out->mesh_st.push_back(0xFFFFFFFFu);
    if (!curves && d.meshes[k].st) {
      out->mesh_st.back() = (uint32_t)(out->st.size() / 2);
      out->st.insert(out->st.end(), d.meshes[k].st, d.meshes[k].st + 2 * d.meshes[k].num_vertices);
    }
    if (curves) {
      const auto& cr = crecs[k - d.num_meshes];
      out->csegs.insert(out->csegs.end(), cr.begin(), cr.end());
      mi.idx_base = 0;
      mi.nrm_base = 0xFFFFFFFFu;
      continue;
    }
    out->btris.insert(out->btris.end(), blas[k].tris.begin(), blas[k].tris.end());
    const rrhip_mesh_desc& m = d.meshes[k];
    for (const TriRec& r : blas[k].tris) {  // shutter-close records (rest copies for static meshes)
      TriRec r1 = r;
      if (m.positions1) {
        const uint32_t* ix = m.indices + 3 * (size_t)r.prim;
        const float* a = m.positions1 + 3 * (size_t)ix[0];
        const float* b = m.positions1 + 3 * (size_t)ix[1];
        const float* c = m.positions1 + 3 * (size_t)ix[2];
        r1.v0x = a[0], r1.v0y = a[1], r1.v0z = a[2];
        r1.v1x = b[0], r1.v1y = b[1], r1.v1z = b[2];
        r1.v2x = c[0], r1.v2y = c[1], r1.v2z = c[2];
      }
      out->btris1.push_back(r1);
    }
    out->deform |= m.positions1 ? 1u : 0u;
    mi.idx_base = (uint32_t)out->mesh_idx.size();
    out->mesh_idx.insert(out->mesh_idx.end(), m.indices, m.indices + 3 * m.num_triangles);
    if (m.normals) {
      mi.nrm_base = (uint32_t)(out->mesh_nrm.size() / 3);
      out->mesh_nrm.insert(out->mesh_nrm.end(), m.normals, m.normals + 3 * m.num_vertices);
    } else {
      mi.nrm_base = 0xFFFFFFFFu;
    }
  }
  if (out->nodes.size() >= (1ull << 31) || out->btris.size() >= (1ull << 28) || out->csegs.size() >= (1ull << 28))
    return *err = "instanced scene too large", false;
  if (out->deform)
    out->motion = 1;  // scene_builder.rs:239-241: deforming meshes make the scene a motion scene
  else
    out->btris1.clear();
  // One entry per TLAS level (two at the last), one per BLAS level, and
  // the remainder of a TLAS leaf re-pushed while its first instance runs.
  out->stack_need = (uint32_t)(tlas_levels + blas_levels + 2);
  return true;
}

// Textures, pattern nodes and bindings (Scene::patterns,
// Material::pattern_bindings): the mip chains (tex.rs:83-108, the 2x2 box
// filter with odd sizes clamping the second tap, f32 sums in its order),
// node records, and per binding the program the shade kernels run: the
// node's input tree in post-order, as pattern::eval walks it (an input
// deeper than 16 levels or naming no node evaluates to magenta).
bool build_patterns(const rrhip_scene_desc& d, PtSceneHost* out, std::string* err) {
  if (d.num_bindings == 0) return true;
  if (!d.bindings) return *err = "num_bindings > 0 with null bindings", false;
  if ((d.num_patterns && !d.patterns) || (d.num_textures && !d.textures))
    return *err = "null pattern or texture arrays", false;
  for (uint32_t i = 0; i < d.num_textures; i++) {
    const rrhip_texture_desc& t = d.textures[i];
    if (t.width == 0 || t.height == 0 || !t.texels || (uint64_t)t.width * t.height > (1ull << 28))
      return *err = "texture needs 1..2^28 texels and data", false;
    PtTexture g{};
    g.first_level = (uint32_t)out->tlev.size();
    g.w = t.width;
    g.h = t.height;
    std::vector<float> cur(t.texels, t.texels + 3 * (size_t)t.width * t.height), next;
    uint32_t w = t.width, h = t.height;
    while (true) {
      out->tlev.push_back(PtTexLevel{w, h, (uint64_t)out->texels.size()});
      for (size_t k = 0; k < (size_t)w * h; k++)
        out->texels.push_back(make_float4(cur[3 * k], cur[3 * k + 1], cur[3 * k + 2], 0.f));
      if (w == 1 && h == 1) break;
      const uint32_t nw = std::max(w / 2, 1u), nh = std::max(h / 2, 1u);
      next.assign(3 * (size_t)nw * nh, 0.f);
      for (uint32_t y = 0; y < nh; y++) {
        const uint32_t y0 = std::min(2 * y, h - 1), y1 = std::min(2 * y + 1, h - 1);
        for (uint32_t x = 0; x < nw; x++) {
          const uint32_t x0 = std::min(2 * x, w - 1), x1 = std::min(2 * x + 1, w - 1);
          for (int c = 0; c < 3; c++) {
            const float sum = cur[3 * ((size_t)y0 * w + x0) + c] + cur[3 * ((size_t)y0 * w + x1) + c] +
                              cur[3 * ((size_t)y1 * w + x0) + c] + cur[3 * ((size_t)y1 * w + x1) + c];
            next[3 * ((size_t)y * nw + x) + c] = sum * 0.25f;
          }
        }
      }
      cur.swap(next);
      w = nw;
      h = nh;
    }
    g.nlevels = (uint32_t)out->tlev.size() - g.first_level;
    out->texs.push_back(g);
  }
  const int32_t ntex = (int32_t)d.num_textures;
  for (uint32_t i = 0; i < d.num_patterns; i++) {
    const rrhip_pattern_desc& p = d.patterns[i];
    if (p.kind > RRHIP_PATTERN_TRIPLANAR) return *err = "unknown pattern kind", false;
    if (p.wrap > RRHIP_WRAP_BLACK) return *err = "unknown texture wrap mode", false;
    if (p.texture < -1 || p.texture >= ntex) return *err = "pattern texture out of range", false;
    if (p.num_stops && (!p.positions || !p.colors)) return *err = "ramp stops need positions and colors", false;
    if (p.kind == RRHIP_PATTERN_FRACTAL && p.octaves > 64)  // each octave is a Perlin lookup per shading point
      return *err = "fractal octaves must be at most 64", false;
    PtPattern g{};
    g.kind = p.kind;
    g.wrap = p.wrap;
    g.texture = p.texture;
    g.udim = -1;
    if (p.udim) {
      g.udim = (int32_t)out->udim.size();
      for (int k = 0; k < 100; k++) {
        if (p.udim[k] < -1 || p.udim[k] >= ntex) return *err = "UDIM tile texture out of range", false;
        out->udim.push_back(p.udim[k]);
      }
    }
    for (int k = 0; k < 2; k++) g.scale[k] = (float)p.scale[k];
    g.frequency = (float)p.frequency;
    g.gain = (float)p.gain;
    g.lacunarity = (float)p.lacunarity;
    g.gamma = (float)p.gamma;
    g.saturation = (float)p.saturation;
    g.octaves = p.octaves;
    put(g.color_a, d3(p.color_a));
    put(g.color_b, d3(p.color_b));
    put(g.cc_gain, d3(p.cc_gain));
    put(g.cc_offset, d3(p.cc_offset));
    g.use_t = p.use_t ? 1u : 0u;
    g.ramp = (uint32_t)out->ramp.size();
    g.nstops = p.num_stops;
    for (uint32_t k = 0; k < p.num_stops; k++)
      out->ramp.push_back(make_float4((float)p.positions[k], (float)p.colors[3 * k], (float)p.colors[3 * k + 1],
                                      (float)p.colors[3 * k + 2]));
    out->pats.push_back(g);
  }
  // Programs: post-order over the input tree with the value-stack depth.
  const auto flatten = [&](uint32_t root, std::vector<PtPatOp>* prog, int* max_sp) {
    int sp = 0;
    const auto push_const = [&](const double c[3]) {
      prog->push_back(PtPatOp{-1, {(float)c[0], (float)c[1], (float)c[2]}});
      *max_sp = std::max(*max_sp, ++sp);
    };
    const double magenta[3] = {1.0, 0.0, 1.0};
    std::function<void(int32_t, int)> emit = [&](int32_t idx, int depth) {
      if (prog->size() > kPatProgMax) return;
      if (depth > 16 || idx < 0 || idx >= (int32_t)d.num_patterns) return push_const(magenta);
      const rrhip_pattern_desc& p = d.patterns[idx];
      const int nin = p.kind == RRHIP_PATTERN_MIX ? 3 : p.kind == RRHIP_PATTERN_COLOR_CORRECT ? 1 : 0;
      for (int k = 0; k < nin; k++) {
        if (p.input[k] >= 0)
          emit(p.input[k], depth + 1);
        else
          push_const(p.constant[k]);
      }
      prog->push_back(PtPatOp{idx, {0.f, 0.f, 0.f}});
      sp += 1 - nin;
      *max_sp = std::max(*max_sp, sp);
    };
    emit((int32_t)root, 0);
  };
  std::vector<std::vector<uint32_t>> per_mat(d.num_materials);
  for (uint32_t i = 0; i < d.num_bindings; i++) {
    const rrhip_binding_desc& b = d.bindings[i];
    if (b.material >= d.num_materials) return *err = "binding material out of range", false;
    if (b.field > RRHIP_BIND_PRESENCE) return *err = "unknown binding field", false;
    if (b.pattern >= d.num_patterns) return *err = "binding pattern out of range", false;
    if (b.field == RRHIP_BIND_PRESENCE) out->patterns |= 2u;  // shadow rays walk and evaluate it at blockers
    per_mat[b.material].push_back(i);
  }
  for (uint32_t m = 0; m < d.num_materials; m++) {
    if (per_mat[m].empty()) continue;
    if (per_mat[m].size() > 15) return *err = "UNSUPPORTED: more than 15 pattern bindings on one material", false;
    out->materials[m].bind = ((uint32_t)out->binds.size() << 4) | (uint32_t)per_mat[m].size();
    for (uint32_t i : per_mat[m]) {
      const rrhip_binding_desc& b = d.bindings[i];
      std::vector<PtPatOp> prog;
      int max_sp = 0;
      flatten(b.pattern, &prog, &max_sp);
      if (prog.size() > kPatProgMax || max_sp > kPatStack)
        return *err = "UNSUPPORTED: pattern graph too large for the device evaluator (" +
                      std::to_string(kPatProgMax) + " nodes, " + std::to_string(kPatStack) + " pending inputs)",
               false;
      out->binds.push_back(PtBinding{b.field, (uint32_t)out->prog.size(), (uint32_t)prog.size(), 0});
      out->prog.insert(out->prog.end(), prog.begin(), prog.end());
    }
  }
  if (out->binds.size() >= (1u << 27)) return *err = "too many pattern bindings", false;
  for (uint32_t m = 0; m < d.num_materials; m++) out->raw_pbr.push_back(raw_pbr<float>(d.materials[m]));
  out->pixel_spread = (float)(std::tan(d.camera.fov_deg * kPi / 180.0 / 2.0) * 2.0 / (double)d.camera.height);
  out->patterns |= 1u;
  if (out->patterns & 2u) {  // a bound presence makes the scene transparent (pt/mod.rs:702-710, 1294)
    out->transparent = 1;
    if (d.num_media > 255) return *err = "at most 255 media with interior media or cutouts (8-bit medium stack)", false;
  }
  return true;
}

}  // namespace

PtMaterial pt_material(const rrhip_material_desc& m) { return make_material(m); }

bool build_pt_scene(const rrhip_scene_desc& d, PtSceneHost* out, std::string* err) {
  *out = PtSceneHost{};
  if (d.num_materials && !d.materials) return *err = "null materials", false;
  if (d.num_lights && !d.lights) return *err = "null lights", false;
  if (d.num_triangles && !d.triangles) return *err = "null triangles", false;
  if (d.num_spheres && !d.spheres) return *err = "null spheres", false;
  for (uint32_t i = 0; i < d.num_materials; i++) {
    const rrhip_material_desc& m = d.materials[i];
    if (m.area_light >= (int32_t)d.num_lights) return *err = "material area_light out of range", false;
    if (m.interior > d.num_media) return *err = "material interior medium out of range", false;
    if (!(m.cutout >= 0.0 && m.cutout <= 1.0)) return *err = "material cutout must be in [0, 1]", false;
    if (m.hair > 1) return *err = "material hair must be 0 or 1", false;
    if (m.hair && (!finite3(m.hair_sigma_a) || !(m.hair_sigma_a[0] >= 0.0 && m.hair_sigma_a[1] >= 0.0 &&
                                                 m.hair_sigma_a[2] >= 0.0) ||
                   !std::isfinite(m.hair_beta_m) || !std::isfinite(m.hair_beta_n) || !(m.hair_eta > 1.0) ||
                   !std::isfinite(m.hair_eta)))
      return *err = "hair material: sigma_a must be finite and >= 0, roughness finite, eta > 1", false;
    out->materials.push_back(make_material(m));
    out->hair |= m.hair ? 1u : 0u;
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
    // Rounding bound of the light points the kernels sample in f32 (the
    // point p + e1 u + e2 v with u, v in [0, 1], or a sphere's surface
    // point): gamma(4) of the summed magnitudes along the normal, gamma(8)
    // of the centre and radius for spheres.
    const auto gamma = [](int n) { return n * 0x1p-24 / (1.0 - n * 0x1p-24); };
    if (l.kind == RRHIP_LIGHT_RECT || l.kind == RRHIP_LIGHT_DISK) {
      double e = 0.0;
      for (int a = 0; a < 3; a++)
        e += std::fabs((double)g.n[a]) *
             (std::fabs((double)g.p[a]) + std::fabs((double)g.e1[a]) + std::fabs((double)g.e2[a]));
      g.err = std::nextafter((float)(gamma(4) * e), INFINITY);
    } else if (l.kind == RRHIP_LIGHT_SPHERE) {
      const double c = std::max({std::fabs((double)g.p[0]), std::fabs((double)g.p[1]), std::fabs((double)g.p[2])});
      g.err = std::nextafter((float)(gamma(8) * (c + g.radius)), INFINITY);
    }
    out->lights.push_back(g);
  }
  // Dome environment map: CDFs as EnvMap::from_data (envmap.rs:36-54).
  if (d.env_rgb) {
    if (out->dome < 0) return *err = "an environment map needs a dome light", false;
    if (d.env_width == 0 || d.env_height == 0 || (uint64_t)d.env_width * d.env_height > (1ull << 28))
      return *err = "environment map size must be 1..2^28 texels", false;
    const uint32_t w = d.env_width, h = d.env_height;
    out->env_w = w;
    out->env_h = h;
    out->env_rgb.assign(d.env_rgb, d.env_rgb + 3ull * w * h);
    for (float v : out->env_rgb)
      if (!(v >= 0.f) || !std::isfinite(v)) return *err = "environment map texels must be finite and >= 0", false;
    out->env_marg.assign(h + 1, 0.0);
    out->env_cond.assign((size_t)(w + 1) * h, 0.0);
    for (uint32_t y = 0; y < h; y++) {
      const double sin_t = std::sin(kPi * (y + 0.5) / h);
      const float* row = &out->env_rgb[3ull * y * w];
      const size_t base = (size_t)y * (w + 1);
      double acc = 0.0;
      for (uint32_t x = 0; x < w; x++) {
        acc += (0.2126 * row[3 * x] + 0.7152 * row[3 * x + 1] + 0.0722 * row[3 * x + 2]) * sin_t;
        out->env_cond[base + x + 1] = acc;
      }
      out->env_marg[y + 1] = out->env_marg[y] + acc;
    }
    out->env_total = std::max(out->env_marg[h], 1e-12);
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
      // a dome with a map: its sin-weighted mean luminance times the
      // intensity (light_sampler.rs:111-118, envmap.rs:61-66)
      const bool mapped = (int32_t)i == out->dome && !out->env_rgb.empty();
      const double base = mapped ? out->env_total / ((double)out->env_w * out->env_h * (2.0 / kPi)) *
                                       d.lights[i].intensity
                                 : lum(mul(d3(d.lights[i].color), d.lights[i].intensity));
      inf_power.push_back(std::max(base, 1e-9) * 4.0 * kPi);
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
  if (d.num_quadrics && !d.quadrics) return *err = "null quadrics", false;
  for (uint32_t i = 0; i < d.num_quadrics; i++) {
    const rrhip_quadric_desc& q = d.quadrics[i];
    if (q.kind > RRHIP_QUADRIC_TORUS) return *err = "unknown quadric kind", false;
    if (!finite3(q.p1) || !finite3(q.p2)) return *err = "quadric parameters must be finite", false;
    if (q.material >= d.num_materials) return *err = "quadric material out of range", false;
    const double par[5] = {q.radius, q.zmin, q.zmax, q.height, q.thetamax};
    for (double v : par)
      if (!std::isfinite(v)) return *err = "quadric parameters must be finite", false;
    for (double v : q.object_to_world)
      if (!std::isfinite(v)) return *err = "quadric transform must be finite", false;
    if (!(q.thetamax > 0.0)) return *err = "quadric thetamax must be > 0", false;
    if (q.kind == RRHIP_QUADRIC_CONE && !(q.height != 0.0)) return *err = "cone height must be nonzero", false;
    const double* m = q.object_to_world;
    // A whole sphere (no z or theta cut) under a similarity transform
    // (rotation, uniform scale, translation) is the loose sphere with the
    // translation as centre and radius * scale: the same hit, outward normal
    // and front face (dot(d, n) < 0), tested without the quadric solvers,
    // so scenes whose quadrics are all such spheres (RIB's Sphere request)
    // keep the trace kernels without them. Only the order of exact ties
    // with other loose objects can differ.
    if (q.kind == RRHIP_QUADRIC_SPHERE && q.radius > 0.0 && q.zmin <= -q.radius && q.zmax >= q.radius &&
        q.thetamax >= 360.0) {
      double gram[3][3];
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) gram[x][y] = m[x] * m[y] + m[4 + x] * m[4 + y] + m[8 + x] * m[8 + y];
      const double s2 = gram[0][0];
      bool similar = s2 > 0.0;
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) similar &= std::fabs(gram[x][y] - (x == y ? s2 : 0.0)) <= 1e-12 * s2;
      if (similar) {
        PtSphere g{};
        g.c[0] = (float)m[3];
        g.c[1] = (float)m[7];
        g.c[2] = (float)m[11];
        g.r = (float)(q.radius * std::sqrt(s2));
        g.material = q.material;
        out->spheres.push_back(g);
        continue;
      }
    }
    // Matrix4::inverse().unwrap_or(identity) (sphere.rs / cylinder.rs constructors)
    const double a = m[0], b = m[1], c = m[2], dd = m[4], e = m[5], f = m[6], g = m[8], h = m[9], ii = m[10];
    const double A = e * ii - f * h, B = -(dd * ii - f * g), C = dd * h - e * g;
    const double det = a * A + b * B + c * C;
    double inv[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    if (std::fabs(det) > 1e-300) {
      const double r = 1.0 / det;
      inv[0] = A * r;
      inv[1] = -(b * ii - c * h) * r;
      inv[2] = (b * f - c * e) * r;
      inv[4] = B * r;
      inv[5] = (a * ii - c * g) * r;
      inv[6] = -(a * f - c * dd) * r;
      inv[8] = C * r;
      inv[9] = -(a * h - b * g) * r;
      inv[10] = (a * e - b * dd) * r;
      for (int row = 0; row < 3; row++)
        inv[4 * row + 3] = -(inv[4 * row] * m[3] + inv[4 * row + 1] * m[7] + inv[4 * row + 2] * m[11]);
    }
    PtQuadric g2{};
    g2.kind = q.kind;
    g2.material = q.material;
    g2.radius = (float)q.radius;
    g2.zmin = (float)q.zmin;
    g2.zmax = (float)q.zmax;
    g2.height = (float)q.height;
    g2.thetamax = (float)q.thetamax;
    if (q.kind == RRHIP_QUADRIC_CONE) g2.k = (float)((q.radius / q.height) * (q.radius / q.height));
    if (q.kind == RRHIP_QUADRIC_PARABOLOID && q.zmax != 0.0) g2.k = (float)(q.radius * q.radius / q.zmax);
    if (q.kind == RRHIP_QUADRIC_HYPERBOLOID) {  // extra.rs Hyperboloid::profile
      const double dz = q.p2[2] - q.p1[2];
      if (std::fabs(dz) >= 1e-6) {
        const double sx = (q.p2[0] - q.p1[0]) / dz, sy = (q.p2[1] - q.p1[1]) / dz;
        const double cx = q.p1[0] - q.p1[2] * sx, cy = q.p1[1] - q.p1[2] * sy;
        g2.pa = (float)(sx * sx + sy * sy);
        g2.pb = (float)(2.0 * (cx * sx + cy * sy));
        g2.pc = (float)(cx * cx + cy * cy);
      }
      g2.z1 = (float)q.p1[2];
      g2.z2 = (float)q.p2[2];
      g2.r1 = (float)(q.p1[0] * q.p1[0] + q.p1[1] * q.p1[1]);
      g2.r2 = (float)(q.p2[0] * q.p2[0] + q.p2[1] * q.p2[1]);
    }
    for (int k = 0; k < 12; k++) {
      g2.inv[k] = (float)inv[k];
      g2.fwd[k] = (float)m[k];
    }
    out->quadrics.push_back(g2);
    const double sx = q.p2[0] - q.p1[0], sy = q.p2[1] - q.p1[1], sz = q.p2[2] - q.p1[2];
    out->quad_st.push_back(make_float2((float)approx_scale(m), (float)std::max(std::sqrt(sx * sx + sy * sy + sz * sz), 1e-12)));
  }
  put(out->background, d3(d.background));
  // Media (medium.rs:28-65)
  if (d.num_media > 0 && !d.media) return *err = "num_media > 0 with null media", false;
  if (d.atmosphere >= (int32_t)d.num_media || d.atmosphere < -1) return *err = "atmosphere index out of range", false;
  for (uint32_t i = 0; i < d.num_media; i++) {
    const rrhip_medium_desc& md = d.media[i];
    if (!finite3(md.sigma_a) || !finite3(md.sigma_s) || !finite3(md.emission) || !std::isfinite(md.g))
      return *err = "medium parameters must be finite", false;
    if (!(md.g > -1.0 && md.g < 1.0)) return *err = "medium g must be in (-1, 1)", false;
    for (int c = 0; c < 3; c++)
      if (md.sigma_a[c] < 0.0 || md.sigma_s[c] < 0.0) return *err = "medium coefficients must be >= 0", false;
    PtMedium g{};
    put(g.sigma_a, d3(md.sigma_a));
    put(g.sigma_s, d3(md.sigma_s));
    const D3 st = add(d3(md.sigma_a), d3(md.sigma_s));
    put(g.sigma_t, st);
    g.majorant = (float)std::max(std::max(st.x, st.y), st.z);
    g.g = (float)md.g;
    put(g.emission, d3(md.emission));
    g.max_distance = md.max_distance > 0.0 ? (float)md.max_distance : INFINITY;
    std::vector<float> grid;
    if (md.procedural > RRHIP_MEDIUM_FBM) return *err = "unknown medium procedural density", false;
    if (md.procedural == RRHIP_MEDIUM_FBM) {
      if (md.nx > 0) return *err = "a medium has either a density grid or a procedural field", false;
      if (md.octaves > 64) return *err = "medium fBm: at most 64 octaves", false;
      const double f[5] = {md.frequency, md.gain, md.lacunarity, md.coverage, md.sharpness};
      for (double x : f)
        if (!std::isfinite(x)) return *err = "medium fBm parameters must be finite", false;
      g.fbm_octaves = std::max(md.octaves, 1u);  // fbm runs at least one octave (displace.rs:86)
      g.fbm_frequency = (float)md.frequency;
      g.fbm_gain = (float)md.gain;
      g.fbm_lacunarity = (float)md.lacunarity;
      g.fbm_coverage = (float)md.coverage;
      g.fbm_sharpness = (float)md.sharpness;
    }
    if (md.nx > 0) {
      if (md.nx < 2 || md.ny < 2 || md.nz < 2 || !md.density)
        return *err = "medium density grid needs >= 2 samples per axis and data", false;
      if ((uint64_t)md.nx * md.ny * md.nz > (1ull << 31)) return *err = "medium density grid too large", false;
      for (int a = 0; a < 3; a++)
        if (!(md.grid_hi[a] > md.grid_lo[a]) || !std::isfinite(md.grid_lo[a]) || !std::isfinite(md.grid_hi[a]))
          return *err = "medium grid bounds must be finite and non-empty", false;
      grid.assign(md.density, md.density + (size_t)md.nx * md.ny * md.nz);
      for (float& v : grid) {
        if (!std::isfinite(v)) return *err = "medium density must be finite", false;
        v = std::min(std::max(v, 0.f), 1.f);  // density fields are clamped to [0, 1] (medium.rs:48-52)
      }
      g.nx = md.nx;
      g.ny = md.ny;
      g.nz = md.nz;
      const uint32_t n[3] = {md.nx, md.ny, md.nz};
      for (int a = 0; a < 3; a++) {
        g.lo[a] = (float)md.grid_lo[a];
        g.hi[a] = (float)md.grid_hi[a];
        g.scale[a] = (float)((double)(n[a] - 1) / (md.grid_hi[a] - md.grid_lo[a]));
      }
    }
    out->media.push_back(g);
    out->media_grids.push_back(std::move(grid));
  }
  out->atmosphere = d.num_media > 0 ? d.atmosphere : -1;
  for (const PtMaterial& m : out->materials) out->transparent |= (m.interior >= 0 || m.presence < 1.f) ? 1u : 0u;
  if (out->transparent && d.num_media > 255)
    return *err = "at most 255 media with interior media or cutouts (8-bit medium stack)", false;
  for (int a = 0; a < 3; a++) {
    out->lo[a] = INFINITY;
    out->hi[a] = -INFINITY;
  }
  auto grow = [&](D3 p, double r) {
    const double v[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; a++) {
      out->lo[a] = std::min(out->lo[a], (float)(v[a] - r));
      out->hi[a] = std::max(out->hi[a], (float)(v[a] + r));
    }
  };
  for (uint32_t i = 0; i < d.num_triangles; i++) {
    grow(d3(d.triangles[i].v0), 0.0);
    grow(d3(d.triangles[i].v1), 0.0);
    grow(d3(d.triangles[i].v2), 0.0);
  }
  for (uint32_t i = 0; i < d.num_spheres; i++) grow(d3(d.spheres[i].center), d.spheres[i].radius);
  for (uint32_t i = 0; i < d.num_quadrics; i++) {  // the object-space extent's corners, placed
    const rrhip_quadric_desc& q = d.quadrics[i];
    double r = std::fabs(q.radius);
    double zlo = q.zmin, zhi = q.zmax;
    if (q.kind == RRHIP_QUADRIC_CONE || q.kind == RRHIP_QUADRIC_DISK) {
      zlo = std::min(0.0, q.height);
      zhi = std::max(0.0, q.height);
    }
    if (q.kind == RRHIP_QUADRIC_HYPERBOLOID) {
      r = std::sqrt(std::max(q.p1[0] * q.p1[0] + q.p1[1] * q.p1[1], q.p2[0] * q.p2[0] + q.p2[1] * q.p2[1]));
      zlo = std::min(q.p1[2], q.p2[2]);
      zhi = std::max(q.p1[2], q.p2[2]);
    }
    if (q.kind == RRHIP_QUADRIC_TORUS) {
      r = std::fabs(q.radius) + std::fabs(q.height);
      zlo = -std::fabs(q.height);
      zhi = std::fabs(q.height);
    }
    if (q.kind == RRHIP_QUADRIC_SPHERE) {
      zlo = std::max(zlo, -r);
      zhi = std::min(zhi, r);
    }
    const double* m = q.object_to_world;
    for (int cx = 0; cx < 2; cx++)
      for (int cy = 0; cy < 2; cy++)
        for (int cz = 0; cz < 2; cz++) {
          const double x = cx ? r : -r, y = cy ? r : -r, z = cz ? zhi : zlo;
          grow(D3{m[0] * x + m[1] * y + m[2] * z + m[3], m[4] * x + m[5] * y + m[6] * z + m[7],
                  m[8] * x + m[9] * y + m[10] * z + m[11]},
               0.0);
        }
  }
  if (!build_instances(d, out, err)) return false;
  for (int a = 0; a < 3; a++) {
    if (!(out->hi[a] >= out->lo[a])) out->lo[a] = out->hi[a] = 0.f;
    if (out->hi[a] - out->lo[a] < 1e-6f) {  // flat scenes: give the axis a unit extent
      out->lo[a] -= 0.5f;
      out->hi[a] += 0.5f;
    }
  }

  // Camera: RenderMan frame of Camera::new (camera.rs:96-119), aimed.
  const rrhip_camera_desc& c = d.camera;
  if (c.width == 0 || c.height == 0) return *err = "camera needs width and height", false;
  if (c.projection > 1) return *err = "camera projection must be 0 (perspective) or 1 (orthographic)", false;
  if (c.projection == 0 && !(c.fov_deg > 0.0 && c.fov_deg < 180.0)) return *err = "camera fov must be in (0, 180)", false;
  if (!(c.lens_radius >= 0.0 && c.lens_radius < 1e30)) return *err = "camera lens_radius must be finite and >= 0", false;
  if (c.lens_radius > 0.0 && !(c.focal_distance > 0.0 && c.focal_distance < 1e30))
    return *err = "camera focal_distance must be > 0 with a lens", false;
  if (c.projection == 1 && !(c.ortho_half[0] > 0.0 && c.ortho_half[1] > 0.0))
    return *err = "orthographic camera needs positive ortho_half", false;
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
  cam.projection = c.projection;
  cam.lens_radius = (float)c.lens_radius;
  cam.focal_distance = (float)c.focal_distance;
  cam.ortho_w = (float)c.ortho_half[0];
  cam.ortho_h = (float)c.ortho_half[1];
  cam.has_motion_inv = c.has_motion_inv ? 1u : 0u;
  for (int k = 0; k < 12; k++) {
    if (c.has_motion_inv && !std::isfinite(c.motion_inv[k])) return *err = "camera motion_inv must be finite", false;
    cam.motion_inv[k] = (float)c.motion_inv[k];
  }
  if (cam.has_motion_inv) out->motion = 1;
  return build_patterns(d, out, err);
}

}  // namespace rrhip
