// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// f64 restatement of render.rs's CPU path tracer for the subset the
// backend renders: src/raytracer/pt/mod.rs (trace_full with media, volume
// hulls, presence cutouts and hair fibers; no subsurface or textures),
// src/raytracer/pt/{bxdf,volume,hair}.rs, src/scene/medium.rs,
// src/scene/{pbr,light,light_sampler,camera}.rs, the Sphere / Triangle /
// quadric / curve / Instance intersectors and Scene::intersect.
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "../include/rrhip.h"
#include "orc_accel.hpp"
#include "orc_hair.hpp"
#include "orc_pattern.hpp"
#include "orc_sampling.hpp"

namespace orc {

// pbr.rs:11-75 (fields used by the lobes; glow/presence/subsurface off)
struct PbrParams {
  double diffuse_gain = 1.0;
  Vec3 diffuse_color{0.5, 0.5, 0.5};
  double diffuse_roughness = 0.0;
  Vec3 specular_face_color{0, 0, 0};
  Vec3 specular_edge_color{1, 1, 1};
  double specular_roughness = 0.2;
  double specular_ior = 1.5;
  double clearcoat_gain = 0.0;
  double clearcoat_roughness = 0.1;
  double fuzz_gain = 0.0;
  Vec3 fuzz_color{1, 1, 1};
  double glass_gain = 0.0;
  double glass_ior = 1.5;
  double glass_roughness = 0.05;
  Vec3 refraction_color{1, 1, 1};
  Vec3 glow{0, 0, 0};

  Vec3 specular_f0() const;
  bool has_specular() const;
  bool lobe_weights(double w[5]) const;
};

struct Material {
  PbrParams pbr;
  Vec3 emission{0, 0, 0};
  int area_light = -1;
  double presence = 1.0;  // PbrParams::presence (pbr.rs:39)
  int interior = -1;      // Material::interior (material.rs:37)
  uint32_t id = 0;        // Material::id, the id AOV's object id (material.rs:39)
  bool has_hair = false;  // Material::hair (material.rs): curve hits scatter with the fiber BSDF
  HairParams hair;
  // Material::pattern_bindings (material.rs:96-102): (BoundField, node)
  std::vector<std::pair<int, int>> bindings;
  // pt/mod.rs:662-667: an interior medium and no lobes or emission.
  bool hull() const;
};

// pt/mod.rs:33-69: the media a path is inside of, innermost on top; the
// atmosphere is the implicit bottom.
struct MediumStack {
  int stack[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int depth = 0;
  int atmosphere = -1;
  int current() const { return depth > 0 ? stack[depth - 1] : atmosphere; }
  void push(int m) {
    if (depth < 8) stack[depth++] = m;
  }
  void pop() {
    if (depth > 0) depth--;
  }
};

// bxdf.rs:16-39
struct Frame {
  Vec3 t, b, n;
  explicit Frame(const Vec3& n_);
  Vec3 to_local(const Vec3& w) const { return {w.dot(t), w.dot(b), w.dot(n)}; }
  Vec3 to_world(const Vec3& w) const { return t * w.x + b * w.y + n * w.z; }
};

// film.rs:11-24: the first real hit's albedo, normal, distance travelled
// and object id, and the radiance whose first continuation was not
// specular-like.
struct AuxSample {
  Vec3 albedo{0, 0, 0}, normal{0, 0, 0};
  double depth = 0.0;
  uint32_t id = 0;
  Vec3 diffuse{0, 0, 0};
};

struct BsdfSample {
  Vec3 wi, f;
  double pdf = 0;
  bool transmitted = false;
  bool specular_lobe = false;
};

void bsdf_eval_pdf(const PbrParams& p, const Vec3& wo, const Vec3& wi, double eta_rel, Vec3* f,
                   double* pdf);
bool bsdf_sample(const PbrParams& p, const Vec3& wo, double eta_rel, Pcg32& rng, BsdfSample* s);

// light.rs:7-46 (no IES, no HDRI)
// envmap.rs:10-148: a lat-long environment map with 2D-CDF importance
// sampling of sin-weighted luminance (nearest texel).
struct EnvMap {
  uint32_t width = 0, height = 0;
  std::vector<float> data;  // rgb, row-major from the +y pole
  std::vector<double> marginal, conditional;
  double total = 0.0;
  void from_data(uint32_t w, uint32_t h, const float* rgb);
  double mean_luminance() const;
  Vec3 eval(const Vec3& dir) const;
  // (dir, radiance, pdf_sa)
  void sample(double u1, double u2, Vec3* dir, Vec3* radiance, double* pdf) const;
  double pdf(const Vec3& dir) const;

 private:
  void texel_of(const Vec3& dir, uint32_t* x, uint32_t* y) const;
  Vec3 texel(uint32_t x, uint32_t y) const;
};

struct Light {
  int kind = RRHIP_LIGHT_POINT;
  double intensity = 1.0;
  Vec3 color{1, 1, 1};
  Vec3 position;                    // point
  Vec3 direction;                   // distant (normalised travel direction)
  double angular_radius = 0;        // distant
  Vec3 corner, edge1, edge2;        // rect
  Vec3 center, e1, e2;              // sphere / disk
  Vec3 normal;                      // rect / disk
  double area = 0;                  // rect / disk
  double radius = 0;                // sphere
  std::shared_ptr<const EnvMap> env;  // dome: lat-long map (light.rs:41)
  Vec3 radiance() const { return color * intensity; }
};

struct LightBvhNode {  // light_sampler.rs:19-29 (f32 storage)
  float min[3];
  float power;
  float max[3];
  uint32_t parent;
  uint32_t a, b;
  uint32_t pad[2];
};

class LightSampler {  // light_sampler.rs:33-312
 public:
  std::vector<LightBvhNode> nodes;
  std::vector<uint32_t> light_leaf, infinite;
  std::vector<double> infinite_power;
  double infinite_total = 0, p_infinite = 0;

  void build(const std::vector<Light>& lights);
  bool sample(const Vec3& p, double u, size_t* light, double* pmf) const;
  double pmf(const Vec3& p, size_t light) const;
};

struct Camera {  // camera.rs:69-179 (perspective, pinhole)
  uint32_t width = 0, height = 0;
  double fov = 45;
  Vec3 eye, forward, up, right;
  int filter = 0;  // 0 box, 1 triangle, 2 gaussian
  double filter_width = 1.0;
  int projection = 0;  // 0 perspective, 1 orthographic
  double lens_radius = 0, focal_distance = 1;
  double ortho_half[2] = {1, 1};
  bool has_motion_inv = false;
  Mat4 motion_inv;
  double filter_1d(double u) const;
  Ray generate_ray_lens(double px, double py, double lu, double lv) const;
  Ray apply_motion(const Ray& r, double t) const;  // camera.rs:134-146
};

struct Intersection {
  double t = 0;
  Vec3 point, normal;
  bool front_face = true;
  uint32_t material = 0;
  Vec3 tangent;  // curve hits: the world tangent (zero for surfaces)
  // intersection.rs:18-24: surface parameters, st units per world unit,
  // world-space gradients of s and t (zero where the primitive has none)
  double st[2] = {0, 0};
  double st_density = 0;
  Vec3 st_grad[2];
};

struct TriangleObj {
  Vec3 v0, v1, v2;
  uint32_t material;
  bool intersect(const Ray& ray, Intersection* out) const;
};

struct SphereObj {
  Vec3 center;
  double radius;
  uint32_t material;
  bool intersect(const Ray& ray, Intersection* out) const;
};

// RenderMan quadrics (sphere.rs, cylinder.rs, cone.rs, extra.rs Disk /
// Paraboloid) with their object-to-world transform.
struct QuadricObj {
  uint32_t kind = 0, material = 0;
  double scale = 1.0;  // transform.approx_scale(), for st density
  double radius = 0, zmin = 0, zmax = 0, height = 0, thetamax = 360;
  double p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};  // hyperboloid
  Mat4 transform, inverse;
  bool intersect(const Ray& ray, Intersection* out) const;
  void quadric_st(const Vec3& p, Intersection* out) const;
};

struct MeshObj {
  Mesh mesh;
  std::vector<float> normals;  // empty: geometric
  std::vector<float> st;       // per-vertex st pairs, empty: none (mesh.rs:20-34)
};

struct InstanceObj {
  uint32_t mesh, material;  // mesh: index of the mesh, or of the curve set when `curves`
  bool curves = false;      // GeomKind::Curves (mesh.rs:223-240)
  Mat4 transform, inverse;
  bool moving = false;
  Mat4 transform1;  // at shutter close
  Aabb world_bounds;
  double scale = 1.0;  // Instance::scale = transform.approx_scale() (mesh.rs:236-238, matrix.rs:246-251)
  // mesh.rs:294-305: the inverse transform at shutter time `time`.
  Mat4 inverse_at(double time) const;
};

// medium.rs:28-65 with the density as a vertex grid, or the procedural
// fBm cloud field DensityField::Fbm (fbm_octaves > 0), or neither
// (homogeneous).
struct Medium {
  Vec3 sigma_a, sigma_s, emission;
  double g = 0;
  double max_distance = 1e300;
  uint32_t nx = 0, ny = 0, nz = 0;
  std::vector<float> grid;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  uint32_t fbm_octaves = 0;
  double fbm_frequency = 0, fbm_gain = 0, fbm_lacunarity = 0, fbm_coverage = 0, fbm_sharpness = 0;
  bool heterogeneous() const { return nx > 0 || fbm_octaves > 0; }
  Vec3 sigma_t() const { return sigma_a + sigma_s; }
  double majorant() const;
  double density_at(const Vec3& p) const;
};

// volume.rs:13-20
struct MediumEvent {
  bool scatter = false;
  double t = 0;
  Vec3 weight{1, 1, 1};
};
MediumEvent sample_distance(const Medium& m, const Vec3& o, const Vec3& d, double t_max, const Vec3& beta,
                            Pcg32& rng);
Vec3 medium_transmittance(const Medium& m, const Vec3& o, const Vec3& d, double dist, Pcg32& rng);
double hg_phase(double g, double cos_theta);
Vec3 hg_sample(double g, const Vec3& wo, double u1, double u2, double* pdf);
// volume.rs:195-240 (Kulla-Fajardo equiangular sampling towards a point,
// its pdf, and the approximate pdf of distance sampling used against it).
double equiangular_sample(const Vec3& o, const Vec3& d, double t_max, const Vec3& light, double u, double* pdf);
double equiangular_pdf(const Vec3& o, const Vec3& d, double t_max, const Vec3& light, double t);
double distance_pdf_approx(const Medium& m, double t, double t_max);

struct RenderParams {
  uint32_t max_depth = 10;
  uint32_t rr_start = 3;
  double firefly_clamp = 60.0;
  uint64_t sample_base = 0;
  double ray_offset = 1e-4;  // pt/mod.rs:30
};

// One pattern node from its ABI description (pattern.rs:42-68 fields).
PatternNode pattern_from_desc(const rrhip_pattern_desc& pd);

class Scene {
 public:
  Camera camera;
  std::vector<TriangleObj> triangles;
  std::vector<SphereObj> spheres;
  std::vector<QuadricObj> quadrics;
  std::vector<MeshObj> meshes;
  std::vector<CurveSet> curve_sets;
  std::vector<InstanceObj> instances;
  Bvh tlas;
  std::vector<Light> lights;
  std::vector<Material> materials;
  LightSampler light_sampler;
  Vec3 background{0, 0, 0};
  int dome = -1;
  std::vector<Medium> media;
  int atmosphere = -1;
  std::vector<Texture> textures;
  std::vector<PatternNode> patterns;
  bool has_motion = false;  // scene_builder.rs:239-241 (moving instances or camera)
  // Push of ray origins off surfaces (pt/mod.rs:30: 1e-4). Tests set a
  // negligible push to model error-bounded origins (f64 rounding of hit
  // points near unit scale is ~1e-16).
  double ray_offset = 1e-4;

  void load(const rrhip_scene_desc& d);
  bool intersect(const Ray& ray, Intersection* out) const;
  // Closest-hit and shadow-ray counts accumulate into these when non-null;
  // aux (optional) receives the sample's AOVs (film.rs:11-24).
  Vec3 trace(Ray ray, Pcg32& rng, const RenderParams& rp, uint64_t* rays, AuxSample* aux = nullptr) const;
  Ray camera_ray(uint32_t x, uint32_t y, Pcg32& rng) const;

 private:
  bool instance_hit(const InstanceObj& inst, const Ray& ray, double t_max, double* t_param,
                    Vec3* world_normal, Vec3* world_tangent = nullptr, Intersection* st_out = nullptr) const;
  // shade_params (pt/mod.rs:391-439): the material's PbrParams with its
  // pattern bindings applied at the hit (cone_width: the ray cone's width
  // there), and the presence they give.
  PbrParams shade_params(const Material& m, const Intersection& hit, const Vec3& ray_dir, double cone_width,
                         double* presence) const;
  // eval(wi, f, pdf): f * cos (or the phase value) towards world wi and the
  // pdf of the scattering strategy, for MIS.
  using EvalFn = std::function<void(const Vec3&, Vec3*, double*)>;
  // `medium`: the medium the shading point is in (-1: none); `time`: the
  // path's shutter time.
  Vec3 sample_light(const Light& light, const Vec3& p, const Vec3& n, const EvalFn& eval, double pick_pmf,
                    Pcg32& rng, uint64_t* rays, int medium, double time) const;
  Vec3 area_contribution(const Vec3& p, const Vec3& n, const Vec3& sample_point,
                         const Vec3& light_normal, double area, const Light& light, const EvalFn& eval,
                         double pick_pmf, Pcg32& rng, uint64_t* rays, int medium, double time) const;
  // Fractional visibility (pt/mod.rs:1238-1306): walks the blockers,
  // multiplying each segment's medium transmittance; hulls push / pop
  // media, cutouts attenuate by 1 - presence, anything else blocks.
  Vec3 visibility(const Vec3& origin, const Vec3& dir, double remaining, Pcg32& rng, uint64_t* rays,
                  int medium, double time) const;
};

}  // namespace orc
