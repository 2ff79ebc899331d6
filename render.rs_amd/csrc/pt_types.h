// Device layout of path-tracing scenes and wavefront path state.
//
// Materials are render.rs's PbrParams with every per-material constant of
// the lobes precomputed (the GPU export precomputes the same kind of
// values, src/raytracer/metal/gpu_scene.rs:323-387); lights carry their
// derived normal / area; the light BVH node is render.rs's 48-byte
// LightBvhNode (src/scene/light_sampler.rs:19-29), built on the host with
// f64 power sums.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace rrhip {

struct alignas(16) PtMaterial {
  float w[5];        // lobe weights: diffuse, specular, clearcoat, fuzz, glass (pbr.rs:191-202)
  uint32_t has_lobes;
  float diffuse[3];  // diffuse_color * diffuse_gain * under_layer_scale / pi
  float on_a, on_b;  // Oren-Nayar A, B from diffuse_roughness
  float spec_f0[3];
  float spec_f90[3];
  float spec_alpha;
  float coat_f0, coat_f90, coat_alpha;
  float fuzz[3];     // fuzz_color * fuzz_gain / pi
  float glass_scale; // glass_gain * under_layer_scale
  float refraction[3];
  float glass_alpha;
  float glass_ior;
  float emission[3]; // Material::emission + glow
  int32_t area_light;
  float nrc_albedo[3];  // radiance-cache input: summed lobe albedo
  float nrc_rough;      // radiance-cache input: roughness of the strongest glossy lobe
  uint32_t sort_key;    // shading class for sorted queues: 1 + bitmask of the active lobes
  int32_t interior;     // medium inside closed surfaces of this material (-1: none)
  float presence;       // probability that a ray stops at the surface (pbr.rs:39)
  uint32_t hull;        // interior medium, no lobes, no emission: an invisible volume boundary
  uint32_t id;          // object id of the id AOV (Material::id), 0 = none
  uint32_t bind;        // pattern-bound fields: first binding << 4 | count (0: none)
  float film_albedo[3]; // albedo AOV: diffuse + primary specular F0 + refraction, each <= 1
  float pad2;
  // Hair fibers (hair.rs HairParams): curve hits scatter with the fiber
  // BSDF instead of the lobes above. Longitudinal variances per lobe,
  // azimuthal logistic scale, absorption, index of refraction.
  uint32_t hair;
  float hair_s;
  float hair_eta;
  float pad3;
  float hair_v[4];
  float hair_sigma_a[3];
  float pad4;
};

// The PbrParams fields the lobes read (render.rs src/scene/pbr.rs:11-45)
// and the material's own emission: f64 in the host build, f32 per material
// on the device for pattern-bound materials.
template <class T>
struct RawPbr {
  T diffuse_gain, diffuse_color[3], diffuse_roughness;
  T specular_face_color[3], specular_edge_color[3], specular_roughness, specular_ior;
  T clearcoat_gain, clearcoat_roughness;
  T fuzz_gain, fuzz_color[3];
  T glass_gain, glass_ior, glass_roughness, refraction_color[3];
  T glow[3], emission[3];
  T pad;
};
static_assert(sizeof(RawPbr<float>) == 128, "RawPbr<float> is 128 bytes");

// Pattern graph (render.rs src/texture/pattern.rs:42-68), one record per
// node: kind (RRHIP_PATTERN_*) and the parameters of every kind.
struct alignas(16) PtPattern {
  uint32_t kind, wrap;
  int32_t texture;    // texture / triplanar: texture index, -1 = missing
  int32_t udim;       // first of the node's 100 UDIM entries in `udim`, -1 = none
  float scale[2];     // texture scaleS / scaleT, checker sScale / tScale
  float frequency, gain;
  float lacunarity, gamma, saturation;
  uint32_t octaves;
  float color_a[3];
  uint32_t use_t;     // ramp over t
  float color_b[3];
  uint32_t ramp;      // first ramp stop
  float cc_gain[3];
  uint32_t nstops;
  float cc_offset[3];
  float pad;
};
// One step of a binding's evaluation program (its node tree in post-order,
// every input expanded in place): node >= 0 evaluates that node (mix and
// colour correct take their inputs from the value stack), -1 pushes c.
struct alignas(16) PtPatOp {
  int32_t node;
  float c[3];
};
struct PtBinding {
  uint32_t field;       // RRHIP_BIND_*
  uint32_t first, len;  // program in `prog`
  uint32_t pad;
};
// A texture's mip chain: its levels in `tlev`; texels rgb (w unused) from
// the top row, level after level.
struct PtTexture {
  uint32_t first_level, nlevels, w, h;
};
struct PtTexLevel {
  uint32_t w, h;
  uint64_t offset;  // first texel in `texels`
};
constexpr int kPatStack = 8;     // value-stack depth of a binding program
constexpr uint32_t kPatProgMax = 256;  // ops of one binding program

enum : uint32_t { kLightPoint = 0, kLightDistant = 1, kLightRect = 2, kLightSphere = 3, kLightDisk = 4, kLightDome = 5 };

struct alignas(16) PtLight {
  uint32_t kind;
  float radiance[3];  // color * intensity
  float p[3];         // position / corner / centre; distant: travel direction
  float radius;       // sphere radius; distant angular radius
  float e1[3];
  float area;
  float e2[3];
  float cos_max;      // distant: cos(angular_radius)
  float n[3];         // rect / disk normal
  float err;          // rounding bound of a sampled light point: along n (rect / disk), or of the
                      // sphere's surface points; shadow rays with error-bounded origins stop short by it
};

struct LightNode {  // light_sampler.rs:19-29
  float min[3];
  float power;
  float max[3];
  uint32_t parent;
  uint32_t a, b;
  uint32_t pad[2];
};
static_assert(sizeof(LightNode) == 48, "LightNode must match render.rs's 48-byte layout");

struct alignas(16) PtTriangle {
  float v0[3], e1[3], e2[3];
  uint32_t material;
  float n[3];  // normalize(e1 x e2)
  float pad;
};

struct alignas(16) PtSphere {
  float c[3], r;
  uint32_t material;
  float pad[3];
};

// A loose quadric: kind (RRHIP_QUADRIC_*), its parameters, and the
// world-to-object / object-to-world affine maps (row-major 3x4).
struct alignas(16) PtQuadric {
  uint32_t kind, material;
  float radius, zmin;
  float zmax, height, thetamax, k;  // k: cone (radius / height)^2, paraboloid rmax^2 / zmax
  float inv[12];
  float fwd[12];
  // hyperboloid: the profile x^2 + y^2 = pa z^2 + pb z + pc, the end points'
  // z and squared radii; torus: radius = major, height = minor, zmin / zmax =
  // phimin / phimax
  float pa, pb, pc, z1, z2, r1, r2, pad;
};

// A placed mesh (render.rs Instance, src/geometry/mesh.rs:223-305): the
// world-to-object affine map (row-major 3x4) and what it references.
struct alignas(16) PtInstance {
  float inv[12];
  uint32_t mesh;
  uint32_t material;
  uint32_t motion;  // index of its PtInstMotion, ~0u: static
  uint32_t deform;  // 1: its mesh deforms (vertices lerped towards btris1 at the shutter time)
};

// A moving instance (Instance::transform, transform1): object-to-world at
// shutter open and close; rays at time t use the inverse of their lerp.
struct alignas(16) PtInstMotion {
  float fwd0[12];
  float fwd1[12];
};

// A participating medium (medium.rs:28-65) with its density as a vertex
// grid over [lo, hi], or render.rs's procedural fBm cloud field
// (fbm_octaves > 0), or neither (homogeneous).
struct alignas(16) PtMedium {
  float sigma_a[3], g;
  float sigma_s[3], majorant;  // channel-max extinction
  float sigma_t[3], max_distance;
  float emission[3];
  uint32_t nx;
  float lo[3];
  uint32_t ny;
  float hi[3];
  uint32_t nz;
  float scale[3];              // (n - 1) / (hi - lo) per axis
  uint32_t fbm_octaves;        // 0: no procedural field
  const float* grid;
  float fbm_frequency, fbm_gain;
  float fbm_lacunarity, fbm_coverage, fbm_sharpness, pad;
};
static_assert(sizeof(PtMedium) == 144, "PtMedium layout");

struct PtMeshInfo {
  uint32_t root;      // BLAS root in the shared node array
  uint32_t idx_base;  // first index of the mesh in mesh_idx
  uint32_t nrm_base;  // first normal (vertex) of the mesh in mesh_nrm; ~0u = geometric normals
  uint32_t curves;    // 1: a curve set, its BLAS leaves index curve segment records
};

// One curve segment (curves.rs CurveSet): a rounded cone from p0 (radius
// r0) to p1 (radius r1), the curve parameter at both ends; 48 bytes in
// BLAS leaf order like the triangle records.
struct alignas(16) CurveRec {
  float p0[3], r0;
  float p1[3], r1;
  float v0, v1;
  uint32_t prim;  // segment index in its set
  uint32_t pad;
};
static_assert(sizeof(CurveRec) == 48, "CurveRec must be 48 bytes");

struct PtCamera {
  uint32_t width, height;
  uint32_t filter;
  float filter_width;
  float eye[3];
  float forward[3], right[3], up[3];
  float half_w, half_h;
  uint32_t projection;          // 0 perspective, 1 orthographic
  float lens_radius, focal_distance;
  float ortho_w, ortho_h;       // orthographic screen-window half extents
  uint32_t has_motion_inv;      // camera motion: rays at time t mapped by lerp(I, motion_inv, t)
  float motion_inv[12];
};

// Read-only scene view passed to the kernels by value.
struct PtSceneView {
  const PtTriangle* tris;
  const PtSphere* spheres;
  const PtMaterial* materials;
  const PtLight* lights;
  const LightNode* lnodes;
  const uint32_t* light_leaf;   // per light: leaf node or ~0u
  const uint32_t* inf_lights;   // infinite light ids
  const double* inf_cdf;        // inclusive power prefix sums (f64)
  // Instanced meshes: TLAS nodes [0, ...) then every mesh's BLAS nodes;
  // BLAS triangle records; instances in TLAS leaf order.
  const Node2* nodes;
  const TriRec* btris;
  const TriRec* btris1;         // scenes with deforming meshes: every record at shutter close (null without)
  const PtInstance* insts;
  const PtMeshInfo* meshes;
  const uint32_t* mesh_idx;
  const float* mesh_nrm;
  const PtMedium* media;
  int32_t atmosphere;           // medium every ray travels in (-1: none)
  uint32_t stack_need;          // traversal stack entries a lane can need
  uint32_t transparent;         // some material has an interior medium or presence < 1: paths keep
                                // a medium stack, shadow rays walk their blockers
  uint32_t motion;              // moving instances or camera: rays carry a shutter time
  const PtInstMotion* inst_motion;
  const float* inst_fwd;        // object-to-world of each instance (3x4 row-major, TLAS leaf order)
  // Ray origins: 0 = error-bounded (pushed outside the hit point's rounding
  // box); > 0 = render.rs's fixed push of this length along the shading
  // normal (pt/mod.rs:30, 1e-4). Set per session.
  float ray_offset;
  // Dome environment map (null: constant dome): rgb texels, row-marginal
  // and per-row conditional CDFs of sin-weighted luminance, their total,
  // and W * H / (2 pi^2 total), the factor of a texel's solid-angle pdf.
  const float* env_rgb;
  const double* env_marg;
  const double* env_cond;
  uint32_t env_w, env_h;
  double env_total;
  double env_pdf_k;
  const PtQuadric* quadrics;
  uint32_t num_quadrics;
  uint32_t num_insts;
  // Curve segments of every curve set in BLAS leaf order; `curves`: some
  // instance places a curve set (the trace kernels then take the variants
  // with the quadric and curve solvers).
  const CurveRec* csegs;
  uint32_t curves;
  uint32_t hair;                // some material is a hair fiber
  // Pattern-bound materials (all null / 0 without): raw PbrParams per
  // material, bindings, their programs, nodes, ramp stops (position, rgb),
  // UDIM tables, textures; texture coordinates of hits: per-mesh first st
  // pair (~0u: none) and the pairs, the scale of each instance (TLAS leaf
  // order) and of each quadric with a hyperboloid's sweep length; the
  // camera's pixel angle that starts every ray cone.
  const RawPbr<float>* raw_pbr;
  const PtBinding* binds;
  const PtPatOp* prog;
  const PtPattern* pats;
  const float4* ramp;
  const int32_t* udim;
  const PtTexture* texs;
  const PtTexLevel* tlev;
  const float4* texels;
  const uint32_t* mesh_st;
  const float* st;
  const float* inst_scale;
  const float2* quad_st;
  float pixel_spread;
  uint32_t patterns;
  uint32_t num_tris, num_spheres, num_lights, num_lnodes, num_inf;
  int32_t dome;
  double p_infinite, inf_total;
  float background[3];
  PtCamera cam;
};

// Wavefront path state, structure of arrays over path slots.
struct PathBuffers {
  float4* o;        // origin.xyz, prev_pdf
  float4* d;        // direction.xyz, primary footprint a0 (radiance cache) or, in pattern scenes, the
                    //   ray cone's spread beyond one pixel
  float4* beta;     // throughput.xyz, unused
  float4* L;        // radiance.xyz, unused
  float4* prev;     // previous vertex xyz, area-spread sum (radiance cache) or the ray cone's width
  uint4* rng;       // PCG32 state lo/hi, inc lo/hi
  uint32_t* flags;  // depth (bits 0-7), from_camera (bit 8), training path (9), suffix (10)
  float4* hit_n;    // hit normal.xyz, t (t < 0: miss)
  uint32_t* hit_m;  // material | front_face << 31
  float4* hit_front; // error-bounded origins: the hit rebuilt from the surface and pushed outside its
  float4* hit_back;  //   rounding box on the ray's side (shading point), and on the far side; the two
                     //   w lanes hold a curve hit's world tangent (octahedral), 4 on other hits
  float4* hit_st;    // pattern scenes: the hit's st, st density; world gradients of s and t
  float4* hit_gs;    //   (null in other scenes)
  float4* hit_gt;
  uint32_t* q0;
  uint32_t* q1;
  uint32_t* sq;     // shadow queue: slot ids
  float4* so;       // shadow origin.xyz, tmax
  float4* sd;       // shadow direction.xyz
  float4* sc;       // shadow contribution rgb (already times beta), w = 1: the record has a second ray
  float4* so2;      // second ray of a record (equiangular connections in media): origin, tmax;
  float4* sd2;      //   direction; contribution (all three null when no scene needs them)
  float4* sc2;
  uint2* mstk;      // transparent scenes: each path's medium stack (8 x 8-bit medium indices)
  uint2* srng;      // transparent scenes: 64-bit seed of each shadow record's blocker walk
  float* stime;     // scenes with motion: shutter time of each shadow record (paths keep theirs in beta.w)
  uint32_t* counts; // [0] live paths, [1] next live, [2] shadow rays, [3] unused
  unsigned long long* stats;  // [0] extension rays, [1] shadow rays, [2] cache queries, [3] training samples
  uint32_t* job;    // batches: job of each slot (null in sessions)
  uint32_t* qs;     // sorted-queue mode: the live queue grouped by shading class
  uint32_t* sort_hist;  // sorted-queue mode: per-block class counts
  // Film (AOV) mode, null otherwise: radiance whose first continuation was
  // not specular-like, and the first real hit's data
  float4* ld;       // diffuse-split radiance rgb
  float4* aux0;     // albedo rgb, depth (0 until set)
  float4* aux1;     // shading normal facing the viewer xyz, distance travelled before the first real hit
  uint32_t* aux_id; // object id of the first real hit
  // Profiling (null otherwise): traversal fetches of extend [0..2] and
  // shadow [3..5] rays: nodes, triangle records, instances entered.
  unsigned long long* prof;
};

// One radiance-cache query / training sample (the 64-byte rrhip_nrc_query
// record): position, roughness, direction towards the viewer, shading
// normal on that side, albedo.
struct alignas(16) NrcQuery {
  float pos[3], roughness;
  float dir[3], pad0;
  float normal[3], pad1;
  float albedo[3], pad2;
};
static_assert(sizeof(NrcQuery) == 64, "query record is 64 bytes");

// Radiance-cache state of a rendering pass. Render paths stop at the
// vertex where their area spread exceeds `spread` times the primary
// footprint and queue a cache query there; training paths (slots
// [train_base, train_base + ntrain)) record every vertex, continue past
// that point for a suffix of the same spread, and queue the cache at the
// suffix's end as its tail.
struct NrcPathBuffers {
  NrcQuery* q;          // queries
  uint32_t* q_slot;     // path slot of each query
  float4* q_beta;       // throughput at the query vertex
  uint32_t* counts;     // [0] queries, [1] training samples built
  NrcQuery* tq;         // training records [path][vertex]
  float4* tb;           // throughput arriving at the vertex
  float4* tl;           // path radiance after the vertex's emission
  uint32_t* tn;         // recorded vertices per training path
  uint32_t train_base, ntrain, max_vertices;
  float spread;
};

struct PtParams {
  uint32_t max_depth;
  uint32_t rr_start;
  float firefly_clamp;
  uint32_t npix;          // pixel columns of this pass (slot = sample * npix + column)
  uint32_t spp_pass;      // samples per pixel in this pass
  uint64_t sample0;       // sample index of the pass's first sample (incl. base)
  const uint32_t* live;   // pixel of each column (adaptive sampling); null: column = pixel
};

// One job of a batch pass: its scene, its pass parameters, where its
// slots and pixels start in the batch, and its accumulator.
struct PtJob {
  PtSceneView S;
  PtParams prm;
  uint32_t slot_begin, pix_begin;
  double* accum;
};

}  // namespace rrhip
