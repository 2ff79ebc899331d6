// Host build of a path-tracing scene into the device layout of pt_types.h.
#pragma once
#include <string>
#include <vector>

#include "../../include/rrhip.h"
#include "pt_types.h"

namespace rrhip {

struct PtSceneHost {
  std::vector<PtTriangle> tris;
  std::vector<PtSphere> spheres;
  std::vector<PtMaterial> materials;
  std::vector<PtLight> lights;
  std::vector<LightNode> lnodes;
  std::vector<uint32_t> light_leaf;
  std::vector<uint32_t> inf_lights;
  std::vector<double> inf_cdf;
  double p_infinite = 0.0, inf_total = 0.0;
  int32_t dome = -1;
  float background[3] = {0, 0, 0};
  PtCamera cam{};
  // Instanced meshes (see PtSceneView).
  std::vector<Node2> nodes;
  std::vector<TriRec> btris;
  std::vector<TriRec> btris1;  // deforming meshes: the records at shutter close (rest copies elsewhere)
  std::vector<PtInstance> insts;
  std::vector<PtMeshInfo> meshes;
  std::vector<uint32_t> mesh_idx;
  std::vector<float> mesh_nrm;
  uint32_t stack_need = 0;  // traversal stack entries the instance path needs
  uint32_t transparent = 0; // a material has an interior medium or presence < 1
  uint32_t motion = 0;      // moving instances, deforming meshes or camera motion
  uint32_t deform = 0;      // some mesh deforms (btris1 holds the shutter-close records)
  std::vector<PtInstMotion> inst_motion;
  std::vector<float> inst_fwd;  // object-to-world of each instance in insts' order (12 floats)
  std::vector<CurveRec> csegs;  // curve segments of every curve set, BLAS leaf order
  uint32_t curves = 0;          // some instance places a curve set
  uint32_t hair = 0;            // some material is a hair fiber
  std::vector<PtMedium> media;                // grid pointers filled at upload
  std::vector<std::vector<float>> media_grids;
  int32_t atmosphere = -1;
  float lo[3], hi[3];       // bounds of all geometry (radiance-cache input range)
  // Dome environment map (envmap.rs): rgb texels, the row-marginal CDF
  // (height + 1) and per-row conditional CDFs ((width + 1) x height) of
  // sin-weighted luminance, in f64, and their total.
  std::vector<PtQuadric> quadrics;
  std::vector<float2> quad_st;    // per quadric: approximate scale, hyperboloid sweep length
  std::vector<float> inst_scale;  // per instance (TLAS leaf order): approximate scale
  // Pattern-bound materials (PtSceneView's pattern fields); empty without.
  uint32_t patterns = 0;
  std::vector<RawPbr<float>> raw_pbr;
  std::vector<PtBinding> binds;
  std::vector<PtPatOp> prog;
  std::vector<PtPattern> pats;
  std::vector<float4> ramp;
  std::vector<int32_t> udim;
  std::vector<PtTexture> texs;
  std::vector<PtTexLevel> tlev;
  std::vector<float4> texels;
  std::vector<uint32_t> mesh_st;  // per geometry: first st pair, ~0u none
  std::vector<float> st;
  float pixel_spread = 0.f;
  uint32_t env_w = 0, env_h = 0;
  std::vector<float> env_rgb;
  std::vector<double> env_marg, env_cond;
  double env_total = 0.0;
};

// Validates and converts the path-tracing part of `d` (loose triangles,
// spheres, meshes and their instances with per-mesh BLAS and a TLAS,
// materials, lights, camera) and builds render.rs's light BVH
// (src/scene/light_sampler.rs:97-142, 242-312: median split on the widest
// axis, power sums in f64 stored as f32). Returns false with `err` set.
bool build_pt_scene(const rrhip_scene_desc& d, PtSceneHost* out, std::string* err);

// One material's device record (lobe constants, hair parameters).
PtMaterial pt_material(const rrhip_material_desc& m);

}  // namespace rrhip
