// Definitions of the opaque ABI handles, shared by the C-ABI translation
// units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/rrhip.h"
#include "layout.h"
#include "pt_types.h"
#include "trace.h"

struct rrhip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string name, arch;
};

struct rrhip_scene {
  rrhip_ctx* ctx = nullptr;
  // Ray-query mesh (one mesh, no instances).
  bool has_query = false;
  rrhip::Node2* d_nodes = nullptr;
  rrhip::TriRec* d_tris = nullptr;
  rrhip::Node4* d_nodes4 = nullptr;  // the query tree collapsed four-wide
  rrhip::DeviceBvh bvh;
  rrhip::TraceTuning tuning;      // closest-hit launches
  rrhip::TraceTuning tuning_any;  // any-hit launches
  rrhip_scene_info info{};
  // Staging for the host-buffer entry points (grown on demand).
  void* d_in = nullptr;
  void* d_out = nullptr;
  uint64_t in_cap = 0, out_cap = 0;
  // Path-tracing scene (camera given): device arrays behind `pt`.
  bool has_pt = false;
  rrhip::PtSceneView pt{};
  void* pt_mem[36] = {};
  std::vector<void*> media_mem;  // density grids
  float pt_lo[3] = {0, 0, 0}, pt_hi[3] = {0, 0, 0};  // bounds of all geometry
  bool pt_eq = false;  // point lights in an atmosphere: shadow records carry a second ray
};
