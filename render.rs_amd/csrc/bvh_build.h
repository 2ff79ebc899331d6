// Binned-SAH BVH builder for triangle meshes (host, multithreaded).
//
// Plays the role of render.rs's `Bvh::build` (src/accel/bvh.rs:96-108,
// 256-345) plus the GPU export (src/raytracer/metal/gpu_scene.rs:290-299,
// 473-491), redesigned: f32 bounds taken exactly from the f32 vertices (no
// f64->f32 rounding, so node boxes never shrink), SAH binned over all three
// axes, subtrees built in parallel, output flattened to child-pair nodes
// in depth-first order with triangles gathered into leaf order.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "layout.h"

namespace rrhip {

struct BuildParams {
  int bins = 32;      // SAH bins per axis
  int max_leaf = 4;   // leaves hold at most this many triangles
  int threads = 0;    // 0 = hardware concurrency
};

struct BuiltBvh {
  std::vector<Node2> nodes;  // nodes[0] is the root
  std::vector<TriRec> tris;  // leaf order
  float root_lo[3], root_hi[3];
  int depth = 0;  // edges on the longest root-to-leaf path
  uint64_t leaves = 0;
};

// Collapses a binary tree into four-wide nodes with the same leaves: each
// node takes its two children and repeatedly opens the interior child of
// largest surface area until it has four (Wald et al. 2008's collapse).
// depth4: edges on the longest root-to-leaf path of the result.
void collapse_bvh4(const std::vector<Node2>& nodes, std::vector<Node4>* out, int* depth4);

// Builds over `ntris` triangles given by `indices` into `positions`
// (xyz f32 triples). Returns false with `err` set on invalid input.
// positions1: a deformation endpoint (same vertex count) whose triangles
// the boxes also cover, or null.
bool build_triangle_bvh(const float* positions, uint64_t nverts, const uint32_t* indices,
                        uint64_t ntris, const BuildParams& params, BuiltBvh* out,
                        std::string* err, const float* positions1 = nullptr);

// Builds over `n` axis-aligned boxes (6 floats each: lo xyz, hi xyz), e.g.
// instance world bounds for a TLAS. Leaves reference slots of `order`,
// which receives the primitive ids in leaf order; out->tris stays empty.
bool build_box_bvh(const float* boxes, uint64_t n, const BuildParams& params, BuiltBvh* out,
                   std::vector<uint32_t>* order, std::string* err);

}  // namespace rrhip
