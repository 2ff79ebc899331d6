// Device data layout shared by the host builder and the HIP kernels.
//
// Ray-query data contract. render.rs flattens its BVH into 32-byte nodes
// holding one box each and permutes triangles into leaf order
// (src/raytracer/metal/gpu_scene.rs:18-28, 473-491). Here every node holds
// the boxes of BOTH children (64 bytes, one 64-byte-aligned line), so one
// node fetch decides the two slab tests of an interior visit, and
// triangles are stored pre-gathered in leaf order (48 bytes: the three f32
// vertices exactly as given, so the f64 re-test sees the caller's bits,
// plus the original triangle index).
#pragma once
#include <cstdint>

namespace rrhip {

// One interior visit: both children's boxes, SoA-packed so the slab test
// reads three float4s.
struct alignas(64) Node2 {
  float c0_lox, c0_hix, c0_loy, c0_hiy;  // child 0 x/y slabs
  float c1_lox, c1_hix, c1_loy, c1_hiy;  // child 1 x/y slabs
  float c0_loz, c0_hiz, c1_loz, c1_hiz;  // both z slabs
  // ref: node index (cnt == 0) or first triangle slot (cnt > 0).
  uint32_t ref0, ref1, cnt0, cnt1;
};
static_assert(sizeof(Node2) == 64, "Node2 must be one 64-byte line");

// Four-wide node (the binary tree collapsed, same leaves): the four
// children's boxes as per-axis float4s (lane k = child k), their
// references (node index, or first triangle slot of a leaf) and leaf
// triangle counts (0: interior child, kEmpty4: no child). One 128-byte
// line, so a visit is one cache line and decides four slab tests.
struct alignas(128) Node4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  uint32_t ref[4];
  uint32_t cnt[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");
constexpr uint32_t kEmpty4 = 0xFFFFFFFFu;
// Four-wide stack entries: a node (bits 0-26) and the mask of its hit
// children not yet taken (bits 27-30), so a visit pushes at most one entry
// (the node itself) and the stack never holds more entries than the tree
// has levels.
constexpr uint32_t kNode4Mask = (1u << 27) - 1u;

struct alignas(16) TriRec {
  float v0x, v0y, v0z;
  float v1x, v1y, v1z;
  float v2x, v2y, v2z;
  uint32_t prim;  // original triangle index
  uint32_t pad0, pad1;
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

// Deepest traversal stack the kernels carry (in LDS). The builder refuses
// trees with more node levels than this.
constexpr int kMaxStack = 64;

}  // namespace rrhip
