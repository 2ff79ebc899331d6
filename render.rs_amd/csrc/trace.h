// Ray-query kernels: launch interface used by the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace rrhip {

// Query ray / hit records (byte-identical to rrhip_ray / rrhip_hit).
struct RayRec {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
};
struct HitRec {
  float t;
  uint32_t prim;
  float u, v;
};
static_assert(sizeof(RayRec) == 32, "RayRec layout");
static_assert(sizeof(HitRec) == 16, "HitRec layout");

constexpr uint32_t kNoHit = 0xFFFFFFFFu;

struct DeviceBvh {
  const Node2* nodes = nullptr;
  const TriRec* tris = nullptr;
  const Node4* nodes4 = nullptr;  // the same tree collapsed four-wide
  uint64_t num_nodes = 0;
  uint64_t num_nodes4 = 0;
  uint64_t num_tris = 0;
  int depth = 0;   // binary tree: edges on the longest path (the stack holds that many entries)
  int depth4 = 0;  // four-wide tree: edges on the longest path
};

// Launch configuration of the traversal kernels.
struct TraceTuning {
  uint32_t block = 128;  // threads per block: 64, 128 or 256
  uint32_t stack = 0;    // LDS stack entries per lane: 16, 24, 32, 48 or 64 (>= stack_need)
  uint32_t chunk = 0;    // rays owned by each wave (multiple of 64); 0: fill the device's resident waves
  uint32_t width = 2;    // tree width: 2 (Node2) or 4 (Node4)
};

enum class TraceMode : int;
// Wave slots of the current device that hold the query kernel of this
// tuning and mode at once (0 if unknown).
uint64_t trace_resident_waves(const TraceTuning& tu, TraceMode mode);
// Rays per wave that spread `n` rays over `slots` resident waves (whole
// wavefronts of rays, at least 64, at most 2^20; 256 when slots is 0).
uint32_t fill_chunk(uint64_t n, uint64_t slots);

// Smallest supported stack size holding `depth` entries (0 if none does).
int stack_for_depth(int depth);
// Stack entries a lane can need: one per level in both trees (one push
// per visit: a binary sibling, or a four-wide node with its remaining
// children).
int stack_need(const DeviceBvh& bvh, uint32_t width);
bool tuning_valid(const TraceTuning& tu, const DeviceBvh& bvh);

enum class TraceMode : int { kClosest, kAny };

// Traces `n` rays. kClosest writes `hits`; kAny writes `occluded` (0/1).
// `counts` (optional, 2 x u32 per ray) receives node fetches and triangle
// fetches per ray, for the traffic model.
hipError_t launch_trace(const DeviceBvh& bvh, const RayRec* rays, uint64_t n, TraceMode mode,
                        HitRec* hits, uint8_t* occluded, uint32_t* counts,
                        const TraceTuning& tuning, hipStream_t stream);

}  // namespace rrhip
