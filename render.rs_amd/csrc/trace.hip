// Closest-hit and any-hit BVH traversal for gfx950 (CDNA4).
//
// Semantics follow render.rs's CPU path: `Bvh::traverse` / `Bvh::any_hit`
// (src/accel/bvh.rs:141-245) over `Mesh::intersect_triangle`
// (src/geometry/mesh.rs:113-142). Hit decisions are made so the returned
// triangle ids and t values are the ones that f64 code computes:
//
//  * node boxes are the exact f32 extents of the f32 vertices, and the
//    f32 slab test widens t_far by (1 + 2^-21) >= 1 + 2*gamma_3 (Ize 2013,
//    "Robust BVH Ray Traversal"), so no box the f64 ray enters is culled;
//  * each triangle first gets an f32 Moller-Trumbore evaluation with a
//    forward error bound (64 ulp of the summed product magnitudes); only a
//    candidate the bound cannot rule out is re-tested in f64 with the
//    operation order of mesh.rs:117-142 and FMA contraction off, which
//    yields the same bits as the f64 CPU code;
//  * equal-t candidates resolve to the lower triangle id.
//
// Execution (wave64): every wave owns a contiguous chunk of rays, by
// default sized so that the launch's waves all fit the device at once
// (one round, no tail of late waves; a shared atomic ray counter measured
// slower). The kernel is bound by vector-instruction issue, and per-ray work varies by
// an order of magnitude between rays, so the wave schedules phases rather
// than letting lanes wait on each other's loops: each iteration, three
// ballots count the lanes that walk interior nodes, the lanes that hold a
// pending leaf and the free lanes (finished or empty), and the wave runs
// one phase for the lanes in it — a refill once at least kRefill lanes are
// free (finished lanes write their hit, free lanes take the next rays of
// the chunk by mbcnt rank), else one triangle test for
// the leaf lanes when they are at least as many as the walking lanes, else
// one interior step (pop, then visit) for the walking lanes. Each branch
// is wave-uniform. Each lane's stack of node indices lives in LDS, laid
// out [depth][thread] so every lane hits its own bank; leaf entries carry
// the top bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "trace.h"
#include "trace_math.h"

namespace rrhip {
namespace {

// Free lanes that trigger a refill of the wave.
constexpr uint32_t kRefill = 8;

// Phase-vote kernel over static ray chunks: wave w owns rays
// [w * chunk, (w + 1) * chunk). WIDTH selects the binary or the
// four-wide tree.
template <int BLOCK, int STACK, bool ANY, bool COUNT, int WIDTH>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void trace_kernel(
    const Node2* __restrict__ nodes, const Node4* __restrict__ nodes4, const TriRec* __restrict__ tris,
    const RayRec* __restrict__ rays, HitRec* __restrict__ hits, uint8_t* __restrict__ occluded,
    uint32_t* __restrict__ counts, uint32_t n, uint32_t chunk) {
  __shared__ uint32_t s_stack[STACK * BLOCK];
  const uint32_t tid = threadIdx.x;
  uint32_t* stk = s_stack + tid;
  const uint32_t wave = (blockIdx.x * BLOCK + tid) >> 6;
  const uint64_t begin64 = (uint64_t)wave * chunk;
  const uint32_t begin = begin64 < n ? (uint32_t)begin64 : n;
  const uint32_t end = (uint64_t)begin + chunk < n ? begin + chunk : n;
  uint32_t cursor = begin;  // wave-uniform: the next ray of the chunk not yet taken

  Lane L;
  bool have = false;
  uint32_t rid = 0;
  for (;;) {
    const bool leafy = have && L.lcnt > 0;
    const bool walking = have && L.lcnt == 0 && L.node != kDone;
    const bool finished = have && L.lcnt == 0 && L.node == kDone;
    const uint32_t n_leaf = (uint32_t)__popcll(__ballot(leafy));
    const uint32_t n_walk = (uint32_t)__popcll(__ballot(walking));
    const uint64_t free_mask = __ballot(!have || finished);
    const uint32_t n_free = (uint32_t)__popcll(free_mask);
    if (n_leaf + n_walk == 0 && cursor >= end) {
      // No lane has work and the chunk is used up.
      if (finished) lane_finish<ANY, COUNT>(L, rid, hits, occluded, counts);
      break;
    }
    if (cursor < end && (n_free >= kRefill || n_leaf + n_walk == 0)) {
      // Refill: finished lanes write their hit, free lanes take new rays.
      if (finished) {
        lane_finish<ANY, COUNT>(L, rid, hits, occluded, counts);
        have = false;
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(free_mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)free_mask, 0u));
      if (!have && rank < end - cursor) {
        rid = cursor + rank;
        have = true;
        lane_start(L, rays, rid);
      }
      cursor = n_free < end - cursor ? cursor + n_free : end;
    } else if (n_leaf != 0 && n_leaf >= n_walk) {
      // Leaf phase: one triangle of each pending leaf.
      if (leafy) {
        const TriRec tr = tris[L.lref];
        if (COUNT) L.n_tris++;
        test_triangle<ANY>(L, tr);
        L.lref++;
        L.lcnt--;
        if (ANY && L.occ) {
          L.lcnt = 0;
          L.node = kDone;
        }
      }
    } else if (walking) {
      if constexpr (WIDTH == 4)
        interior_step4<BLOCK, COUNT>(L, nodes4, stk);
      else
        interior_step<BLOCK, COUNT>(L, nodes, stk);
    }
  }
}

// Wave slots of the current device that can hold `fn` at once (blocks per
// CU from the kernel's registers and LDS, times CUs); 0 if unknown.
uint64_t resident_waves(const void* fn, int block) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0) != hipSuccess) return 0;
  return (uint64_t)per_cu * (uint64_t)(block / 64) * (uint64_t)cus;
}

template <int BLOCK, int STACK, bool ANY, bool COUNT, int WIDTH>
uint64_t slots_of() {
  static const uint64_t slots =
      resident_waves((const void*)trace_kernel<BLOCK, STACK, ANY, COUNT, WIDTH>, BLOCK);
  return slots;
}

template <int BLOCK, int STACK, bool ANY, bool COUNT, int WIDTH>
hipError_t launch_cfg(const DeviceBvh& bvh, const RayRec* rays, uint32_t n, HitRec* hits,
                      uint8_t* occ, uint32_t* counts, uint32_t chunk, hipStream_t stream) {
  if (chunk == 0) {
    // Fill: split the rays evenly over the wave slots the device holds at
    // once, so the launch runs as a single round of resident waves.
    chunk = fill_chunk(n, slots_of<BLOCK, STACK, ANY, COUNT, WIDTH>());
  }
  const uint64_t waves = ((uint64_t)n + chunk - 1) / chunk;
  const uint64_t blocks = (waves * 64 + BLOCK - 1) / BLOCK;
  hipLaunchKernelGGL((trace_kernel<BLOCK, STACK, ANY, COUNT, WIDTH>), dim3((uint32_t)blocks), dim3(BLOCK),
                     0, stream, bvh.nodes, bvh.nodes4, bvh.tris, rays, hits, occ, counts, n, chunk);
  return hipGetLastError();
}

template <int BLOCK, bool ANY, bool COUNT, int WIDTH>
hipError_t launch_stack(const DeviceBvh& bvh, const RayRec* rays, uint32_t n, HitRec* hits,
                        uint8_t* occ, uint32_t* counts, const TraceTuning& tu,
                        hipStream_t stream) {
  switch (tu.stack) {
    case 16: return launch_cfg<BLOCK, 16, ANY, COUNT, WIDTH>(bvh, rays, n, hits, occ, counts, tu.chunk, stream);
    case 24: return launch_cfg<BLOCK, 24, ANY, COUNT, WIDTH>(bvh, rays, n, hits, occ, counts, tu.chunk, stream);
    case 32: return launch_cfg<BLOCK, 32, ANY, COUNT, WIDTH>(bvh, rays, n, hits, occ, counts, tu.chunk, stream);
    case 48: return launch_cfg<BLOCK, 48, ANY, COUNT, WIDTH>(bvh, rays, n, hits, occ, counts, tu.chunk, stream);
    case 64: return launch_cfg<BLOCK, 64, ANY, COUNT, WIDTH>(bvh, rays, n, hits, occ, counts, tu.chunk, stream);
    default: return hipErrorInvalidValue;
  }
}

template <bool ANY, bool COUNT, int WIDTH>
hipError_t launch_block(const DeviceBvh& bvh, const RayRec* rays, uint32_t n, HitRec* hits,
                        uint8_t* occ, uint32_t* counts, const TraceTuning& tu,
                        hipStream_t stream) {
  switch (tu.block) {
    case 64: return launch_stack<64, ANY, COUNT, WIDTH>(bvh, rays, n, hits, occ, counts, tu, stream);
    case 128: return launch_stack<128, ANY, COUNT, WIDTH>(bvh, rays, n, hits, occ, counts, tu, stream);
    case 256: return launch_stack<256, ANY, COUNT, WIDTH>(bvh, rays, n, hits, occ, counts, tu, stream);
    default: return hipErrorInvalidValue;
  }
}

template <bool ANY, bool COUNT>
hipError_t launch_width(const DeviceBvh& bvh, const RayRec* rays, uint32_t n, HitRec* hits,
                        uint8_t* occ, uint32_t* counts, const TraceTuning& tu,
                        hipStream_t stream) {
  if (tu.width == 4) return launch_block<ANY, COUNT, 4>(bvh, rays, n, hits, occ, counts, tu, stream);
  return launch_block<ANY, COUNT, 2>(bvh, rays, n, hits, occ, counts, tu, stream);
}

template <bool ANY, int WIDTH, int BLOCK>
uint64_t slots_stack(uint32_t stack) {
  switch (stack) {
    case 16: return slots_of<BLOCK, 16, ANY, false, WIDTH>();
    case 24: return slots_of<BLOCK, 24, ANY, false, WIDTH>();
    case 32: return slots_of<BLOCK, 32, ANY, false, WIDTH>();
    case 48: return slots_of<BLOCK, 48, ANY, false, WIDTH>();
    case 64: return slots_of<BLOCK, 64, ANY, false, WIDTH>();
    default: return 0;
  }
}

template <bool ANY, int WIDTH>
uint64_t slots_block(const TraceTuning& tu) {
  switch (tu.block) {
    case 64: return slots_stack<ANY, WIDTH, 64>(tu.stack);
    case 128: return slots_stack<ANY, WIDTH, 128>(tu.stack);
    case 256: return slots_stack<ANY, WIDTH, 256>(tu.stack);
    default: return 0;
  }
}

}  // namespace

uint64_t trace_resident_waves(const TraceTuning& tu, TraceMode mode) {
  const bool any = mode == TraceMode::kAny;
  if (tu.width == 4) return any ? slots_block<true, 4>(tu) : slots_block<false, 4>(tu);
  return any ? slots_block<true, 2>(tu) : slots_block<false, 2>(tu);
}

uint32_t fill_chunk(uint64_t n, uint64_t slots) {
  if (slots == 0) return 256;
  const uint64_t per = (n + slots - 1) / slots;
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>((per + 63) / 64 * 64, 64), 1u << 20);
}

int stack_for_depth(int depth) {
  const int sizes[] = {16, 24, 32, 48, 64};
  for (int s : sizes)
    if (depth <= s) return s;
  return 0;
}

int stack_need(const DeviceBvh& bvh, uint32_t width) {
  return width == 4 ? bvh.depth4 + 1 : bvh.depth;
}

bool tuning_valid(const TraceTuning& tu, const DeviceBvh& bvh) {
  const bool block_ok = tu.block == 64 || tu.block == 128 || tu.block == 256;
  const bool width_ok = tu.width == 2 || (tu.width == 4 && bvh.nodes4 != nullptr);
  const bool stack_ok = stack_for_depth(tu.stack) == (int)tu.stack && (int)tu.stack >= stack_need(bvh, tu.width);
  const bool chunk_ok = tu.chunk == 0 || (tu.chunk % 64 == 0 && tu.chunk <= (1u << 20));
  return block_ok && width_ok && stack_ok && chunk_ok;
}

hipError_t launch_trace(const DeviceBvh& bvh, const RayRec* rays, uint64_t n, TraceMode mode,
                        HitRec* hits, uint8_t* occluded, uint32_t* counts,
                        const TraceTuning& tu, hipStream_t stream) {
  if (!tuning_valid(tu, bvh)) return hipErrorInvalidValue;
  const uint64_t kChunk = 1ull << 30;
  for (uint64_t off = 0; off < n; off += kChunk) {
    const uint32_t m = (uint32_t)((n - off) < kChunk ? (n - off) : kChunk);
    const RayRec* r = rays + off;
    HitRec* h = hits ? hits + off : nullptr;
    uint8_t* o = occluded ? occluded + off : nullptr;
    uint32_t* c = counts ? counts + 2 * off : nullptr;
    hipError_t e;
    if (mode == TraceMode::kAny) {
      e = c ? launch_width<true, true>(bvh, r, m, h, o, c, tu, stream)
            : launch_width<true, false>(bvh, r, m, h, o, c, tu, stream);
    } else {
      e = c ? launch_width<false, true>(bvh, r, m, h, o, c, tu, stream)
            : launch_width<false, false>(bvh, r, m, h, o, c, tu, stream);
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace rrhip
