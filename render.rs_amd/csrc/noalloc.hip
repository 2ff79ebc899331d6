// Closest-hit query over host buffers that allocates nothing: the rays
// and hits are staged through a fixed device array instead of buffers the
// library grows with hipMalloc. For callers that may not allocate device
// memory (tight serving loops, harnesses that forbid hipMalloc).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "api_internal.h"
#include "scene_internal.h"
#include "trace.h"

namespace rrhip {

namespace {

constexpr uint64_t kMaxRays = RRHIP_NOALLOC_MAX_RAYS;
constexpr uint64_t kRayBytes = kMaxRays * sizeof(RayRec);
constexpr uint64_t kScratchBytes = kMaxRays * (sizeof(RayRec) + sizeof(HitRec));
static_assert(kRayBytes % 256 == 0, "hit staging stays 256-byte aligned");

// Staging of rays (first kRayBytes) then hits. Each device holds its own
// copy at its own address; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (the array's name is
// a host-side stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

// Resolves `device`'s copy of the staging array once; `device` must be
// current. Called by rrhip_init.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(RRHIP_ERR_UNSUPPORTED, "staging array: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return RRHIP_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(RRHIP_ERR_INTERNAL, "staging array has no device address");
  g_scratch[device] = p;
  return RRHIP_OK;
}

}  // namespace rrhip

using namespace rrhip;

extern "C" int rrhip_trace_closest_noalloc(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, rrhip_hit* hits) {
  if (!s) return fail(RRHIP_ERR_INVALID_ARGUMENT, "null scene");
  if (!s->has_query) return fail(RRHIP_ERR_UNSUPPORTED, "scene has no ray-query mesh");
  if (n && (!rays || !hits)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "null ray or output buffer");
  if (n > kMaxRays)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_trace_closest_noalloc: " + std::to_string(n) +
                                                " rays exceed the staging limit of " +
                                                std::to_string(kMaxRays) + " (split the batch)");
  if (n == 0) return RRHIP_OK;
  const int dev = s->ctx->device;
  RRHIP_HIP_TRY(hipSetDevice(dev));
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    base = g_scratch[dev];
  }
  if (!base) {
    if (int st = noalloc_setup(dev)) return st;
    std::lock_guard<std::mutex> lock(g_mu);
    base = g_scratch[dev];
  }
  if (s->bvh.num_nodes == 0) {
    for (uint64_t i = 0; i < n; i++) hits[i] = rrhip_hit{INFINITY, RRHIP_NO_HIT, 0.f, 0.f};
    return RRHIP_OK;
  }
  auto* d_rays = (RayRec*)base;
  auto* d_hits = (HitRec*)((unsigned char*)base + kRayBytes);
  hipStream_t st = s->ctx->stream;
  RRHIP_HIP_TRY(hipMemcpyAsync(d_rays, rays, n * sizeof(RayRec), hipMemcpyHostToDevice, st));
  RRHIP_HIP_TRY(launch_trace(s->bvh, d_rays, n, TraceMode::kClosest, d_hits, nullptr, nullptr, s->tuning, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(hits, d_hits, n * sizeof(HitRec), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  return RRHIP_OK;
}
