// Progressive path-tracing sessions (C ABI): the HIP counterpart of
// render.rs's PtSession / WfSession (src/raytracer/metal/renderer.rs:260-763).
//
// A session owns the wavefront path buffers for `samples_per_pass` samples
// of every pixel and an f64 accumulator (RRHIP_ACCUM_STRIDE doubles per
// pixel). rrhip_render_samples(start, count) renders samples
// [start, start + count) in passes: raygen, then per bounce extend / shade
// / shadow over device-resident queues (no host synchronisation inside a
// pass), then a deterministic resolve into the accumulator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <vector>

#include "api_internal.h"
#include "pt_kernels.h"
#include "scene_internal.h"

using namespace rrhip;

struct rrhip_session {
  rrhip_scene* scene = nullptr;
  PtParams prm{};
  uint64_t sample_base = 0;
  uint32_t spp_pass = 1;
  uint64_t slots = 0;
  PathBuffers P{};
  std::vector<void*> mem;
  double* d_accum = nullptr;
  double* relmse_partial = nullptr;
  uint64_t samples_done = 0;
  double last_ms = 0.0;
};

namespace {

template <class T>
hipError_t alloc_n(rrhip_session* s, T** p, uint64_t n) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, std::max<uint64_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) s->mem.push_back(v);
  *p = (T*)v;
  return e;
}

void free_session(rrhip_session* s) {
  for (void* p : s->mem) (void)hipFree(p);
  s->mem.clear();
  delete s;
}

int check_session(rrhip_session* s) {
  if (!s) return fail(RRHIP_ERR_INVALID_ARGUMENT, "null session");
  return RRHIP_OK;
}

}  // namespace

extern "C" {

int rrhip_session_create(rrhip_scene* scene, const rrhip_render_params* params, rrhip_session** out) {
  if (!scene || !params || !out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_create: null argument");
  *out = nullptr;
  if (params->struct_size != sizeof(rrhip_render_params))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_create: struct_size mismatch (ABI version)");
  if (!scene->has_pt) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_create: scene has no camera");
  if (params->max_depth == 0 || params->max_depth > 255)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_create: max_depth must be in 1..255");
  const uint64_t npix = (uint64_t)scene->pt.cam.width * scene->pt.cam.height;
  uint32_t spp = params->samples_per_pass;
  if (spp == 0) spp = (uint32_t)std::min<uint64_t>(32, std::max<uint64_t>(1, (4ull << 20) / npix));
  if ((uint64_t)spp * npix >= (1ull << 31))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_create: samples_per_pass x pixels exceeds 2^31");
  if (hipSetDevice(scene->ctx->device) != hipSuccess) return fail(RRHIP_ERR_DEVICE, "hipSetDevice");
  auto* s = new (std::nothrow) rrhip_session();
  if (!s) return fail(RRHIP_ERR_OUT_OF_MEMORY, "rrhip_session_create: host allocation failed");
  s->scene = scene;
  s->spp_pass = spp;
  s->slots = (uint64_t)spp * npix;
  s->sample_base = params->sample_base;
  s->prm.max_depth = params->max_depth;
  s->prm.rr_start = params->rr_start;
  s->prm.firefly_clamp = (float)params->firefly_clamp;
  s->prm.npix = (uint32_t)npix;
  PathBuffers& P = s->P;
  const uint64_t n = s->slots;
  hipError_t e = alloc_n(s, &P.o, n);
  if (e == hipSuccess) e = alloc_n(s, &P.d, n);
  if (e == hipSuccess) e = alloc_n(s, &P.beta, n);
  if (e == hipSuccess) e = alloc_n(s, &P.L, n);
  if (e == hipSuccess) e = alloc_n(s, &P.prev, n);
  if (e == hipSuccess) e = alloc_n(s, &P.rng, n);
  if (e == hipSuccess) e = alloc_n(s, &P.flags, n);
  if (e == hipSuccess) e = alloc_n(s, &P.hit_n, n);
  if (e == hipSuccess) e = alloc_n(s, &P.hit_m, n);
  if (e == hipSuccess) e = alloc_n(s, &P.q0, n);
  if (e == hipSuccess) e = alloc_n(s, &P.q1, n);
  if (e == hipSuccess) e = alloc_n(s, &P.sq, n);
  if (e == hipSuccess) e = alloc_n(s, &P.so, n);
  if (e == hipSuccess) e = alloc_n(s, &P.sd, n);
  if (e == hipSuccess) e = alloc_n(s, &P.sc, n);
  if (e == hipSuccess) e = alloc_n(s, &P.counts, 4);
  if (e == hipSuccess) e = alloc_n(s, &P.stats, 2);
  if (e == hipSuccess) e = alloc_n(s, &s->d_accum, npix * RRHIP_ACCUM_STRIDE);
  if (e == hipSuccess) e = hipMemset(s->d_accum, 0, npix * RRHIP_ACCUM_STRIDE * sizeof(double));
  if (e == hipSuccess) e = hipMemset(P.stats, 0, 2 * sizeof(unsigned long long));
  if (e != hipSuccess) {
    free_session(s);
    return fail_hip(e, "rrhip_session_create");
  }
  *out = s;
  return RRHIP_OK;
}

int rrhip_session_destroy(rrhip_session* s) {
  if (!s) return RRHIP_OK;
  (void)hipSetDevice(s->scene->ctx->device);
  (void)hipStreamSynchronize(s->scene->ctx->stream);
  free_session(s);
  return RRHIP_OK;
}

int rrhip_render_samples(rrhip_session* s, uint32_t start, uint32_t count) {
  if (int st = check_session(s)) return st;
  rrhip_scene* sc = s->scene;
  RRHIP_HIP_TRY(hipSetDevice(sc->ctx->device));
  hipStream_t st = sc->ctx->stream;
  RRHIP_HIP_TRY(hipEventRecord(sc->ctx->ev0, st));
  for (uint32_t done = 0; done < count;) {
    const uint32_t n = std::min(s->spp_pass, count - done);
    PtParams prm = s->prm;
    prm.spp_pass = n;
    prm.sample0 = s->sample_base + (uint64_t)start + done;
    const uint32_t total = n * prm.npix;
    RRHIP_HIP_TRY(pt_launch_raygen(sc->pt, s->P, prm, total, st));
    for (uint32_t b = 0; b < prm.max_depth; b++) {
      const uint32_t* q = (b & 1) ? s->P.q1 : s->P.q0;
      uint32_t* qn = (b & 1) ? s->P.q0 : s->P.q1;
      RRHIP_HIP_TRY(pt_launch_bounce(sc->pt, s->P, prm, q, qn, total, st));
    }
    RRHIP_HIP_TRY(pt_launch_resolve(s->P, prm, s->d_accum, st));
    done += n;
  }
  RRHIP_HIP_TRY(hipEventRecord(sc->ctx->ev1, st));
  RRHIP_HIP_TRY(hipEventSynchronize(sc->ctx->ev1));
  float ms = 0.f;
  RRHIP_HIP_TRY(hipEventElapsedTime(&ms, sc->ctx->ev0, sc->ctx->ev1));
  s->last_ms = ms;
  s->samples_done += count;
  return RRHIP_OK;
}

int rrhip_read_accum(rrhip_session* s, double* out) {
  if (int st = check_session(s)) return st;
  if (!out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_read_accum: null output");
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  const uint64_t bytes = (uint64_t)s->prm.npix * RRHIP_ACCUM_STRIDE * sizeof(double);
  RRHIP_HIP_TRY(hipMemcpyAsync(out, s->d_accum, bytes, hipMemcpyDeviceToHost, s->scene->ctx->stream));
  RRHIP_HIP_TRY(hipStreamSynchronize(s->scene->ctx->stream));
  return RRHIP_OK;
}

int rrhip_load_accum(rrhip_session* s, const double* in) {
  if (int st = check_session(s)) return st;
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  const uint64_t bytes = (uint64_t)s->prm.npix * RRHIP_ACCUM_STRIDE * sizeof(double);
  if (in) {
    RRHIP_HIP_TRY(hipMemcpyAsync(s->d_accum, in, bytes, hipMemcpyHostToDevice, s->scene->ctx->stream));
  } else {
    RRHIP_HIP_TRY(hipMemsetAsync(s->d_accum, 0, bytes, s->scene->ctx->stream));
  }
  RRHIP_HIP_TRY(hipStreamSynchronize(s->scene->ctx->stream));
  return RRHIP_OK;
}

int rrhip_relmse(rrhip_session* s, rrhip_session* ref, double eps, double* out) {
  if (int st = check_session(s)) return st;
  if (int st = check_session(ref)) return st;
  if (!out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_relmse: null output");
  if (s->prm.npix != ref->prm.npix || s->scene->ctx != ref->scene->ctx)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_relmse: sessions differ in size or context");
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  hipStream_t st = s->scene->ctx->stream;
  const uint32_t blocks = 512;
  double* partial = s->relmse_partial;
  if (!partial) {
    RRHIP_HIP_TRY(alloc_n(s, &s->relmse_partial, blocks));
    partial = s->relmse_partial;
  }
  RRHIP_HIP_TRY(pt_launch_relmse(s->d_accum, ref->d_accum, s->prm.npix, eps, partial, blocks, st));
  std::vector<double> h(blocks);
  RRHIP_HIP_TRY(hipMemcpyAsync(h.data(), partial, blocks * sizeof(double), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  double sum = 0.0;
  for (double v : h) sum += v;
  *out = sum / (3.0 * (double)s->prm.npix);
  return RRHIP_OK;
}

int rrhip_session_get_stats(rrhip_session* s, rrhip_session_stats* out) {
  if (int st = check_session(s)) return st;
  if (!out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_get_stats: null output");
  unsigned long long st[2] = {0, 0};
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  RRHIP_HIP_TRY(hipMemcpyAsync(st, s->P.stats, sizeof(st), hipMemcpyDeviceToHost, s->scene->ctx->stream));
  RRHIP_HIP_TRY(hipStreamSynchronize(s->scene->ctx->stream));
  *out = rrhip_session_stats{};
  out->extension_rays = st[0];
  out->shadow_rays = st[1];
  out->samples_rendered = s->samples_done;
  out->path_slots = s->slots;
  out->samples_per_pass = s->spp_pass;
  out->last_render_ms = s->last_ms;
  return RRHIP_OK;
}

}  // extern "C"
