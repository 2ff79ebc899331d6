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
#include "nrc_internal.h"
#include "pt_kernels.h"
#include "pt_scene.h"
#include "scene_internal.h"

using namespace rrhip;

struct rrhip_session {
  rrhip_scene* scene = nullptr;
  PtSceneView S{};  // the scene's view with this session's ray-origin rule
  PtParams prm{};
  uint64_t sample_base = 0;
  uint32_t spp_pass = 1;
  uint64_t slots = 0;       // rendering path slots (spp_pass x pixels)
  uint64_t path_cap = 0;    // slots of the path buffers (rendering + training)
  PathBuffers P{};
  std::vector<void*> mem;       // accumulator and small buffers
  std::vector<void*> path_mem;  // path buffers and radiance-cache pass state
  double* d_accum = nullptr;
  double* relmse_partial = nullptr;
  uint64_t samples_done = 0;
  double last_ms = 0.0;
  // Radiance cache (null: plain path tracing).
  rrhip_nrc* nrc = nullptr;
  NrcPathBuffers N{};
  uint32_t* nrc_scan = nullptr;
  NrcQuery* nrc_bq = nullptr;
  float* nrc_bt = nullptr;
  uint32_t nrc_bcap = 0;
  // Launch mode bit 0: one hipGraph per pass size, captured on first use
  // and replayed with the first sample index written to d_sample0; bit 1:
  // the live queue sorted by shading class before shade; bit 2: the cache's
  // optimiser steps stay on the rendering stream (no overlap).
  uint32_t launch_mode = 0;
  uint64_t* d_sample0 = nullptr;
  std::vector<std::pair<uint32_t, hipGraphExec_t>> graphs;
  // Adaptive sampling (tolerance 0: off): ping-pong lists of the pixels
  // still sampling, per-block offsets for their compaction and the device
  // count. Until the first check (`pruned` false) every pixel samples.
  double ad_tol = 0.0;
  uint32_t* ad_live[2] = {nullptr, nullptr};
  uint32_t* ad_bcount = nullptr;
  uint32_t* ad_nlive = nullptr;
  uint32_t ad_cur = 0;
  uint32_t nlive = 0;
  bool pruned = false;
  // Film mode: per-slot AOV buffers live with the path buffers; the film
  // accumulator (kFilmStride doubles per pixel) here.
  bool film = false;
  double* d_film = nullptr;
  // Profiling (rrhip_session_set_profiling): counting trace kernels whose
  // fetch totals land in prof_counts, events around each bounce's kernels
  // and the cache's inference and optimiser steps, device times summed per
  // pass; ray and cache counters are reported relative to prof_base.
  bool profiling = false;
  unsigned long long* prof_counts = nullptr;
  std::vector<hipEvent_t> prof_ev;  // 4 per bounce round, then inference and step begin / end
  unsigned long long prof_base[4] = {0, 0, 0, 0};
  double prof_ms[5] = {0, 0, 0, 0, 0};  // extend, shade, shadow, cache inference, cache steps
  uint64_t prof_passes = 0, prof_rounds = 0;
};

namespace {

template <class T>
hipError_t alloc_in(std::vector<void*>& mem, T** p, uint64_t n) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, std::max<uint64_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) mem.push_back(v);
  *p = (T*)v;
  return e;
}

template <class T>
hipError_t alloc_n(rrhip_session* s, T** p, uint64_t n) {
  return alloc_in(s->mem, p, n);
}

void free_paths(rrhip_session* s) {
  for (void* p : s->path_mem) (void)hipFree(p);
  s->path_mem.clear();
  unsigned long long* stats = s->P.stats;  // lives in s->mem
  s->P = PathBuffers{};
  s->P.stats = stats;
  s->N = NrcPathBuffers{};
  s->path_cap = 0;
}

void drop_graphs(rrhip_session* s) {
  for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.second);
  s->graphs.clear();
}

void free_profile_events(rrhip_session* s) {
  for (hipEvent_t e : s->prof_ev) (void)hipEventDestroy(e);
  s->prof_ev.clear();
}

void free_session(rrhip_session* s) {
  drop_graphs(s);
  free_paths(s);
  free_profile_events(s);
  if (s->prof_counts) (void)hipFree(s->prof_counts);
  if (s->d_film) (void)hipFree(s->d_film);
  for (void* p : s->mem) (void)hipFree(p);
  s->mem.clear();
  delete s;
}

// (Re)allocates the path buffers for n slots.
hipError_t alloc_paths(rrhip_session* s, uint64_t n) {
  free_paths(s);
  std::vector<void*>& m = s->path_mem;
  PathBuffers& P = s->P;
  hipError_t e = alloc_in(m, &P.o, n);
  if (e == hipSuccess) e = alloc_in(m, &P.d, n);
  if (e == hipSuccess) e = alloc_in(m, &P.beta, n);
  if (e == hipSuccess) e = alloc_in(m, &P.L, n);
  if (e == hipSuccess) e = alloc_in(m, &P.prev, n);
  if (e == hipSuccess) e = alloc_in(m, &P.rng, n);
  if (e == hipSuccess) e = alloc_in(m, &P.flags, n);
  if (e == hipSuccess) e = alloc_in(m, &P.hit_n, n);
  if (e == hipSuccess) e = alloc_in(m, &P.hit_m, n);
  if (e == hipSuccess) e = alloc_in(m, &P.hit_front, n);
  if (e == hipSuccess) e = alloc_in(m, &P.hit_back, n);
  if (e == hipSuccess) e = alloc_in(m, &P.q0, n);
  if (e == hipSuccess) e = alloc_in(m, &P.q1, n);
  if (e == hipSuccess) e = alloc_in(m, &P.sq, n);
  if (e == hipSuccess) e = alloc_in(m, &P.so, n);
  if (e == hipSuccess) e = alloc_in(m, &P.sd, n);
  if (e == hipSuccess) e = alloc_in(m, &P.sc, n);
  if (e == hipSuccess && s->scene->pt_eq) {
    e = alloc_in(m, &P.so2, n);
    if (e == hipSuccess) e = alloc_in(m, &P.sd2, n);
    if (e == hipSuccess) e = alloc_in(m, &P.sc2, n);
  }
  if (e == hipSuccess && s->scene->pt.transparent) {
    e = alloc_in(m, &P.mstk, n);
    if (e == hipSuccess) e = alloc_in(m, &P.srng, n);
  }
  if (e == hipSuccess && s->scene->pt.motion) e = alloc_in(m, &P.stime, n);
  if (e == hipSuccess && s->scene->pt.patterns) {
    e = alloc_in(m, &P.hit_st, n);
    if (e == hipSuccess) e = alloc_in(m, &P.hit_gs, n);
    if (e == hipSuccess) e = alloc_in(m, &P.hit_gt, n);
  }
  if (e == hipSuccess && s->film) {
    e = alloc_in(m, &P.ld, n);
    if (e == hipSuccess) e = alloc_in(m, &P.aux0, n);
    if (e == hipSuccess) e = alloc_in(m, &P.aux1, n);
    if (e == hipSuccess) e = alloc_in(m, &P.aux_id, n);
  }
  if (e == hipSuccess) e = alloc_in(m, &P.counts, 4);
  if (e == hipSuccess) s->path_cap = n;
  return e;
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
  if (!(params->ray_offset >= 0.0 && params->ray_offset < 1e30))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_create: ray_offset must be finite and >= 0");
  const uint64_t npix = (uint64_t)scene->pt.cam.width * scene->pt.cam.height;
  uint32_t spp = params->samples_per_pass;
  if (spp == 0) spp = (uint32_t)std::min<uint64_t>(32, std::max<uint64_t>(1, (4ull << 20) / npix));
  if ((uint64_t)spp * npix >= (1ull << 31))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_create: samples_per_pass x pixels exceeds 2^31");
  if (hipSetDevice(scene->ctx->device) != hipSuccess) return fail(RRHIP_ERR_DEVICE, "hipSetDevice");
  auto* s = new (std::nothrow) rrhip_session();
  if (!s) return fail(RRHIP_ERR_OUT_OF_MEMORY, "rrhip_session_create: host allocation failed");
  s->scene = scene;
  s->S = scene->pt;
  s->S.ray_offset = (float)params->ray_offset;
  s->spp_pass = spp;
  s->slots = (uint64_t)spp * npix;
  s->sample_base = params->sample_base;
  s->prm.max_depth = params->max_depth;
  s->prm.rr_start = params->rr_start;
  s->prm.firefly_clamp = (float)params->firefly_clamp;
  s->prm.npix = (uint32_t)npix;
  hipError_t e = alloc_paths(s, s->slots);
  unsigned long long* stats = nullptr;
  if (e == hipSuccess) e = alloc_n(s, &stats, 4);
  s->P.stats = stats;
  if (e == hipSuccess) e = alloc_n(s, &s->d_accum, npix * RRHIP_ACCUM_STRIDE);
  // cleared on the context's stream (non-blocking, so not ordered after a
  // null-stream memset) and finished before the session is handed out
  hipStream_t zst = scene->ctx->stream;
  if (e == hipSuccess) e = hipMemsetAsync(s->d_accum, 0, npix * RRHIP_ACCUM_STRIDE * sizeof(double), zst);
  if (e == hipSuccess) e = hipMemsetAsync(stats, 0, 4 * sizeof(unsigned long long), zst);
  if (e == hipSuccess) e = hipStreamSynchronize(zst);
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
  if (s->nrc) (void)nrc_drain(s->nrc);  // its steps may still read this session's buffers
  free_session(s);
  return RRHIP_OK;
}

}  // extern "C"

namespace {

// Enqueues one pass of n samples of each of ncol pixels (raygen, bounces,
// cache work, resolve): the pixels of list `live`, or all when it is null.
// sample0_dev: null = sample0 is used as given.
hipError_t enqueue_pass(rrhip_session* s, uint32_t n, uint32_t ncol, const uint32_t* live, uint64_t sample0,
                        const uint64_t* sample0_dev, hipStream_t st) {
  rrhip_scene* sc = s->scene;
  rrhip_nrc* nrc = s->nrc;
  NrcPathBuffers N = nrc ? s->N : NrcPathBuffers{};
  PtParams prm = s->prm;
  prm.spp_pass = n;
  prm.sample0 = sample0;
  prm.npix = ncol;
  prm.live = live;
  const uint32_t total = n * ncol;
  N.train_base = total;  // training paths follow this pass's rendering paths
  const uint32_t paths = total + N.ntrain;
  PathBuffers P = s->P;
  if (!(s->launch_mode & 2)) P.qs = nullptr;
  const uint32_t rounds = prm.max_depth + (sc->pt.transparent ? kPtMaxSkips : 0);
  // profiling (never under graph capture): the counting kernels and events
  const bool prof = s->profiling && !sample0_dev;
  hipEvent_t* ev = nullptr;
  hipError_t e = hipSuccess;
  if (prof) {
    P.prof = s->prof_counts;
    while (e == hipSuccess && s->prof_ev.size() < 4 * (size_t)rounds + 4) {
      hipEvent_t x = nullptr;
      e = hipEventCreate(&x);
      if (e == hipSuccess) s->prof_ev.push_back(x);
    }
    ev = s->prof_ev.data();
  }
  if (e == hipSuccess) e = pt_launch_raygen(s->S, P, prm, total, N, sample0_dev, st);
  for (uint32_t b = 0; b < rounds && e == hipSuccess; b++) {
    const uint32_t* q = (b & 1) ? P.q1 : P.q0;
    uint32_t* qn = (b & 1) ? P.q0 : P.q1;
    e = pt_launch_bounce(s->S, P, prm, q, qn, paths, N, st, prof ? ev + 4 * b : nullptr);
  }
  hipEvent_t* ev_nrc = prof ? ev + 4 * rounds : nullptr;
  bool nrc_timed = false, steps_timed = false;
  hipStream_t ts_used = st;
  if (nrc && e == hipSuccess) {
    // cache radiance at the terminal vertices (and training tails) once the
    // previous pass's optimiser steps are done, then this pass's training
    // samples and optimiser steps: on the cache's training stream, where
    // they overlap the next pass's tracing (also across calls), unless
    // launch mode bit 2 or graph capture keeps them on this stream
    const bool overlap = !sample0_dev && !(s->launch_mode & 4);
    e = nrc_join(nrc, st);
    if (ev_nrc && e == hipSuccess) e = hipEventRecord(ev_nrc[0], st);
    if (e == hipSuccess) e = nrc_launch_infer_paths(nrc_infer_view(nrc), N.q, &N.counts[0], paths, N.q_slot, N.q_beta, s->P.L, st);
    if (ev_nrc && e == hipSuccess) {
      e = hipEventRecord(ev_nrc[1], st);
      nrc_timed = e == hipSuccess;
    }
    if (e == hipSuccess)
      e = pt_launch_nrc_samples(s->P, N, prm.firefly_clamp, s->nrc_scan, s->nrc_bcap, s->nrc_bq, s->nrc_bt, st);
    const uint32_t batch = nrc->cfg.batch;
    const bool steps = N.ntrain > 0 && nrc->cfg.steps_per_pass > 0;
    hipStream_t ts = st;
    if (steps && overlap && e == hipSuccess) {
      e = hipEventRecord(nrc->ev_samples, st);
      if (e == hipSuccess) e = hipStreamWaitEvent(nrc->train_st, nrc->ev_samples, 0);
      ts = nrc->train_st;
    }
    if (ev_nrc && steps && e == hipSuccess) e = hipEventRecord(ev_nrc[2], ts);
    for (uint32_t k = 0; k < nrc->cfg.steps_per_pass && steps && e == hipSuccess; k++) {
      if ((uint64_t)k * batch >= s->nrc_bcap) break;
      e = pt_launch_nrc_set_batch(N, k, batch, &nrc->d_ctl->n, ts);
      if (e == hipSuccess)
        e = nrc_step_device(nrc, s->nrc_bq + (size_t)k * batch, s->nrc_bt + 3 * (size_t)k * batch, batch, ts);
    }
    if (ev_nrc && steps && e == hipSuccess) {
      e = hipEventRecord(ev_nrc[3], ts);
      steps_timed = e == hipSuccess;
      ts_used = ts;
    }
    if (steps && overlap && e == hipSuccess) {
      e = hipEventRecord(nrc->ev_trained, ts);
      nrc->pending = e == hipSuccess;
    }
  }
  if (e == hipSuccess) e = pt_launch_resolve(s->P, prm, s->d_accum, s->film ? s->d_film : nullptr, st);
  if (prof && e == hipSuccess) {
    // wait for the pass (and its optimiser steps) and sum the kernel times
    e = hipStreamSynchronize(st);
    if (e == hipSuccess && ts_used != st) e = hipStreamSynchronize(ts_used);
    float ms = 0.f;
    for (uint32_t b = 0; b < rounds && e == hipSuccess; b++)
      for (int k = 0; k < 3 && e == hipSuccess; k++) {
        e = hipEventElapsedTime(&ms, ev[4 * b + k], ev[4 * b + k + 1]);
        s->prof_ms[k] += ms;
      }
    if (e == hipSuccess && nrc_timed) {
      e = hipEventElapsedTime(&ms, ev_nrc[0], ev_nrc[1]);
      s->prof_ms[3] += ms;
    }
    if (e == hipSuccess && steps_timed) {
      e = hipEventElapsedTime(&ms, ev_nrc[2], ev_nrc[3]);
      s->prof_ms[4] += ms;
    }
    s->prof_passes++;
    s->prof_rounds += rounds;
  }
  return e;
}

// The pass graph of size n (captured on first use).
hipError_t pass_graph(rrhip_session* s, uint32_t n, hipStream_t st, hipGraphExec_t* out) {
  for (auto& g : s->graphs)
    if (g.first == n) {
      *out = g.second;
      return hipSuccess;
    }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t pe = enqueue_pass(s, n, s->prm.npix, nullptr, 0, s->d_sample0, st);
  e = hipStreamEndCapture(st, &graph);
  if (pe != hipSuccess) e = pe;
  hipGraphExec_t exec = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return e;
  s->graphs.push_back({n, exec});
  *out = exec;
  return hipSuccess;
}

// Adaptive sampling check: drops the pixels that met the stopping rule
// from the live list (in order) and reads the new count.
hipError_t adaptive_prune(rrhip_session* s, hipStream_t st) {
  const uint32_t* in = s->pruned ? s->ad_live[s->ad_cur] : nullptr;
  const uint32_t n = s->pruned ? s->nlive : s->prm.npix;
  if (n == 0) return hipSuccess;
  uint32_t* out = s->ad_live[s->ad_cur ^ 1];
  hipError_t e = pt_launch_adaptive_prune(s->d_accum, in, n, s->ad_tol, s->ad_bcount, out, s->ad_nlive, st);
  uint32_t live = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&live, s->ad_nlive, sizeof(live), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  s->ad_cur ^= 1;
  s->nlive = live;
  s->pruned = true;
  return hipSuccess;
}

}  // namespace

extern "C" {

int rrhip_render_samples(rrhip_session* s, uint32_t start, uint32_t count) {
  if (int st = check_session(s)) return st;
  rrhip_scene* sc = s->scene;
  RRHIP_HIP_TRY(hipSetDevice(sc->ctx->device));
  hipStream_t st = sc->ctx->stream;
  RRHIP_HIP_TRY(hipEventRecord(sc->ctx->ev0, st));
  for (uint32_t done = 0; done < count;) {
    uint32_t n = std::min(s->spp_pass, count - done);
    const uint64_t sample0 = s->sample_base + (uint64_t)start + done;
    if (s->ad_tol > 0.0) {
      // pt/mod.rs:306-327: pixels are checked before samples 32, 48, 64, ...
      // (sample index relative to the session's base); passes end there.
      const uint64_t k = (uint64_t)start + done;
      if (k >= 32 && k % 16 == 0) RRHIP_HIP_TRY(adaptive_prune(s, st));
      if (s->pruned && s->nlive == 0) break;
      const uint64_t next = k < 32 ? 32 : (k / 16 + 1) * 16;
      n = (uint32_t)std::min<uint64_t>(n, next - k);
      if (s->pruned)
        RRHIP_HIP_TRY(enqueue_pass(s, n, s->nlive, s->ad_live[s->ad_cur], sample0, nullptr, st));
      else
        RRHIP_HIP_TRY(enqueue_pass(s, n, s->prm.npix, nullptr, sample0, nullptr, st));
    } else if ((s->launch_mode & 1) && !s->nrc && !s->profiling) {
      hipGraphExec_t exec = nullptr;
      RRHIP_HIP_TRY(pass_graph(s, n, st, &exec));
      RRHIP_HIP_TRY(pt_launch_set_u64(s->d_sample0, sample0, st));
      RRHIP_HIP_TRY(hipGraphLaunch(exec, st));
    } else {
      RRHIP_HIP_TRY(enqueue_pass(s, n, s->prm.npix, nullptr, sample0, nullptr, st));
    }
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

int rrhip_session_set_launch_mode(rrhip_session* s, uint32_t mode) {
  if (int st = check_session(s)) return st;
  if (mode > 7) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_set_launch_mode: mode must be 0..7");
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  RRHIP_HIP_TRY(hipStreamSynchronize(s->scene->ctx->stream));
  if ((mode & 1) && !s->d_sample0) RRHIP_HIP_TRY(alloc_n(s, &s->d_sample0, 1));
  drop_graphs(s);
  s->launch_mode = mode;
  return RRHIP_OK;
}

int rrhip_session_autotune(rrhip_session* s, uint32_t samples, uint32_t* best_mode, float* best_ms) {
  if (int st = check_session(s)) return st;
  if (samples == 0) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_autotune: samples must be > 0");
  rrhip_scene* sc = s->scene;
  RRHIP_HIP_TRY(hipSetDevice(sc->ctx->device));
  hipStream_t st = sc->ctx->stream;
  // Keep the accumulator, the film accumulator and the statistics: the
  // trial renders use sample indices past 2^40 and are discarded, on every
  // exit path (a cache is detached, so the trials never train it).
  const uint64_t abytes = (uint64_t)s->prm.npix * RRHIP_ACCUM_STRIDE * sizeof(double);
  const uint64_t fbytes = s->film ? (uint64_t)s->prm.npix * kFilmStride * sizeof(double) : 0;
  std::vector<double> keep(abytes / sizeof(double)), keep_film(fbytes / sizeof(double));
  unsigned long long stats[4];
  RRHIP_HIP_TRY(hipMemcpyAsync(keep.data(), s->d_accum, abytes, hipMemcpyDeviceToHost, st));
  if (fbytes) RRHIP_HIP_TRY(hipMemcpyAsync(keep_film.data(), s->d_film, fbytes, hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(stats, s->P.stats, sizeof(stats), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  const uint64_t base = s->sample_base;
  const uint64_t done = s->samples_done;
  const uint32_t mode0 = s->launch_mode;
  rrhip_nrc* nrc = s->nrc;
  const double tol = s->ad_tol;
  s->nrc = nullptr;  // trials never train the cache nor prune pixels
  s->ad_tol = 0.0;
  s->sample_base = base + (1ull << 40);
  uint32_t best = 0;
  float best_t = 0.f;
  int status = RRHIP_OK;
  for (uint32_t mode = 0; mode < 4 && status == RRHIP_OK; mode++) {
    status = rrhip_session_set_launch_mode(s, mode | (mode0 & 4));
    if (status == RRHIP_OK) status = rrhip_render_samples(s, 0, std::min(samples, s->spp_pass));  // warm-up / capture
    if (status == RRHIP_OK) status = rrhip_render_samples(s, 0, samples);
    if (status == RRHIP_OK && (mode == 0 || s->last_ms < best_t)) {
      best = mode;
      best_t = (float)s->last_ms;
    }
  }
  s->nrc = nrc;
  s->ad_tol = tol;
  s->sample_base = base;
  s->samples_done = done;
  // Restore whatever the trials touched, also after a failed trial (then
  // the launch mode goes back to the caller's; the first error is kept).
  int rs = rrhip_session_set_launch_mode(s, status == RRHIP_OK ? best | (mode0 & 4) : mode0);
  hipError_t e = hipMemcpyAsync(s->d_accum, keep.data(), abytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && fbytes) e = hipMemcpyAsync(s->d_film, keep_film.data(), fbytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->P.stats, stats, sizeof(stats), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (status == RRHIP_OK) status = rs;
  if (status == RRHIP_OK && e != hipSuccess) status = fail_hip(e, "rrhip_session_autotune");
  if (status != RRHIP_OK) return status;
  if (best_mode) *best_mode = best | (mode0 & 4);
  if (best_ms) *best_ms = best_t;
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
  if (s->film) {
    const uint64_t fbytes = (uint64_t)s->prm.npix * kFilmStride * sizeof(double);
    RRHIP_HIP_TRY(hipMemsetAsync(s->d_film, 0, fbytes, s->scene->ctx->stream));
  }
  RRHIP_HIP_TRY(hipStreamSynchronize(s->scene->ctx->stream));
  s->pruned = false;  // every pixel samples again until the next check
  return RRHIP_OK;
}

int rrhip_session_set_film(rrhip_session* s, int enable) {
  if (int st = check_session(s)) return st;
  if (enable && s->nrc)
    return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_session_set_film: not with a radiance cache attached");
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  RRHIP_HIP_TRY(hipStreamSynchronize(s->scene->ctx->stream));
  drop_graphs(s);  // captured passes hold the old buffers
  if (s->d_film) {
    (void)hipFree(s->d_film);
    s->d_film = nullptr;
  }
  s->film = enable != 0;
  hipError_t e = alloc_paths(s, s->path_cap);
  const uint64_t fbytes = (uint64_t)s->prm.npix * kFilmStride * sizeof(double);
  if (e == hipSuccess && s->film) e = hipMalloc((void**)&s->d_film, fbytes);
  if (e == hipSuccess && s->film) e = hipMemsetAsync(s->d_film, 0, fbytes, s->scene->ctx->stream);
  if (e == hipSuccess && s->film) e = hipStreamSynchronize(s->scene->ctx->stream);
  if (e != hipSuccess) {
    // leave a usable plain session behind
    (void)hipGetLastError();
    s->film = false;
    if (s->d_film) (void)hipFree(s->d_film);
    s->d_film = nullptr;
    if (alloc_paths(s, s->path_cap) != hipSuccess) return fail_hip(e, "rrhip_session_set_film (session unusable)");
    return fail_hip(e, "rrhip_session_set_film");
  }
  return RRHIP_OK;
}

int rrhip_read_film(rrhip_session* s, double* out) {
  if (int st = check_session(s)) return st;
  if (!out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_read_film: null output");
  if (!s->film) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_read_film: film mode is off");
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  const uint64_t npix = s->prm.npix;
  std::vector<double> acc(npix * RRHIP_ACCUM_STRIDE), film(npix * kFilmStride);
  hipStream_t st = s->scene->ctx->stream;
  RRHIP_HIP_TRY(hipMemcpyAsync(acc.data(), s->d_accum, acc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(film.data(), s->d_film, film.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  // pt/mod.rs:223-274: sample means; specular = max(beauty - diffuse, 0);
  // depth over the samples with a first real hit; the most voted id (the
  // first of equal counts) and its share of the samples.
  for (uint64_t p = 0; p < npix; p++) {
    const double* a = &acc[p * RRHIP_ACCUM_STRIDE];
    const double* f = &film[p * kFilmStride];
    double* o = out + p * RRHIP_FILM_CHANNELS;
    const double n = a[6] > 0.0 ? a[6] : 1.0;
    for (int c = 0; c < 3; c++) {
      o[c] = a[c] / n;
      o[3 + c] = f[c] / n;
      o[6 + c] = std::max(o[c] - o[3 + c], 0.0);
      o[9 + c] = f[3 + c] / n;
      o[12 + c] = f[6 + c] / n;
    }
    o[15] = f[10] > 0.0 ? f[9] / f[10] : 0.0;
    double best_id = 0.0, best_n = 0.0;
    for (uint32_t j = 0; j < kFilmIds; j++) {
      if (f[11 + kFilmIds + j] > best_n) {
        best_n = f[11 + kFilmIds + j];
        best_id = f[11 + j];
      }
    }
    o[16] = best_id;
    o[17] = best_n / n;
  }
  return RRHIP_OK;
}

int rrhip_session_set_adaptive(rrhip_session* s, double tolerance) {
  if (int st = check_session(s)) return st;
  if (!(tolerance >= 0.0) || tolerance > 1e30)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_set_adaptive: tolerance must be finite and >= 0");
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  RRHIP_HIP_TRY(hipStreamSynchronize(s->scene->ctx->stream));
  if (tolerance > 0.0 && !s->ad_nlive) {
    const uint64_t npix = s->prm.npix;
    RRHIP_HIP_TRY(alloc_n(s, &s->ad_live[0], npix));
    RRHIP_HIP_TRY(alloc_n(s, &s->ad_live[1], npix));
    RRHIP_HIP_TRY(alloc_n(s, &s->ad_bcount, (npix + 255) / 256));
    RRHIP_HIP_TRY(alloc_n(s, &s->ad_nlive, 1));
  }
  s->ad_tol = tolerance;
  s->pruned = false;
  return RRHIP_OK;
}

int rrhip_session_average_spp(rrhip_session* s, double* out) {
  if (int st = check_session(s)) return st;
  if (!out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_average_spp: null output");
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  hipStream_t st = s->scene->ctx->stream;
  const uint32_t blocks = 512;
  if (!s->relmse_partial) RRHIP_HIP_TRY(alloc_n(s, &s->relmse_partial, blocks));
  RRHIP_HIP_TRY(pt_launch_count_sum(s->d_accum, s->prm.npix, s->relmse_partial, blocks, st));
  std::vector<double> h(blocks);
  RRHIP_HIP_TRY(hipMemcpyAsync(h.data(), s->relmse_partial, blocks * sizeof(double), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  double sum = 0.0;
  for (double v : h) sum += v;
  *out = sum / (double)s->prm.npix;
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

int rrhip_pcg32_draws(rrhip_ctx* ctx, const uint64_t* pixels, const uint64_t* samples, uint32_t n, uint32_t count,
                      uint32_t* out) {
  if (!ctx) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_pcg32_draws: null context");
  if (n && (!pixels || !samples || !out)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_pcg32_draws: null buffer");
  if (count == 0 || count > 4096) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_pcg32_draws: count must be 1..4096");
  if (n == 0) return RRHIP_OK;
  RRHIP_HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t ib = (size_t)n * sizeof(uint64_t), ob = (size_t)n * count * sizeof(uint32_t);
  char* d = nullptr;
  RRHIP_HIP_TRY(hipMalloc((void**)&d, 2 * ib + ob));
  uint64_t* dp = (uint64_t*)d;
  uint64_t* ds = (uint64_t*)(d + ib);
  uint32_t* dout = (uint32_t*)(d + 2 * ib);
  hipError_t e = hipMemcpyAsync(dp, pixels, ib, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(ds, samples, ib, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = pt_launch_pcg32_draws(dp, ds, n, count, dout, st);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  if (e != hipSuccess) return fail_hip(e, "rrhip_pcg32_draws");
  return RRHIP_OK;
}

int rrhip_hair_bsdf(rrhip_ctx* ctx, const rrhip_material_desc* material, uint32_t n, const float* wo,
                    const float* wi, const float* h, const float* u, float* f, float* pdf, float* ws, float* fs,
                    float* pdfs) {
  if (!ctx || !material) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_hair_bsdf: null argument");
  if (n && (!wo || !wi || !h || !u || !f || !pdf || !ws || !fs || !pdfs))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_hair_bsdf: null buffer");
  if (material->hair != 1) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_hair_bsdf: not a hair material");
  if (n > (1u << 24)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_hair_bsdf: at most 2^24 queries");
  if (n == 0) return RRHIP_OK;
  const PtMaterial m = pt_material(*material);
  RRHIP_HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  // inputs: wo, wi (3 each), h, u (4); outputs: f (3), pdf, ws (3), fs (3), pdfs
  const size_t nn = n;
  const size_t in_f = nn * 11, out_f = nn * 11;
  float* d = nullptr;
  RRHIP_HIP_TRY(hipMalloc((void**)&d, (in_f + out_f) * sizeof(float)));
  float *dwo = d, *dwi = d + 3 * nn, *dh = d + 6 * nn, *du = d + 7 * nn;
  float *df = d + in_f, *dpdf = df + 3 * nn, *dws = dpdf + nn, *dfs = dws + 3 * nn, *dpdfs = dfs + 3 * nn;
  hipError_t e = hipMemcpyAsync(dwo, wo, 3 * nn * sizeof(float), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dwi, wi, 3 * nn * sizeof(float), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dh, h, nn * sizeof(float), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(du, u, 4 * nn * sizeof(float), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = pt_launch_hair_bsdf(m, n, dwo, dwi, dh, du, df, dpdf, dws, dfs, dpdfs, st);
  if (e == hipSuccess) e = hipMemcpyAsync(f, df, 3 * nn * sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(pdf, dpdf, nn * sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(ws, dws, 3 * nn * sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(fs, dfs, 3 * nn * sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(pdfs, dpdfs, nn * sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  if (e != hipSuccess) return fail_hip(e, "rrhip_hair_bsdf");
  return RRHIP_OK;
}

int rrhip_session_set_profiling(rrhip_session* s, int enable) {
  if (int st = check_session(s)) return st;
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  hipStream_t st = s->scene->ctx->stream;
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  if (s->nrc) RRHIP_HIP_TRY(nrc_drain(s->nrc));
  if (!enable) {
    s->profiling = false;
    free_profile_events(s);
    if (s->prof_counts) (void)hipFree(s->prof_counts);
    s->prof_counts = nullptr;
    return RRHIP_OK;
  }
  if (!s->prof_counts) RRHIP_HIP_TRY(hipMalloc(&s->prof_counts, 8 * sizeof(unsigned long long)));
  RRHIP_HIP_TRY(hipMemsetAsync(s->prof_counts, 0, 8 * sizeof(unsigned long long), st));
  RRHIP_HIP_TRY(hipMemcpyAsync(s->prof_base, s->P.stats, sizeof(s->prof_base), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  for (double& m : s->prof_ms) m = 0.0;
  s->prof_passes = s->prof_rounds = 0;
  s->profiling = true;
  return RRHIP_OK;
}

int rrhip_session_get_profile(rrhip_session* s, rrhip_pt_profile* out) {
  if (int st = check_session(s)) return st;
  if (!out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_get_profile: null output");
  if (!s->profiling) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_get_profile: profiling is off");
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  hipStream_t st = s->scene->ctx->stream;
  unsigned long long c[6], r[4];
  RRHIP_HIP_TRY(hipMemcpyAsync(c, s->prof_counts, sizeof(c), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(r, s->P.stats, sizeof(r), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  *out = rrhip_pt_profile{};
  out->extend_rays = r[0] - s->prof_base[0];
  out->extend_nodes = c[0];
  out->extend_triangles = c[1];
  out->extend_instances = c[2];
  out->shadow_rays = r[1] - s->prof_base[1];
  out->shadow_nodes = c[3];
  out->shadow_triangles = c[4];
  out->shadow_instances = c[5];
  out->cache_queries = r[2] - s->prof_base[2];
  out->cache_samples = r[3] - s->prof_base[3];
  out->passes = s->prof_passes;
  out->bounce_rounds = s->prof_rounds;
  out->extend_ms = s->prof_ms[0];
  out->shade_ms = s->prof_ms[1];
  out->shadow_ms = s->prof_ms[2];
  out->cache_infer_ms = s->prof_ms[3];
  out->cache_steps_ms = s->prof_ms[4];
  return RRHIP_OK;
}

int rrhip_session_get_stats(rrhip_session* s, rrhip_session_stats* out) {
  if (int st = check_session(s)) return st;
  if (!out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_get_stats: null output");
  unsigned long long st[4] = {0, 0, 0, 0};
  RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
  RRHIP_HIP_TRY(hipMemcpyAsync(st, s->P.stats, sizeof(st), hipMemcpyDeviceToHost, s->scene->ctx->stream));
  RRHIP_HIP_TRY(hipStreamSynchronize(s->scene->ctx->stream));
  *out = rrhip_session_stats{};
  out->extension_rays = st[0];
  out->shadow_rays = st[1];
  out->cache_queries = st[2];
  out->cache_samples = st[3];
  out->samples_rendered = s->samples_done;
  out->path_slots = s->slots;
  out->samples_per_pass = s->spp_pass;
  out->last_render_ms = s->last_ms;
  return RRHIP_OK;
}

}  // extern "C"

extern "C" int rrhip_session_set_nrc(rrhip_session* s, rrhip_nrc* nrc) {
  if (int st = check_session(s)) return st;
  if (!nrc) {
    RRHIP_HIP_TRY(hipSetDevice(s->scene->ctx->device));
    if (s->nrc) RRHIP_HIP_TRY(nrc_drain(s->nrc));
    drop_graphs(s);
    s->nrc = nullptr;
    return RRHIP_OK;
  }
  rrhip_scene* sc = s->scene;
  if (nrc->ctx != sc->ctx) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_set_nrc: cache and session differ in context");
  if (sc->pt.atmosphere >= 0 || sc->pt.transparent)
    return fail(RRHIP_ERR_UNSUPPORTED,
                "rrhip_session_set_nrc: the radiance cache does not model participating media or cutouts");
  if (s->film) return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_session_set_nrc: not in film mode");
  if (sc->pt.hair) return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_session_set_nrc: the radiance cache does not model hair fibers");
  if (sc->pt.patterns)
    return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_session_set_nrc: the radiance cache does not model pattern-bound materials");
  const uint64_t ntrain = nrc->cfg.train_paths;
  const uint64_t cap = s->slots + ntrain;
  if (cap >= (1ull << 31)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_session_set_nrc: path slots exceed 2^31");
  RRHIP_HIP_TRY(hipSetDevice(sc->ctx->device));
  RRHIP_HIP_TRY(hipStreamSynchronize(sc->ctx->stream));
  if (s->nrc) RRHIP_HIP_TRY(nrc_drain(s->nrc));
  RRHIP_HIP_TRY(nrc_drain(nrc));  // its steps may still read another session's buffers
  drop_graphs(s);  // captured passes hold the old buffers
  if (nrc->bounds_from_scene) nrc_set_bounds(nrc, sc->pt_lo, sc->pt_hi);
  s->nrc = nullptr;
  hipError_t e = alloc_paths(s, cap);
  std::vector<void*>& m = s->path_mem;
  NrcPathBuffers& N = s->N;
  const uint32_t D = s->prm.max_depth;
  if (e == hipSuccess) e = alloc_in(m, &N.q, cap);
  if (e == hipSuccess) e = alloc_in(m, &N.q_slot, cap);
  if (e == hipSuccess) e = alloc_in(m, &N.q_beta, cap);
  if (e == hipSuccess) e = alloc_in(m, &N.counts, 2);
  if (e == hipSuccess) e = alloc_in(m, &N.tq, ntrain * D);
  if (e == hipSuccess) e = alloc_in(m, &N.tb, ntrain * D);
  if (e == hipSuccess) e = alloc_in(m, &N.tl, ntrain * D);
  if (e == hipSuccess) e = alloc_in(m, &N.tn, ntrain);
  if (e == hipSuccess) e = alloc_in(m, &s->nrc_scan, ntrain);
  const uint64_t bcap = std::min<uint64_t>(ntrain * D, (uint64_t)nrc->cfg.steps_per_pass * nrc->cfg.batch);
  s->nrc_bcap = (uint32_t)bcap;
  if (e == hipSuccess) e = alloc_in(m, &s->nrc_bq, bcap);
  if (e == hipSuccess) e = alloc_in(m, &s->nrc_bt, 3 * bcap);
  if (e != hipSuccess) {
    // leave a usable plain session behind
    (void)hipGetLastError();
    free_paths(s);
    if (alloc_paths(s, s->slots) != hipSuccess) return fail_hip(e, "rrhip_session_set_nrc (session unusable)");
    return fail_hip(e, "rrhip_session_set_nrc");
  }
  N.train_base = (uint32_t)s->slots;
  N.ntrain = (uint32_t)ntrain;
  N.max_vertices = D;
  N.spread = nrc->cfg.spread;
  s->nrc = nrc;
  return RRHIP_OK;
}
