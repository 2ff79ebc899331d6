// Batched path-tracing jobs (C ABI): many scenes advanced through one
// launch sequence per pass (pt_launch_batch_pass).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "api_internal.h"
#include "pt_kernels.h"
#include "scene_internal.h"

using namespace rrhip;

struct rrhip_batch {
  rrhip_ctx* ctx = nullptr;
  std::vector<rrhip_batch_job> jobs;
  std::vector<uint32_t> npix;
  uint32_t spp_pass = 1;
  bool auto_schedule = false;
  uint64_t slot_budget = 0;  // auto schedule: path slots per pass
  uint64_t cap = 0;          // path slots
  uint64_t total_pix = 0;
  PathBuffers P{};
  std::vector<void*> mem;
  std::vector<double*> accum;
  PtJob* d_jobs = nullptr;  // one table per pass, each `jobs.size()` entries
  uint32_t passes = 0;
  unsigned long long last_rays[2] = {0, 0};
  // One persistent launch (rrhip_batch_render_persistent): a path slot per
  // resident lane, the job table in item order, the work counter, per-job
  // pixels left and completion clocks (+ the launch's first block start).
  uint32_t p_flags = 0, p_lanes = 0;
  PathBuffers PP{};
  PtJob* d_pjobs = nullptr;
  uint32_t* d_next = nullptr;
  uint32_t* d_left = nullptr;
  unsigned long long* d_clock = nullptr;
};

namespace {

template <class T>
hipError_t alloc_n(rrhip_batch* b, T** p, uint64_t n) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, std::max<uint64_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) b->mem.push_back(v);
  *p = (T*)v;
  return e;
}

void free_batch(rrhip_batch* b) {
  for (void* p : b->mem) (void)hipFree(p);
  delete b;
}

// The pass schedule: for each pass, the jobs rendering in it with their
// sample count, first sample, slot and pixel offsets. With a fixed
// samples_per_pass every unfinished job renders that many samples per
// pass; with samples_per_pass = 0 (auto) passes are filled up to a slot
// budget shortest job first (fewest samples x pixels), so small jobs
// finish in the first passes while large ones keep the device full.
std::vector<std::vector<PtJob>> schedule(const rrhip_batch* b) {
  const size_t nj = b->jobs.size();
  std::vector<uint64_t> done(nj, 0);
  std::vector<size_t> order(nj);
  for (size_t j = 0; j < nj; j++) order[j] = j;
  if (b->auto_schedule)
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
      return (uint64_t)b->jobs[x].spp * b->npix[x] < (uint64_t)b->jobs[y].spp * b->npix[y];
    });
  std::vector<std::vector<PtJob>> out;
  for (;;) {
    std::vector<std::pair<size_t, uint32_t>> take;  // (job, samples this pass)
    uint64_t left = b->slot_budget;
    for (size_t j : order) {
      const uint64_t rem = b->jobs[j].spp - done[j];
      if (rem == 0) continue;
      uint64_t n;
      if (b->auto_schedule) {
        n = std::min<uint64_t>(rem, left / b->npix[j]);
        if (n == 0 && take.empty()) n = 1;  // a job larger than the budget still progresses
      } else {
        n = std::min<uint64_t>(rem, b->spp_pass);
      }
      if (n == 0) continue;
      take.push_back({j, (uint32_t)n});
      left -= std::min<uint64_t>(left, n * b->npix[j]);
    }
    if (take.empty()) break;
    std::sort(take.begin(), take.end());  // slots in job order
    std::vector<PtJob> pass;
    uint32_t slot = 0, pix = 0;
    for (const auto& tj : take) {
      const size_t j = tj.first;
      const rrhip_batch_job& J = b->jobs[j];
      PtJob t{};
      t.S = J.scene->pt;
      t.S.ray_offset = (float)J.params.ray_offset;
      t.prm.max_depth = J.params.max_depth;
      t.prm.rr_start = J.params.rr_start;
      t.prm.firefly_clamp = (float)J.params.firefly_clamp;
      t.prm.npix = b->npix[j];
      t.prm.spp_pass = tj.second;
      t.prm.sample0 = J.params.sample_base + done[j];
      t.slot_begin = slot;
      t.pix_begin = pix;
      t.accum = b->accum[j];
      slot += t.prm.spp_pass * t.prm.npix;
      pix += t.prm.npix;
      done[j] += tj.second;
      pass.push_back(t);
    }
    out.push_back(std::move(pass));
  }
  return out;
}

}  // namespace

extern "C" {

int rrhip_batch_create(rrhip_ctx* ctx, uint32_t num_jobs, const rrhip_batch_job* jobs, uint32_t samples_per_pass,
                       rrhip_batch** out) {
  if (!ctx || !jobs || !out || num_jobs == 0) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: null argument or no jobs");
  *out = nullptr;
  uint64_t total_pix = 0;
  for (uint32_t j = 0; j < num_jobs; j++) {
    const rrhip_batch_job& J = jobs[j];
    if (!J.scene || !J.scene->has_pt) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: job without a path-tracing scene");
    if (J.scene->ctx != ctx) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: job scene from another context");
    if (J.params.struct_size != sizeof(rrhip_render_params))
      return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: render params struct_size mismatch");
    if (J.params.max_depth == 0 || J.params.max_depth > 255)
      return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: max_depth must be in 1..255");
    if (J.spp == 0) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: spp must be > 0");
    if (!(J.params.ray_offset >= 0.0 && J.params.ray_offset < 1e30))
      return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: ray_offset must be finite and >= 0");
    if (J.scene->pt.hair)
      return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_batch_create: hair fibers render in sessions, not in batches");
    if (J.scene->pt.patterns)
      return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_batch_create: pattern-bound materials render in sessions, not in batches");
    total_pix += (uint64_t)J.scene->pt.cam.width * J.scene->pt.cam.height;
  }
  uint32_t spp = samples_per_pass;
  if (spp > 0 && (uint64_t)spp * total_pix >= (1ull << 31))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: samples_per_pass x pixels exceeds 2^31");
  RRHIP_HIP_TRY(hipSetDevice(ctx->device));
  auto* b = new (std::nothrow) rrhip_batch();
  if (!b) return fail(RRHIP_ERR_OUT_OF_MEMORY, "rrhip_batch_create: host allocation failed");
  b->ctx = ctx;
  b->jobs.assign(jobs, jobs + num_jobs);
  b->spp_pass = spp;
  b->auto_schedule = spp == 0;
  b->slot_budget = 1ull << 24;
  b->total_pix = total_pix;
  hipError_t e = hipSuccess;
  for (uint32_t j = 0; j < num_jobs && e == hipSuccess; j++) {
    const uint32_t np = jobs[j].scene->pt.cam.width * jobs[j].scene->pt.cam.height;
    b->npix.push_back(np);
    double* a = nullptr;
    e = alloc_n(b, &a, (uint64_t)np * RRHIP_ACCUM_STRIDE);
    b->accum.push_back(a);
  }
  // capacity: the largest pass of the schedule
  std::vector<std::vector<PtJob>> sched;
  if (e == hipSuccess) {
    sched = schedule(b);
    for (const auto& pass : sched) {
      uint64_t n = 0;
      for (const PtJob& t : pass) n += (uint64_t)t.prm.spp_pass * t.prm.npix;
      b->cap = std::max(b->cap, n);
    }
    if (b->cap >= (1ull << 31)) {
      free_batch(b);
      return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_create: a pass exceeds 2^31 path slots");
    }
  }
  PathBuffers& P = b->P;
  const uint64_t n = b->cap;
  if (e == hipSuccess) e = alloc_n(b, &P.o, n);
  if (e == hipSuccess) e = alloc_n(b, &P.d, n);
  if (e == hipSuccess) e = alloc_n(b, &P.beta, n);
  if (e == hipSuccess) e = alloc_n(b, &P.L, n);
  if (e == hipSuccess) e = alloc_n(b, &P.prev, n);
  if (e == hipSuccess) e = alloc_n(b, &P.rng, n);
  if (e == hipSuccess) e = alloc_n(b, &P.flags, n);
  if (e == hipSuccess) e = alloc_n(b, &P.hit_n, n);
  if (e == hipSuccess) e = alloc_n(b, &P.hit_m, n);
  if (e == hipSuccess) e = alloc_n(b, &P.hit_front, n);
  if (e == hipSuccess) e = alloc_n(b, &P.hit_back, n);
  if (e == hipSuccess) e = alloc_n(b, &P.q0, n);
  if (e == hipSuccess) e = alloc_n(b, &P.q1, n);
  if (e == hipSuccess) e = alloc_n(b, &P.sq, n);
  if (e == hipSuccess) e = alloc_n(b, &P.so, n);
  if (e == hipSuccess) e = alloc_n(b, &P.sd, n);
  if (e == hipSuccess) e = alloc_n(b, &P.sc, n);
  bool eq = false, transparent = false, motion = false;
  for (uint32_t j = 0; j < num_jobs; j++) {
    eq |= jobs[j].scene->pt_eq;
    transparent |= jobs[j].scene->pt.transparent != 0;
    motion |= jobs[j].scene->pt.motion != 0;
  }
  if (e == hipSuccess && eq) {
    e = alloc_n(b, &P.so2, n);
    if (e == hipSuccess) e = alloc_n(b, &P.sd2, n);
    if (e == hipSuccess) e = alloc_n(b, &P.sc2, n);
  }
  if (e == hipSuccess && transparent) {
    e = alloc_n(b, &P.mstk, n);
    if (e == hipSuccess) e = alloc_n(b, &P.srng, n);
  }
  if (e == hipSuccess && motion) e = alloc_n(b, &P.stime, n);
  if (e == hipSuccess) e = alloc_n(b, &P.job, n);
  if (e == hipSuccess) e = alloc_n(b, &P.counts, 8);
  if (e == hipSuccess) e = alloc_n(b, &P.stats, 4);
  // The whole schedule's job tables, uploaded once.
  if (e == hipSuccess) {
    b->passes = (uint32_t)sched.size();
    e = alloc_n(b, &b->d_jobs, (uint64_t)b->passes * num_jobs);
  }
  if (e == hipSuccess) {
    std::vector<PtJob> flat((size_t)b->passes * num_jobs);
    for (uint32_t k = 0; k < b->passes; k++) std::copy(sched[k].begin(), sched[k].end(), flat.begin() + (size_t)k * num_jobs);
    e = hipMemcpy(b->d_jobs, flat.data(), flat.size() * sizeof(PtJob), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    free_batch(b);
    return fail_hip(e, "rrhip_batch_create");
  }
  *out = b;
  return RRHIP_OK;
}

int rrhip_batch_destroy(rrhip_batch* b) {
  if (!b) return RRHIP_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  free_batch(b);
  return RRHIP_OK;
}

int rrhip_batch_render(rrhip_batch* b, double* job_ms, double* total_ms) {
  if (!b) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_render: null batch");
  RRHIP_HIP_TRY(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  const std::vector<std::vector<PtJob>> sched = schedule(b);
  const uint32_t nj = (uint32_t)b->jobs.size();
  for (uint32_t j = 0; j < nj; j++)
    RRHIP_HIP_TRY(hipMemsetAsync(b->accum[j], 0, (size_t)b->npix[j] * RRHIP_ACCUM_STRIDE * sizeof(double), st));
  RRHIP_HIP_TRY(hipMemsetAsync(b->P.stats, 0, 4 * sizeof(unsigned long long), st));
  std::vector<hipEvent_t> ev(sched.size() + 1, nullptr);
  for (auto& x : ev) RRHIP_HIP_TRY(hipEventCreate(&x));
  std::vector<uint32_t> last_pass(nj, 0);
  int status = RRHIP_OK;
  hipError_t e = hipEventRecord(ev[0], st);
  for (size_t k = 0; k < sched.size() && e == hipSuccess; k++) {
    const std::vector<PtJob>& pass = sched[k];
    uint32_t total = 0, pix = 0, depth = 0;
    bool quadrics = false;
    for (const PtJob& t : pass) {
      quadrics |= t.S.num_quadrics > 0 || t.S.curves;
      total += t.prm.spp_pass * t.prm.npix;
      pix += t.prm.npix;
      depth = std::max(depth, t.prm.max_depth + (t.S.transparent ? kPtMaxSkips : 0));
    }
    e = pt_launch_batch_pass(b->d_jobs + k * nj, (uint32_t)pass.size(), b->P, total, pix, depth, quadrics, st);
    if (e == hipSuccess) e = hipEventRecord(ev[k + 1], st);
    for (const PtJob& t : pass) {
      const size_t j = (size_t)(std::find(b->accum.begin(), b->accum.end(), t.accum) - b->accum.begin());
      last_pass[j] = (uint32_t)k + 1;
    }
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev.back());
  if (e == hipSuccess) {
    float ms = 0.f;
    for (uint32_t j = 0; j < nj && e == hipSuccess; j++) {
      e = hipEventElapsedTime(&ms, ev[0], ev[last_pass[j]]);
      if (job_ms) job_ms[j] = ms;
    }
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev.back());
    if (total_ms) *total_ms = ms;
  }
  if (e == hipSuccess) e = hipMemcpy(b->last_rays, b->P.stats, sizeof(b->last_rays), hipMemcpyDeviceToHost);
  if (e != hipSuccess) status = fail_hip(e, "rrhip_batch_render");
  for (auto& x : ev) (void)hipEventDestroy(x);
  return status;
}

}  // extern "C"

namespace {

// Path slots and work tables of the persistent launch (allocated on first
// use, kept with the batch).
hipError_t persistent_setup(rrhip_batch* b) {
  if (b->d_pjobs) return hipSuccess;
  const uint32_t nj = (uint32_t)b->jobs.size();
  bool eq = false;
  for (const rrhip_batch_job& J : b->jobs) {
    b->p_flags |= (J.scene->pt.transparent ? 1u : 0u) | (J.scene->pt.motion ? 2u : 0u) |
                  (J.scene->pt.num_quadrics || J.scene->pt.curves ? 4u : 0u);
    eq |= J.scene->pt_eq;
  }
  b->p_lanes = pt_batch_persistent_lanes(b->p_flags);
  if (b->p_lanes == 0) return hipErrorInvalidValue;
  const uint64_t n = b->p_lanes;
  PathBuffers& P = b->PP;
  hipError_t e = alloc_n(b, &P.o, n);
  if (e == hipSuccess) e = alloc_n(b, &P.d, n);
  if (e == hipSuccess) e = alloc_n(b, &P.beta, n);
  if (e == hipSuccess) e = alloc_n(b, &P.L, n);
  if (e == hipSuccess) e = alloc_n(b, &P.prev, n);
  if (e == hipSuccess) e = alloc_n(b, &P.rng, n);
  if (e == hipSuccess) e = alloc_n(b, &P.flags, n);
  if (e == hipSuccess) e = alloc_n(b, &P.hit_n, n);
  if (e == hipSuccess) e = alloc_n(b, &P.hit_m, n);
  if (e == hipSuccess) e = alloc_n(b, &P.hit_front, n);
  if (e == hipSuccess) e = alloc_n(b, &P.hit_back, n);
  if (e == hipSuccess) e = alloc_n(b, &P.q0, n);
  if (e == hipSuccess) e = alloc_n(b, &P.sq, n);
  if (e == hipSuccess) e = alloc_n(b, &P.so, n);
  if (e == hipSuccess) e = alloc_n(b, &P.sd, n);
  if (e == hipSuccess) e = alloc_n(b, &P.sc, n);
  if (e == hipSuccess && eq) {
    e = alloc_n(b, &P.so2, n);
    if (e == hipSuccess) e = alloc_n(b, &P.sd2, n);
    if (e == hipSuccess) e = alloc_n(b, &P.sc2, n);
  }
  if (e == hipSuccess && (b->p_flags & 1u)) {
    e = alloc_n(b, &P.mstk, n);
    if (e == hipSuccess) e = alloc_n(b, &P.srng, n);
  }
  if (e == hipSuccess && (b->p_flags & 2u)) e = alloc_n(b, &P.stime, n);
  if (e == hipSuccess) e = alloc_n(b, &P.stats, 4);
  if (e == hipSuccess) e = alloc_n(b, &b->d_next, 1);
  if (e == hipSuccess) e = alloc_n(b, &b->d_left, nj);
  if (e == hipSuccess) e = alloc_n(b, &b->d_clock, (uint64_t)nj + 1);
  if (e == hipSuccess) e = alloc_n(b, &b->d_pjobs, nj);
  return e;
}

}  // namespace

extern "C" {

int rrhip_batch_render_persistent(rrhip_batch* b, double* job_ms, double* total_ms) {
  if (!b) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_render_persistent: null batch");
  for (const rrhip_batch_job& J : b->jobs)
    if (J.params.ray_offset > 0.0)
      return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_batch_render_persistent: a job uses the fixed ray-origin push");
  RRHIP_HIP_TRY(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  RRHIP_HIP_TRY(persistent_setup(b));
  const uint32_t nj = (uint32_t)b->jobs.size();
  // items: every pixel of every job, jobs shortest first (samples x pixels)
  std::vector<uint32_t> order(nj);
  for (uint32_t j = 0; j < nj; j++) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    return (uint64_t)b->jobs[x].spp * b->npix[x] < (uint64_t)b->jobs[y].spp * b->npix[y];
  });
  std::vector<PtJob> table(nj);
  std::vector<uint32_t> left(nj), pos(nj);
  uint64_t items = 0;
  for (uint32_t k = 0; k < nj; k++) {
    const uint32_t j = order[k];
    const rrhip_batch_job& J = b->jobs[j];
    PtJob& t = table[k];
    t.S = J.scene->pt;
    t.S.ray_offset = 0.f;
    t.prm.max_depth = J.params.max_depth;
    t.prm.rr_start = J.params.rr_start;
    t.prm.firefly_clamp = (float)J.params.firefly_clamp;
    t.prm.npix = b->npix[j];
    t.prm.spp_pass = J.spp;
    t.prm.sample0 = J.params.sample_base;
    t.pix_begin = (uint32_t)items;
    t.accum = b->accum[j];
    left[k] = b->npix[j];
    pos[j] = k;
    items += b->npix[j];
  }
  if (items >= (1ull << 32)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_render_persistent: too many pixels");
  for (uint32_t j = 0; j < nj; j++)
    RRHIP_HIP_TRY(hipMemsetAsync(b->accum[j], 0, (size_t)b->npix[j] * RRHIP_ACCUM_STRIDE * sizeof(double), st));
  const std::vector<unsigned long long> clock0(nj + 1, ~0ull);
  const uint32_t zero = 0;
  RRHIP_HIP_TRY(hipMemsetAsync(b->PP.stats, 0, 4 * sizeof(unsigned long long), st));
  RRHIP_HIP_TRY(hipMemcpyAsync(b->d_pjobs, table.data(), nj * sizeof(PtJob), hipMemcpyHostToDevice, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(b->d_left, left.data(), nj * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(b->d_clock, clock0.data(), (nj + 1) * sizeof(unsigned long long),
                               hipMemcpyHostToDevice, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(b->d_next, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
  RRHIP_HIP_TRY(hipEventRecord(b->ctx->ev0, st));
  RRHIP_HIP_TRY(pt_launch_batch_persistent(b->d_pjobs, nj, b->PP, (uint32_t)items, b->p_lanes, b->p_flags, b->d_next,
                                           b->d_left, b->d_clock, st));
  RRHIP_HIP_TRY(hipEventRecord(b->ctx->ev1, st));
  RRHIP_HIP_TRY(hipEventSynchronize(b->ctx->ev1));
  float ms = 0.f;
  RRHIP_HIP_TRY(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
  if (total_ms) *total_ms = ms;
  std::vector<unsigned long long> clock(nj + 1);
  RRHIP_HIP_TRY(hipMemcpy(clock.data(), b->d_clock, (nj + 1) * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  RRHIP_HIP_TRY(hipMemcpy(b->last_rays, b->PP.stats, sizeof(b->last_rays), hipMemcpyDeviceToHost));
  if (job_ms) {
    int khz = 0;
    RRHIP_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->ctx->device));
    if (khz <= 0) return fail(RRHIP_ERR_DEVICE, "rrhip_batch_render_persistent: no device clock rate");
    for (uint32_t j = 0; j < nj; j++) job_ms[j] = (double)(clock[pos[j]] - clock[nj]) / (double)khz;
  }
  return RRHIP_OK;
}

int rrhip_batch_read_accum(rrhip_batch* b, uint32_t job, double* out) {
  if (!b || !out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_read_accum: null argument");
  if (job >= b->jobs.size()) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_read_accum: job out of range");
  RRHIP_HIP_TRY(hipSetDevice(b->ctx->device));
  RRHIP_HIP_TRY(hipMemcpyAsync(out, b->accum[job], (size_t)b->npix[job] * RRHIP_ACCUM_STRIDE * sizeof(double),
                               hipMemcpyDeviceToHost, b->ctx->stream));
  RRHIP_HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return RRHIP_OK;
}

int rrhip_batch_get_stats(rrhip_batch* b, rrhip_session_stats* out) {
  if (!b || !out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_batch_get_stats: null argument");
  *out = rrhip_session_stats{};
  out->extension_rays = b->last_rays[0];
  out->shadow_rays = b->last_rays[1];
  out->path_slots = b->cap;
  out->samples_per_pass = b->auto_schedule ? 0 : b->spp_pass;
  out->reserved = b->passes;
  return RRHIP_OK;
}

}  // extern "C"
