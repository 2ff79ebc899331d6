// Neural radiance cache: C ABI (creation, parameters, host-array inference
// and training steps). The kernels are in nrc.hip; sessions drive them
// from pt_session.cpp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <vector>

#include "api_internal.h"
#include "nrc_internal.h"
#include "sampling.h"
#include "scene_internal.h"

using namespace rrhip;

namespace {

template <class T>
hipError_t alloc_n(rrhip_nrc* c, T** p, uint64_t n) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, (n < 1 ? 1 : n) * sizeof(T));
  if (e == hipSuccess) c->mem.push_back(v);
  *p = (T*)v;
  return e;
}

void free_nrc(rrhip_nrc* c) {
  if (c->train_st) (void)hipStreamDestroy(c->train_st);
  if (c->ev_samples) (void)hipEventDestroy(c->ev_samples);
  if (c->ev_trained) (void)hipEventDestroy(c->ev_trained);
  for (void* p : c->mem) (void)hipFree(p);
  delete c;
}

rrhip_nrc_config default_config() {
  rrhip_nrc_config c{};
  c.struct_size = sizeof(rrhip_nrc_config);
  c.levels = 16;
  c.log2_table = 19;
  c.seed = 1;
  c.base_res = 16.f;
  c.max_res = 2048.f;
  c.learning_rate = 1e-2f;
  c.beta1 = 0.9f;
  c.beta2 = 0.99f;
  c.epsilon = 1e-15f;
  c.batch = 4096;
  c.steps_per_pass = 16;
  c.train_paths = 65536;
  c.spread = 1e-4f;
  return c;
}

int check_config(const rrhip_nrc_config& c) {
  if (c.struct_size != sizeof(rrhip_nrc_config))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: struct_size mismatch (ABI version)");
  if (c.levels < 1 || c.levels > 16) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: levels must be 1..16");
  if (c.log2_table < 8 || c.log2_table > 24)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: log2_table must be 8..24");
  if (!(c.base_res >= 1.f) || !(c.max_res >= c.base_res) || !(c.max_res <= 1e6f))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: need 1 <= base_res <= max_res <= 1e6");
  if (!(c.learning_rate > 0.f) || !(c.beta1 >= 0.f && c.beta1 < 1.f) || !(c.beta2 >= 0.f && c.beta2 < 1.f) ||
      !(c.epsilon >= 0.f))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: invalid Adam hyper-parameters");
  if (c.batch < 64 || c.batch > kNrcMaxBatch)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: batch must be 64..65536");
  if (c.steps_per_pass > 64) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: steps_per_pass must be <= 64");
  if (c.train_paths > (1u << 22))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: train_paths must be <= 2^22");
  if (!(c.spread > 0.f)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: spread must be > 0");
  if (!(c.ema >= 0.f && c.ema < 1.f)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: ema must be in [0, 1)");
  return RRHIP_OK;
}

// Cells per axis of each level: floor(base * b^l), b the geometric growth
// from base_res to max_res.
uint32_t level_res(const rrhip_nrc_config& c, uint32_t l) {
  if (c.levels <= 1) return (uint32_t)std::floor((double)c.base_res + 1e-6);
  const double b = std::exp((std::log((double)c.max_res) - std::log((double)c.base_res)) / (double)(c.levels - 1));
  return (uint32_t)std::floor((double)c.base_res * std::pow(b, (double)l) + 1e-6);
}

hipError_t ensure_staging(rrhip_nrc* c, uint64_t n) {
  if (n <= c->cap) return hipSuccess;
  for (void* p : {(void*)c->d_q, (void*)c->d_t, (void*)c->d_out}) {
    if (!p) continue;
    (void)hipFree(p);
    for (auto& m : c->mem)
      if (m == p) m = nullptr;
  }
  c->d_q = nullptr;
  c->d_t = nullptr;
  c->d_out = nullptr;
  c->cap = 0;
  hipError_t e = alloc_n(c, &c->d_q, n);
  if (e == hipSuccess) e = alloc_n(c, &c->d_t, 3 * n);
  if (e == hipSuccess) e = alloc_n(c, &c->d_out, 3 * n);
  if (e == hipSuccess) c->cap = n;
  return e;
}

// Null check, then the cache's own stream waits for optimiser steps a
// session may still be running on it.
int check_nrc(rrhip_nrc* c) {
  if (!c) return fail(RRHIP_ERR_INVALID_ARGUMENT, "null cache");
  if (c->pending) {
    RRHIP_HIP_TRY(hipSetDevice(c->ctx->device));
    RRHIP_HIP_TRY(nrc_join(c, c->ctx->stream));
  }
  return RRHIP_OK;
}

}  // namespace

namespace rrhip {

hipError_t nrc_join(rrhip_nrc* c, hipStream_t st) {
  if (!c->pending) return hipSuccess;
  const hipError_t e = hipStreamWaitEvent(st, c->ev_trained, 0);
  if (e == hipSuccess) c->pending = false;
  return e;
}

hipError_t nrc_drain(rrhip_nrc* c) {
  c->pending = false;
  return c->train_st ? hipStreamSynchronize(c->train_st) : hipSuccess;
}

void nrc_set_bounds(rrhip_nrc* c, const float lo[3], const float hi[3]) {
  for (int a = 0; a < 3; a++) {
    c->G.lo[a] = lo[a];
    const float ext = hi[a] - lo[a];
    c->G.inv_ext[a] = ext > 0.f ? 1.f / ext : 1.f;
  }
}

hipError_t nrc_step_device(rrhip_nrc* c, const NrcQuery* q, const float* targets, uint32_t capacity,
                           hipStream_t st) {
  hipError_t e = nrc_launch_train(c->G, q, targets, c->d_ctl, capacity, c->d_part, c->d_genc, c->d_tg, st);
  if (e == hipSuccess)
    e = nrc_launch_reduce_mlp(c->d_part, c->d_ctl, capacity, c->d_w, c->d_wm, c->d_wv, nullptr, c->hp, true, st);
  if (e == hipSuccess)
    e = nrc_launch_adam_grid(c->d_table, c->d_tm, c->d_tv, c->d_tg, c->grid_n, c->d_ctl, nullptr, c->hp, true, st,
                             c->d_ema_table, c->d_inf_table, c->cfg.ema);
  if (e == hipSuccess && c->d_inf_table) {
    e = nrc_launch_ema_mlp(c->d_w, c->d_ema_w, c->d_inf_w, c->d_ctl, c->cfg.ema, st);
    if (e == hipSuccess) e = nrc_launch_pack(c->d_inf_w, c->d_inf_wf, nullptr, c->d_ctl, false, st);
  }
  if (e == hipSuccess) e = nrc_launch_pack(c->d_w, c->d_wf, c->d_wb, c->d_ctl, true, st);
  return e;
}

}  // namespace rrhip

extern "C" {

int rrhip_nrc_create(rrhip_ctx* ctx, const rrhip_nrc_config* config, rrhip_nrc** out) {
  if (!ctx || !out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_create: null argument");
  *out = nullptr;
  const rrhip_nrc_config cfg = config ? *config : default_config();
  if (int s = check_config(cfg)) return s;
  RRHIP_HIP_TRY(hipSetDevice(ctx->device));
  auto* c = new (std::nothrow) rrhip_nrc();
  if (!c) return fail(RRHIP_ERR_OUT_OF_MEMORY, "rrhip_nrc_create: host allocation failed");
  c->ctx = ctx;
  c->cfg = cfg;
  c->hp = NrcHyper{cfg.learning_rate, cfg.beta1, cfg.beta2, cfg.epsilon};
  const uint64_t T = 1ull << cfg.log2_table;
  c->grid_n = (uint64_t)cfg.levels * T * 2;
  NrcGridView& G = c->G;
  G.levels = cfg.levels;
  G.log2_table = cfg.log2_table;
  G.dense_mask = 0;
  for (uint32_t l = 0; l < cfg.levels; l++) {
    G.res[l] = level_res(cfg, l);
    const uint64_t n1 = (uint64_t)G.res[l] + 1;
    if (n1 * n1 * n1 <= T) G.dense_mask |= 1u << l;
  }
  bool scene_bounds = false;
  for (int a = 0; a < 3; a++) scene_bounds |= !(cfg.hi[a] > cfg.lo[a]);
  c->bounds_from_scene = scene_bounds;
  const float unit_lo[3] = {0.f, 0.f, 0.f}, unit_hi[3] = {1.f, 1.f, 1.f};
  nrc_set_bounds(c, scene_bounds ? unit_lo : cfg.lo, scene_bounds ? unit_hi : cfg.hi);

  hipError_t e = hipStreamCreateWithFlags(&c->train_st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_samples, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_trained, hipEventDisableTiming);
  if (e == hipSuccess) e = alloc_n(c, &c->d_table, c->grid_n);
  if (e == hipSuccess) e = alloc_n(c, &c->d_tm, c->grid_n);
  if (e == hipSuccess) e = alloc_n(c, &c->d_tv, c->grid_n);
  if (e == hipSuccess) e = alloc_n(c, &c->d_tg, c->grid_n);
  if (e == hipSuccess) e = alloc_n(c, &c->d_w, (uint64_t)kNrcMlpParams);
  if (e == hipSuccess) e = alloc_n(c, &c->d_wm, (uint64_t)kNrcMlpParams);
  if (e == hipSuccess) e = alloc_n(c, &c->d_wv, (uint64_t)kNrcMlpParams);
  if (e == hipSuccess) e = alloc_n(c, &c->d_wgrad, (uint64_t)kNrcMlpParams);
  if (e == hipSuccess) e = alloc_n(c, &c->d_wf, (uint64_t)kNrcFwdSize);
  if (e == hipSuccess) e = alloc_n(c, &c->d_wb, (uint64_t)kNrcBwdSize);
  if (e == hipSuccess) e = alloc_n(c, &c->d_part, (uint64_t)(kNrcMaxBatch / kNrcTrainBlock) * kNrcPartStride);
  if (e == hipSuccess) e = alloc_n(c, &c->d_genc, (uint64_t)kNrcMaxBatch * 32);
  if (e == hipSuccess) e = alloc_n(c, &c->d_ctl, 1);
  if (cfg.ema > 0.f) {
    if (e == hipSuccess) e = alloc_n(c, &c->d_ema_table, c->grid_n);
    if (e == hipSuccess) e = alloc_n(c, &c->d_inf_table, c->grid_n);
    if (e == hipSuccess) e = alloc_n(c, &c->d_ema_w, (uint64_t)kNrcMlpParams);
    if (e == hipSuccess) e = alloc_n(c, &c->d_inf_w, (uint64_t)kNrcMlpParams);
    if (e == hipSuccess) e = alloc_n(c, &c->d_inf_wf, (uint64_t)kNrcFwdSize);
  }
  if (e != hipSuccess) {
    free_nrc(c);
    return fail_hip(e, "rrhip_nrc_create");
  }
  G.table = c->d_table;
  G.wf = c->d_wf;
  G.wb = c->d_wb;
  // initial parameters
  Pcg32 rng = Pcg32::make(cfg.seed, 0x4e5243u);
  std::vector<float> table(c->grid_n);
  for (auto& v : table) v = (rng.next_f32() * 2.f - 1.f) * 1e-4f;
  std::vector<float> w(kNrcMlpParams, 0.f);
  const uint32_t rows[6] = {64, 64, 64, 64, 64, 3}, cols[6] = {48, 64, 64, 64, 64, 64};
  const uint32_t fan[6] = {43, 64, 64, 64, 64, 64};
  for (int l = 0; l < 6; l++) {
    const float s = std::sqrt(6.f / (float)fan[l]);
    for (uint32_t r = 0; r < rows[l]; r++)
      for (uint32_t k = 0; k < cols[l]; k++) {
        const float u = rng.next_f32() * 2.f - 1.f;
        w[kNrcOffW[l] + r * cols[l] + k] = k < fan[l] ? u * s : 0.f;
      }
  }
  *out = c;
  const int st = rrhip_nrc_set_params(c, w.data(), table.data());
  if (st != RRHIP_OK) {
    *out = nullptr;
    free_nrc(c);
    return st;
  }
  return RRHIP_OK;
}

int rrhip_nrc_destroy(rrhip_nrc* c) {
  if (!c) return RRHIP_OK;
  (void)hipSetDevice(c->ctx->device);
  (void)hipStreamSynchronize(c->ctx->stream);
  (void)nrc_drain(c);
  free_nrc(c);
  return RRHIP_OK;
}

int rrhip_nrc_get_info(rrhip_nrc* c, rrhip_nrc_info* info) {
  if (int s = check_nrc(c)) return s;
  if (!info) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_get_info: null output");
  NrcControl ctl{};
  RRHIP_HIP_TRY(hipSetDevice(c->ctx->device));
  RRHIP_HIP_TRY(hipMemcpyAsync(&ctl, c->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, c->ctx->stream));
  RRHIP_HIP_TRY(hipStreamSynchronize(c->ctx->stream));
  *info = rrhip_nrc_info{};
  info->levels = c->G.levels;
  info->log2_table = c->G.log2_table;
  for (uint32_t l = 0; l < c->G.levels; l++) info->res[l] = c->G.res[l];
  info->mlp_params = kNrcMlpParams;
  info->grid_params = c->grid_n;
  info->steps = ctl.step;
  info->last_loss = ctl.last_loss;
  for (int a = 0; a < 3; a++) {
    info->lo[a] = c->G.lo[a];
    info->hi[a] = c->G.lo[a] + 1.f / c->G.inv_ext[a];
  }
  return RRHIP_OK;
}

int rrhip_nrc_get_params(rrhip_nrc* c, float* mlp, float* grid) {
  if (int s = check_nrc(c)) return s;
  RRHIP_HIP_TRY(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  if (mlp) RRHIP_HIP_TRY(hipMemcpyAsync(mlp, c->d_w, kNrcMlpParams * sizeof(float), hipMemcpyDeviceToHost, st));
  if (grid) RRHIP_HIP_TRY(hipMemcpyAsync(grid, c->d_table, c->grid_n * sizeof(float), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  return RRHIP_OK;
}

int rrhip_nrc_set_params(rrhip_nrc* c, const float* mlp, const float* grid) {
  if (int s = check_nrc(c)) return s;
  RRHIP_HIP_TRY(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  if (mlp) RRHIP_HIP_TRY(hipMemcpyAsync(c->d_w, mlp, kNrcMlpParams * sizeof(float), hipMemcpyHostToDevice, st));
  if (grid) RRHIP_HIP_TRY(hipMemcpyAsync(c->d_table, grid, c->grid_n * sizeof(float), hipMemcpyHostToDevice, st));
  RRHIP_HIP_TRY(hipMemsetAsync(c->d_wm, 0, kNrcMlpParams * sizeof(float), st));
  RRHIP_HIP_TRY(hipMemsetAsync(c->d_wv, 0, kNrcMlpParams * sizeof(float), st));
  RRHIP_HIP_TRY(hipMemsetAsync(c->d_tm, 0, c->grid_n * sizeof(float), st));
  RRHIP_HIP_TRY(hipMemsetAsync(c->d_tv, 0, c->grid_n * sizeof(float), st));
  RRHIP_HIP_TRY(hipMemsetAsync(c->d_tg, 0, c->grid_n * sizeof(unsigned long long), st));
  RRHIP_HIP_TRY(hipMemsetAsync(c->d_ctl, 0, sizeof(NrcControl), st));
  RRHIP_HIP_TRY(nrc_launch_pack(c->d_w, c->d_wf, c->d_wb, c->d_ctl, false, st));
  if (c->d_inf_table) {  // averages restart from the parameters set
    RRHIP_HIP_TRY(hipMemsetAsync(c->d_ema_table, 0, c->grid_n * sizeof(float), st));
    RRHIP_HIP_TRY(hipMemsetAsync(c->d_ema_w, 0, kNrcMlpParams * sizeof(float), st));
    RRHIP_HIP_TRY(hipMemcpyAsync(c->d_inf_table, c->d_table, c->grid_n * sizeof(float), hipMemcpyDeviceToDevice, st));
    RRHIP_HIP_TRY(hipMemcpyAsync(c->d_inf_w, c->d_w, kNrcMlpParams * sizeof(float), hipMemcpyDeviceToDevice, st));
    RRHIP_HIP_TRY(nrc_launch_pack(c->d_inf_w, c->d_inf_wf, nullptr, c->d_ctl, false, st));
  }
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  return RRHIP_OK;
}

int rrhip_nrc_infer(rrhip_nrc* c, const rrhip_nrc_query* queries, uint64_t n, float* rgb) {
  if (int s = check_nrc(c)) return s;
  if (n == 0) return RRHIP_OK;
  if (!queries || !rgb) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_infer: null array");
  if (n >= (1ull << 31)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_infer: n must be < 2^31");
  RRHIP_HIP_TRY(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  RRHIP_HIP_TRY(ensure_staging(c, n));
  RRHIP_HIP_TRY(hipMemcpyAsync(c->d_q, queries, n * sizeof(NrcQuery), hipMemcpyHostToDevice, st));
  RRHIP_HIP_TRY(nrc_launch_infer(nrc_infer_view(c), c->d_q, nullptr, (uint32_t)n, c->d_out, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(rgb, c->d_out, 3 * n * sizeof(float), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  return RRHIP_OK;
}

namespace {
int stage_batch(rrhip_nrc* c, const rrhip_nrc_query* q, const float* t, uint64_t n, const char* who) {
  if (!q || !t) return fail(RRHIP_ERR_INVALID_ARGUMENT, std::string(who) + ": null array");
  if (n == 0 || n > kNrcMaxBatch) return fail(RRHIP_ERR_INVALID_ARGUMENT, std::string(who) + ": n must be 1..65536");
  RRHIP_HIP_TRY(hipSetDevice(c->ctx->device));
  hipStream_t st = c->ctx->stream;
  RRHIP_HIP_TRY(ensure_staging(c, n));
  RRHIP_HIP_TRY(hipMemcpyAsync(c->d_q, q, n * sizeof(NrcQuery), hipMemcpyHostToDevice, st));
  RRHIP_HIP_TRY(hipMemcpyAsync(c->d_t, t, 3 * n * sizeof(float), hipMemcpyHostToDevice, st));
  const uint32_t n32 = (uint32_t)n;
  RRHIP_HIP_TRY(hipMemcpyAsync(&c->d_ctl->n, &n32, sizeof(n32), hipMemcpyHostToDevice, st));
  return RRHIP_OK;
}
}  // namespace

int rrhip_nrc_train_step(rrhip_nrc* c, const rrhip_nrc_query* queries, const float* targets, uint64_t n,
                         float* loss) {
  if (int s = check_nrc(c)) return s;
  if (int s = stage_batch(c, queries, targets, n, "rrhip_nrc_train_step")) return s;
  hipStream_t st = c->ctx->stream;
  RRHIP_HIP_TRY(nrc_step_device(c, c->d_q, c->d_t, (uint32_t)n, st));
  NrcControl ctl{};
  RRHIP_HIP_TRY(hipMemcpyAsync(&ctl, c->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  if (loss) *loss = ctl.last_loss;
  return RRHIP_OK;
}

int rrhip_nrc_time_steps(rrhip_nrc* c, const rrhip_nrc_query* queries, const float* targets, uint64_t n,
                         uint32_t steps, float* elapsed_ms) {
  if (int s = check_nrc(c)) return s;
  if (!elapsed_ms || steps == 0) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_nrc_time_steps: null output or 0 steps");
  if (int s = stage_batch(c, queries, targets, n, "rrhip_nrc_time_steps")) return s;
  hipStream_t st = c->ctx->stream;
  RRHIP_HIP_TRY(hipEventRecord(c->ctx->ev0, st));
  for (uint32_t k = 0; k < steps; k++) RRHIP_HIP_TRY(nrc_step_device(c, c->d_q, c->d_t, (uint32_t)n, st));
  RRHIP_HIP_TRY(hipEventRecord(c->ctx->ev1, st));
  RRHIP_HIP_TRY(hipEventSynchronize(c->ctx->ev1));
  RRHIP_HIP_TRY(hipEventElapsedTime(elapsed_ms, c->ctx->ev0, c->ctx->ev1));
  return RRHIP_OK;
}

int rrhip_nrc_gradients(rrhip_nrc* c, const rrhip_nrc_query* queries, const float* targets, uint64_t n, float* loss,
                        float* mlp_grad, float* grid_grad) {
  if (int s = check_nrc(c)) return s;
  if (int s = stage_batch(c, queries, targets, n, "rrhip_nrc_gradients")) return s;
  hipStream_t st = c->ctx->stream;
  if (!c->d_gout) RRHIP_HIP_TRY(alloc_n(c, &c->d_gout, c->grid_n));
  RRHIP_HIP_TRY(nrc_launch_train(c->G, c->d_q, c->d_t, c->d_ctl, (uint32_t)n, c->d_part, c->d_genc, c->d_tg, st));
  RRHIP_HIP_TRY(nrc_launch_reduce_mlp(c->d_part, c->d_ctl, (uint32_t)n, c->d_w, c->d_wm, c->d_wv, c->d_wgrad, c->hp,
                                      false, st));
  RRHIP_HIP_TRY(nrc_launch_adam_grid(c->d_table, c->d_tm, c->d_tv, c->d_tg, c->grid_n, c->d_ctl, c->d_gout, c->hp,
                                     false, st));
  if (mlp_grad)
    RRHIP_HIP_TRY(hipMemcpyAsync(mlp_grad, c->d_wgrad, kNrcMlpParams * sizeof(float), hipMemcpyDeviceToHost, st));
  if (grid_grad)
    RRHIP_HIP_TRY(hipMemcpyAsync(grid_grad, c->d_gout, c->grid_n * sizeof(float), hipMemcpyDeviceToHost, st));
  NrcControl ctl{};
  RRHIP_HIP_TRY(hipMemcpyAsync(&ctl, c->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  if (loss) *loss = ctl.last_loss;
  return RRHIP_OK;
}

}  // extern "C"
