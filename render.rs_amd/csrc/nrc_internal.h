// Definition of the radiance-cache handle, shared by the C-ABI units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/rrhip.h"
#include "nrc.h"

struct rrhip_nrc {
  rrhip_ctx* ctx = nullptr;
  rrhip_nrc_config cfg{};
  rrhip::NrcGridView G{};
  rrhip::NrcHyper hp{};
  uint64_t grid_n = 0;  // levels * 2^log2_table * 2
  bool bounds_from_scene = false;
  float* d_table = nullptr;
  float* d_tm = nullptr;
  float* d_tv = nullptr;
  unsigned long long* d_tg = nullptr;
  float* d_w = nullptr;
  float* d_wm = nullptr;
  float* d_wv = nullptr;
  float* d_wgrad = nullptr;
  uint16_t* d_wf = nullptr;
  uint16_t* d_wb = nullptr;
  float* d_part = nullptr;
  float* d_genc = nullptr;  // per-sample encoding gradients of a step (kNrcMaxBatch x 32)
  rrhip::NrcControl* d_ctl = nullptr;
  // Inference parameters when cfg.ema > 0: moving averages of the grid and
  // MLP weights, their debiased values, and the bf16 forward copy.
  float* d_ema_table = nullptr;
  float* d_inf_table = nullptr;
  float* d_ema_w = nullptr;
  float* d_inf_w = nullptr;
  uint16_t* d_inf_wf = nullptr;
  // staging for the host-array entry points (grown on demand)
  rrhip::NrcQuery* d_q = nullptr;
  float* d_t = nullptr;
  float* d_out = nullptr;
  float* d_gout = nullptr;
  uint64_t cap = 0;
  std::vector<void*> mem;
  // Optimiser steps that a path-tracing session runs for its passes go to
  // train_st; `pending` means work recorded in ev_trained may still run.
  // Every use of the cache on another stream first waits for it.
  hipStream_t train_st = nullptr;
  hipEvent_t ev_samples = nullptr, ev_trained = nullptr;
  bool pending = false;
};

namespace rrhip {
// The parameters inference reads: the moving averages when enabled, else
// the trained ones.
inline NrcGridView nrc_infer_view(const rrhip_nrc* c) {
  NrcGridView v = c->G;
  if (c->d_inf_table) {
    v.table = c->d_inf_table;
    v.wf = c->d_inf_wf;
  }
  return v;
}
// Makes stream st wait for the cache's pending optimiser steps.
hipError_t nrc_join(rrhip_nrc* nrc, hipStream_t st);
// Blocks the host until the cache's optimiser steps are done.
hipError_t nrc_drain(rrhip_nrc* nrc);
// Sets the position bounds of a cache (normalisation of the grid input).
void nrc_set_bounds(rrhip_nrc* nrc, const float lo[3], const float hi[3]);
// One optimiser step on the device batch q / targets with ctl->n samples
// (already on the device); `capacity` bounds n.
hipError_t nrc_step_device(rrhip_nrc* nrc, const NrcQuery* q, const float* targets, uint32_t capacity,
                           hipStream_t st);
}  // namespace rrhip
