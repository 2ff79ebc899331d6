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
  rrhip::NrcControl* d_ctl = nullptr;
  // staging for the host-array entry points (grown on demand)
  rrhip::NrcQuery* d_q = nullptr;
  float* d_t = nullptr;
  float* d_out = nullptr;
  float* d_gout = nullptr;
  uint64_t cap = 0;
  std::vector<void*> mem;
};

namespace rrhip {
// Sets the position bounds of a cache (normalisation of the grid input).
void nrc_set_bounds(rrhip_nrc* nrc, const float lo[3], const float hi[3]);
// One optimiser step on the device batch q / targets with ctl->n samples
// (already on the device); `capacity` bounds n.
hipError_t nrc_step_device(rrhip_nrc* nrc, const NrcQuery* q, const float* targets, uint32_t capacity,
                           hipStream_t st);
}  // namespace rrhip
