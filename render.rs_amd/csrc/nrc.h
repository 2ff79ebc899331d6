// Neural radiance cache: device layout and kernel launchers.
//
// render.rs has no radiance cache; this is the learned part of the HIP
// backend. A multiresolution hash grid (up to 16 levels x 2 features,
// 2^log2_table entries per level, f32) encodes position; the network maps
// 48 inputs (32 encoding + direction, roughness, normal, 1, albedo, zero
// padding) through 5 hidden layers of 64 ReLU units to RGB radiance. The
// weights live as an f32 master copy and two bf16 copies laid out for the
// 32x32x16 bf16 matrix-core instruction (forward, and transposed for
// back-propagation); every product runs on the matrix cores with f32
// accumulation.
//
// Training is deterministic: weight gradients are written per block and
// summed in block order, grid gradients are added as 2^-36 fixed-point
// 64-bit integers (integer addition is associative).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pt_types.h"

namespace rrhip {

constexpr int kNrcIn = 48;
constexpr int kNrcWidth = 64;
constexpr int kNrcMaxLevels = 16;
// Natural f32 parameter layout: W0 [64][48], W1..W4 [64][64], W5 [3][64].
constexpr uint32_t kNrcOffW[6] = {0, 3072, 7168, 11264, 15360, 19456};
constexpr uint32_t kNrcMlpParams = 19456 + 3 * 64;
// bf16 copies: forward W0p [64][48], W1p..W4p [64][64], W5p [32][64]
// (rows 3..31 zero); backward W0Tp [32][64] (input dims 0..31), W1Tp..W4Tp
// [64][64], W5Tp [64][16] (K rows 3..15 zero).
constexpr uint32_t kNrcFwdOff[6] = {0, 3072, 7168, 11264, 15360, 19456};
constexpr uint32_t kNrcFwdSize = 19456 + 32 * 64;
constexpr uint32_t kNrcBwdOff[6] = {0, 2048, 6144, 10240, 14336, 18432};
constexpr uint32_t kNrcBwdSize = 18432 + 64 * 16;
// Training: samples per block, and the per-block partial record (weight
// gradients then the block's loss).
constexpr uint32_t kNrcTrainBlock = 64;
constexpr uint32_t kNrcPartStride = kNrcMlpParams + 64;
constexpr uint32_t kNrcMaxBatch = 1u << 16;
constexpr double kNrcFixedScale = 68719476736.0;  // 2^36

struct NrcGridView {
  const float* table;      // [levels][T][2]
  const uint16_t* wf;      // forward bf16 weights
  const uint16_t* wb;      // backward bf16 weights
  uint32_t levels, log2_table;
  uint32_t res[kNrcMaxLevels];  // cells per axis
  uint32_t dense_mask;          // bit l: level l indexes densely
  float lo[3], inv_ext[3];
};

struct NrcHyper {
  float lr, b1, b2, eps;
};

// Device-side control block.
struct NrcControl {
  uint32_t n;        // samples in the pending batch
  uint32_t step;     // optimiser steps taken
  float last_loss;   // loss of the last batch
  uint32_t pad;
};

// y (n x 3 floats) for n queries (count read from `count` when non-null).
hipError_t nrc_launch_infer(const NrcGridView& G, const NrcQuery* q, const uint32_t* count, uint32_t n,
                            float* out, hipStream_t st);
// For the path integrator: L[slot[i]] += beta[i] * max(y_i, 0).
hipError_t nrc_launch_infer_paths(const NrcGridView& G, const NrcQuery* q, const uint32_t* count,
                                  uint32_t capacity, const uint32_t* slots, const float4* beta, float4* L,
                                  hipStream_t st);
// Forward + backward over the batch (ctl->n samples, <= capacity): block
// partials into `part`, each sample's 32 encoding gradients into `genc`
// (capacity x 32 floats), then the grid gradients into `grid_grad`.
hipError_t nrc_launch_train(const NrcGridView& G, const NrcQuery* q, const float* targets,
                            const NrcControl* ctl, uint32_t capacity, float* part, float* genc,
                            unsigned long long* grid_grad, hipStream_t st);
// Sums the partials; with `apply`, Adam on the MLP master weights (step
// ctl->step + 1), else the summed gradient into `grad_out`; loss into ctl.
hipError_t nrc_launch_reduce_mlp(const float* part, NrcControl* ctl, uint32_t capacity, float* w, float* m,
                                 float* v, float* grad_out, NrcHyper hp, bool apply, hipStream_t st);
// Adam on the grid (or, without `apply`, grid gradient as f32 into
// grad_out); clears the fixed-point gradients. With `ema` set, also the
// moving average ema = alpha ema + (1 - alpha) table and its debiased
// value (step ctl->step + 1) into `inf`.
hipError_t nrc_launch_adam_grid(float* table, float* m, float* v, unsigned long long* grid_grad,
                                uint64_t n, const NrcControl* ctl, float* grad_out, NrcHyper hp, bool apply,
                                hipStream_t st, float* ema = nullptr, float* inf = nullptr, float alpha = 0.f);
// The same moving average over the MLP master weights (when ctl->n > 0).
hipError_t nrc_launch_ema_mlp(const float* w, float* ema, float* inf, const NrcControl* ctl, float alpha,
                              hipStream_t st);
// bf16 copies from the master weights (wb null: the forward copy only);
// counts the step when count_step and ctl->n > 0.
hipError_t nrc_launch_pack(const float* w, uint16_t* wf, uint16_t* wb, NrcControl* ctl, bool count_step,
                           hipStream_t st);

}  // namespace rrhip
