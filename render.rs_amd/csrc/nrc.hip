// Neural radiance cache kernels for gfx950: fused hash-grid encoding + MLP
// inference, fused forward + backward training, deterministic gradient
// reduction, Adam.
//
// Matrix-core mapping (v_mfma_f32_32x32x16_bf16, wave64): a wave evaluates
// 32 queries at once with the batch in the N dimension, so a layer is
// out[64 x 32] = W[64 x K] * in[K x 32] as two 32-row M tiles and K/16
// K-steps. With the batch in N, the accumulator layout of one layer (lane
// l: column l % 32, rows 8q + 4(l / 32) + j) is, register for register,
// the B-operand layout of the next layer once K is permuted within each
// 16-wide step (kperm below), so activations never leave registers
// between layers; the weights are stored pre-permuted to match.
// Weight gradients contract over the batch and go through LDS
// ([neuron][sample], rows padded to 144 bytes so 16-byte reads of 16
// consecutive lanes hit distinct banks).
#include <hip/hip_runtime.h>

#include "nrc.h"

namespace rrhip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kInferBlock = 256;
constexpr uint32_t kLdsStride = 72;   // bf16 per padded 64-wide row
constexpr uint32_t kLdsStride0 = 56;  // bf16 per padded 48-wide row (W0)
constexpr uint32_t kLdsW[6] = {0, 3584, 3584 + 4608, 3584 + 2 * 4608, 3584 + 3 * 4608, 3584 + 4 * 4608};
constexpr uint32_t kLdsWSize = 3584 + 4 * 4608 + 32 * 72;

// Natural K index of permuted K slot `ks` (16-wide steps; slot 8h + j of a
// step holds what lane half h has in accumulator row 8(j/4) + 4h + j%4).
__host__ __device__ constexpr uint32_t kperm(uint32_t ks) {
  const uint32_t c = ks >> 4, s = ks & 15u, h = s >> 3, j = s & 7u;
  return 16u * c + 8u * (j >> 2) + 4u * h + (j & 3u);
}

__device__ __forceinline__ uint32_t bf16_bits(float f) {  // round to nearest even
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_round(float f) { return __uint_as_float(bf16_bits(f) << 16); }
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 pack8(const float* v) {
  u32x4 u;
  u.x = bf16_bits(v[0]) | (bf16_bits(v[1]) << 16);
  u.y = bf16_bits(v[2]) | (bf16_bits(v[3]) << 16);
  u.z = bf16_bits(v[4]) | (bf16_bits(v[5]) << 16);
  u.w = bf16_bits(v[6]) | (bf16_bits(v[7]) << 16);
  return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ bf16x8 load8(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *(const u32x4*)p); }
__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int i = 0; i < 16; i++) z[i] = 0.f;
  return z;
}

// ---- encoding -----------------------------------------------------------------

__device__ __forceinline__ void normalise(const NrcGridView& G, const NrcQuery& q, float pn[3]) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float v = (q.pos[a] - G.lo[a]) * G.inv_ext[a];
    pn[a] = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
  }
}

struct Cell {
  uint32_t x, y, z;
  float fx, fy, fz;
};

__device__ __forceinline__ Cell cell_of(const NrcGridView& G, const float pn[3], uint32_t l) {
  const float nf = (float)G.res[l];
  float s[3], fl[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    s[a] = pn[a] * nf;
    fl[a] = fminf(floorf(s[a]), nf - 1.f);
  }
  return Cell{(uint32_t)fl[0], (uint32_t)fl[1], (uint32_t)fl[2], s[0] - fl[0], s[1] - fl[1], s[2] - fl[2]};
}

// Entry of corner k (bit 0: +x, bit 1: +y, bit 2: +z) and its weight.
__device__ __forceinline__ uint32_t corner(const NrcGridView& G, uint32_t l, const Cell& c, int k, float& w) {
  const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
  w = (dx ? c.fx : 1.f - c.fx) * (dy ? c.fy : 1.f - c.fy) * (dz ? c.fz : 1.f - c.fz);
  const uint32_t x = c.x + dx, y = c.y + dy, z = c.z + dz;
  if ((G.dense_mask >> l) & 1u) {
    const uint32_t n1 = G.res[l] + 1;
    return x + n1 * (y + n1 * z);
  }
  return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & ((1u << G.log2_table) - 1u);
}

__device__ __forceinline__ float2 encode_level(const NrcGridView& G, const float pn[3], uint32_t l) {
  if (l >= G.levels) return make_float2(0.f, 0.f);
  const Cell c = cell_of(G, pn, l);
  const float2* tab = (const float2*)G.table + ((size_t)l << G.log2_table);
  float f0 = 0.f, f1 = 0.f;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    float w;
    const uint32_t id = corner(G, l, c, k, w);
    const float2 e = tab[id];
    f0 += w * e.x;
    f1 += w * e.y;
  }
  return make_float2(f0, f1);
}

// The 24 input values lane half h contributes, in K-slot order: chunks 0
// and 1 hold the encoding of levels 8c + 2h + {0, 1, 4, 5}, chunk 2 the
// direction, roughness, albedo (h = 0) or normal and the constant 1 (h = 1).
__device__ __forceinline__ void make_inputs(const NrcGridView& G, const NrcQuery& q, uint32_t h, float v[3][8]) {
  float pn[3];
  normalise(G, q, pn);
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const uint32_t la = 8u * c + 2u * h;
    const uint32_t lv[4] = {la, la + 1, la + 4, la + 5};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float2 f = encode_level(G, pn, lv[i]);
      v[c][2 * i] = f.x;
      v[c][2 * i + 1] = f.y;
      // Keep one level's gathers in flight per lane: without the fence the
      // scheduler hoists all 64 and the register count caps occupancy.
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  if (h == 0) {
    v[2][0] = q.dir[0];
    v[2][1] = q.dir[1];
    v[2][2] = q.dir[2];
    v[2][3] = q.roughness;
    v[2][4] = q.albedo[0];
    v[2][5] = q.albedo[1];
    v[2][6] = q.albedo[2];
    v[2][7] = 0.f;
  } else {
    v[2][0] = q.normal[0];
    v[2][1] = q.normal[1];
    v[2][2] = q.normal[2];
    v[2][3] = 1.f;
    v[2][4] = 0.f;
    v[2][5] = 0.f;
    v[2][6] = 0.f;
    v[2][7] = 0.f;
  }
}

// ReLU of a 64-row accumulator pair as the four B-operand K-steps of the
// next layer.
__device__ __forceinline__ void relu_pack(const f32x16 acc[2], bf16x8 out[4]) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = fmaxf(acc[c >> 1][8 * (c & 1) + j], 0.f);
    out[c] = pack8(v);
  }
}

// ---- inference -------------------------------------------------------------------

__device__ __forceinline__ void stage_weights(const uint16_t* __restrict__ wf, uint16_t* s_w) {
  // W0: 64 rows x 6 chunks; W1..W4: 64 x 8; W5: 32 x 8 (16-byte chunks)
  constexpr uint32_t n0 = 64 * 6, n1 = 4 * 64 * 8, n5 = 32 * 8;
  for (uint32_t i = threadIdx.x; i < n0 + n1 + n5; i += blockDim.x) {
    const uint16_t* src;
    uint16_t* dst;
    if (i < n0) {
      const uint32_t m = i / 6, c = i % 6;
      src = wf + kNrcFwdOff[0] + m * 48 + 8 * c;
      dst = s_w + kLdsW[0] + m * kLdsStride0 + 8 * c;
    } else if (i < n0 + n1) {
      const uint32_t r = i - n0, l = 1 + r / 512, m = (r % 512) / 8, c = r % 8;
      src = wf + kNrcFwdOff[l] + m * 64 + 8 * c;
      dst = s_w + kLdsW[l] + m * kLdsStride + 8 * c;
    } else {
      const uint32_t r = i - n0 - n1, m = r / 8, c = r % 8;
      src = wf + kNrcFwdOff[5] + m * 64 + 8 * c;
      dst = s_w + kLdsW[5] + m * kLdsStride + 8 * c;
    }
    *(u32x4*)dst = *(const u32x4*)src;
  }
}

// Forward pass of one wave's 32 queries; lanes of half 0 return y.
__device__ __forceinline__ f32x16 forward_lds(const uint16_t* s_w, const bf16x8 in[3], uint32_t col, uint32_t h) {
  f32x16 acc[2] = {zero16(), zero16()};
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
      acc[mt] = mfma(load8(s_w + kLdsW[0] + (32 * mt + col) * kLdsStride0 + 16 * c + 8 * h), in[c], acc[mt]);
  bf16x8 a[4];
  relu_pack(acc, a);
#pragma unroll
  for (int l = 1; l < 5; l++) {
    acc[0] = zero16();
    acc[1] = zero16();
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int mt = 0; mt < 2; mt++)
        acc[mt] = mfma(load8(s_w + kLdsW[l] + (32 * mt + col) * kLdsStride + 16 * c + 8 * h), a[c], acc[mt]);
    relu_pack(acc, a);
  }
  f32x16 y = zero16();
#pragma unroll
  for (int c = 0; c < 4; c++) y = mfma(load8(s_w + kLdsW[5] + col * kLdsStride + 16 * c + 8 * h), a[c], y);
  return y;
}

template <bool PATHS>
__global__ __launch_bounds__(kInferBlock) void nrc_infer_kernel(NrcGridView G, const NrcQuery* __restrict__ q,
                                                               const uint32_t* __restrict__ count, uint32_t n_static,
                                                               float* __restrict__ out,
                                                               const uint32_t* __restrict__ slots,
                                                               const float4* __restrict__ beta, float4* L) {
  __shared__ __align__(16) uint16_t s_w[kLdsWSize];
  stage_weights(G.wf, s_w);
  __syncthreads();
  const uint32_t n = count ? *count : n_static;
  const uint32_t lane = threadIdx.x & 63u, col = lane & 31u, h = lane >> 5;
  const uint32_t nwaves = gridDim.x * (kInferBlock / 64);
  for (uint32_t tile = blockIdx.x * (kInferBlock / 64) + threadIdx.x / 64; tile * 32u < n; tile += nwaves) {
    const uint32_t qi = tile * 32u + col;
    const bool valid = qi < n;
    NrcQuery qq = {};
    if (valid) qq = q[qi];
    float v[3][8];
    make_inputs(G, qq, h, v);
    bf16x8 in[3];
#pragma unroll
    for (int c = 0; c < 3; c++) in[c] = pack8(v[c]);
    const f32x16 y = forward_lds(s_w, in, col, h);
    if (h == 0 && valid) {
      if (PATHS) {
        const uint32_t s = slots[qi];
        const float4 b = beta[qi];
        float4 l = L[s];
        l.x += b.x * fmaxf(y[0], 0.f);
        l.y += b.y * fmaxf(y[1], 0.f);
        l.z += b.z * fmaxf(y[2], 0.f);
        L[s] = l;
      } else {
        out[3 * (size_t)qi + 0] = y[0];
        out[3 * (size_t)qi + 1] = y[1];
        out[3 * (size_t)qi + 2] = y[2];
      }
    }
  }
}

// ---- training --------------------------------------------------------------------

constexpr uint32_t kActRows = 64 * kLdsStride;  // one layer's [64][72] bf16

// relu(acc) in bf16 into rows [neuron][sample] of one layer's LDS block.
__device__ __forceinline__ void store_act(uint16_t* layer, const f32x16 acc[2], uint32_t sl, uint32_t h) {
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const uint32_t row = 32 * mt + 8 * (r >> 2) + 4 * h + (r & 3);
      layer[row * kLdsStride + sl] = (uint16_t)bf16_bits(fmaxf(acc[mt][r], 0.f));
    }
}

// delta = grad * [act > 0] in accumulator layout, rounded to bf16: stored
// to LDS and returned as the four B-operand K-steps.
__device__ __forceinline__ void mask_delta(f32x16 acc[2], const uint16_t* act, uint16_t* dl, uint32_t sl, uint32_t h,
                                           bf16x8 d[4]) {
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const uint32_t row = 32 * mt + 8 * (r >> 2) + 4 * h + (r & 3);
      const float a = bf16_to_f32(act[row * kLdsStride + sl]);
      const float g = a > 0.f ? bf16_round(acc[mt][r]) : 0.f;
      acc[mt][r] = g;
      dl[row * kLdsStride + sl] = (uint16_t)bf16_bits(g);
    }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = acc[c >> 1][8 * (c & 1) + j];
    d[c] = pack8(v);
  }
}

// dW[out][in] += sum over the block's 64 samples of delta[out][s] act[in][s]
// for this wave's 32-row out tile; `cols` = in width (48 or 64).
__device__ __forceinline__ void weight_grad(const uint16_t* dl, const uint16_t* act, uint32_t mt, uint32_t cols,
                                            uint32_t col, uint32_t h, float* part) {
#pragma unroll
  for (uint32_t nt = 0; nt < 2; nt++) {
    f32x16 acc = zero16();
#pragma unroll
    for (int c = 0; c < 4; c++)
      acc = mfma(load8(dl + (32 * mt + col) * kLdsStride + 16 * c + 8 * h),
                 load8(act + (32 * nt + col) * kLdsStride + 16 * c + 8 * h), acc);
    const uint32_t in = 32 * nt + col;
    if (in < cols) {
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const uint32_t out = 32 * mt + 8 * (r >> 2) + 4 * h + (r & 3);
        part[out * cols + in] = acc[r];
      }
    }
  }
}

__device__ __forceinline__ unsigned long long to_fixed(float v) {
  v = fminf(fmaxf(v, -1000.f), 1000.f);
  return (unsigned long long)(long long)rintf(v * 68719476736.f);
}

__global__ __launch_bounds__(128) void nrc_train_kernel(NrcGridView G, const NrcQuery* __restrict__ q,
                                                       const float* __restrict__ targets,
                                                       const NrcControl* __restrict__ ctl, float* __restrict__ part,
                                                       float* __restrict__ genc) {
  __shared__ __align__(16) uint16_t s_act[6 * kActRows];
  __shared__ __align__(16) uint16_t s_dlt[2 * kActRows];
  __shared__ float s_dy[3 * 64];
  __shared__ float s_loss[64];
  const uint32_t n = ctl->n;
  const uint32_t base = blockIdx.x * kNrcTrainBlock;
  if (base >= n) return;  // uniform over the block
  const uint32_t tid = threadIdx.x, lane = tid & 63u, col = lane & 31u, h = lane >> 5, w = tid >> 6;
  const uint32_t sl = 32 * w + col;
  const uint32_t si = base + sl;
  const bool valid = si < n;
  float* pb = part + (size_t)blockIdx.x * kNrcPartStride;
  for (uint32_t i = tid; i < 16 * kLdsStride; i += blockDim.x) s_act[48 * kLdsStride + i] = 0;

  NrcQuery qq = {};
  if (valid) qq = q[si];
  float v[3][8];
  make_inputs(G, qq, h, v);
  bf16x8 a[4];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    a[c] = pack8(v[c]);
#pragma unroll
    for (int j = 0; j < 8; j++) s_act[kperm(16 * c + 8 * h + j) * kLdsStride + sl] = (uint16_t)bf16_bits(v[c][j]);
  }
  // forward, keeping every layer's activations in LDS
  f32x16 acc[2] = {zero16(), zero16()};
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
      acc[mt] = mfma(load8(G.wf + kNrcFwdOff[0] + (32 * mt + col) * 48 + 16 * c + 8 * h), a[c], acc[mt]);
#pragma unroll
  for (int l = 1; l < 5; l++) {
    relu_pack(acc, a);
    store_act(s_act + l * kActRows, acc, sl, h);
    acc[0] = zero16();
    acc[1] = zero16();
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int mt = 0; mt < 2; mt++)
        acc[mt] = mfma(load8(G.wf + kNrcFwdOff[l] + (32 * mt + col) * 64 + 16 * c + 8 * h), a[c], acc[mt]);
  }
  relu_pack(acc, a);
  store_act(s_act + 5 * kActRows, acc, sl, h);
  f32x16 y = zero16();
#pragma unroll
  for (int c = 0; c < 4; c++) y = mfma(load8(G.wf + kNrcFwdOff[5] + col * 64 + 16 * c + 8 * h), a[c], y);

  // relative L2 loss, stop-gradient denominator
  const float inv = 1.f / (3.f * (float)n);
  float dy[3] = {0.f, 0.f, 0.f};
  float lsum = 0.f;
  if (h == 0 && valid) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float yk = y[k], t = targets[3 * (size_t)si + k];
      const float den = yk * yk + 0.01f;
      lsum += (yk - t) * (yk - t) / den * inv;
      dy[k] = bf16_round(2.f * (yk - t) / den * inv);
    }
  }
  if (h == 0) {
    s_dy[0 * 64 + sl] = dy[0];
    s_dy[1 * 64 + sl] = dy[1];
    s_dy[2 * 64 + sl] = dy[2];
    s_loss[sl] = lsum;
  }
  // output rows 0..2 sit in K slots 0..2 of lane half 0
  const float dyv[8] = {dy[0], dy[1], dy[2], 0.f, 0.f, 0.f, 0.f, 0.f};
  const bf16x8 dyb = pack8(dyv);
#pragma unroll
  for (int mt = 0; mt < 2; mt++)
    acc[mt] = mfma(load8(G.wb + kNrcBwdOff[5] + (32 * mt + col) * 16 + 8 * h), dyb, zero16());
  bf16x8 d[4];
  mask_delta(acc, s_act + 5 * kActRows, s_dlt + 1 * kActRows, sl, h, d);
  __syncthreads();
  // dW5 = dy (x) A5 over the block's samples, in sample order
  for (uint32_t e = tid; e < 192; e += blockDim.x) {
    const uint32_t r = e / 64, k = e % 64;
    float s = 0.f;
    for (uint32_t j = 0; j < 64; j++) s += s_dy[r * 64 + j] * bf16_to_f32(s_act[5 * kActRows + k * kLdsStride + j]);
    pb[kNrcOffW[5] + e] = s;
  }
  if (tid == 0) {
    float s = 0.f;
    for (uint32_t j = 0; j < 64; j++) s += s_loss[j];
    pb[kNrcMlpParams] = s;
  }
  // layers 4 .. 0: dW_l = delta_{l+1} (x) A_l, delta_l = (W_l^T delta_{l+1}) [A_l > 0]
#pragma unroll
  for (int l = 4; l >= 0; l--) {
    const uint16_t* dl = s_dlt + ((l + 1) & 1) * kActRows;
    weight_grad(dl, s_act + l * kActRows, w, l == 0 ? 48u : 64u, col, h, pb + kNrcOffW[l]);
    if (l > 0) {
      acc[0] = zero16();
      acc[1] = zero16();
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
          acc[mt] = mfma(load8(G.wb + kNrcBwdOff[l] + (32 * mt + col) * 64 + 16 * c + 8 * h), d[c], acc[mt]);
      mask_delta(acc, s_act + l * kActRows, s_dlt + (l & 1) * kActRows, sl, h, d);
      __syncthreads();
    }
  }
  // input gradient of the 32 encoding dims (level l, feature f at 2l + f),
  // scattered into the grid by nrc_grid_scatter_kernel
  f32x16 gx = zero16();
#pragma unroll
  for (int c = 0; c < 4; c++) gx = mfma(load8(G.wb + kNrcBwdOff[0] + col * 64 + 16 * c + 8 * h), d[c], gx);
  if (!valid) return;
  float2* gs = (float2*)(genc + (size_t)si * 32);
#pragma unroll
  for (int qd = 0; qd < 4; qd++)
#pragma unroll
    for (int e = 0; e < 2; e++) gs[4 * qd + 2 * h + e] = make_float2(gx[4 * qd + 2 * e], gx[4 * qd + 2 * e + 1]);
}

// Grid gradients: each (sample, level) adds the level's two encoding
// gradients times each corner's trilinear weight to the eight corner
// entries, as 2^-36 fixed-point 64-bit integers (the sum is exact and
// independent of the order the atomics land in). Sixteen adjacent lanes
// take one (sample, level): lane k the feature k & 1 of corner k >> 1, so a
// wave's atomic touches the few cache lines of four cells (a corner's two
// features share a line, and at densely indexed levels so do the +x
// neighbours) instead of 64 unrelated lines, and the atomics of every
// sample-level pair are in flight at once.
__global__ __launch_bounds__(256) void nrc_grid_scatter_kernel(NrcGridView G, const NrcQuery* __restrict__ q,
                                                              const NrcControl* __restrict__ ctl,
                                                              const float* __restrict__ genc,
                                                              unsigned long long* __restrict__ gg) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t pair = i >> 4, k = i & 15u;
  const uint32_t s = pair / kNrcMaxLevels, l = pair % kNrcMaxLevels;
  if (s >= ctl->n || l >= G.levels) return;
  NrcQuery qq;
  qq.pos[0] = q[s].pos[0];
  qq.pos[1] = q[s].pos[1];
  qq.pos[2] = q[s].pos[2];
  float pn[3];
  normalise(G, qq, pn);
  const float gk = genc[(size_t)s * 32 + 2 * l + (k & 1u)];
  const Cell c = cell_of(G, pn, l);
  float wk;
  const uint32_t id = corner(G, l, c, (int)(k >> 1), wk);
  atomicAdd(gg + (((size_t)l << G.log2_table) + id) * 2 + (k & 1u), to_fixed(gk * wk));
}

__device__ __forceinline__ void adam_update(float& p, float g, float& m, float& v, const NrcHyper& hp, float c1,
                                            float c2) {
  m = hp.b1 * m + (1.f - hp.b1) * g;
  v = hp.b2 * v + (1.f - hp.b2) * g * g;
  p -= hp.lr * (m / c1) / (sqrtf(v / c2) + hp.eps);
}

__device__ __forceinline__ void bias_corrections(const NrcControl* ctl, const NrcHyper& hp, float& c1, float& c2) {
  const double t = (double)ctl->step + 1.0;
  c1 = (float)(1.0 - pow((double)hp.b1, t));
  c2 = (float)(1.0 - pow((double)hp.b2, t));
}

// Sums the training blocks' partials of each parameter: kReduceSeg threads
// per parameter each add a contiguous run of blocks in block order, then the
// runs are added in run order (a fixed order, so results are reproducible).
constexpr uint32_t kReduceSeg = 8;
constexpr uint32_t kReduceParams = 256 / kReduceSeg;  // parameters per block

__global__ __launch_bounds__(256) void nrc_reduce_mlp_kernel(const float* __restrict__ part, NrcControl* ctl,
                                                            float* __restrict__ w, float* __restrict__ m,
                                                            float* __restrict__ v, float* __restrict__ grad_out,
                                                            NrcHyper hp, int apply) {
  __shared__ float s_run[kReduceSeg][kReduceParams];
  const uint32_t n = ctl->n;
  if (n == 0) return;
  const uint32_t pl = threadIdx.x % kReduceParams, seg = threadIdx.x / kReduceParams;
  const uint32_t i = blockIdx.x * kReduceParams + pl;
  const uint32_t nb = (n + kNrcTrainBlock - 1) / kNrcTrainBlock;
  const uint32_t per = (nb + kReduceSeg - 1) / kReduceSeg;
  float r = 0.f;
  if (i <= kNrcMlpParams)
    for (uint32_t b = seg * per; b < min(nb, (seg + 1) * per); b++) r += part[(size_t)b * kNrcPartStride + i];
  s_run[seg][pl] = r;
  __syncthreads();
  if (seg != 0 || i > kNrcMlpParams) return;
  float s = 0.f;
#pragma unroll
  for (uint32_t k = 0; k < kReduceSeg; k++) s += s_run[k][pl];
  if (i == kNrcMlpParams) {
    ctl->last_loss = s;
    return;
  }
  if (!apply) {
    grad_out[i] = s;
    return;
  }
  float c1, c2;
  bias_corrections(ctl, hp, c1, c2);
  float p = w[i], mm = m[i], vv = v[i];
  adam_update(p, s, mm, vv, hp, c1, c2);
  w[i] = p;
  m[i] = mm;
  v[i] = vv;
}

// Moving average of the parameters for inference: e_t = a e_{t-1} +
// (1 - a) p_t, used as e_t / (1 - a^t) (t counts steps since the
// parameters were set).
__device__ __forceinline__ float ema_debias(const NrcControl* ctl, float a) {
  return (float)(1.0 / (1.0 - pow((double)a, (double)ctl->step + 1.0)));
}

__global__ __launch_bounds__(256) void nrc_adam_grid_kernel(float* __restrict__ table, float* __restrict__ m,
                                                           float* __restrict__ v, unsigned long long* __restrict__ gg,
                                                           uint64_t n, const NrcControl* __restrict__ ctl,
                                                           float* __restrict__ grad_out, NrcHyper hp, int apply,
                                                           float* __restrict__ ema, float* __restrict__ inf,
                                                           float alpha) {
  if (ctl->n == 0) return;
  float c1 = 1.f, c2 = 1.f;
  if (apply) bias_corrections(ctl, hp, c1, c2);
  const float corr = ema ? ema_debias(ctl, alpha) : 1.f;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const auto one = [&](uint64_t i, long long fx) {
    const float g = (float)fx * (1.f / 68719476736.f);
    if (!apply) {
      grad_out[i] = g;
      return;
    }
    float p = table[i], mm = m[i], vv = v[i];
    adam_update(p, g, mm, vv, hp, c1, c2);
    table[i] = p;
    m[i] = mm;
    v[i] = vv;
    if (ema) {
      const float e = alpha * ema[i] + (1.f - alpha) * p;
      ema[i] = e;
      inf[i] = e * corr;
    }
  };
  if (apply && !ema) {
    // four parameters per thread with 16-byte accesses (the tables are
    // 256-byte aligned); the gradients are cleared where any was set
    const uint64_t n4 = n / 4;
    for (uint64_t j = t0; j < n4; j += stride) {
      ulonglong2* g2 = (ulonglong2*)gg + 2 * j;
      const ulonglong2 ga = g2[0], gb = g2[1];
      const long long fx[4] = {(long long)ga.x, (long long)ga.y, (long long)gb.x, (long long)gb.y};
      if (fx[0] | fx[1] | fx[2] | fx[3]) {
        g2[0] = make_ulonglong2(0ull, 0ull);
        g2[1] = make_ulonglong2(0ull, 0ull);
      }
      float4 p = ((float4*)table)[j], mm = ((float4*)m)[j], vv = ((float4*)v)[j];
      adam_update(p.x, (float)fx[0] * (1.f / 68719476736.f), mm.x, vv.x, hp, c1, c2);
      adam_update(p.y, (float)fx[1] * (1.f / 68719476736.f), mm.y, vv.y, hp, c1, c2);
      adam_update(p.z, (float)fx[2] * (1.f / 68719476736.f), mm.z, vv.z, hp, c1, c2);
      adam_update(p.w, (float)fx[3] * (1.f / 68719476736.f), mm.w, vv.w, hp, c1, c2);
      ((float4*)table)[j] = p;
      ((float4*)m)[j] = mm;
      ((float4*)v)[j] = vv;
    }
    for (uint64_t i = 4 * n4 + t0; i < n; i += stride) {
      const long long fx = (long long)gg[i];
      if (fx != 0) gg[i] = 0ull;
      one(i, fx);
    }
    return;
  }
  for (uint64_t i = t0; i < n; i += stride) {
    const long long fx = (long long)gg[i];
    if (fx != 0) gg[i] = 0ull;
    one(i, fx);
  }
}

__global__ __launch_bounds__(256) void nrc_ema_mlp_kernel(const float* __restrict__ w, float* __restrict__ ema,
                                                         float* __restrict__ inf, const NrcControl* __restrict__ ctl,
                                                         float alpha) {
  if (ctl->n == 0) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNrcMlpParams) return;
  const float e = alpha * ema[i] + (1.f - alpha) * w[i];
  ema[i] = e;
  inf[i] = e * ema_debias(ctl, alpha);
}

__global__ __launch_bounds__(256) void nrc_pack_kernel(const float* __restrict__ w, uint16_t* __restrict__ wf,
                                                      uint16_t* __restrict__ wb, NrcControl* ctl, int count_step) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && count_step && ctl->n > 0) ctl->step += 1;
  if (i < kNrcFwdSize) {
    float val;
    if (i < kNrcFwdOff[1]) {
      const uint32_t mm = i / 48, ks = i % 48;
      val = w[kNrcOffW[0] + mm * 48 + kperm(ks)];
    } else if (i < kNrcFwdOff[5]) {
      const uint32_t l = 1 + (i - kNrcFwdOff[1]) / 4096, r = (i - kNrcFwdOff[1]) % 4096;
      val = w[kNrcOffW[l] + (r / 64) * 64 + kperm(r % 64)];
    } else {
      const uint32_t r = i - kNrcFwdOff[5], mm = r / 64;
      val = mm < 3 ? w[kNrcOffW[5] + mm * 64 + kperm(r % 64)] : 0.f;
    }
    wf[i] = (uint16_t)bf16_bits(val);
    return;
  }
  const uint32_t j = i - kNrcFwdSize;
  if (!wb || j >= kNrcBwdSize) return;
  float val;
  if (j < kNrcBwdOff[1]) {
    const uint32_t mm = j / 64, ks = j % 64;  // input dim mm < 32
    val = w[kNrcOffW[0] + kperm(ks) * 48 + mm];
  } else if (j < kNrcBwdOff[5]) {
    const uint32_t l = 1 + (j - kNrcBwdOff[1]) / 4096, r = (j - kNrcBwdOff[1]) % 4096;
    val = w[kNrcOffW[l] + kperm(r % 64) * 64 + r / 64];
  } else {
    const uint32_t r = j - kNrcBwdOff[5], mm = r / 16, o = kperm(r % 16);
    val = o < 3 ? w[kNrcOffW[5] + o * 64 + mm] : 0.f;
  }
  wb[j] = (uint16_t)bf16_bits(val);
}

uint32_t infer_grid(uint32_t n) {
  // persistent blocks: enough waves to cover the queries, at most 8 blocks per CU
  const uint32_t tiles = (n + 31) / 32;
  const uint32_t blocks = (tiles + 3) / 4;
  return blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
}

}  // namespace

hipError_t nrc_launch_infer(const NrcGridView& G, const NrcQuery* q, const uint32_t* count, uint32_t n, float* out,
                            hipStream_t st) {
  hipLaunchKernelGGL(nrc_infer_kernel<false>, dim3(infer_grid(n)), dim3(kInferBlock), 0, st, G, q, count, n, out,
                     nullptr, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t nrc_launch_infer_paths(const NrcGridView& G, const NrcQuery* q, const uint32_t* count, uint32_t capacity,
                                  const uint32_t* slots, const float4* beta, float4* L, hipStream_t st) {
  hipLaunchKernelGGL(nrc_infer_kernel<true>, dim3(infer_grid(capacity)), dim3(kInferBlock), 0, st, G, q, count,
                     capacity, nullptr, slots, beta, L);
  return hipGetLastError();
}

hipError_t nrc_launch_train(const NrcGridView& G, const NrcQuery* q, const float* targets, const NrcControl* ctl,
                            uint32_t capacity, float* part, float* genc, unsigned long long* grid_grad,
                            hipStream_t st) {
  const uint32_t blocks = (capacity + kNrcTrainBlock - 1) / kNrcTrainBlock;
  hipLaunchKernelGGL(nrc_train_kernel, dim3(blocks < 1 ? 1 : blocks), dim3(128), 0, st, G, q, targets, ctl, part,
                     genc);
  const uint64_t lanes = (uint64_t)(capacity < 1 ? 1 : capacity) * kNrcMaxLevels * 16;
  hipLaunchKernelGGL(nrc_grid_scatter_kernel, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, st, G, q, ctl,
                     genc, grid_grad);
  return hipGetLastError();
}

hipError_t nrc_launch_reduce_mlp(const float* part, NrcControl* ctl, uint32_t capacity, float* w, float* m, float* v,
                                 float* grad_out, NrcHyper hp, bool apply, hipStream_t st) {
  (void)capacity;
  hipLaunchKernelGGL(nrc_reduce_mlp_kernel, dim3((kNrcMlpParams + 1 + kReduceParams - 1) / kReduceParams), dim3(256),
                     0, st, part, ctl, w, m, v, grad_out, hp, apply ? 1 : 0);
  return hipGetLastError();
}

hipError_t nrc_launch_adam_grid(float* table, float* m, float* v, unsigned long long* grid_grad, uint64_t n,
                                const NrcControl* ctl, float* grad_out, NrcHyper hp, bool apply, hipStream_t st,
                                float* ema, float* inf, float alpha) {
  const uint64_t b = (n + 255) / 256;
  const uint32_t blocks = (uint32_t)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
  hipLaunchKernelGGL(nrc_adam_grid_kernel, dim3(blocks), dim3(256), 0, st, table, m, v, grid_grad, n, ctl, grad_out,
                     hp, apply ? 1 : 0, ema, inf, alpha);
  return hipGetLastError();
}

hipError_t nrc_launch_pack(const float* w, uint16_t* wf, uint16_t* wb, NrcControl* ctl, bool count_step,
                           hipStream_t st) {
  const uint32_t n = wb ? kNrcFwdSize + kNrcBwdSize : kNrcFwdSize;
  hipLaunchKernelGGL(nrc_pack_kernel, dim3((n + 255) / 256), dim3(256), 0, st, w, wf, wb, ctl, count_step ? 1 : 0);
  return hipGetLastError();
}

hipError_t nrc_launch_ema_mlp(const float* w, float* ema, float* inf, const NrcControl* ctl, float alpha,
                              hipStream_t st) {
  hipLaunchKernelGGL(nrc_ema_mlp_kernel, dim3((kNrcMlpParams + 255) / 256), dim3(256), 0, st, w, ema, inf, ctl, alpha);
  return hipGetLastError();
}

}  // namespace rrhip
