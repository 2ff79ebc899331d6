// PCG32 and hash-gradient Perlin noise, usable on host and device.
//
// PCG32 XSH-RR with render.rs's seeding (src/raytracer/pt/sampler.rs:6-58):
// the integer stream is identical on every backend, so per-(pixel, sample)
// draws agree bit for bit with the CPU path. Perlin/fBm follow
// src/geometry/displace.rs:36-97 (fixed hash, 12 cube-edge gradients,
// quintic fade).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rrhip {

struct Pcg32 {
  uint64_t state;
  uint64_t inc;

  __host__ __device__ static inline uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  __host__ __device__ inline uint32_t next_u32() {
    const uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    const uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  __host__ __device__ static inline Pcg32 make(uint64_t seed, uint64_t stream) {
    Pcg32 r{0, (stream << 1) | 1u};
    r.next_u32();
    r.state += seed;
    r.next_u32();
    return r;
  }
  __host__ __device__ static inline Pcg32 for_pixel_sample(uint64_t pixel, uint64_t sample) {
    return make(mix(pixel * 0x9e3779b97f4a7c15ull ^ sample), mix(sample) | 1u);
  }
  // u32 / 2^32 in f64, as the CPU path draws (sampler.rs:45-47).
  __host__ __device__ inline double next_f64() { return (double)next_u32() * (1.0 / 4294967296.0); }
  // 24-bit uniform in [0, 1) for f32 arithmetic.
  __host__ __device__ inline float next_f32() {
    return (float)(next_u32() >> 8) * (1.0f / 16777216.0f);
  }
};

__host__ __device__ inline uint32_t perlin_hash(int64_t x) {
  uint64_t u = (uint64_t)x * 0x9e3779b97f4a7c15ull;
  u ^= (uint64_t)((int64_t)u >> 29);
  u *= 0xbf58476d1ce4e5b9ull;
  u ^= (uint64_t)((int64_t)u >> 32);
  return (uint32_t)u;
}

__host__ __device__ inline double perlin_corner(int64_t ix, int64_t iy, int64_t iz, double fx,
                                               double fy, double fz) {
  const uint64_t h0 = (uint64_t)ix * 73856093ull ^ (uint64_t)iy * 19349663ull ^
                      (uint64_t)iz * 83492791ull;
  const uint32_t g = perlin_hash((int64_t)h0) % 12u;
  // 12 cube-edge gradients: (±1,±1,0), (±1,0,±1), (0,±1,±1).
  const double sa = (g & 1u) ? -1.0 : 1.0;
  const double sb = (g & 2u) ? -1.0 : 1.0;
  if (g < 4u) return sa * fx + sb * fy + 0.0 * fz;
  if (g < 8u) return sa * fx + 0.0 * fy + sb * fz;
  return 0.0 * fx + sa * fy + sb * fz;
}

__host__ __device__ inline double perlin3(double px, double py, double pz) {
  const double cx = floor(px), cy = floor(py), cz = floor(pz);
  const double fx = px - cx, fy = py - cy, fz = pz - cz;
  const int64_t ix = (int64_t)cx, iy = (int64_t)cy, iz = (int64_t)cz;
  double c[8];
  for (int k = 0; k < 8; k++) {
    const int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    c[k] = perlin_corner(ix + dx, iy + dy, iz + dz, fx - dx, fy - dy, fz - dz);
  }
  auto fade = [](double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); };
  auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
  const double u = fade(fx), v = fade(fy), w = fade(fz);
  const double y0 = lerp(lerp(c[0], c[1], u), lerp(c[2], c[3], u), v);
  const double y1 = lerp(lerp(c[4], c[5], u), lerp(c[6], c[7], u), v);
  return lerp(y0, y1, w);
}

__host__ __device__ inline double fbm3(double px, double py, double pz, double frequency,
                                       uint32_t octaves, double gain, double lacunarity,
                                       const double off[3]) {
  double sum = 0.0, amp = 1.0, freq = frequency;
  const uint32_t n = octaves < 1 ? 1 : octaves;
  for (uint32_t i = 0; i < n; i++) {
    sum += amp * perlin3(px * freq + off[0], py * freq + off[1], pz * freq + off[2]);
    amp *= gain;
    freq *= lacunarity;
  }
  return sum;
}

}  // namespace rrhip
