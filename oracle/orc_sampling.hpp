// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// PCG32 (src/raytracer/pt/sampler.rs:6-58) and the hash-gradient Perlin /
// fBm of src/geometry/displace.rs:36-97.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

struct Pcg32 {
  uint64_t state = 0, inc = 0;
  // sampler.rs:12-21 (pcg32_srandom_r)
  static Pcg32 make(uint64_t seed, uint64_t stream) {
    Pcg32 r;
    r.state = 0;
    r.inc = (stream << 1) | 1u;
    r.next_u32();
    r.state += seed;
    r.next_u32();
    return r;
  }
  static uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  // sampler.rs:24-32
  static Pcg32 for_pixel_sample(uint64_t pixel, uint64_t sample) {
    return make(mix(pixel * 0x9e3779b97f4a7c15ull ^ sample), mix(sample) | 1u);
  }
  // sampler.rs:34-42 (XSH-RR)
  uint32_t next_u32() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
  }
  // sampler.rs:45-47
  double next_f64() { return (double)next_u32() * (1.0 / 4294967296.0); }
};

// displace.rs:37-43
inline uint32_t perlin_hash(int64_t x) {
  uint64_t u = (uint64_t)x;
  u = u * 0x9e3779b97f4a7c15ull;
  u ^= (uint64_t)((int64_t)u >> 29);
  u = u * 0xbf58476d1ce4e5b9ull;
  u ^= (uint64_t)((int64_t)u >> 32);
  return (uint32_t)u;
}

// displace.rs:45-54
inline void perlin_gradient(int64_t ix, int64_t iy, int64_t iz, double g[3]) {
  static const double G[12][3] = {{1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0},
                                  {1, 0, 1},  {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
                                  {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
  uint64_t a = (uint64_t)ix * 73856093ull, b = (uint64_t)iy * 19349663ull,
           c = (uint64_t)iz * 83492791ull;
  uint32_t h = perlin_hash((int64_t)(a ^ b ^ c));
  const double* s = G[h % 12];
  g[0] = s[0];
  g[1] = s[1];
  g[2] = s[2];
}

inline double perlin_fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

// displace.rs:60-80
inline double perlin(double px, double py, double pz) {
  double cx = std::floor(px), cy = std::floor(py), cz = std::floor(pz);
  double f[3] = {px - cx, py - cy, pz - cz};
  int64_t ix = (int64_t)cx, iy = (int64_t)cy, iz = (int64_t)cz;
  double corner[8];
  for (int k = 0; k < 8; k++) {
    int64_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    double g[3];
    perlin_gradient(ix + dx, iy + dy, iz + dz, g);
    double d0 = f[0] - (double)dx, d1 = f[1] - (double)dy, d2 = f[2] - (double)dz;
    corner[k] = g[0] * d0 + g[1] * d1 + g[2] * d2;
  }
  double u = perlin_fade(f[0]), v = perlin_fade(f[1]), w = perlin_fade(f[2]);
  auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
  double x00 = lerp(corner[0], corner[1], u);
  double x10 = lerp(corner[2], corner[3], u);
  double x01 = lerp(corner[4], corner[5], u);
  double x11 = lerp(corner[6], corner[7], u);
  double y0 = lerp(x00, x10, v);
  double y1 = lerp(x01, x11, v);
  return lerp(y0, y1, w);
}

struct FbmParams {
  double frequency = 1.0;
  uint32_t octaves = 4;
  double gain = 0.5;
  double lacunarity = 2.0;
  double offset[3] = {0, 0, 0};
};

// displace.rs:82-97
inline double fbm(double px, double py, double pz, const FbmParams& p) {
  double sum = 0.0, amp = 1.0, freq = p.frequency;
  uint32_t oct = p.octaves < 1 ? 1 : p.octaves;
  for (uint32_t i = 0; i < oct; i++) {
    sum += amp * perlin(px * freq + p.offset[0], py * freq + p.offset[1], pz * freq + p.offset[2]);
    amp *= p.gain;
    freq *= p.lacunarity;
  }
  return sum;
}

}  // namespace orc
