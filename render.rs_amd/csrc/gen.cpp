// Deterministic workload generators (host): the displaced heightfield of
// the ray-query benchmark and its query rays.
#include <cmath>
#include <thread>
#include <vector>

#include "api_internal.h"
#include "sampling.h"

using namespace rrhip;

extern "C" int rrhip_gen_heightfield(uint32_t nx, uint32_t ny, uint64_t seed, float amplitude,
                                     float frequency, uint32_t octaves, float* positions,
                                     uint32_t* indices) {
  if (nx == 0 || ny == 0 || !positions || !indices)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_gen_heightfield: empty grid or null output");
  if ((uint64_t)(nx + 1) * (ny + 1) >= 0xffffffffull)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_gen_heightfield: grid too large");
  Pcg32 r = Pcg32::make(seed, 0x5eed);
  const double off[3] = {1000.0 * r.next_f64(), 1000.0 * r.next_f64(), 1000.0 * r.next_f64()};
  const uint32_t vx = nx + 1, vy = ny + 1;
  auto rows = [&](uint32_t j0, uint32_t j1) {
    for (uint32_t j = j0; j < j1; j++) {
      const double y = (double)j / (double)ny;
      for (uint32_t i = 0; i < vx; i++) {
        const double x = (double)i / (double)nx;
        const double z = (double)amplitude *
                         fbm3(x, y, 0.0, (double)frequency, octaves, 0.5, 2.0, off);
        float* p = positions + 3 * ((size_t)j * vx + i);
        p[0] = (float)x;
        p[1] = (float)y;
        p[2] = (float)z;
      }
    }
  };
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; t++)
    pool.emplace_back(rows, (uint32_t)((uint64_t)vy * t / nt), (uint32_t)((uint64_t)vy * (t + 1) / nt));
  for (auto& th : pool) th.join();
  for (uint32_t j = 0; j < ny; j++) {
    for (uint32_t i = 0; i < nx; i++) {
      const uint32_t v00 = j * vx + i, v10 = v00 + 1, v01 = v00 + vx, v11 = v01 + 1;
      uint32_t* q = indices + 6 * ((size_t)j * nx + i);
      q[0] = v00;
      q[1] = v10;
      q[2] = v11;
      q[3] = v00;
      q[4] = v11;
      q[5] = v01;
    }
  }
  return RRHIP_OK;
}

extern "C" int rrhip_gen_query_rays(const float lo[3], const float hi[3], float inflate,
                                    uint64_t seed, uint64_t n, rrhip_ray* rays) {
  if (!lo || !hi || (!rays && n))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_gen_query_rays: null argument");
  double c[3], h[3];
  for (int a = 0; a < 3; a++) {
    c[a] = 0.5 * ((double)lo[a] + (double)hi[a]);
    h[a] = 0.5 * ((double)hi[a] - (double)lo[a]) * (1.0 + (double)inflate);
  }
  const double two_pi = 6.283185307179586;
  auto work = [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; i++) {
      Pcg32 r = Pcg32::for_pixel_sample(i, seed);
      const double ox = c[0] + h[0] * (2.0 * r.next_f64() - 1.0);
      const double oy = c[1] + h[1] * (2.0 * r.next_f64() - 1.0);
      const double oz = c[2] + h[2] * (2.0 * r.next_f64() - 1.0);
      const double u = r.next_f64(), v = r.next_f64();
      const double z = 1.0 - 2.0 * u;
      const double s = std::sqrt(std::fmax(0.0, 1.0 - z * z));
      const double phi = two_pi * v;
      rrhip_ray& ray = rays[i];
      ray.ox = (float)ox;
      ray.oy = (float)oy;
      ray.oz = (float)oz;
      ray.tmin = 1e-9f;
      ray.dx = (float)(s * std::cos(phi));
      ray.dy = (float)(s * std::sin(phi));
      ray.dz = (float)z;
      ray.tmax = INFINITY;
    }
  };
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 65536) nt = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; t++) pool.emplace_back(work, n * t / nt, n * (t + 1) / nt);
  for (auto& th : pool) th.join();
  return RRHIP_OK;
}
