// Deterministic workload generators (host): the displaced heightfield of
// the ray-query benchmark and its query rays.
#include <algorithm>
#include <cmath>
#include <thread>
#include <unordered_map>
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

namespace {

// Area-weighted vertex normals (render.rs smooth_normals, src/geometry/subdiv.rs:329-358).
void smooth_normals(const std::vector<float>& pos, const std::vector<uint32_t>& idx, float* out) {
  const size_t nv = pos.size() / 3;
  std::vector<double> acc(3 * nv, 0.0);
  for (size_t t = 0; t + 2 < idx.size(); t += 3) {
    const float* a = &pos[3 * (size_t)idx[t]];
    const float* b = &pos[3 * (size_t)idx[t + 1]];
    const float* c = &pos[3 * (size_t)idx[t + 2]];
    const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
    const double e2[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    for (int k = 0; k < 3; k++)
      for (int a2 = 0; a2 < 3; a2++) acc[3 * (size_t)idx[t + k] + a2] += n[a2];
  }
  for (size_t v = 0; v < nv; v++) {
    const double* n = &acc[3 * v];
    const double l = std::fmax(std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]), 1e-12);
    for (int a = 0; a < 3; a++) out[3 * v + a] = (float)(n[a] / l);
  }
}

}  // namespace

// Unit icosphere subdivided `levels` times (20 * 4^levels triangles), its
// vertices displaced along their smooth normals by amplitude * fBm
// (render.rs displace_mesh, src/geometry/displace.rs:99-123) and normals
// rebuilt from the displaced faces.
extern "C" int rrhip_gen_icosphere(uint32_t levels, uint64_t seed, float amplitude, float frequency,
                                   uint32_t octaves, float* positions, uint32_t* indices, float* normals) {
  if (levels > 8 || !positions || !indices || !normals)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_gen_icosphere: levels must be <= 8 and outputs non-null");
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                           0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
  std::vector<uint32_t> f = {0, 11, 5, 0, 5,  1, 0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                             4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4,  2, 3, 2, 6, 3,
                             6, 8, 3, 8,  9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  auto unit = [&](size_t i) {
    double* p = &v[3 * i];
    const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    p[0] /= l;
    p[1] /= l;
    p[2] /= l;
  };
  for (size_t i = 0; i < 12; i++) unit(i);
  for (uint32_t lv = 0; lv < levels; lv++) {
    std::vector<uint32_t> nf;
    nf.reserve(f.size() * 4);
    std::unordered_map<uint64_t, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      const uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      const uint32_t id = (uint32_t)(v.size() / 3);
      for (int k = 0; k < 3; k++) v.push_back(0.5 * (v[3 * (size_t)a + k] + v[3 * (size_t)b + k]));
      unit(id);
      mid.emplace(key, id);
      return id;
    };
    for (size_t k = 0; k < f.size(); k += 3) {
      const uint32_t a = f[k], b = f[k + 1], c = f[k + 2];
      const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      const uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nf.insert(nf.end(), tri, tri + 12);
    }
    f.swap(nf);
  }
  const size_t nv = v.size() / 3;
  std::vector<float> pos(3 * nv);
  for (size_t i = 0; i < 3 * nv; i++) pos[i] = (float)v[i];
  std::vector<float> base(3 * nv);
  smooth_normals(pos, f, base.data());
  Pcg32 r = Pcg32::make(seed, 0x5eed);
  const double off[3] = {1000.0 * r.next_f64(), 1000.0 * r.next_f64(), 1000.0 * r.next_f64()};
  for (size_t i = 0; i < nv; i++) {
    const double p[3] = {(double)pos[3 * i], (double)pos[3 * i + 1], (double)pos[3 * i + 2]};
    const double d = fbm3(p[0], p[1], p[2], (double)frequency, octaves, 0.5, 2.0, off) * (double)amplitude;
    for (int k = 0; k < 3; k++) positions[3 * i + k] = (float)(p[k] + (double)base[3 * i + k] * d);
  }
  std::vector<float> disp(positions, positions + 3 * nv);
  smooth_normals(disp, f, normals);
  std::copy(f.begin(), f.end(), indices);
  return RRHIP_OK;
}

// medium.rs:55-63 (DensityField::Fbm) sampled on an n^3 vertex grid.
extern "C" int rrhip_gen_density_grid(uint32_t n, uint64_t seed, float frequency, uint32_t octaves,
                                      float coverage, float sharpness, const float lo[3], const float hi[3],
                                      float* out) {
  if (n < 2 || n > 1024 || !lo || !hi || !out)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_gen_density_grid: need 2 <= n <= 1024 and non-null arrays");
  for (int a = 0; a < 3; a++)
    if (!(hi[a] > lo[a])) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_gen_density_grid: empty bounds");
  Pcg32 r = Pcg32::make(seed, 0x5eed);
  const double off[3] = {1000.0 * r.next_f64(), 1000.0 * r.next_f64(), 1000.0 * r.next_f64()};
  auto slabs = [&](uint32_t k0, uint32_t k1) {
    for (uint32_t k = k0; k < k1; k++)
      for (uint32_t j = 0; j < n; j++)
        for (uint32_t i = 0; i < n; i++) {
          const uint32_t idx[3] = {i, j, k};
          double p[3];
          for (int a = 0; a < 3; a++)
            p[a] = (double)lo[a] + ((double)hi[a] - (double)lo[a]) * (double)idx[a] / (double)(n - 1);
          const double v = fbm3(p[0], p[1], p[2], (double)frequency, octaves, 0.5, 2.0, off) * 0.5 + 0.5;
          const double rho = (v - (1.0 - (double)coverage)) * (double)sharpness;
          out[((size_t)k * n + j) * n + i] = (float)std::min(std::max(rho, 0.0), 1.0);
        }
  };
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; t++)
    pool.emplace_back(slabs, (uint32_t)((uint64_t)n * t / nt), (uint32_t)((uint64_t)n * (t + 1) / nt));
  for (auto& th : pool) th.join();
  return RRHIP_OK;
}

extern "C" int rrhip_gen_fbm(uint64_t n, const double* points, double frequency, uint32_t octaves, double gain,
                             double lacunarity, const double offset[3], double* out) {
  if (n && (!points || !out)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_gen_fbm: null buffer");
  if (octaves > 64) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_gen_fbm: at most 64 octaves");
  const double zero[3] = {0.0, 0.0, 0.0};
  const double* off = offset ? offset : zero;
  for (uint64_t i = 0; i < n; i++)
    out[i] = fbm3(points[3 * i], points[3 * i + 1], points[3 * i + 2], frequency, octaves, gain, lacunarity, off);
  return RRHIP_OK;
}
