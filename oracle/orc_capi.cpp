// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// C entry points of the oracle, loaded by the tests (ctypes) as the
// checker and by bench.py as the CPU baseline ("port" of render.rs's
// multithreaded f64 CPU path). Nothing here is linked into the product.
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "orc_accel.hpp"
#include "orc_sampling.hpp"

using namespace orc;

namespace {

// Query ray layout shared with the product ABI: 8 floats
// {ox, oy, oz, tmin, dx, dy, dz, tmax}.
inline Ray ray_from_f32(const float* r) {
  return Ray::raw(Vec3{(double)r[0], (double)r[1], (double)r[2]},
                  Vec3{(double)r[4], (double)r[5], (double)r[6]});
}

template <class F>
void parallel_for(uint64_t n, int nthreads, F&& fn) {
  if (nthreads <= 1 || n < 1024) {
    for (uint64_t i = 0; i < n; i++) fn(i);
    return;
  }
  std::atomic<uint64_t> next{0};
  const uint64_t chunk = 256;
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; t++) {
    pool.emplace_back([&]() {
      for (;;) {
        uint64_t b = next.fetch_add(chunk);
        if (b >= n) break;
        uint64_t e = b + chunk < n ? b + chunk : n;
        for (uint64_t i = b; i < e; i++) fn(i);
      }
    });
  }
  for (auto& th : pool) th.join();
}

struct OrcMesh {
  Mesh mesh;
};

// Closest hit of one query ray against the mesh BLAS (the C1 query: a
// world-space mesh traversed by Bvh::traverse with Mesh::intersect_triangle,
// mesh.rs:346-353; the lower t bound is the ray's own tmin).
bool mesh_closest(const Mesh& m, const float* rf, uint32_t* prim, double* t, double* u, double* v,
                  uint64_t* nn, uint64_t* np) {
  Ray ray = ray_from_f32(rf);
  double tmin = (double)rf[3], tmax = (double)rf[7];
  auto isect = [&](uint32_t tri, double t_cur, double* t_out) {
    double tt, uu, vv;
    if (!m.intersect_triangle(ray, tri, tmin, t_cur, &tt, &uu, &vv)) return false;
    if (!(tt < tmax)) return false;  // the ray's own bound stays strict
    *t_out = tt;
    return true;
  };
  uint32_t wp;
  double wt;
  if (!m.blas.traverse(ray, tmax, isect, &wp, &wt, nn, np)) return false;
  // The winner's barycentrics (the same deterministic f64 evaluation).
  double tt, uu = 0, vv = 0;
  m.intersect_triangle(ray, wp, tmin, INF, &tt, &uu, &vv);
  *prim = wp;
  *t = wt;
  *u = uu;
  *v = vv;
  return true;
}

}  // namespace

extern "C" {

// ---- PCG32 ---------------------------------------------------------------
void orc_pcg32_seq(uint64_t seed, uint64_t stream, uint64_t n, uint32_t* out) {
  Pcg32 r = Pcg32::make(seed, stream);
  for (uint64_t i = 0; i < n; i++) out[i] = r.next_u32();
}

void orc_pcg32_pixel_sample(uint64_t pixel, uint64_t sample, uint64_t n, uint32_t* out) {
  Pcg32 r = Pcg32::for_pixel_sample(pixel, sample);
  for (uint64_t i = 0; i < n; i++) out[i] = r.next_u32();
}

// ---- Perlin / fBm ----------------------------------------------------------
void orc_perlin(const double* xyz, uint64_t n, double* out) {
  for (uint64_t i = 0; i < n; i++) out[i] = perlin(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

void orc_fbm(const double* xyz, uint64_t n, double frequency, uint32_t octaves, double gain,
             double lacunarity, const double* offset3, double* out) {
  FbmParams p;
  p.frequency = frequency;
  p.octaves = octaves;
  p.gain = gain;
  p.lacunarity = lacunarity;
  for (int k = 0; k < 3; k++) p.offset[k] = offset3[k];
  for (uint64_t i = 0; i < n; i++) out[i] = fbm(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], p);
}

// ---- Mesh BVH (closest / any hit) -------------------------------------------
void* orc_mesh_create(const float* positions, uint64_t nverts, const uint32_t* indices,
                      uint64_t ntris, int bins, int max_leaf, int sah_leaf_mult) {
  auto* h = new OrcMesh();
  BvhParams p;
  p.bins = bins;
  p.max_leaf = max_leaf;
  p.sah_leaf_mult = sah_leaf_mult;
  h->mesh.build(positions, nverts, indices, ntris, p);
  return h;
}

void orc_mesh_destroy(void* h) { delete (OrcMesh*)h; }

uint64_t orc_mesh_node_count(void* h) { return ((OrcMesh*)h)->mesh.blas.nodes.size(); }

// rays: n x 8 floats. Outputs: prim (0xFFFFFFFF = miss), t/u/v in f64.
// stats (optional, may be NULL): [0] = nodes popped, [1] = triangles tested.
void orc_mesh_trace_closest(void* h, const float* rays, uint64_t n, uint32_t* prim, double* t,
                            double* u, double* v, int nthreads, uint64_t* stats) {
  const Mesh& m = ((OrcMesh*)h)->mesh;
  std::atomic<uint64_t> tot_nodes{0}, tot_prims{0};
  parallel_for(n, nthreads, [&](uint64_t i) {
    uint64_t nn = 0, np = 0;
    uint32_t p;
    double tt, uu, vv;
    if (mesh_closest(m, rays + 8 * i, &p, &tt, &uu, &vv, stats ? &nn : nullptr,
                     stats ? &np : nullptr)) {
      prim[i] = p;
      t[i] = tt;
      if (u) u[i] = uu;
      if (v) v[i] = vv;
    } else {
      prim[i] = 0xFFFFFFFFu;
      t[i] = 0;
      if (u) u[i] = 0;
      if (v) v[i] = 0;
    }
    if (stats) {
      tot_nodes += nn;
      tot_prims += np;
    }
  });
  if (stats) {
    stats[0] = tot_nodes.load();
    stats[1] = tot_prims.load();
  }
}

// Any hit with t in (tmin, tmax) (bvh.rs:208-245 + mesh.rs:384-407).
void orc_mesh_trace_any(void* h, const float* rays, uint64_t n, uint8_t* occluded, int nthreads) {
  const Mesh& m = ((OrcMesh*)h)->mesh;
  parallel_for(n, nthreads, [&](uint64_t i) {
    const float* rf = rays + 8 * i;
    Ray ray = ray_from_f32(rf);
    double tmin = (double)rf[3], tmax = (double)rf[7];
    bool occ = m.blas.any_hit(ray, tmax, [&](uint32_t tri, double lim) {
      double tt, uu, vv;
      return m.intersect_triangle(ray, tri, tmin, lim, &tt, &uu, &vv) && tt < lim;
    });
    occluded[i] = occ ? 1 : 0;
  });
}

// O(N) brute force over every triangle (the bvh.rs:403-453 oracle pattern).
void orc_brute_closest(const float* positions, const uint32_t* indices, uint64_t ntris,
                       const float* rays, uint64_t n, uint32_t* prim, double* t, int nthreads) {
  Mesh m;  // geometry only: no BVH is built
  uint64_t maxv = 0;
  for (uint64_t k = 0; k < 3 * ntris; k++) maxv = indices[k] > maxv ? indices[k] : maxv;
  m.positions.assign(positions, positions + 3 * (maxv + 1));
  m.indices.assign(indices, indices + 3 * ntris);
  parallel_for(n, nthreads, [&](uint64_t i) {
    const float* rf = rays + 8 * i;
    Ray ray = ray_from_f32(rf);
    double tmin = (double)rf[3], tmax = (double)rf[7];
    uint32_t bp = 0xFFFFFFFFu;
    double bt = tmax;
    for (uint32_t tri = 0; tri < (uint32_t)ntris; tri++) {
      double tt, uu, vv;
      if (m.intersect_triangle(ray, tri, tmin, bt, &tt, &uu, &vv) && tt < tmax) {
        if (bp == 0xFFFFFFFFu || tt < bt || (tt == bt && tri < bp)) {
          bp = tri;
          bt = tt;
        }
      }
    }
    prim[i] = bp;
    t[i] = bp == 0xFFFFFFFFu ? 0.0 : bt;
  });
}

// Distance, measured perpendicular to the ray, from the ray's line to the
// nearest edge line of triangle `prim` (the "near an edge" measure used to
// characterise f32 vs f64 hit-ID disagreements).
void orc_edge_distance(const float* positions, const uint32_t* indices, const float* rays,
                       const uint32_t* prims, uint64_t n, double* out) {
  for (uint64_t i = 0; i < n; i++) {
    if (prims[i] == 0xFFFFFFFFu) {
      out[i] = -1.0;
      continue;
    }
    const float* rf = rays + 8 * i;
    Vec3 o{(double)rf[0], (double)rf[1], (double)rf[2]};
    Vec3 d{(double)rf[4], (double)rf[5], (double)rf[6]};
    Vec3 vtx[3];
    for (int k = 0; k < 3; k++) {
      uint32_t vi = indices[3 * (uint64_t)prims[i] + k];
      vtx[k] = {(double)positions[3 * (uint64_t)vi], (double)positions[3 * (uint64_t)vi + 1],
                (double)positions[3 * (uint64_t)vi + 2]};
    }
    double best = INF;
    for (int k = 0; k < 3; k++) {
      Vec3 a = vtx[k], e = vtx[(k + 1) % 3] - vtx[k];
      Vec3 c = d.cross(e);
      double cl = c.length();
      double dist = cl > 0 ? std::fabs((a - o).dot(c)) / cl : 0.0;
      best = std::fmin(best, dist);
    }
    out[i] = best;
  }
}

}  // extern "C"
