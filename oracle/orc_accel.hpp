// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// f64 restatement of render.rs's binned-SAH BVH (src/accel/bvh.rs) and of
// the mesh triangle test (src/geometry/mesh.rs:113-142).
#pragma once
#include <cstdint>
#include <limits>
#include <vector>

#include "orc_math.hpp"

namespace orc {

constexpr double INF = std::numeric_limits<double>::infinity();

// bvh.rs:9-73
struct Aabb {
  Vec3 min{INF, INF, INF};
  Vec3 max{-INF, -INF, -INF};
  void grow_point(const Vec3& p) {
    min = {std::fmin(min.x, p.x), std::fmin(min.y, p.y), std::fmin(min.z, p.z)};
    max = {std::fmax(max.x, p.x), std::fmax(max.y, p.y), std::fmax(max.z, p.z)};
  }
  void grow(const Aabb& o) {
    grow_point(o.min);
    grow_point(o.max);
  }
  Vec3 centroid() const { return (min + max) * 0.5; }
  // bvh.rs:45-51
  double surface_area() const {
    Vec3 d = max - min;
    if (d.x < 0.0 || d.y < 0.0 || d.z < 0.0) return 0.0;
    return 2.0 * (d.x * d.y + d.y * d.z + d.z * d.x);
  }
  // bvh.rs:55-72 (slab test; returns false on miss, entry distance in *t).
  bool hit(const Vec3& o, const Vec3& inv, double t_max, double* t_entry) const;
};

// Build parameters. render.rs hard-codes BINS=16, MAX_LEAF=4 and lets the
// SAH stop at up to 4*MAX_LEAF primitives (bvh.rs:91-92, 330-334); the
// ray-query workload asks for 32 bins and leaves of at most 4, which is
// `{32, 4, 1}` here.
struct BvhParams {
  int bins = 16;
  int max_leaf = 4;
  int sah_leaf_mult = 4;  // SAH-preferred leaf allowed when count <= max_leaf*mult
};

struct BvhNode {
  Aabb bounds;
  uint32_t first = 0;  // interior: left child (right = left+1); leaf: first slot
  uint32_t count = 0;  // 0 = interior
};

class Bvh {
 public:
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> prim_order;

  void build(const std::vector<Aabb>& bounds, const BvhParams& params);

  // bvh.rs:141-204. `isect(prim, t_max, &t)` must return true only for a
  // hit strictly closer than t_max. Equal-t candidates keep the lower prim
  // id (render.rs keeps the first one found; the two agree except at exact
  // ties, which the tie rule makes order-independent).
  template <class F>
  bool traverse(const Ray& ray, double t_max, F&& isect, uint32_t* out_prim, double* out_t,
                uint64_t* n_nodes = nullptr, uint64_t* n_prims = nullptr) const;

  // bvh.rs:208-245
  template <class F>
  bool any_hit(const Ray& ray, double t_limit, F&& hit_prim) const;

 private:
  void subdivide(const std::vector<Aabb>& bounds, const std::vector<Vec3>& centroids,
                 const BvhParams& p, size_t node_idx, size_t start, size_t count);
};

inline bool Aabb::hit(const Vec3& o, const Vec3& inv, double t_max, double* t_entry) const {
  double t0x = (min.x - o.x) * inv.x, t1x = (max.x - o.x) * inv.x;
  double t_near = std::fmin(t0x, t1x), t_far = std::fmax(t0x, t1x);
  double t0y = (min.y - o.y) * inv.y, t1y = (max.y - o.y) * inv.y;
  t_near = std::fmax(t_near, std::fmin(t0y, t1y));
  t_far = std::fmin(t_far, std::fmax(t0y, t1y));
  double t0z = (min.z - o.z) * inv.z, t1z = (max.z - o.z) * inv.z;
  t_near = std::fmax(t_near, std::fmin(t0z, t1z));
  t_far = std::fmin(t_far, std::fmax(t0z, t1z));
  if (t_near <= t_far && t_far > 0.0 && t_near < t_max) {
    *t_entry = std::fmax(t_near, 0.0);
    return true;
  }
  return false;
}

template <class F>
bool Bvh::traverse(const Ray& ray, double t_max, F&& isect, uint32_t* out_prim, double* out_t,
                   uint64_t* n_nodes, uint64_t* n_prims) const {
  if (nodes.empty()) return false;
  const Vec3& o = ray.origin;
  Vec3 inv{1.0 / ray.direction.x, 1.0 / ray.direction.y, 1.0 / ray.direction.z};
  bool found = false;
  uint32_t best_prim = 0;
  double best_t = 0;
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const BvhNode& node = nodes[stack[--sp]];
    if (n_nodes) ++*n_nodes;
    double te;
    if (!node.bounds.hit(o, inv, t_max, &te)) continue;
    if (node.count > 0) {
      for (uint32_t i = node.first; i < node.first + node.count; i++) {
        uint32_t prim = prim_order[i];
        if (n_prims) ++*n_prims;
        double t;
        // Accept t <= t_max only for the equal-t tie rule; strictly closer
        // hits behave exactly as bvh.rs:169-174.
        if (isect(prim, t_max, &t)) {
          if (!found || t < best_t || (t == best_t && prim < best_prim)) {
            t_max = t;
            best_t = t;
            best_prim = prim;
            found = true;
          }
        }
      }
    } else {
      uint32_t l = node.first, r = l + 1;
      double dl, dr;
      bool hl = nodes[l].bounds.hit(o, inv, t_max, &dl);
      bool hr = nodes[r].bounds.hit(o, inv, t_max, &dr);
      if (hl && hr) {
        uint32_t nearer = dl <= dr ? l : r, farther = dl <= dr ? r : l;
        stack[sp++] = farther;
        stack[sp++] = nearer;
      } else if (hl) {
        stack[sp++] = l;
      } else if (hr) {
        stack[sp++] = r;
      }
    }
  }
  if (found) {
    *out_prim = best_prim;
    *out_t = best_t;
  }
  return found;
}

template <class F>
bool Bvh::any_hit(const Ray& ray, double t_limit, F&& hit_prim) const {
  if (nodes.empty()) return false;
  const Vec3& o = ray.origin;
  Vec3 inv{1.0 / ray.direction.x, 1.0 / ray.direction.y, 1.0 / ray.direction.z};
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const BvhNode& node = nodes[stack[--sp]];
    double te;
    if (!node.bounds.hit(o, inv, t_limit, &te)) continue;
    if (node.count > 0) {
      for (uint32_t i = node.first; i < node.first + node.count; i++)
        if (hit_prim(prim_order[i], t_limit)) return true;
    } else {
      stack[sp++] = node.first;
      stack[sp++] = node.first + 1;
    }
  }
  return false;
}

// Triangle mesh with f32 storage and f64 intersection (mesh.rs:20-142).
struct Mesh {
  std::vector<float> positions;   // xyz triples
  std::vector<float> positions1;  // deformation endpoint at shutter close (mesh.rs:31), empty: none
  std::vector<uint32_t> indices;
  Bvh blas;
  Aabb local_bounds;

  // mesh.rs:48-84: triangle boxes cover both endpoints when pos1 is given
  void build(const float* pos, size_t nverts, const uint32_t* idx, size_t ntris,
             const BvhParams& params, const float* pos1 = nullptr);
  Vec3 vertex(uint32_t i) const {
    return {(double)positions[3 * i], (double)positions[3 * i + 1], (double)positions[3 * i + 2]};
  }
  // mesh.rs:96-111: the vertex at shutter time, lerped in f64 when the
  // mesh deforms and time > 0
  Vec3 vertex_at(uint32_t i, double time) const {
    if (positions1.empty() || !(time > 0.0)) return vertex(i);
    const float* a = &positions[3 * i];
    const float* b = &positions1[3 * i];
    return {(double)a[0] + ((double)b[0] - (double)a[0]) * time, (double)a[1] + ((double)b[1] - (double)a[1]) * time,
            (double)a[2] + ((double)b[2] - (double)a[2]) * time};
  }
  size_t triangle_count() const { return indices.size() / 3; }
  // mesh.rs:117-142 with the lower t bound as a parameter (render.rs uses
  // 1e-9 there), at the ray's shutter time.
  bool intersect_triangle(const Ray& ray, uint32_t tri, double t_min, double t_max, double* t,
                          double* u, double* v) const;
};

}  // namespace orc
