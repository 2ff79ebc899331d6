// ORACLE — test infrastructure only (see orc_math.hpp header).
#include "orc_accel.hpp"

#include <algorithm>

namespace orc {

static inline double axis_get(const Vec3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// bvh.rs:96-108
void Bvh::build(const std::vector<Aabb>& bounds, const BvhParams& params) {
  nodes.clear();
  prim_order.resize(bounds.size());
  for (size_t i = 0; i < bounds.size(); i++) prim_order[i] = (uint32_t)i;
  if (bounds.empty()) return;
  std::vector<Vec3> centroids(bounds.size());
  for (size_t i = 0; i < bounds.size(); i++) centroids[i] = bounds[i].centroid();
  nodes.reserve(bounds.size() * 2);
  nodes.push_back(BvhNode{});
  subdivide(bounds, centroids, params, 0, 0, bounds.size());
}

// bvh.rs:256-345 (recursive binned SAH along the widest centroid axis)
void Bvh::subdivide(const std::vector<Aabb>& bounds, const std::vector<Vec3>& centroids,
                    const BvhParams& p, size_t node_idx, size_t start, size_t count) {
  Aabb nb, cb;
  for (size_t k = start; k < start + count; k++) {
    uint32_t q = prim_order[k];
    nb.grow(bounds[q]);
    cb.grow_point(centroids[q]);
  }
  nodes[node_idx].bounds = nb;
  if (count <= (size_t)p.max_leaf) {
    nodes[node_idx].first = (uint32_t)start;
    nodes[node_idx].count = (uint32_t)count;
    return;
  }
  Vec3 extent = cb.max - cb.min;
  int axis = (extent.x >= extent.y && extent.x >= extent.z) ? 0 : (extent.y >= extent.z ? 1 : 2);
  double axis_min = axis_get(cb.min, axis);
  double axis_extent = axis_get(extent, axis);
  size_t split_at = start + count / 2;  // median fallback
  if (axis_extent > 1e-12) {
    const int B = p.bins;
    std::vector<Aabb> bin_bounds(B);
    std::vector<size_t> bin_counts(B, 0);
    double scale = (double)B / axis_extent;
    for (size_t k = start; k < start + count; k++) {
      uint32_t q = prim_order[k];
      size_t b = (size_t)((axis_get(centroids[q], axis) - axis_min) * scale);
      if (b > (size_t)(B - 1)) b = B - 1;
      bin_counts[b]++;
      bin_bounds[b].grow(bounds[q]);
    }
    std::vector<Aabb> right_acc(B);
    Aabb acc;
    for (int i = B - 1; i >= 1; i--) {
      acc.grow(bin_bounds[i]);
      right_acc[i] = acc;
    }
    double best_cost = INF;
    int best_bin = 0;
    Aabb left;
    size_t left_count = 0;
    for (int i = 0; i < B - 1; i++) {
      left.grow(bin_bounds[i]);
      left_count += bin_counts[i];
      size_t right_count = count - left_count;
      if (left_count == 0 || right_count == 0) continue;
      double cost = left.surface_area() * (double)left_count +
                    right_acc[i + 1].surface_area() * (double)right_count;
      if (cost < best_cost) {
        best_cost = cost;
        best_bin = i;
      }
    }
    double leaf_cost = nb.surface_area() * (double)count;
    if (std::isfinite(best_cost) && best_cost < leaf_cost) {
      double boundary = axis_min + axis_extent * (double)(best_bin + 1) / (double)B;
      // bvh.rs:359-368 in-place partition (stable for the 'true' side order).
      size_t i = 0;
      uint32_t* s = prim_order.data() + start;
      for (size_t j = 0; j < count; j++) {
        if (axis_get(centroids[s[j]], axis) < boundary) {
          std::swap(s[i], s[j]);
          i++;
        }
      }
      if (i > 0 && i < count) split_at = start + i;
    } else if (count <= (size_t)(p.max_leaf * p.sah_leaf_mult)) {
      nodes[node_idx].first = (uint32_t)start;
      nodes[node_idx].count = (uint32_t)count;
      return;
    }
  }
  size_t left_idx = nodes.size();
  nodes.push_back(BvhNode{});
  nodes.push_back(BvhNode{});
  nodes[node_idx].first = (uint32_t)left_idx;
  nodes[node_idx].count = 0;
  subdivide(bounds, centroids, p, left_idx, start, split_at - start);
  subdivide(bounds, centroids, p, left_idx + 1, split_at, start + count - split_at);
}

// mesh.rs:48-84 (no deformation endpoint)
void Mesh::build(const float* pos, size_t nverts, const uint32_t* idx, size_t ntris,
                 const BvhParams& params, const float* pos1) {
  positions.assign(pos, pos + 3 * nverts);
  positions1.clear();
  if (pos1) positions1.assign(pos1, pos1 + 3 * nverts);
  indices.assign(idx, idx + 3 * ntris);
  std::vector<Aabb> bounds(ntris);
  local_bounds = Aabb{};
  for (size_t t = 0; t < ntris; t++) {
    Aabb b;
    b.grow_point(vertex(indices[3 * t]));
    b.grow_point(vertex(indices[3 * t + 1]));
    b.grow_point(vertex(indices[3 * t + 2]));
    if (pos1)
      for (int k = 0; k < 3; k++) {
        const float* q = pos1 + 3 * (size_t)indices[3 * t + k];
        b.grow_point(Vec3{(double)q[0], (double)q[1], (double)q[2]});
      }
    bounds[t] = b;
    local_bounds.grow(b);
  }
  blas.build(bounds, params);
}

// mesh.rs:117-142. Operation order is kept so the f64 results are the
// same bits render.rs computes (built with -ffp-contract=off).
bool Mesh::intersect_triangle(const Ray& ray, uint32_t tri, double t_min, double t_max,
                              double* t_out, double* u_out, double* v_out) const {
  uint32_t i0 = indices[3 * tri], i1 = indices[3 * tri + 1], i2 = indices[3 * tri + 2];
  Vec3 v0 = vertex_at(i0, ray.time);
  Vec3 e1 = vertex_at(i1, ray.time) - v0;
  Vec3 e2 = vertex_at(i2, ray.time) - v0;
  Vec3 p = ray.direction.cross(e2);
  double det = e1.dot(p);
  if (std::fabs(det) < 1e-14) return false;
  double inv = 1.0 / det;
  Vec3 tv = ray.origin - v0;
  double u = tv.dot(p) * inv;
  if (!(u >= 0.0 && u <= 1.0)) return false;
  Vec3 q = tv.cross(e1);
  double v = ray.direction.dot(q) * inv;
  if (v < 0.0 || u + v > 1.0) return false;
  double t = e2.dot(q) * inv;
  if (t > t_min && t <= t_max) {
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return true;
  }
  return false;
}

}  // namespace orc
