// Binned-SAH BVH builder (see bvh_build.h).
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <functional>
#include <cmath>
#include <limits>
#include <thread>

namespace rrhip {
namespace {

struct Box {
  float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                 std::numeric_limits<float>::infinity()};
  float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                 -std::numeric_limits<float>::infinity()};
  void grow(const Box& b) {
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  void grow(const float p[3]) {
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  // Surface area in double: only used to compare split costs.
  double area() const {
    double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
};

struct BNode {
  Box box;
  uint32_t left = 0;   // interior: left child (right child = left + 1)
  uint32_t first = 0;  // leaf: first slot in ids
  uint32_t count = 0;  // 0 = interior
};

class Builder {
 public:
  Builder(const std::vector<Box>& pb, const std::vector<float>& cen, const BuildParams& p)
      : pbox_(pb), cen_(cen), p_(p) {
    size_t n = pb.size();
    ids_.resize(n);
    for (size_t i = 0; i < n; i++) ids_[i] = (uint32_t)i;
    nodes_.resize(std::max<size_t>(1, 2 * n));
    next_node_.store(1);
    int hw = p.threads > 0 ? p.threads : (int)std::thread::hardware_concurrency();
    max_threads_ = std::max(1, hw);
  }

  void run() { subdivide(0, 0, (uint32_t)ids_.size()); }

  std::vector<BNode>& nodes() { return nodes_; }
  std::vector<uint32_t>& ids() { return ids_; }
  uint32_t node_count() const { return next_node_.load(); }

 private:
  void subdivide(uint32_t ni, uint32_t start, uint32_t count) {
    Box nb, cb;
    for (uint32_t k = start; k < start + count; k++) {
      uint32_t q = ids_[k];
      nb.grow(pbox_[q]);
      cb.grow(&cen_[3 * (size_t)q]);
    }
    nodes_[ni].box = nb;
    if (count <= (uint32_t)p_.max_leaf) {
      nodes_[ni].first = start;
      nodes_[ni].count = count;
      return;
    }
    const int B = p_.bins;
    int best_axis = -1, best_bin = 0;
    double best_cost = std::numeric_limits<double>::infinity();
    std::vector<Box> bb(B), racc(B);
    std::vector<uint32_t> bc(B);
    for (int axis = 0; axis < 3; axis++) {
      float cmin = cb.lo[axis], ext = cb.hi[axis] - cb.lo[axis];
      if (!(ext > 0.0f)) continue;
      for (int b = 0; b < B; b++) {
        bb[b] = Box();
        bc[b] = 0;
      }
      const float scale = (float)B / ext;
      for (uint32_t k = start; k < start + count; k++) {
        uint32_t q = ids_[k];
        int b = (int)((cen_[3 * (size_t)q + axis] - cmin) * scale);
        b = std::min(std::max(b, 0), B - 1);
        bc[b]++;
        bb[b].grow(pbox_[q]);
      }
      Box acc;
      for (int b = B - 1; b >= 1; b--) {
        acc.grow(bb[b]);
        racc[b] = acc;
      }
      Box left;
      uint32_t lc = 0;
      for (int b = 0; b < B - 1; b++) {
        left.grow(bb[b]);
        lc += bc[b];
        uint32_t rc = count - lc;
        if (lc == 0 || rc == 0) continue;
        double cost = left.area() * lc + racc[b + 1].area() * rc;
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_bin = b;
        }
      }
    }
    uint32_t mid = 0;
    if (best_axis >= 0) {
      const int axis = best_axis;
      const float cmin = cb.lo[axis], ext = cb.hi[axis] - cb.lo[axis];
      const float scale = (float)B / ext;
      uint32_t* beg = ids_.data() + start;
      uint32_t* split = std::partition(beg, beg + count, [&](uint32_t q) {
        int b = (int)((cen_[3 * (size_t)q + axis] - cmin) * scale);
        b = std::min(std::max(b, 0), B - 1);
        return b <= best_bin;
      });
      mid = (uint32_t)(split - beg);
    }
    if (mid == 0 || mid == count) mid = count / 2;  // coincident centroids: halve

    uint32_t l = next_node_.fetch_add(2);
    nodes_[ni].left = l;
    nodes_[ni].count = 0;
    if (count > 65536 && active_.fetch_add(1) < max_threads_) {
      std::thread th([this, l, start, mid]() { subdivide(l, start, mid); });
      subdivide(l + 1, start + mid, count - mid);
      th.join();
      active_.fetch_sub(1);
    } else {
      if (count > 65536) active_.fetch_sub(1);
      subdivide(l, start, mid);
      subdivide(l + 1, start + mid, count - mid);
    }
  }

  const std::vector<Box>& pbox_;
  const std::vector<float>& cen_;
  BuildParams p_;
  std::vector<uint32_t> ids_;
  std::vector<BNode> nodes_;
  std::atomic<uint32_t> next_node_{1};
  std::atomic<int> active_{1};
  int max_threads_ = 1;
};

struct Flattener {
  const std::vector<BNode>& bn;
  const std::vector<uint32_t>& ids;
  const float* pos;       // triangle build: vertices and indices
  const uint32_t* idx;
  BuiltBvh* out;
  std::vector<uint32_t>* order = nullptr;  // box build: leaf-ordered primitive ids
  int max_depth = 0;
  uint64_t leaves = 0;

  uint32_t leaf_slot() const { return order ? (uint32_t)order->size() : (uint32_t)out->tris.size(); }

  void put_child(Node2& n, int k, uint32_t child, int depth) {
    const BNode& c = bn[child];
    float* xy = k == 0 ? &n.c0_lox : &n.c1_lox;
    xy[0] = c.box.lo[0];
    xy[1] = c.box.hi[0];
    xy[2] = c.box.lo[1];
    xy[3] = c.box.hi[1];
    if (k == 0) {
      n.c0_loz = c.box.lo[2];
      n.c0_hiz = c.box.hi[2];
    } else {
      n.c1_loz = c.box.lo[2];
      n.c1_hiz = c.box.hi[2];
    }
    uint32_t ref, cnt;
    if (c.count > 0) {
      ref = leaf_slot();
      cnt = c.count;
      leaves++;
      for (uint32_t s = c.first; s < c.first + c.count; s++) {
        if (order)
          order->push_back(ids[s]);
        else
          emit_tri(ids[s]);
      }
    } else {
      ref = flatten(child, depth + 1);
      cnt = 0;
    }
    if (k == 0) {
      n.ref0 = ref;
      n.cnt0 = cnt;
    } else {
      n.ref1 = ref;
      n.cnt1 = cnt;
    }
  }

  void emit_tri(uint32_t t) {
    TriRec r{};
    const float* v[3];
    for (int k = 0; k < 3; k++) v[k] = pos + 3 * (size_t)idx[3 * (size_t)t + k];
    r.v0x = v[0][0];
    r.v0y = v[0][1];
    r.v0z = v[0][2];
    r.v1x = v[1][0];
    r.v1y = v[1][1];
    r.v1z = v[1][2];
    r.v2x = v[2][0];
    r.v2y = v[2][1];
    r.v2z = v[2][2];
    r.prim = t;
    out->tris.push_back(r);
  }

  // Emits the child-pair node of interior BNode `b`; returns its index.
  uint32_t flatten(uint32_t b, int depth) {
    max_depth = std::max(max_depth, depth);
    uint32_t me = (uint32_t)out->nodes.size();
    out->nodes.emplace_back();
    Node2 n{};
    put_child(n, 0, bn[b].left, depth);
    put_child(n, 1, bn[b].left + 1, depth);
    out->nodes[me] = n;
    return me;
  }
};

}  // namespace

bool build_triangle_bvh(const float* positions, uint64_t nverts, const uint32_t* indices,
                        uint64_t ntris, const BuildParams& params, BuiltBvh* out,
                        std::string* err, const float* positions1) {
  out->nodes.clear();
  out->tris.clear();
  out->depth = 0;
  out->leaves = 0;
  // Leaf stack entries of the kernels pack (count - 1) in 3 bits and the
  // first triangle slot in 28 bits.
  if (params.bins < 2 || params.bins > 256 || params.max_leaf < 1 || params.max_leaf > 8) {
    *err = "build params out of range (bins 2..256, max_leaf 1..8)";
    return false;
  }
  if (ntris >= (1ull << 28) || nverts >= 0xffffffffull) {
    *err = "mesh too large (at most 2^28 - 1 triangles)";
    return false;
  }
  for (uint64_t i = 0; i < 3 * nverts; i++) {
    if (!std::isfinite(positions[i]) || (positions1 && !std::isfinite(positions1[i]))) {
      *err = "mesh positions must be finite";
      return false;
    }
  }
  std::vector<Box> pb(ntris);
  std::vector<float> cen(3 * ntris);
  Box root;
  for (uint64_t t = 0; t < ntris; t++) {
    Box b;
    for (int k = 0; k < 3; k++) {
      uint32_t vi = indices[3 * t + k];
      if (vi >= nverts) {
        *err = "triangle index out of range";
        return false;
      }
      b.grow(positions + 3 * (size_t)vi);
      if (positions1) b.grow(positions1 + 3 * (size_t)vi);  // boxes cover both shutter ends
    }
    pb[t] = b;
    for (int a = 0; a < 3; a++) cen[3 * t + a] = 0.5f * (b.lo[a] + b.hi[a]);
    root.grow(b);
  }
  for (int a = 0; a < 3; a++) {
    out->root_lo[a] = root.lo[a];
    out->root_hi[a] = root.hi[a];
  }
  if (ntris == 0) return true;

  Builder builder(pb, cen, params);
  builder.run();
  const std::vector<BNode>& bn = builder.nodes();
  out->nodes.reserve(builder.node_count() / 2 + 1);
  out->tris.reserve(ntris);
  Flattener f{bn, builder.ids(), positions, indices, out};
  if (bn[0].count > 0) {
    // A single leaf: both children of the root reference it (a duplicate
    // test of the same triangles is harmless for closest and any hit).
    Node2 n{};
    out->nodes.push_back(n);
    f.put_child(n, 0, 0, 0);
    n.c1_lox = n.c0_lox;
    n.c1_hix = n.c0_hix;
    n.c1_loy = n.c0_loy;
    n.c1_hiy = n.c0_hiy;
    n.c1_loz = n.c0_loz;
    n.c1_hiz = n.c0_hiz;
    n.ref1 = n.ref0;
    n.cnt1 = n.cnt0;
    out->nodes[0] = n;
  } else {
    f.flatten(0, 0);
  }
  out->depth = f.max_depth;
  out->leaves = f.leaves;
  if (out->depth + 1 > kMaxStack) {
    *err = "BVH too deep for the traversal stack";
    return false;
  }
  return true;
}

namespace {

struct Child4 {
  float lo[3], hi[3];
  uint32_t ref, cnt;
  float area() const {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
  }
};

Child4 child_of(const Node2& n, int k) {
  Child4 c;
  if (k == 0) {
    c.lo[0] = n.c0_lox; c.hi[0] = n.c0_hix; c.lo[1] = n.c0_loy; c.hi[1] = n.c0_hiy;
    c.lo[2] = n.c0_loz; c.hi[2] = n.c0_hiz; c.ref = n.ref0; c.cnt = n.cnt0;
  } else {
    c.lo[0] = n.c1_lox; c.hi[0] = n.c1_hix; c.lo[1] = n.c1_loy; c.hi[1] = n.c1_hiy;
    c.lo[2] = n.c1_loz; c.hi[2] = n.c1_hiz; c.ref = n.ref1; c.cnt = n.cnt1;
  }
  return c;
}

struct Collapser {
  const std::vector<Node2>& n2;
  std::vector<Node4>* out;
  int depth = 0;

  uint32_t build(uint32_t i, int level) {
    depth = std::max(depth, level);
    const uint32_t idx = (uint32_t)out->size();
    out->push_back(Node4{});
    Child4 c[4];
    int m = 2;
    c[0] = child_of(n2[i], 0);
    c[1] = child_of(n2[i], 1);
    while (m < 4) {
      int best = -1;
      float best_area = -1.f;
      for (int k = 0; k < m; k++)
        if (c[k].cnt == 0 && c[k].area() > best_area) {
          best = k;
          best_area = c[k].area();
        }
      if (best < 0) break;
      const Node2& open = n2[c[best].ref];
      c[best] = child_of(open, 0);
      c[m++] = child_of(open, 1);
    }
    Node4 nd;
    for (int k = 0; k < 4; k++) {
      if (k < m) {
        nd.lox[k] = c[k].lo[0]; nd.hix[k] = c[k].hi[0];
        nd.loy[k] = c[k].lo[1]; nd.hiy[k] = c[k].hi[1];
        nd.loz[k] = c[k].lo[2]; nd.hiz[k] = c[k].hi[2];
        nd.cnt[k] = c[k].cnt;
        nd.ref[k] = c[k].cnt ? c[k].ref : build(c[k].ref, level + 1);
      } else {
        nd.lox[k] = nd.loy[k] = nd.loz[k] = 0.f;
        nd.hix[k] = nd.hiy[k] = nd.hiz[k] = 0.f;
        nd.cnt[k] = kEmpty4;
        nd.ref[k] = 0;
      }
    }
    (*out)[idx] = nd;
    return idx;
  }
};

}  // namespace

void collapse_bvh4(const std::vector<Node2>& nodes, std::vector<Node4>* out, int* depth4) {
  out->clear();
  *depth4 = 0;
  if (nodes.empty()) return;
  out->reserve(nodes.size() / 2 + 1);
  Collapser c{nodes, out};
  c.build(0, 0);
  *depth4 = c.depth;
}

bool build_box_bvh(const float* boxes, uint64_t n, const BuildParams& params, BuiltBvh* out,
                   std::vector<uint32_t>* order, std::string* err) {
  out->nodes.clear();
  out->tris.clear();
  out->depth = 0;
  out->leaves = 0;
  order->clear();
  if (params.bins < 2 || params.bins > 256 || params.max_leaf < 1 || params.max_leaf > 8) {
    *err = "build params out of range (bins 2..256, max_leaf 1..8)";
    return false;
  }
  if (n >= (1ull << 28)) {
    *err = "too many primitives (at most 2^28 - 1)";
    return false;
  }
  std::vector<Box> pb(n);
  std::vector<float> cen(3 * n);
  Box root;
  for (uint64_t i = 0; i < n; i++) {
    Box b;
    for (int a = 0; a < 3; a++) {
      b.lo[a] = boxes[6 * i + a];
      b.hi[a] = boxes[6 * i + 3 + a];
      if (!std::isfinite(b.lo[a]) || !std::isfinite(b.hi[a]) || b.lo[a] > b.hi[a]) {
        *err = "primitive boxes must be finite with lo <= hi";
        return false;
      }
      cen[3 * i + a] = 0.5f * (b.lo[a] + b.hi[a]);
    }
    pb[i] = b;
    root.grow(b);
  }
  for (int a = 0; a < 3; a++) {
    out->root_lo[a] = root.lo[a];
    out->root_hi[a] = root.hi[a];
  }
  if (n == 0) return true;
  Builder builder(pb, cen, params);
  builder.run();
  const std::vector<BNode>& bn = builder.nodes();
  out->nodes.reserve(builder.node_count() / 2 + 1);
  order->reserve(n);
  Flattener f{bn, builder.ids(), nullptr, nullptr, out};
  f.order = order;
  if (bn[0].count > 0) {
    Node2 nd{};
    out->nodes.push_back(nd);
    f.put_child(nd, 0, 0, 0);
    nd.c1_lox = nd.c0_lox;
    nd.c1_hix = nd.c0_hix;
    nd.c1_loy = nd.c0_loy;
    nd.c1_hiy = nd.c0_hiy;
    nd.c1_loz = nd.c0_loz;
    nd.c1_hiz = nd.c0_hiz;
    nd.ref1 = nd.ref0;
    nd.cnt1 = nd.cnt0;
    out->nodes[0] = nd;
  } else {
    f.flatten(0, 0);
  }
  out->depth = f.max_depth;
  out->leaves = f.leaves;
  if (out->depth + 1 > kMaxStack) {
    *err = "BVH too deep for the traversal stack";
    return false;
  }
  return true;
}

}  // namespace rrhip
