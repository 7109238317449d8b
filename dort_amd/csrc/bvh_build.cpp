// bvh_build.cpp -- host BVH construction and flattening to the GPU node layout.
//
// Replaces dort's Bvh<Traits> build (src/dort/bvh.cpp:9-416), which splits at the
// centroid middle with 6-triangle leaves.  Here: top-down binned SAH (Wald 2007,
// "On fast construction of SAH-based bounding volume hierarchies") over triangle
// centroids, subtrees built in parallel with std::thread, depth bounded so the
// GPU traversal stack (64 entries) can never overflow: below depth 32 every split
// falls back to an object-median split, which halves the range.  The binary tree
// is then flattened to BVH2 (child boxes stored in the parent) or collapsed to a
// 4- or 8-wide BVH, and triangles are copied in leaf order as 3 x float4.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <future>
#include <limits>
#include <stdexcept>
#include <thread>
#include <vector>

#include "bvh_build.hpp"

namespace dgpu_host {

namespace {

struct Aabb {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const Aabb& b) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
  }
  void grow(const float* p) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
  }
  float area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.f)) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
  }
};

struct BNode {
  Aabb box;
  int32_t left = -1, right = -1;  // -1: leaf
  uint32_t first = 0, count = 0;  // leaf range into the prim index array
};

constexpr float C_ISECT = 1.0f;
constexpr int SAH_DEPTH_LIMIT = 32;
constexpr uint32_t PAR_MIN = 1u << 16;

struct Builder {
  const float* cent;    // 3 per prim
  const Aabb* pbox;
  uint32_t* idx;
  uint32_t max_leaf;
  uint32_t bins;
  int par_levels;
  float c_trav;         // SAH cost of a node visit (C_ISECT = 1 per triangle)

  Aabb range_box(uint32_t b, uint32_t e, Aabb* cbox) const {
    Aabb r, c;
    for (uint32_t i = b; i < e; ++i) { r.grow(pbox[idx[i]]); c.grow(cent + 3 * idx[i]); }
    if (cbox) *cbox = c;
    return r;
  }

  // Returns the split position (b < mid < e) or 0 to make a leaf.
  uint32_t split(uint32_t b, uint32_t e, const Aabb& box, const Aabb& cbox, int depth, bool force) {
    uint32_t n = e - b;
    int axis = 0;
    float ext[3];
    for (int k = 0; k < 3; ++k) ext[k] = cbox.hi[k] - cbox.lo[k];
    if (ext[1] > ext[axis]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    auto median = [&]() -> uint32_t {
      uint32_t mid = b + n / 2;
      std::nth_element(idx + b, idx + mid, idx + e, [&](uint32_t x, uint32_t y) {
        return cent[3 * x + axis] < cent[3 * y + axis];
      });
      return mid;
    };
    if (depth >= SAH_DEPTH_LIMIT) return n <= max_leaf && !force ? 0 : median();
    // binned SAH over the three axes
    float best_cost = INFINITY;
    int best_axis = -1;
    uint32_t best_bin = 0;
    std::vector<uint32_t> cnt(bins);
    std::vector<Aabb> bb(bins);
    std::vector<float> right_area(bins);
    std::vector<uint32_t> right_cnt(bins);
    for (int k = 0; k < 3; ++k) {
      if (!(ext[k] > 0.f)) continue;
      std::fill(cnt.begin(), cnt.end(), 0u);
      std::fill(bb.begin(), bb.end(), Aabb());
      float scale = float(bins) / ext[k];
      for (uint32_t i = b; i < e; ++i) {
        uint32_t p = idx[i];
        int bi = (int)((cent[3 * p + k] - cbox.lo[k]) * scale);
        bi = std::min(std::max(bi, 0), (int)bins - 1);
        cnt[bi]++;
        bb[bi].grow(pbox[p]);
      }
      Aabb acc;
      uint32_t ac = 0;
      for (int i = (int)bins - 1; i > 0; --i) {
        acc.grow(bb[i]); ac += cnt[i];
        right_area[i] = acc.area(); right_cnt[i] = ac;
      }
      acc = Aabb(); ac = 0;
      for (uint32_t i = 0; i + 1 < bins; ++i) {
        acc.grow(bb[i]); ac += cnt[i];
        uint32_t rc = right_cnt[i + 1];
        if (ac == 0 || rc == 0) continue;
        float c = acc.area() * float(ac) + right_area[i + 1] * float(rc);
        if (c < best_cost) { best_cost = c; best_axis = k; best_bin = i; }
      }
    }
    float parent_area = box.area();
    float leaf_cost = C_ISECT * float(n);
    if (best_axis < 0) return (n <= max_leaf && !force) ? 0 : median();
    float split_cost = c_trav + C_ISECT * (parent_area > 0.f ? best_cost / parent_area : float(n));
    if (n <= max_leaf && leaf_cost <= split_cost && !force) return 0;
    int k = best_axis;
    float scale = float(bins) / ext[k];
    uint32_t* mid = std::partition(idx + b, idx + e, [&](uint32_t p) {
      int bi = (int)((cent[3 * p + k] - cbox.lo[k]) * scale);
      bi = std::min(std::max(bi, 0), (int)bins - 1);
      return (uint32_t)bi <= best_bin;
    });
    uint32_t m = (uint32_t)(mid - idx);
    if (m == b || m == e) return median();
    return m;
  }

  // Builds the subtree of [b, e) into `out` (indices local to `out`), returns its root.
  int32_t build(std::vector<BNode>& out, uint32_t b, uint32_t e, const Aabb& box, const Aabb& cbox,
                int depth, uint32_t& max_depth) {
    max_depth = std::max(max_depth, (uint32_t)depth);
    int32_t me = (int32_t)out.size();
    out.push_back(BNode());
    out[me].box = box;
    uint32_t n = e - b;
    uint32_t mid = n >= 2 ? split(b, e, box, cbox, depth, depth == 0) : 0;
    if (mid == 0) {
      out[me].first = b;
      out[me].count = n;
      return me;
    }
    Aabb lc, rc;
    Aabb lb = range_box(b, mid, &lc), rb = range_box(mid, e, &rc);
    if (depth < par_levels && n >= PAR_MIN) {
      std::vector<BNode> lv, rv;
      uint32_t ld = 0, rd = 0;
      auto fut = std::async(std::launch::async, [&]() { build(lv, b, mid, lb, lc, depth + 1, ld); });
      build(rv, mid, e, rb, rc, depth + 1, rd);
      fut.get();
      max_depth = std::max(max_depth, std::max(ld, rd));
      int32_t loff = (int32_t)out.size();
      for (auto nd : lv) {
        if (nd.left >= 0) { nd.left += loff; nd.right += loff; }
        out.push_back(nd);
      }
      int32_t roff = (int32_t)out.size();
      for (auto nd : rv) {
        if (nd.left >= 0) { nd.left += roff; nd.right += roff; }
        out.push_back(nd);
      }
      out[me].left = loff;
      out[me].right = roff;
    } else {
      int32_t l = build(out, b, mid, lb, lc, depth + 1, max_depth);
      int32_t r = build(out, mid, e, rb, rc, depth + 1, max_depth);
      out[me].left = l;
      out[me].right = r;
    }
    return me;
  }
};

void put_tri(std::vector<float>& tris, uint32_t slot, const float* pos, const uint32_t* ind, uint32_t t) {
  float* o = tris.data() + 12 * (size_t)slot;
  for (int k = 0; k < 3; ++k) {
    const float* p = pos + 3 * (size_t)ind[3 * (size_t)t + k];
    o[4 * k + 0] = p[0];
    o[4 * k + 1] = p[1];
    o[4 * k + 2] = p[2];
    o[4 * k + 3] = 0.f;
  }
  std::memcpy(&o[3], &t, 4);  // v0.w = scene triangle id
}

}  // namespace

BuiltBvh build_bvh(const float* positions, uint32_t num_vertices, const uint32_t* indices,
                   uint32_t num_tris, const BvhParams& prm) {
  (void)num_vertices;
  if (num_tris == 0) throw std::invalid_argument("scene has no triangles");
  if (num_tris >= (1u << 27)) throw std::invalid_argument("2^27 triangles or more");
  BuiltBvh out;
  out.width = prm.width == 8 ? 8 : (prm.width == 4 ? 4 : 2);
  uint32_t max_leaf = std::min(std::max(prm.max_leaf_size, 1u), 16u);
  std::vector<float> cent(3 * (size_t)num_tris);
  std::vector<Aabb> pbox(num_tris);
  std::vector<uint32_t> idx(num_tris);
  Aabb root, rootc;
  for (uint32_t t = 0; t < num_tris; ++t) {
    Aabb b;
    for (int k = 0; k < 3; ++k) b.grow(positions + 3 * (size_t)indices[3 * (size_t)t + k]);
    pbox[t] = b;
    for (int k = 0; k < 3; ++k) cent[3 * (size_t)t + k] = 0.5f * (b.lo[k] + b.hi[k]);
    idx[t] = t;
    root.grow(b);
    rootc.grow(&cent[3 * (size_t)t]);
  }
  Builder bld{cent.data(), pbox.data(), idx.data(), max_leaf, std::max(4u, std::min(prm.sah_bins, 64u)), 4,
              prm.node_cost > 0.f ? prm.node_cost : 1.f};
  std::vector<BNode> nodes;
  nodes.reserve(2 * (size_t)num_tris / std::max(1u, max_leaf) + 16);
  uint32_t depth = 0;
  bld.build(nodes, 0, num_tris, root, rootc, 0, depth);
  out.depth = depth;

  // triangles in leaf order (the prim index array order)
  std::vector<float> tris(12 * (size_t)num_tris);
  for (uint32_t i = 0; i < num_tris; ++i) put_tri(tris, i, positions, indices, idx[i]);
  out.tri_leaf_of.assign(num_tris, 0);
  for (uint32_t i = 0; i < num_tris; ++i) out.tri_leaf_of[idx[i]] = i;
  out.tris = std::move(tris);

  auto child_code = [&](int32_t c) -> uint32_t {
    const BNode& n = nodes[c];
    if (n.left < 0) return leaf_code(n.first, n.count);
    return 0;  // patched with the inner node's flat index
  };

  if (nodes[0].left < 0) {
    // a single leaf (one triangle): make an inner root with the leaf twice
    BNode l = nodes[0];
    nodes.clear();
    BNode r;
    r.box = l.box;
    r.left = 1; r.right = 2;
    nodes.push_back(r); nodes.push_back(l); nodes.push_back(l);
  }

  if (out.width == 2) {
    // DFS order of inner nodes
    std::vector<int32_t> flat_of(nodes.size(), -1), order;
    std::vector<int32_t> st{0};
    while (!st.empty()) {
      int32_t c = st.back(); st.pop_back();
      if (nodes[c].left < 0) continue;
      flat_of[c] = (int32_t)order.size();
      order.push_back(c);
      st.push_back(nodes[c].right);
      st.push_back(nodes[c].left);
    }
    out.nodes.assign(16 * order.size(), 0.f);
    uint32_t leaves = 0;
    for (size_t i = 0; i < order.size(); ++i) {
      const BNode& n = nodes[order[i]];
      const Aabb& a = nodes[n.left].box;
      const Aabb& b = nodes[n.right].box;
      float* o = out.nodes.data() + 16 * i;
      const float p = prm.box_pad;
      o[0] = a.lo[0] - p; o[1] = a.hi[0] + p; o[2] = a.lo[1] - p; o[3] = a.hi[1] + p;
      o[4] = b.lo[0] - p; o[5] = b.hi[0] + p; o[6] = b.lo[1] - p; o[7] = b.hi[1] + p;
      o[8] = a.lo[2] - p; o[9] = a.hi[2] + p; o[10] = b.lo[2] - p; o[11] = b.hi[2] + p;
      uint32_t c0 = nodes[n.left].left < 0 ? child_code(n.left) : (uint32_t)flat_of[n.left];
      uint32_t c1 = nodes[n.right].left < 0 ? child_code(n.right) : (uint32_t)flat_of[n.right];
      leaves += (nodes[n.left].left < 0) + (nodes[n.right].left < 0);
      std::memcpy(&o[12], &c0, 4);
      std::memcpy(&o[13], &c1, 4);
    }
    out.num_nodes = (uint32_t)order.size();
    out.num_leaves = leaves;
  } else {
    // collapse to W-wide: open the child with the largest surface area until W children
    const size_t W = out.width;
    struct Wn { int32_t src; std::vector<int32_t> kids; };
    std::vector<Wn> wide;
    std::vector<int32_t> wide_of(nodes.size(), -1);
    std::vector<int32_t> st{0};
    while (!st.empty()) {
      int32_t c = st.back(); st.pop_back();
      Wn w; w.src = c;
      std::vector<int32_t> kids{nodes[c].left, nodes[c].right};
      while (kids.size() < W) {
        int best = -1;
        float best_a = -1.f;
        for (size_t i = 0; i < kids.size(); ++i) {
          if (nodes[kids[i]].left < 0) continue;
          float a = nodes[kids[i]].box.area();
          if (a > best_a) { best_a = a; best = (int)i; }
        }
        if (best < 0) break;
        int32_t k = kids[best];
        kids.erase(kids.begin() + best);
        kids.push_back(nodes[k].left);
        kids.push_back(nodes[k].right);
      }
      w.kids = kids;
      wide_of[c] = (int32_t)wide.size();
      wide.push_back(w);
      for (int i = (int)kids.size() - 1; i >= 0; --i)
        if (nodes[kids[i]].left >= 0) st.push_back(kids[i]);
    }
    // Compressed nodes: the child boxes are quantized to 8 bits per plane in the frame
    // of the node's own box (Ylitie, Karras, Laine 2017, "Efficient incoherent ray
    // traversal on GPUs through compressed wide BVHs"), rounded outward.
    // BVH4, 64 bytes (4 x 16 B):
    //   [0]  origin x, y, z (float), exponents ex | ey << 8 | ez << 16 (int8 each)
    //   [1]  lo.x[4] hi.x[4] lo.y[4] hi.y[4]   (one byte per child, 4 per word)
    //   [2]  lo.z[4] hi.z[4] child[0] child[1]
    //   [3]  child[2] child[3] 0 0
    // BVH8, 96 bytes (6 x 16 B):
    //   [0]  origin, exponents
    //   [1]  lo.x[0..3] lo.x[4..7] hi.x[0..3] hi.x[4..7]
    //   [2]  the same for y
    //   [3]  the same for z
    //   [4]  child[0..3]
    //   [5]  child[4..7]
    // A plane is origin + q * 2^e.  The boxes are grown by box_pad before quantizing.
    // An empty slot has an inverted box (lo 255, hi 0) and empty_child_code.
    const size_t NF = W == 8 ? 24 : 16;  // floats per node
    out.nodes.assign(NF * wide.size(), 0.f);
    uint32_t leaves = 0;
    for (size_t i = 0; i < wide.size(); ++i) {
      float* o = out.nodes.data() + NF * i;
      uint32_t* w = reinterpret_cast<uint32_t*>(o);
      const size_t nk = wide[i].kids.size();
      double lo[8][3], hi[8][3], nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (size_t c = 0; c < nk; ++c) {
        const BNode& n = nodes[wide[i].kids[c]];
        for (int a = 0; a < 3; ++a) {
          // padded child box, rounded outward to float
          lo[c][a] = (double)std::nextafter(n.box.lo[a] - prm.box_pad, -INFINITY);
          hi[c][a] = (double)std::nextafter(n.box.hi[a] + prm.box_pad, INFINITY);
          nlo[a] = std::min(nlo[a], lo[c][a]);
          nhi[a] = std::max(nhi[a], hi[c][a]);
        }
      }
      uint32_t exps = 0;
      int ex[3];
      for (int a = 0; a < 3; ++a) {
        // smallest e with 255 * 2^e >= extent, in [-100, 100]
        double ext = nhi[a] - nlo[a];
        int e = -100;
        while (e < 100 && std::ldexp(255.0, e) < ext) ++e;
        ex[a] = e;
        o[a] = (float)nlo[a];  // a float already: nlo is a min of floats
        exps |= (uint32_t)(uint8_t)(int8_t)e << (8 * a);
      }
      w[3] = exps;
      uint32_t ql_b[3][8], qh_b[3][8];  // quantized planes per axis and child
      const uint32_t empty = empty_child_code(num_tris);
      uint32_t code[8] = {empty, empty, empty, empty, empty, empty, empty, empty};
      for (size_t c = 0; c < W; ++c) {
        for (int a = 0; a < 3; ++a) {
          uint32_t ql = 255, qh = 0;  // empty slot: inverted box
          if (c < nk) {
            // (child - origin) * 2^-e is exact in double unless the subtraction rounds;
            // then one more quantum outward keeps the quantized box around the child
            double sc = std::ldexp(1.0, -ex[a]);
            double dl = lo[c][a] - nlo[a], dh = hi[c][a] - nlo[a];
            double fl = std::floor(dl * sc), ce = std::ceil(dh * sc);
            if (dl + nlo[a] != lo[c][a]) fl -= 1.0;
            if (dh + nlo[a] != hi[c][a]) ce += 1.0;
            ql = (uint32_t)std::min(std::max(fl, 0.0), 255.0);
            qh = (uint32_t)std::min(std::max(ce, 0.0), 255.0);
          }
          ql_b[a][c] = ql;
          qh_b[a][c] = qh;
        }
        if (c < nk) {
          int32_t k = wide[i].kids[c];
          const BNode& n = nodes[k];
          if (n.left < 0) { code[c] = leaf_code(n.first, n.count); ++leaves; }
          else code[c] = (uint32_t)wide_of[k];
        }
      }
      auto pack4 = [](const uint32_t* b) { return b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24); };
      if (W == 4) {
        w[4] = pack4(ql_b[0]); w[5] = pack4(qh_b[0]); w[6] = pack4(ql_b[1]); w[7] = pack4(qh_b[1]);
        w[8] = pack4(ql_b[2]); w[9] = pack4(qh_b[2]); w[10] = code[0]; w[11] = code[1];
        w[12] = code[2]; w[13] = code[3]; w[14] = 0; w[15] = 0;
      } else {
        for (int a = 0; a < 3; ++a) {
          w[4 + 4 * a] = pack4(ql_b[a]);
          w[5 + 4 * a] = pack4(ql_b[a] + 4);
          w[6 + 4 * a] = pack4(qh_b[a]);
          w[7 + 4 * a] = pack4(qh_b[a] + 4);
        }
        for (int c = 0; c < 8; ++c) w[16 + c] = code[c];
      }
    }
    out.num_nodes = (uint32_t)wide.size();
    out.num_leaves = leaves;
    out.tris.resize(out.tris.size() + 12, 0.f);  // the empty-slot triangle: degenerate, never hit
    // wide depth
    std::vector<uint32_t> dep(wide.size(), 0);
    uint32_t md = 0;
    for (size_t i = 0; i < wide.size(); ++i) {
      for (int32_t k : wide[i].kids) {
        if (nodes[k].left >= 0) { dep[wide_of[k]] = dep[i] + 1; md = std::max(md, dep[i] + 1); }
      }
    }
    out.depth = md + 1;
  }
  return out;
}

}  // namespace dgpu_host
