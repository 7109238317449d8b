// bvh_layout_check.cpp -- CPU check of the BVH builder's output layout, compiled by
// tests/test_bvh_layout.py with dort_amd/csrc/bvh_build.cpp.
//
// For random and degenerate triangle soups, BVH2 and compressed BVH4, it decodes the
// node arrays independently of the GPU code and checks:
//   * every leaf-order triangle slot is reached exactly once from the root, and every
//     node other than the root is the child of exactly one node;
//   * tri_leaf_of is the inverse of the scene ids stored in v0.w, the vertices of each
//     slot are those of that scene triangle;
//   * every child box (BVH4: decoded from origin + q 2^e) contains every vertex of every
//     triangle below it, grown by the box padding; quantized boxes are no looser than
//     the padded child box plus one quantum on each side;
//   * leaves hold 1..max_leaf triangles, the depth reported is the depth reached;
//   * BVH4 empty slots have inverted boxes and name the degenerate triangle after the
//     scene's triangles.
// Prints "OK <cases>" or the first failure, and exits non-zero on failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh_build.hpp"

using namespace dgpu_host;

namespace {

struct Box {
  double lo[3], hi[3];
};

struct Soup {
  std::vector<float> pos;
  std::vector<uint32_t> ind;
  uint32_t n = 0;
};

Soup random_soup(uint32_t n, float scale, float size, uint32_t seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(-1.f, 1.f);
  Soup s;
  s.n = n;
  s.pos.resize(9 * (size_t)n);
  s.ind.resize(3 * (size_t)n);
  for (uint32_t t = 0; t < n; ++t) {
    float c[3] = {U(rng) * scale, U(rng) * scale, U(rng) * scale};
    for (int v = 0; v < 3; ++v)
      for (int a = 0; a < 3; ++a) s.pos[9 * (size_t)t + 3 * v + a] = c[a] + size * scale * U(rng);
    for (int v = 0; v < 3; ++v) s.ind[3 * (size_t)t + v] = 3 * t + v;
  }
  return s;
}

// A height field of 2 * k * k triangles sharing vertices (flat in y where h = 0).
Soup grid_soup(uint32_t k, float h, uint32_t seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  Soup s;
  for (uint32_t j = 0; j <= k; ++j)
    for (uint32_t i = 0; i <= k; ++i) {
      s.pos.push_back((float)i);
      s.pos.push_back(h * U(rng));
      s.pos.push_back((float)j);
    }
  for (uint32_t j = 0; j < k; ++j)
    for (uint32_t i = 0; i < k; ++i) {
      uint32_t a = j * (k + 1) + i, b = a + 1, c = a + k + 1, d = c + 1;
      uint32_t t[6] = {a, b, d, a, d, c};
      s.ind.insert(s.ind.end(), t, t + 6);
    }
  s.n = (uint32_t)(s.ind.size() / 3);
  return s;
}

struct Checker {
  const Soup& s;
  const BuiltBvh& b;
  BvhParams prm;
  std::vector<int> slot_seen, node_seen;
  uint32_t max_depth = 0;
  std::string err;

  const float* vert(uint32_t slot, int v) const { return &b.tris[12 * (size_t)slot + 4 * v]; }

  void fail(const std::string& m) {
    if (err.empty()) err = m;
  }

  void check_box(const Box& box, uint32_t code, const char* what) {
    // every vertex of every triangle below `code` lies in `box` (boxes already padded)
    std::vector<uint32_t> slots;
    collect(code, slots, 0);
    for (uint32_t sl : slots)
      for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) {
          double x = vert(sl, v)[a];
          if (!(x >= box.lo[a] && x <= box.hi[a])) {
            fail(std::string(what) + ": vertex outside its child box");
            return;
          }
        }
  }

  // slots of the triangles below a child code (no bookkeeping)
  void collect(uint32_t code, std::vector<uint32_t>& out, int depth) {
    if (depth > 200) { fail("cycle"); return; }
    if (code & 0x80000000u) {
      uint32_t f = code & 0x07FFFFFFu, c = ((code >> 27) & 15u) + 1u;
      for (uint32_t k = 0; k < c; ++k) out.push_back(f + k);
      return;
    }
    uint32_t ch[8];
    int nc = children(code, ch, nullptr);
    for (int k = 0; k < nc; ++k) collect(ch[k], out, depth + 1);
  }

  // child codes of node `node` and (optionally) their decoded boxes
  int children(uint32_t node, uint32_t* ch, Box* boxes) {
    if (node >= b.num_nodes) { fail("child index beyond the node array"); return 0; }
    if (b.width == 2) {
      const float* o = &b.nodes[16 * (size_t)node];
      std::memcpy(&ch[0], &o[12], 4);
      std::memcpy(&ch[1], &o[13], 4);
      if (boxes) {
        boxes[0] = Box{{o[0], o[2], o[8]}, {o[1], o[3], o[9]}};
        boxes[1] = Box{{o[4], o[6], o[10]}, {o[5], o[7], o[11]}};
      }
      return 2;
    }
    const int W = (int)b.width;
    const uint32_t* w = node_words(node);
    float org[3];
    std::memcpy(org, w, 12);
    int e[3] = {(int8_t)(w[3] & 0xFF), (int8_t)((w[3] >> 8) & 0xFF), (int8_t)((w[3] >> 16) & 0xFF)};
    int n = 0;
    for (int c = 0; c < W; ++c) {
      uint32_t code = child_code(w, c);
      if (code == empty_child_code(s.n)) {
        // an empty slot: inverted box on every axis
        for (int a = 0; a < 3; ++a)
          if (plane(w, a, c, false) != 255 || plane(w, a, c, true) != 0) fail("empty slot without an inverted box");
        continue;
      }
      ch[n] = code;
      if (boxes)
        for (int a = 0; a < 3; ++a) {
          boxes[n].lo[a] = org[a] + std::ldexp((double)plane(w, a, c, false), e[a]);
          boxes[n].hi[a] = org[a] + std::ldexp((double)plane(w, a, c, true), e[a]);
        }
      ++n;
    }
    if (n == 0) fail("wide node without children");
    return n;
  }

  // the words of wide node `node` (16 floats per BVH4 node, 24 per BVH8 node)
  const uint32_t* node_words(uint32_t node) const {
    return reinterpret_cast<const uint32_t*>(&b.nodes[(b.width == 8 ? 24 : 16) * (size_t)node]);
  }
  // quantized plane of child c on axis a (hi: the upper plane)
  uint32_t plane(const uint32_t* w, int a, int c, bool hi) const {
    if (b.width == 4) return (w[4 + 2 * a + (hi ? 1 : 0)] >> (8 * c)) & 0xFFu;
    return (w[4 + 4 * a + (hi ? 2 : 0) + (c >> 2)] >> (8 * (c & 3))) & 0xFFu;
  }
  uint32_t child_code(const uint32_t* w, int c) const { return b.width == 4 ? w[10 + c] : w[16 + c]; }

  void walk(uint32_t code, uint32_t depth) {
    if (!err.empty()) return;
    if (depth > 200) { fail("cycle"); return; }
    if (code & 0x80000000u) {
      uint32_t f = code & 0x07FFFFFFu, c = ((code >> 27) & 15u) + 1u;
      if (c > prm.max_leaf_size) fail("leaf larger than max_leaf_size");
      for (uint32_t k = 0; k < c; ++k) {
        if (f + k >= s.n) { fail("triangle slot beyond the triangles"); return; }
        slot_seen[f + k]++;
      }
      return;
    }
    node_seen[code]++;
    max_depth = std::max(max_depth, depth);
    uint32_t ch[8];
    Box bx[8];
    int n = children(code, ch, bx);
    for (int k = 0; k < n; ++k) {
      check_box(bx[k], ch[k], b.width == 2 ? "BVH2" : (b.width == 4 ? "BVH4" : "BVH8"));
      walk(ch[k], depth + 1);
    }
  }

  // quantized BVH4 / BVH8 boxes: no looser than the padded box of the triangles below
  // plus one quantum per side (the builder's outward rounding)
  void check_tightness() {
    if (b.width == 2) return;
    for (uint32_t node = 0; node < b.num_nodes && err.empty(); ++node) {
      const uint32_t* w = node_words(node);
      int e[3] = {(int8_t)(w[3] & 0xFF), (int8_t)((w[3] >> 8) & 0xFF), (int8_t)((w[3] >> 16) & 0xFF)};
      uint32_t ch[8];
      Box bx[8];
      int n = children(node, ch, bx);
      for (int k = 0; k < n; ++k) {
        std::vector<uint32_t> slots;
        collect(ch[k], slots, 0);
        for (int a = 0; a < 3; ++a) {
          double lo = INFINITY, hi = -INFINITY;
          for (uint32_t sl : slots)
            for (int v = 0; v < 3; ++v) {
              lo = std::min(lo, (double)vert(sl, v)[a]);
              hi = std::max(hi, (double)vert(sl, v)[a]);
            }
          double q = std::ldexp(1.0, e[a]);
          double slack = 2.0 * q + 4.0 * std::fabs(lo) * 1.2e-7 + prm.box_pad;
          if (bx[k].lo[a] < lo - prm.box_pad - slack || bx[k].hi[a] > hi + prm.box_pad + slack) {
            fail("quantized box looser than one quantum");
            return;
          }
        }
      }
    }
  }

  bool run() {
    slot_seen.assign(s.n, 0);
    node_seen.assign(b.num_nodes, 0);
    // ids and vertices of the leaf-order records
    // BVH4 / BVH8: one more record, the degenerate triangle that empty slots name
    size_t records = s.n + (b.width != 2 ? 1 : 0);
    if (b.tris.size() != 12 * records || b.tri_leaf_of.size() != s.n) fail("triangle arrays have the wrong size");
    if (b.width != 2 && err.empty())
      for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a)
          if (vert(s.n, v)[a] != 0.f) fail("the empty-slot triangle is not degenerate");
    for (uint32_t sl = 0; sl < s.n && err.empty(); ++sl) {
      uint32_t id;
      std::memcpy(&id, &b.tris[12 * (size_t)sl + 3], 4);
      if (id >= s.n || b.tri_leaf_of[id] != sl) { fail("tri_leaf_of is not the inverse of the stored ids"); break; }
      for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a)
          if (vert(sl, v)[a] != s.pos[3 * (size_t)s.ind[3 * (size_t)id + v] + a]) fail("slot vertices differ from the scene's");
    }
    walk(0, 1);
    for (uint32_t sl = 0; sl < s.n && err.empty(); ++sl)
      // (one triangle: the root holds the one leaf twice)
      if (slot_seen[sl] != 1 && s.n > 1) fail("a triangle slot is not reached exactly once");
    for (uint32_t nd = 0; nd < b.num_nodes && err.empty(); ++nd)
      if (node_seen[nd] != 1) fail("a node is not reached exactly once");
    if (b.width != 2 && err.empty() && max_depth != b.depth) fail("reported depth differs from the depth reached");
    check_tightness();
    return err.empty();
  }
};

}  // namespace

// With the argument "threaded" only the threaded-build case runs (the thread-sanitizer
// run of tests/test_sanitizers.py).
int main(int argc, char** argv) {
  const bool threaded_only = argc > 1 && std::strcmp(argv[1], "threaded") == 0;
  struct Case {
    const char* name;
    Soup s;
  };
  std::vector<Case> cases;
  cases.push_back({"one triangle", random_soup(1, 1.f, 0.1f, 1)});
  cases.push_back({"two triangles", random_soup(2, 1.f, 0.1f, 2)});
  cases.push_back({"seven triangles", random_soup(7, 1.f, 0.1f, 3)});
  cases.push_back({"random 20k", random_soup(20000, 1.f, 0.01f, 4)});
  cases.push_back({"random 50k, scale 1000", random_soup(50000, 1000.f, 0.01f, 5)});
  cases.push_back({"random 3k, scale 1e-3", random_soup(3000, 1e-3f, 0.05f, 6)});
  cases.push_back({"flat grid 64", grid_soup(64, 0.f, 7)});
  cases.push_back({"height field 100", grid_soup(100, 5.f, 8)});
  Soup same = random_soup(1, 1.f, 0.1f, 9);
  for (int k = 1; k < 40; ++k) {  // 40 copies of one triangle: no split separates them
    same.pos.insert(same.pos.end(), same.pos.begin(), same.pos.begin() + 9);
    for (int v = 0; v < 3; ++v) same.ind.push_back(3 * k + v);
  }
  same.n = 40;
  cases.push_back({"40 identical triangles", same});
  // large enough that the top subtrees are built on threads of their own (PAR_MIN)
  cases.push_back({"random 300k (threaded build)", random_soup(300000, 10.f, 0.01f, 10)});
  int n = 0;
  for (auto& c : cases) {
    if (threaded_only && c.s.n <= 100000) continue;
    for (uint32_t width : {2u, 4u, 8u})
      for (uint32_t leaf : {1u, 4u, 8u, 16u}) {
        if (c.s.n > 100000 && leaf != 4u) continue;
        BvhParams p;
        p.width = width;
        p.max_leaf_size = leaf;
        float m = 0.f;
        for (float x : c.s.pos) m = std::max(m, std::fabs(x));
        p.box_pad = std::ldexp(m, -21);
        BuiltBvh b = build_bvh(c.s.pos.data(), (uint32_t)(c.s.pos.size() / 3), c.s.ind.data(), c.s.n, p);
        Checker ck{c.s, b, p};
        if (!ck.run()) {
          std::printf("FAIL %s width %u leaf %u: %s\n", c.name, width, leaf, ck.err.c_str());
          return 1;
        }
        ++n;
      }
  }
  std::printf("OK %d\n", n);
  return 0;
}
