// bvh_build.hpp -- host BVH builder interface (see bvh_build.cpp).
#pragma once
#include <cstdint>
#include <vector>

namespace dgpu_host {

struct BvhParams {
  uint32_t width = 2;          // 2, 4 or 8
  uint32_t max_leaf_size = 4;  // 1..16
  uint32_t sah_bins = 16;
  float node_cost = 1.f;       // SAH: cost of one node visit relative to one triangle test
  float box_pad = 0.f;         // every node box is grown by this much on each side
};

struct BuiltBvh {
  std::vector<float> nodes;          // per node: 16 floats (64 B) for BVH2 and compressed BVH4, 24 (96 B) for BVH8
  std::vector<float> tris;           // 12 floats per triangle, leaf order (BVH4/8: + the empty-slot triangle)
  std::vector<uint32_t> tri_leaf_of; // scene triangle id -> leaf-order slot
  uint32_t width = 2;
  uint32_t depth = 0;
  uint32_t num_nodes = 0;
  uint32_t num_leaves = 0;
};


// Leaf child code: bit 31 set, bits 27..30 = count-1, bits 0..26 = first triangle.
inline uint32_t leaf_code(uint32_t first, uint32_t count) {
  return 0x80000000u | ((count - 1u) << 27) | first;
}

// Child code of an empty slot of a wide node: a leaf holding only triangle slot
// num_tris, a degenerate triangle (all vertices at the origin) that the builder
// appends and that no ray hits.  The slot's box is inverted, so a ray visits it only
// through rounding, and then harmlessly.
inline uint32_t empty_child_code(uint32_t num_tris) { return leaf_code(num_tris, 1); }

BuiltBvh build_bvh(const float* positions, uint32_t num_vertices, const uint32_t* indices,
                   uint32_t num_tris, const BvhParams& prm);

}  // namespace dgpu_host
