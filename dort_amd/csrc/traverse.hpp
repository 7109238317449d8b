// traverse.hpp -- closest-hit / any-hit BVH traversal with the watertight FP32
// ray-triangle test, for one ray per lane.
//
// Replaces dort's Bvh::traverse_elems (src/dort/bvh.cpp:418-469) with
// MeshBvhPrimitive::intersect / intersect_p (src/dort/mesh_bvh_primitive.cpp:13-45)
// and the Moller-Trumbore test of TriangleUv::hit (src/dort/triangle.cpp:102-135).
//
// * Triangle test: Woop, Benthin, Wald, "Watertight ray/triangle intersection",
//   JCGT 2(1) 2013, two-sided like dort, with the FP64 re-evaluation of zero
//   edge functions.  Accepts t in [tmin, tmax]; a later hit at equal t replaces
//   the current one, as dort's `t > t_max` rejection does (triangle.cpp:131).
// * Box test: slabs with (lo - o) * inv_dir and the far distance scaled by
//   1 + 2*gamma(3) (Ize 2013, "Robust BVH ray traversal"), so a box the ray
//   touches is never culled by rounding.
// * Stack: the top `lds_cap` entries of each lane's stack live in LDS
//   (column-major, stride = workgroup size, so lanes hit distinct banks); deeper
//   entries spill to a per-thread slice of global memory.  The builder bounds the
//   depth so the total never exceeds the spill capacity.
#pragma once
#include "dgpu_device.hpp"

namespace dgpu {

struct TraceRay {
  V3 o, d;
  float tmin, tmax;
  uint32_t ignore;  // leaf-order triangle index to skip (0xFFFFFFFF: none)
};

struct TraceHit {
  float t, b1, b2;
  uint32_t tri;     // leaf-order index, 0xFFFFFFFF if none
};

struct LaneStack {
  uint32_t* lds;        // this lane's column in the workgroup's LDS stack
  uint32_t lds_stride;  // workgroup size
  uint32_t lds_cap;     // entries kept in LDS (power of two)
  uint32_t* spill;      // this thread's column in the global spill area
  size_t spill_stride;  // threads in the grid
  uint32_t cap;         // total entries (LDS + spill); the builder's depth bound keeps sp below it
  uint32_t sp;

  DG_INLINE void push(uint32_t v) {
    if (sp >= cap) return;  // unreachable for a well-formed BVH; never write past the spill slice
    uint32_t slot = sp & (lds_cap - 1u);
    if (sp >= lds_cap) spill[(size_t)(sp - lds_cap) * spill_stride] = lds[slot * lds_stride];
    lds[slot * lds_stride] = v;
    ++sp;
  }
  DG_INLINE uint32_t pop() {
    --sp;
    uint32_t slot = sp & (lds_cap - 1u);
    uint32_t v = lds[slot * lds_stride];
    if (sp >= lds_cap) lds[slot * lds_stride] = spill[(size_t)(sp - lds_cap) * spill_stride];
    return v;
  }
};

struct WtRay {
  int kx, ky, kz;
  float sx, sy, sz;
  V3 o;
};

DG_INLINE WtRay wt_setup(V3 o, V3 d) {
  WtRay w;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  float dz = comp(d, kz);
  if (dz < 0.f) { int t = kx; kx = ky; ky = t; }
  w.kx = kx; w.ky = ky; w.kz = kz;
  w.sx = comp(d, kx) / dz;
  w.sy = comp(d, ky) / dz;
  w.sz = 1.0f / dz;
  w.o = o;
  return w;
}

// Watertight test; on a hit inside [tmin, tmax] returns true with t, b1, b2.
// Contraction into FMA is off here: an edge shared by two triangles must give edge
// functions of exactly opposite sign, which needs both products rounded the same way.
DG_INLINE bool wt_hit(const WtRay& w, float4 va, float4 vb, float4 vc, float tmin, float tmax,
                      float& t_out, float& b1_out, float& b2_out) {
#pragma clang fp contract(off)
  V3 A = xyz(va) - w.o, B = xyz(vb) - w.o, C = xyz(vc) - w.o;
  float Az = comp(A, w.kz), Bz = comp(B, w.kz), Cz = comp(C, w.kz);
  float Ax = comp(A, w.kx) - w.sx * Az, Ay = comp(A, w.ky) - w.sy * Az;
  float Bx = comp(B, w.kx) - w.sx * Bz, By = comp(B, w.ky) - w.sy * Bz;
  float Cx = comp(C, w.kx) - w.sx * Cz, Cy = comp(C, w.ky) - w.sy * Cz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  float det = U + V + W;
  if (det == 0.f) return false;
  float T = U * (w.sz * Az) + V * (w.sz * Bz) + W * (w.sz * Cz);
  float t = T / det;
  if (!(t >= tmin && t <= tmax)) return false;
  float inv = 1.0f / det;
  t_out = t;
  b1_out = V * inv;
  b2_out = W * inv;
  return true;
}

constexpr float ROBUST_FAR = 1.0000004f;  // 1 + 2 gamma(3), gamma(n) = n u / (1 - n u), u = 2^-24
// Upper bound on node visits per ray: a guard that ends every traversal even on a
// malformed BVH (a well-formed one visits each node at most once).
constexpr uint32_t MAX_VISITS = 1u << 26;

DG_INLINE float robust_inv(float d) {
  const float ooeps = 8.2718061e-25f;  // 2^-80
  return 1.0f / (fabsf(d) > ooeps ? d : copysignf(ooeps, d));
}

// Slab test of one box against [tmin, tmax]; returns the entry distance or +inf on a miss.
DG_INLINE float box_enter(float lox, float hix, float loy, float hiy, float loz, float hiz,
                          V3 o, V3 inv, float tmin, float tmax) {
  float tx0 = (lox - o.x) * inv.x, tx1 = (hix - o.x) * inv.x;
  float ty0 = (loy - o.y) * inv.y, ty1 = (hiy - o.y) * inv.y;
  float tz0 = (loz - o.z) * inv.z, tz1 = (hiz - o.z) * inv.z;
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  return tn <= tf * ROBUST_FAR ? tn : INFINITY;
}

// Traversal state of one ray.
struct TravState {
  V3 o, inv;
  WtRay w;
  float tmin, tmax;
  uint32_t ignore;  // leaf-order triangle to skip
  uint32_t node;    // next node to visit
  TraceHit h;
};

DG_INLINE void trav_begin(TravState& s, const TraceRay& r, LaneStack& st) {
  s.o = r.o;
  s.inv = v3(robust_inv(r.d.x), robust_inv(r.d.y), robust_inv(r.d.z));
  s.w = wt_setup(r.o, r.d);
  s.tmin = r.tmin;
  s.tmax = r.tmax;
  s.ignore = r.ignore;
  s.node = 0;
  s.h.t = INFINITY; s.h.b1 = 0.f; s.h.b2 = 0.f; s.h.tri = 0xFFFFFFFFu;
  st.sp = 0;
}

// Tests the triangles of a leaf; returns true if (ANY) a hit was found.
template <bool ANY>
DG_INLINE bool leaf_test(const float4* __restrict__ tris, uint32_t code, TravState& s) {
  uint32_t first = leaf_first(code), cnt = leaf_count(code);
  for (uint32_t k = 0; k < cnt; ++k) {
    uint32_t ti = first + k;
    if (ti == s.ignore) continue;
    float4 va = tris[3 * ti], vb = tris[3 * ti + 1], vc = tris[3 * ti + 2];
    float t, b1, b2;
    if (wt_hit(s.w, va, vb, vc, s.tmin, s.tmax, t, b1, b2)) {
      s.h.t = t; s.h.b1 = b1; s.h.b2 = b2; s.h.tri = ti;
      if (ANY) return true;
      s.tmax = t;
    }
  }
  return false;
}

constexpr uint32_t TRAV_END = 0xFFFFFFFFu;  // no node left (the builder keeps leaf codes below it)

DG_INLINE void trav_pop(TravState& s, LaneStack& st) { s.node = st.sp ? st.pop() : TRAV_END; }

// Visits inner node s.node.  BVH2: both child boxes in the parent node (64 B);
// BVH4: 128 B nodes.  s.node becomes the nearest child hit (the others pushed far
// to near), or the next stacked node, or TRAV_END.
template <int WIDTH>
DG_INLINE void trav_inner(const float4* __restrict__ nodes, TravState& s, LaneStack& st) {
  if (WIDTH == 4) {
    const float4* n = nodes + 8 * (size_t)s.node;
    float4 lx = n[0], hx = n[1], ly = n[2], hy = n[3], lz = n[4], hz = n[5], cc = n[6];
    float td[4];
    uint32_t ch[4];
    td[0] = box_enter(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, s.o, s.inv, s.tmin, s.tmax);
    td[1] = box_enter(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, s.o, s.inv, s.tmin, s.tmax);
    td[2] = box_enter(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, s.o, s.inv, s.tmin, s.tmax);
    td[3] = box_enter(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, s.o, s.inv, s.tmin, s.tmax);
    ch[0] = __float_as_uint(cc.x); ch[1] = __float_as_uint(cc.y);
    ch[2] = __float_as_uint(cc.z); ch[3] = __float_as_uint(cc.w);
    // sort the four (distance, child) pairs, misses (inf) last: 5-comparator network
#define DG_CSWAP(a, b)                                                      \
    if (td[b] < td[a]) {                                                    \
      float tt = td[a]; td[a] = td[b]; td[b] = tt;                          \
      uint32_t cx = ch[a]; ch[a] = ch[b]; ch[b] = cx;                       \
    }
    DG_CSWAP(0, 1) DG_CSWAP(2, 3) DG_CSWAP(0, 2) DG_CSWAP(1, 3) DG_CSWAP(1, 2)
#undef DG_CSWAP
    if (td[0] != INFINITY) {
      if (td[3] != INFINITY) st.push(ch[3]);
      if (td[2] != INFINITY) st.push(ch[2]);
      if (td[1] != INFINITY) st.push(ch[1]);
      s.node = ch[0];
      return;
    }
  } else {
    const float4* n = nodes + 4 * (size_t)s.node;
    float4 n0 = n[0], n1 = n[1], n2 = n[2], n3 = n[3];
    float t0 = box_enter(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, s.o, s.inv, s.tmin, s.tmax);
    float t1 = box_enter(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, s.o, s.inv, s.tmin, s.tmax);
    uint32_t c0 = __float_as_uint(n3.x), c1 = __float_as_uint(n3.y);
    bool h0 = t0 != INFINITY, h1 = t1 != INFINITY;
    if (h0 && h1) {
      bool swap = t1 < t0;
      s.node = swap ? c1 : c0;
      st.push(swap ? c0 : c1);
      return;
    }
    if (h0 || h1) {
      s.node = h0 ? c0 : c1;
      return;
    }
  }
  trav_pop(s, st);
}

// Whole traversal of one ray; returns true on a hit (closest) / any hit (ANY).
// One loop, one node (inner or leaf) per iteration.  Measured on MI355X against the
// "while-while" loop and speculative traversal of Aila & Laine 2009 ("Understanding
// the efficiency of ray traversal on GPUs"), and against replacing finished rays
// per lane: this plain loop was the fastest of the four on the terrain workload.
template <bool ANY, int WIDTH>
DG_INLINE bool traverse(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                        const TraceRay& r, LaneStack& st, TraceHit& h) {
  TravState s;
  trav_begin(s, r, st);
  uint32_t visits = 0;
  while (s.node != TRAV_END && visits++ < MAX_VISITS) {
    if (!(s.node & LEAF_BIT)) {
      trav_inner<WIDTH>(nodes, s, st);
    } else {
      if (leaf_test<ANY>(tris, s.node, s)) break;
      trav_pop(s, st);
    }
  }
  h = s.h;
  return h.tri != 0xFFFFFFFFu;
}

}  // namespace dgpu
