// traverse.hpp -- closest-hit / any-hit BVH traversal with the watertight FP32
// ray-triangle test, for one ray per lane.
//
// Replaces dort's Bvh::traverse_elems (src/dort/bvh.cpp:418-469) with
// MeshBvhPrimitive::intersect / intersect_p (src/dort/mesh_bvh_primitive.cpp:13-45)
// and the Moller-Trumbore test of TriangleUv::hit (src/dort/triangle.cpp:102-135).
//
// * Triangle test: Woop, Benthin, Wald, "Watertight ray/triangle intersection",
//   JCGT 2(1) 2013, two-sided like dort, with the FP64 re-evaluation of zero
//   edge functions and the ray-space transform applied as three FMA rows.  Accepts t in [tmin, tmax]; a later hit at equal t replaces
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

// Per-ray part of the watertight test (Woop et al.): the dominant axis kz of the
// direction, the axes kx, ky that follow it (swapped when d[kz] < 0, which keeps the
// winding), and the shear sx = d[kx] / d[kz], sy = d[ky] / d[kz], sz = 1 / d[kz].  A vertex
// v goes to ray space as x = fma(-sx, A[kz], A[kx]), y = fma(-sy, A[kz], A[ky]),
// z = sz A[kz] with A = v - o.  Every vertex is transformed the same way in every
// triangle that shares it, which is all that watertightness needs.  The three axis
// indices are packed in one word (2 bits each), so the per-ray state is the origin, three
// factors and one word: eight registers where three transform rows took twelve.
struct WtRay {
  V3 o;
  float sx, sy, sz;
  uint32_t k;  // kx | ky << 2 | kz << 4
};

DG_INLINE WtRay wt_setup(V3 o, V3 d) {
  WtRay w;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  float dz = comp(d, kz);
  if (dz < 0.f) { int t = kx; kx = ky; ky = t; }
  w.sx = comp(d, kx) / dz;
  w.sy = comp(d, ky) / dz;
  w.sz = 1.0f / dz;
  w.o = o;
  w.k = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
  return w;
}

// Component k (0, 1, 2) of a, by two selects.
DG_INLINE float comp_sel(V3 a, uint32_t k) { return k == 0u ? a.x : (k == 1u ? a.y : a.z); }

// Watertight test; on a hit inside [tmin, tmax] returns true with t, b1, b2.
// Contraction into FMA is off here: an edge shared by two triangles must give edge
// functions of exactly opposite sign, which needs both products rounded the same way.
// The range test runs first on T against det * [tmin, tmax] (widened by 2^-20), so
// the reciprocal is taken only for candidates; the final test is on t itself.  The
// reciprocal is the hardware's (1 ulp) rather than a correctly rounded division.
DG_INLINE bool wt_hit(const WtRay& w, float4 va, float4 vb, float4 vc, float tmin, float tmax,
                      float& t_out, float& b1_out, float& b2_out) {
#pragma clang fp contract(off)
  V3 A = xyz(va) - w.o, B = xyz(vb) - w.o, C = xyz(vc) - w.o;
  const uint32_t kx = w.k & 3u, ky = (w.k >> 2) & 3u, kz = w.k >> 4;
  const float Akz = comp_sel(A, kz), Bkz = comp_sel(B, kz), Ckz = comp_sel(C, kz);
  float Ax = fmaf(-w.sx, Akz, comp_sel(A, kx)), Ay = fmaf(-w.sy, Akz, comp_sel(A, ky));
  float Bx = fmaf(-w.sx, Bkz, comp_sel(B, kx)), By = fmaf(-w.sy, Bkz, comp_sel(B, ky));
  float Cx = fmaf(-w.sx, Ckz, comp_sel(C, kx)), Cy = fmaf(-w.sy, Ckz, comp_sel(C, ky));
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if (fminf(fminf(U, V), W) < 0.f && fmaxf(fmaxf(U, V), W) > 0.f) return false;  // mixed signs
  float det = U + V + W;
  if (det == 0.f) return false;
  float Az = w.sz * Akz, Bz = w.sz * Bkz, Cz = w.sz * Ckz;
  float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
  uint32_t sgn = __float_as_uint(det) & 0x80000000u;
  float Tn = __uint_as_float(__float_as_uint(T) ^ sgn), dn = fabsf(det);
  if (Tn > tmax * dn * 1.000001f || Tn < tmin * dn * (tmin > 0.f ? 0.999999f : 1.000001f)) return false;
  float inv = __builtin_amdgcn_rcpf(det);  // 1 ulp (v_rcp_f32); t is compared to 1e-5 relative
  float t = T * inv;
  if (!(t >= tmin && t <= tmax)) return false;
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

// Entry distance of the slab intervals [t?0, t?1] (either order) clipped to [tmin, tmax],
// or +inf when empty; the far distance is scaled by 1 + 2 gamma(3).
DG_INLINE float slab_enter(float tx0, float tx1, float ty0, float ty1, float tz0, float tz1, float tmin,
                           float tmax) {
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  return tn <= tf * ROBUST_FAR ? tn : INFINITY;
}

// Slab test of one box against [tmin, tmax]; returns the entry distance or +inf on a miss.
// Exact form (any ray origin): (b - o) * inv, far distance scaled by 1 + 2 gamma(3).
// FAST form: fma(b, inv, -o * inv), one instruction per plane instead of two.  Its
// error in distance along an axis is below 3u M (u = 2^-24) where M bounds |o| and
// |b|; the builder pads every box by M 2^-21 > 3u M for M covering the scene and the
// camera, so for rays that start inside that range (all rays of the path tracer) the
// FAST test never culls a box that the ray touches.  `o` holds -o * inv for FAST.
template <bool FAST>
DG_INLINE float box_enter(float lox, float hix, float loy, float hiy, float loz, float hiz,
                          V3 o, V3 inv, float tmin, float tmax) {
  float tx0, tx1, ty0, ty1, tz0, tz1;
  if (FAST) {
    tx0 = fmaf(lox, inv.x, o.x); tx1 = fmaf(hix, inv.x, o.x);
    ty0 = fmaf(loy, inv.y, o.y); ty1 = fmaf(hiy, inv.y, o.y);
    tz0 = fmaf(loz, inv.z, o.z); tz1 = fmaf(hiz, inv.z, o.z);
  } else {
    tx0 = (lox - o.x) * inv.x; tx1 = (hix - o.x) * inv.x;
    ty0 = (loy - o.y) * inv.y; ty1 = (hiy - o.y) * inv.y;
    tz0 = (loz - o.z) * inv.z; tz1 = (hiz - o.z) * inv.z;
  }
  return slab_enter(tx0, tx1, ty0, ty1, tz0, tz1, tmin, tmax);
}

// Traversal state of one ray.
struct TravState {
  V3 o, inv;        // origin (FAST: -origin * inv), reciprocal direction
  WtRay w;
  float tmin, tmax;
  uint32_t ignore;  // leaf-order triangle to skip
  uint32_t node;    // next node to visit
  uint32_t visits;  // steps taken
  TraceHit h;
};

template <bool FAST>
DG_INLINE void trav_begin(TravState& s, const TraceRay& r, LaneStack& st) {
  s.inv = v3(robust_inv(r.d.x), robust_inv(r.d.y), robust_inv(r.d.z));
  s.o = FAST ? v3(-r.o.x * s.inv.x, -r.o.y * s.inv.y, -r.o.z * s.inv.z) : r.o;
  s.w = wt_setup(r.o, r.d);
  s.tmin = r.tmin;
  s.tmax = r.tmax;
  s.ignore = r.ignore;
  s.node = 0;
  s.visits = 0;
  s.h.t = INFINITY; s.h.b1 = 0.f; s.h.b2 = 0.f; s.h.tri = 0xFFFFFFFFu;
  st.sp = 0;
}

// Tests the first triangle of leaf `code`, whose vertices are va, vb, vc; returns
// true if (ANY) a hit was found.  A leaf is walked one triangle per traversal step
// (see trav_step).
template <bool ANY>
DG_INLINE bool leaf_tri(float4 va, float4 vb, float4 vc, uint32_t code, TravState& s) {
  uint32_t ti = leaf_first(code);
  float t, b1, b2;
  if (ti == s.ignore || !wt_hit(s.w, va, vb, vc, s.tmin, s.tmax, t, b1, b2)) return false;
  s.h.t = t; s.h.b1 = b1; s.h.b2 = b2; s.h.tri = ti;
  if (!ANY) s.tmax = t;
  return ANY;
}

DG_INLINE uint4 f4_bits(float4 v) {
  return make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w));
}

constexpr uint32_t TRAV_END = 0xFFFFFFFFu;  // no node left (the builder keeps leaf codes below it)

DG_INLINE void trav_pop(TravState& s, LaneStack& st) { s.node = st.sp ? st.pop() : TRAV_END; }

DG_INLINE float byte_f(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xFFu); }
// The int8 exponent in byte k of w (|e| <= 100).
DG_INLINE int exp_i8(uint32_t w, int k) { return (int)(int8_t)((w >> (8 * k)) & 0xFFu); }

// Entry distances of the four children of a compressed node (A: origin + exponents,
// Q: lo.x hi.x lo.y hi.y bytes, R: lo.z hi.z bytes), +inf for a child the ray misses.
// A plane is origin + q 2^e:
//  FAST:  t = fma(q, 2^e inv, fma(origin, inv, -o inv)), two instructions per plane;
//         the box padding covers its rounding (see box_enter), so the far distance is
//         not widened;
//  exact: plane = fma(q, 2^e, origin) (one rounding, covered by the box padding), then
//         (plane - o) * inv, far distance scaled by 1 + 2 gamma(3) as in box_enter.
// Along an axis the entry plane is lo where the direction is positive and hi where it
// is negative (t is monotonic in the plane, also after rounding), so each node picks
// its near and far plane words by the direction's sign instead of a min and a max per
// child and axis.  An empty slot's inverted box (lo 255 > hi 0) then has its entry
// beyond its exit on every axis and is missed without a test of its child code.
template <bool FAST>
DG_INLINE void decode_boxes(uint4 A, uint4 Q, uint4 R, const TravState& s, float& d0, float& d1, float& d2,
                            float& d3) {
  V3 org = v3(__uint_as_float(A.x), __uint_as_float(A.y), __uint_as_float(A.z));
  bool nx = s.inv.x < 0.f, ny = s.inv.y < 0.f, nz = s.inv.z < 0.f;
  uint32_t near_x = nx ? Q.y : Q.x, far_x = nx ? Q.x : Q.y;
  uint32_t near_y = ny ? Q.w : Q.z, far_y = ny ? Q.z : Q.w;
  uint32_t near_z = nz ? R.y : R.x, far_z = nz ? R.x : R.y;
  float d[4];
  if (FAST) {
    // 2^e inv, exact (ldexp); inv is at least 1 in magnitude and 2^e at least 2^-100
    V3 si = v3(ldexpf(s.inv.x, exp_i8(A.w, 0)), ldexpf(s.inv.y, exp_i8(A.w, 1)), ldexpf(s.inv.z, exp_i8(A.w, 2)));
    V3 of = v3(fmaf(org.x, s.inv.x, s.o.x), fmaf(org.y, s.inv.y, s.o.y), fmaf(org.z, s.inv.z, s.o.z));
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float tn = fmaxf(fmaxf(fmaf(byte_f(near_x, c), si.x, of.x), fmaf(byte_f(near_y, c), si.y, of.y)),
                       fmaxf(fmaf(byte_f(near_z, c), si.z, of.z), s.tmin));
      float tf = fminf(fminf(fmaf(byte_f(far_x, c), si.x, of.x), fmaf(byte_f(far_y, c), si.y, of.y)),
                       fminf(fmaf(byte_f(far_z, c), si.z, of.z), s.tmax));
      d[c] = tn <= tf ? tn : INFINITY;
    }
  } else {
    V3 sc = v3(ldexpf(1.f, exp_i8(A.w, 0)), ldexpf(1.f, exp_i8(A.w, 1)), ldexpf(1.f, exp_i8(A.w, 2)));
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float tn = fmaxf(fmaxf((fmaf(byte_f(near_x, c), sc.x, org.x) - s.o.x) * s.inv.x,
                             (fmaf(byte_f(near_y, c), sc.y, org.y) - s.o.y) * s.inv.y),
                       fmaxf((fmaf(byte_f(near_z, c), sc.z, org.z) - s.o.z) * s.inv.z, s.tmin));
      float tf = fminf(fminf((fmaf(byte_f(far_x, c), sc.x, org.x) - s.o.x) * s.inv.x,
                             (fmaf(byte_f(far_y, c), sc.y, org.y) - s.o.y) * s.inv.y),
                       fminf((fmaf(byte_f(far_z, c), sc.z, org.z) - s.o.z) * s.inv.z, s.tmax));
      d[c] = tn <= tf * ROBUST_FAR ? tn : INFINITY;
    }
  }
  d0 = d[0]; d1 = d[1]; d2 = d[2]; d3 = d[3];
}

// Entry distances of the eight children of a compressed BVH8 node (A: origin +
// exponents; X, Y, Z: per axis lo[0..3] lo[4..7] hi[0..3] hi[4..7] bytes), +inf for a
// child the ray misses; the planes as in decode_boxes.
template <bool FAST>
DG_INLINE void decode_boxes8(uint4 A, uint4 X, uint4 Y, uint4 Z, const TravState& s, float d[8]) {
  V3 org = v3(__uint_as_float(A.x), __uint_as_float(A.y), __uint_as_float(A.z));
  bool nx = s.inv.x < 0.f, ny = s.inv.y < 0.f, nz = s.inv.z < 0.f;
  const uint32_t near_x[2] = {nx ? X.z : X.x, nx ? X.w : X.y}, far_x[2] = {nx ? X.x : X.z, nx ? X.y : X.w};
  const uint32_t near_y[2] = {ny ? Y.z : Y.x, ny ? Y.w : Y.y}, far_y[2] = {ny ? Y.x : Y.z, ny ? Y.y : Y.w};
  const uint32_t near_z[2] = {nz ? Z.z : Z.x, nz ? Z.w : Z.y}, far_z[2] = {nz ? Z.x : Z.z, nz ? Z.y : Z.w};
  if (FAST) {
    V3 si = v3(ldexpf(s.inv.x, exp_i8(A.w, 0)), ldexpf(s.inv.y, exp_i8(A.w, 1)), ldexpf(s.inv.z, exp_i8(A.w, 2)));
    V3 of = v3(fmaf(org.x, s.inv.x, s.o.x), fmaf(org.y, s.inv.y, s.o.y), fmaf(org.z, s.inv.z, s.o.z));
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int h = c >> 2, k = c & 3;
      float tn = fmaxf(fmaxf(fmaf(byte_f(near_x[h], k), si.x, of.x), fmaf(byte_f(near_y[h], k), si.y, of.y)),
                       fmaxf(fmaf(byte_f(near_z[h], k), si.z, of.z), s.tmin));
      float tf = fminf(fminf(fmaf(byte_f(far_x[h], k), si.x, of.x), fmaf(byte_f(far_y[h], k), si.y, of.y)),
                       fminf(fmaf(byte_f(far_z[h], k), si.z, of.z), s.tmax));
      d[c] = tn <= tf ? tn : INFINITY;
    }
  } else {
    V3 sc = v3(ldexpf(1.f, exp_i8(A.w, 0)), ldexpf(1.f, exp_i8(A.w, 1)), ldexpf(1.f, exp_i8(A.w, 2)));
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int h = c >> 2, k = c & 3;
      float tn = fmaxf(fmaxf((fmaf(byte_f(near_x[h], k), sc.x, org.x) - s.o.x) * s.inv.x,
                             (fmaf(byte_f(near_y[h], k), sc.y, org.y) - s.o.y) * s.inv.y),
                       fmaxf((fmaf(byte_f(near_z[h], k), sc.z, org.z) - s.o.z) * s.inv.z, s.tmin));
      float tf = fminf(fminf((fmaf(byte_f(far_x[h], k), sc.x, org.x) - s.o.x) * s.inv.x,
                             (fmaf(byte_f(far_y[h], k), sc.y, org.y) - s.o.y) * s.inv.y),
                       fminf((fmaf(byte_f(far_z[h], k), sc.z, org.z) - s.o.z) * s.inv.z, s.tmax));
      d[c] = tn <= tf * ROBUST_FAR ? tn : INFINITY;
    }
  }
}

// Orders (distance, child) pairs a and b by distance with selects, no branches.
DG_INLINE void order_pair(float& da, uint32_t& ca, float& db, uint32_t& cb) {
  bool sw = db < da;
  float d0 = sw ? db : da, d1 = sw ? da : db;
  uint32_t c0 = sw ? cb : ca, c1 = sw ? ca : cb;
  da = d0; db = d1; ca = c0; cb = c1;
}

// Visits inner node s.node.  BVH2: both child boxes in the parent node (64 B);
// BVH4: compressed 64 B nodes (A, Q, R, K), loaded by the caller.
// s.node becomes the nearest child hit (the others pushed far to near), or the next
// stacked node, or TRAV_END.  Any-hit rays (ANY) stop at the first hit anywhere, so
// only the nearest child is picked (a three-comparison tournament) and the other hit
// children are pushed unsorted.
// One 16 B word of the BVH or triangle array.  Both are device allocations, and the load
// says so: a batched scene's BVH pointers are read from LDS, and pointers read from
// memory are otherwise generic, which makes every fetch a flat load.
DG_INLINE float4 bvh_word(const float4* base, size_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(1))) float4*)base)[i];
#else
  return base[i];  // the host pass of the same source (host callers read host arrays)
#endif
}

template <int WIDTH, bool ANY, bool FAST>
DG_INLINE void trav_inner(const float4* __restrict__ nodes, TravState& s, LaneStack& st, uint4 A, uint4 Q,
                          uint4 R, uint4 K, uint4 C0 = make_uint4(0, 0, 0, 0), uint4 C1 = make_uint4(0, 0, 0, 0)) {
  if (WIDTH == 8) {
    // compressed 96 B node: A origin, Q / R / K the x / y / z planes, C0 / C1 the child codes
    float d[8];
    decode_boxes8<FAST>(A, Q, R, K, s, d);
    uint32_t c[8] = {C0.x, C0.y, C0.z, C0.w, C1.x, C1.y, C1.z, C1.w};
    // any-hit rays: the nearest child by a tournament; closest-hit rays: a 19-comparator
    // sorting network (Knuth), so the children are pushed far to near
    order_pair(d[0], c[0], d[1], c[1]);
    order_pair(d[2], c[2], d[3], c[3]);
    order_pair(d[4], c[4], d[5], c[5]);
    order_pair(d[6], c[6], d[7], c[7]);
    order_pair(d[0], c[0], d[2], c[2]);
    order_pair(d[4], c[4], d[6], c[6]);
    order_pair(d[0], c[0], d[4], c[4]);
    if (!ANY) {
      // 15 more comparators complete the sort (22 in all; checked on every 0-1 input)
      order_pair(d[1], c[1], d[3], c[3]);
      order_pair(d[5], c[5], d[7], c[7]);
      order_pair(d[1], c[1], d[5], c[5]);
      order_pair(d[2], c[2], d[6], c[6]);
      order_pair(d[3], c[3], d[7], c[7]);
      order_pair(d[2], c[2], d[3], c[3]);
      order_pair(d[4], c[4], d[5], c[5]);
      order_pair(d[6], c[6], d[7], c[7]);
      order_pair(d[2], c[2], d[4], c[4]);
      order_pair(d[3], c[3], d[5], c[5]);
      order_pair(d[1], c[1], d[4], c[4]);
      order_pair(d[3], c[3], d[6], c[6]);
      order_pair(d[1], c[1], d[2], c[2]);
      order_pair(d[3], c[3], d[4], c[4]);
      order_pair(d[5], c[5], d[6], c[6]);
    }
    if (d[0] != INFINITY) {
#pragma unroll
      for (int k = 7; k >= 1; --k)
        if (d[k] != INFINITY) st.push(c[k]);
      s.node = c[0];
      return;
    }
  } else if (WIDTH == 4) {
    // compressed node (64 B): origin, per-axis scale 2^e, 8-bit child planes
    float d0, d1, d2, d3;
    decode_boxes<FAST>(A, Q, R, s, d0, d1, d2, d3);
    uint32_t c0 = R.z, c1 = R.w, c2 = K.x, c3 = K.y;
    // (0,1) (2,3) then the two winners: the nearest is in slot 0
    order_pair(d0, c0, d1, c1);
    order_pair(d2, c2, d3, c3);
    order_pair(d0, c0, d2, c2);
    if (!ANY) {
      // complete the 5-comparator network: slots 1..3 far to near after (1,3), (1,2)
      order_pair(d1, c1, d3, c3);
      order_pair(d1, c1, d2, c2);
    }
    if (d0 != INFINITY) {
      if (d3 != INFINITY) st.push(c3);
      if (d2 != INFINITY) st.push(c2);
      if (d1 != INFINITY) st.push(c1);
      s.node = c0;
      return;
    }
  } else {
    size_t n = 4 * (size_t)s.node;
    float4 n0 = bvh_word(nodes, n), n1 = bvh_word(nodes, n + 1), n2 = bvh_word(nodes, n + 2),
           n3 = bvh_word(nodes, n + 3);
    float t0 = box_enter<FAST>(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, s.o, s.inv, s.tmin, s.tmax);
    float t1 = box_enter<FAST>(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, s.o, s.inv, s.tmin, s.tmax);
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

// One traversal step of one ray: visits the inner node in s.node, or tests ONE
// triangle of the leaf in s.node and leaves the rest of the leaf (first + 1,
// count - 1) in s.node.  Lanes of a wave diverge between inner nodes and leaves in
// nearly every step (each lane is at a leaf in about one step of eight, and there are
// 64 lanes), so a step costs an inner-node visit plus one triangle test instead of
// an inner-node visit plus the longest leaf of the wave.
// s.node becomes TRAV_END when the ray is finished (stack empty, or any-hit found),
// or when it exceeds MAX_VISITS steps.
// BVH4: the first 48 B of a node and a whole triangle are loaded through the same
// three 16 B loads (each lane from its own address, node or triangle), so a wave
// whose lanes hold both kinds issues 4 loads per step instead of 7.  The node's last
// 8 B (child codes 2, 3) are loaded with them, the leaf lanes reading the root's (a
// valid address every wave shares): loaded in the node branch, they are needed right
// after the box test and their latency shows.  The addresses are computed with
// arithmetic: written as a select of two pointers, the compiler makes it a divergent
// branch.
template <bool ANY, int WIDTH, bool FAST>
DG_INLINE void trav_step(const float4* __restrict__ nodes, const float4* __restrict__ tris, TravState& s,
                         LaneStack& st) {
  if (++s.visits > MAX_VISITS) {
    s.node = TRAV_END;
    return;
  }
  bool leaf = (s.node & LEAF_BIT) != 0u;
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, c = a, k = a, e = a, f = a;
  if (WIDTH == 8) {
    // node: 6 x 16 B at 6 * node, the first three shared with a triangle's 3 x 16 B at
    // 3 * first; the node's z planes and child codes (words 3-5) loaded alongside, the
    // leaf lanes reading the root's
    size_t idx = s.node & 0x07FFFFFFu;
    const float4* p = (leaf ? tris : nodes) + (leaf ? 3 * idx : 6 * idx);
    a = bvh_word(p, 0); b = bvh_word(p, 1); c = bvh_word(p, 2);
    const size_t nb = leaf ? 0u : 6 * idx;
    k = bvh_word(nodes, nb + 3u); e = bvh_word(nodes, nb + 4u); f = bvh_word(nodes, nb + 5u);
    asm volatile("" : : "v"(a.w), "v"(b.w), "v"(c.w), "v"(k.x), "v"(k.y), "v"(e.x), "v"(f.x));
  }
  if (WIDTH == 4) {
    // node: 4 x 16 B at 4 * node; triangle: 3 x 16 B at 3 * first (both indices < 2^27)
    size_t idx = s.node & 0x07FFFFFFu;
    const float4* p = (leaf ? tris : nodes) + ((idx << 2) - (leaf ? idx : 0u));
    a = bvh_word(p, 0); b = bvh_word(p, 1); c = bvh_word(p, 2);
    k = bvh_word(nodes, (leaf ? 0u : (idx << 2)) + 3u);
    // The triangle test reads only xyz and no k; without this the compiler moves the
    // node-only words into the node branch as separate loads.  Takes values, no memory
    // access.
    asm volatile("" : : "v"(a.w), "v"(b.w), "v"(c.w), "v"(k.x), "v"(k.y));
  }
  if (!leaf) {
    trav_inner<WIDTH, ANY, FAST>(nodes, s, st, f4_bits(a), f4_bits(b), f4_bits(c), f4_bits(k), f4_bits(e),
                                 f4_bits(f));
    return;
  }
  if (WIDTH == 2) {
    uint32_t ti = leaf_first(s.node);
    a = bvh_word(tris, 3 * (size_t)ti); b = bvh_word(tris, 3 * (size_t)ti + 1); c = bvh_word(tris, 3 * (size_t)ti + 2);
  }
  if (leaf_tri<ANY>(a, b, c, s.node, s)) s.node = TRAV_END;
  else if (leaf_count(s.node) > 1u) s.node += 1u - (1u << 27);  // next triangle: first + 1, count - 1
  else trav_pop(s, st);
}

// Whole traversal of one ray; returns true on a hit (closest) / any hit (ANY).
// One loop, one step (inner node or one leaf triangle) per iteration.  FAST: the ray
// starts inside the range the box padding covers (see box_enter).
template <bool ANY, int WIDTH, bool FAST>
DG_INLINE bool traverse(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                        const TraceRay& r, LaneStack& st, TraceHit& h) {
  TravState s;
  trav_begin<FAST>(s, r, st);
  while (s.node != TRAV_END) trav_step<ANY, WIDTH, FAST>(nodes, tris, s, st);
  h = s.h;
  return h.tri != 0xFFFFFFFFu;
}

}  // namespace dgpu
