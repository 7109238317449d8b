// kernels.hip -- HIP kernels of the MI355X path tracer (gfx950).
//
// Intersection kernels (dgpu_trace_closest / dgpu_trace_any) and the wavefront
// path tracer that replaces PathRenderer::sample (src/dort/path_renderer.cpp:23-206):
//
//   raygen  -> [ extend (closest hit) -> shade -> connect (any hit) ] x (max_depth+1) -> finalize
//
// * extend / connect are persistent kernels: each wave takes 64 queue entries at a
//   time from a global counter until the queue (whose length is on the device) is
//   drained, so the host never reads queue lengths back.
// * shade compacts surviving paths into the next extension queue and NEE rays into
//   the shadow queue with a wave ballot + mbcnt prefix and one atomic per wave.
// * dort's NEE-through-BSDF ray (path_renderer.cpp:149-182) is a closest-hit ray
//   that counts only if it hits the chosen light; here the light triangle is
//   intersected analytically in shade and the ray becomes an any-hit ray on
//   [eps, t_light] that ignores the light itself (for background lights: [eps, inf)).
//   Both give the same estimate except at exact ties.
// * NEE contributions are staged per path (pend) and added by the path's owner
//   thread once the visibility is known, so per-pixel sums are deterministic.
#include <algorithm>

#include "dgpu_internal.hpp"
#include "light_paths.hpp"
#include "traverse.hpp"
#include "wave_ops.hpp"

namespace dgpu {

// ----------------------------------------------------------------- trace kernels

struct BvhRef {
  const float4* nodes;
  const float4* tris;
};

// FAST: rays of the path tracer (origins inside the scene's padded range, see box_enter);
// the intersection API takes rays from anywhere and uses the exact box test.
template <bool ANY, int WIDTH, bool FAST>
__device__ __forceinline__ bool trace_one(BvhRef b, const TraceRay& r, LaneStack& st, TraceHit& h) {
  return traverse<ANY, WIDTH, FAST>(b.nodes, b.tris, r, st, h);
}

// Trace kernels ask for 8 waves per SIMD (at most 64 VGPRs): traversal is bound by
// instruction issue and memory latency, and the eighth wave measured 2-3% faster than
// seven, also where a few registers spill.  BVH8 traversal keeps twice the boxes and a
// longer sort in registers: at 8 waves it spilled 80 bytes a lane and ran 20% slower on
// the 2M-triangle terrain than at 6 waves (80 VGPRs, no spill); 5 waves was no faster.
#ifndef DGPU_TRACE_WAVES_W8
#define DGPU_TRACE_WAVES_W8 6
#endif
#define DG_TRACE_WAVES(W) __attribute__((amdgpu_waves_per_eu((W) == 8 ? DGPU_TRACE_WAVES_W8 : 8)))

// The wave's batch, traced 64 rays at a time: each part of the batch runs until its
// longest ray ends.
template <bool ANY, int WIDTH, bool FAST, uint32_t NBATCH, class IO>
__device__ __forceinline__ void trace_batch_parts(IO& io, uint32_t n, uint32_t base, LaneStack& st) {
  for (uint32_t j = 0; j < NBATCH; ++j) {
    uint32_t i = base + 64u * j + lane_id();
    if (i < n) {
      TraceRay r;
      BvhRef bvh = io.load(i, r);
      TraceHit h;
      bool found = trace_one<ANY, WIDTH, FAST>(bvh, r, st, h);
      io.store(i, h, found);
    }
  }
}

// The wave's batch with refills: a lane whose ray has ended takes the next entry of the
// batch (a neighbouring ray of the queue) once at least REFILL lanes of the wave are
// idle, so the wave no longer runs at the pace of its longest ray with the other lanes
// masked off.  Each refill hands the idle lanes consecutive entries in lane order.
template <bool ANY, int WIDTH, bool FAST, uint32_t NBATCH, class IO>
__device__ __forceinline__ void trace_batch_refill(IO& io, uint32_t n, uint32_t base, LaneStack& st,
                                                   uint32_t refill) {
  const uint32_t lane = lane_id();
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64u - lane));
  const uint32_t end = min(base + 64u * NBATCH, n);
  uint32_t next = base + 64u;  // first entry of the batch that no lane has taken
  uint32_t i = base + lane;
  bool busy = i < end;
  TravState s;
  BvhRef bvh{nullptr, nullptr};
  if (busy) {
    TraceRay r;
    bvh = io.load(i, r);
    trav_begin<FAST>(s, r, st);
  }
  while (__ballot(busy) != 0ull) {
    if (busy) {
      trav_step<ANY, WIDTH, FAST>(bvh.nodes, bvh.tris, s, st);
      if (s.node == TRAV_END) {
        io.store(i, s.h, s.h.tri != 0xFFFFFFFFu);
        busy = false;
      }
    }
    const unsigned long long idle = __ballot(!busy);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (next < end && n_idle >= refill) {  // uniform across the wave
      if (!busy) {
        uint32_t j = next + (uint32_t)__popcll(idle & below);
        if (j < end) {
          i = j;
          busy = true;
          TraceRay r;
          bvh = io.load(i, r);
          trav_begin<FAST>(s, r, st);
        }
      }
      next += n_idle;
    }
  }
}

// Shared persistent loop: `io.load(i, ray)` returns the ray's BVH, `io.store(i, hit, found)`.
// Each wave takes NBATCH x 64 consecutive queue entries per atomic (the fetch counter is
// one address for the whole chip).  Consecutive entries are neighbouring pixels (queues
// are compacted in order), so the lanes of a wave fetch largely the same nodes and
// triangles.  Any-hit rays refill idle lanes from the wave's batch (DGPU_REFILL_ANY):
// shadow rays end at their first hit, so a wave's rays differ most in length, and
// refilling cut the terrain render's connect time by a fifth.  The path tracer's
// closest-hit rays are refilled after the first bounce (DGPU_REFILL_LATE), which cut
// extend by an eighth; the camera rays of bounce 0 trace each 64-ray part to its end
// (refilled, they are coherent enough that mixing traversal stages cost more than the
// idle lanes).  The intersection API's closest-hit rays refill too (DGPU_REFILL_API):
// +14% on its 16M-ray workload, half of whose rays start anywhere in random directions.
// The threshold is a kernel argument.  Refilling from the global queue instead (Aila & Laine 2009, dynamic fetch)
// was 1.2-3.8x slower: the refilled lanes hold rays from distant parts of the queue and
// that coherence is lost.  Giving each XCD (blockIdx % 8) its own contiguous eighth of
// the queue was also slower (7%): with one queue front the whole chip works on one
// window of neighbouring rays, which the shared 256 MB Infinity Cache serves.
template <bool ANY, int WIDTH, int LDS, int BLOCK, bool FAST, class IO>
__device__ __forceinline__ void trace_loop(IO& io, uint32_t n, uint32_t* fetch, uint32_t* spill,
                                           uint32_t stack_cap, uint32_t refill) {
  constexpr uint32_t NBATCH = ANY ? TRACE_BATCH_ANY : TRACE_BATCH;
  __shared__ uint32_t stk[LDS * BLOCK];
  LaneStack st;
  st.lds = stk + threadIdx.x;
  st.lds_stride = BLOCK;
  st.lds_cap = LDS;
  size_t gthreads = (size_t)gridDim.x * BLOCK;
  st.spill = spill + (size_t)blockIdx.x * BLOCK + threadIdx.x;
  st.spill_stride = gthreads;
  st.cap = stack_cap;
  // Atomics on the fetch counter (one address for the chip) serialise in L2 at ~90M/s:
  // the last one of every wave, which only finds the queue drained, cost a launch ~90 us
  // at its end.  The first wave of a workgroup to find the queue drained says so in LDS,
  // and the workgroup's other waves leave without their atomic; an empty queue costs
  // none.
  __shared__ uint32_t drained;
  if (threadIdx.x == 0) drained = 0u;
  __syncthreads();
  if (n == 0u) return;
  for (;;) {
    uint32_t base = n;
    if (lane_id() == 0 && *(volatile uint32_t*)&drained == 0u) {
      base = atomicAdd(fetch, 64u * NBATCH);
      if (base >= n) *(volatile uint32_t*)&drained = 1u;
    }
    base = __shfl(base, 0);
    if (base >= n) break;
    if (refill != 0u) trace_batch_refill<ANY, WIDTH, FAST, NBATCH>(io, n, base, st, refill);
    else trace_batch_parts<ANY, WIDTH, FAST, NBATCH>(io, n, base, st);
  }
}

struct ApiIO {
  dgpu_rays rays;
  dgpu_hits hits;
  uint32_t* occluded;
  BvhRef bvh;
  __device__ __forceinline__ BvhRef load(uint32_t i, TraceRay& r) const {
    r.o = v3(rays.ox[i], rays.oy[i], rays.oz[i]);
    r.d = v3(rays.dx[i], rays.dy[i], rays.dz[i]);
    r.tmin = rays.tmin[i];
    r.tmax = rays.tmax[i];
    r.ignore = 0xFFFFFFFFu;
    return bvh;
  }
  __device__ __forceinline__ void store(uint32_t i, const TraceHit& h, bool found) const {
    if (occluded) { occluded[i] = found ? 1u : 0u; return; }
    hits.t[i] = found ? h.t : INFINITY;
    hits.prim[i] = found ? (int32_t)__float_as_uint(bvh.tris[3 * h.tri].w) : -1;
    hits.b1[i] = found ? h.b1 : 0.f;
    hits.b2[i] = found ? h.b2 : 0.f;
  }
};

template <bool ANY, int WIDTH, int LDS, int BLOCK, bool FAST>
__global__ __launch_bounds__(BLOCK) DG_TRACE_WAVES(WIDTH) void k_trace_api(ApiIO io, uint32_t n, uint32_t* fetch,
                                                     uint32_t* spill, uint32_t stack_cap) {
  trace_loop<ANY, WIDTH, LDS, BLOCK, FAST>(io, n, fetch, spill, stack_cap, ANY ? DGPU_REFILL_ANY : DGPU_REFILL_API);
}

// Batched queues carry chunk path indices, like a single render's: a path's scene is the
// chunk segment holding it.  The segments of the current chunk (their first path and
// scene) sit in LDS, and a binary search over the starts finds a path's segment.
struct SegTable {
  const uint32_t* begin;  // first path of each segment (ascending; begin[0] = 0)
  const BvhRef* bvh;      // trace kernels: the segment's scene BVH
  uint32_t n;
};

// The segment holding chunk path p: the last one whose first path is <= p.
__device__ __forceinline__ uint32_t seg_of(const uint32_t* begin, uint32_t n, uint32_t p) {
  uint32_t b = 0;
#pragma unroll
  for (uint32_t step = MAX_CHUNK_SEGS / 2u; step > 0u; step >>= 1)
    if (b + step < n && begin[b + step] <= p) b += step;
  return b;
}

// The BVH of the segment holding chunk path p.
__device__ __forceinline__ BvhRef seg_bvh_of(const SegTable& t, uint32_t p) {
  return t.bvh[seg_of(t.begin, t.n, p)];
}

template <bool BATCH>
struct ExtendIO {
  const float4* qo;
  const float4* qd;
  const uint32_t* qpath;
  float4* hit;
  BvhRef bvh;            // single scene
  SegTable segs;         // batch: the chunk's segments (in LDS)
  __device__ __forceinline__ BvhRef load(uint32_t i, TraceRay& r) const {
    float4 o = qo[i], d = qd[i];
    r.o = xyz(o); r.tmin = o.w;
    r.d = xyz(d); r.tmax = d.w;
    r.ignore = 0xFFFFFFFFu;
    if (BATCH) return seg_bvh_of(segs, qpath[i]);
    return bvh;
  }
  __device__ __forceinline__ void store(uint32_t i, const TraceHit& h, bool found) const {
    hit[i] = make_float4(found ? h.t : INFINITY, h.b1, h.b2, __uint_as_float(found ? h.tri : 0xFFFFFFFFu));
  }
};

template <bool BATCH>
struct ConnectIO {
  const float4* so;
  const float4* sd;
  const uint2* meta;   // (pending slot, ignored triangle); batch: see load()
  float* pend_w;        // .w of the pending NEE record: 1 = visible so far
  BvhRef bvh;
  SegTable segs;        // batch: the chunk's segments (in LDS)
  __device__ __forceinline__ BvhRef load(uint32_t i, TraceRay& r) const {
    float4 o = so[i], d = sd[i];
    r.o = xyz(o); r.tmin = o.w;
    r.d = xyz(d); r.tmax = d.w;
    uint2 m = meta[i];
    if (BATCH) {
      // batch: (path, ignored triangle | NEE slot k << 31); a leaf-order index is below
      // 2^27, so "none" (0x7FFFFFFF after the mask) matches no triangle
      r.ignore = m.y & 0x7FFFFFFFu;
      return seg_bvh_of(segs, m.x);
    }
    r.ignore = m.y;
    return bvh;
  }
  __device__ __forceinline__ void store(uint32_t i, const TraceHit&, bool found) const {
    if (!found) return;
    uint2 m = meta[i];
    size_t slot = BATCH ? 2 * (size_t)m.x + (m.y >> 31) : m.x;
    pend_w[4 * slot] = 0.f;
  }
};

// The current chunk's segments into the workgroup's LDS: first paths and scene BVHs.
template <bool BATCH, int BLOCK>
__device__ __forceinline__ SegTable load_seg_table(const BatchEntry* batch, const ChunkDesc* chunks,
                                                   const uint32_t* chunk_cur, uint32_t* begin, BvhRef* bvh) {
  SegTable t{nullptr, nullptr, 0u};
  if (!BATCH) return t;
  const ChunkDesc& cd = chunks[*chunk_cur];
  t.n = cd.n_segs;
  for (uint32_t k = threadIdx.x; k < t.n; k += BLOCK) {
    const ChunkSeg& sg = cd.segs[k];
    begin[k] = sg.slot_begin;
    bvh[k].nodes = batch[sg.scene].S.nodes;
    bvh[k].tris = batch[sg.scene].S.tris;
  }
  __syncthreads();
  t.begin = begin;
  t.bvh = bvh;
  return t;
}

template <int WIDTH, int LDS, int BLOCK, bool BATCH>
__global__ __launch_bounds__(BLOCK) DG_TRACE_WAVES(WIDTH) void k_extend(ExtendIO<BATCH> io, const uint32_t* count,
                                                  uint32_t* fetch, uint32_t* spill, uint32_t stack_cap,
                                                  const BatchEntry* batch, const ChunkDesc* chunks,
                                                  const uint32_t* chunk_cur, uint32_t refill) {
  __shared__ uint32_t seg_begin[BATCH ? MAX_CHUNK_SEGS : 1];
  __shared__ BvhRef seg_bvh[BATCH ? MAX_CHUNK_SEGS : 1];
  io.segs = load_seg_table<BATCH, BLOCK>(batch, chunks, chunk_cur, seg_begin, seg_bvh);
  trace_loop<false, WIDTH, LDS, BLOCK, true>(io, *count, fetch, spill, stack_cap, refill);
}

template <int WIDTH, int LDS, int BLOCK, bool BATCH>
__global__ __launch_bounds__(BLOCK) DG_TRACE_WAVES(WIDTH) void k_connect(ConnectIO<BATCH> io, const uint32_t* count,
                                                   uint32_t* fetch, uint32_t* spill, uint32_t stack_cap,
                                                   const BatchEntry* batch, const ChunkDesc* chunks,
                                                   const uint32_t* chunk_cur) {
  __shared__ uint32_t seg_begin[BATCH ? MAX_CHUNK_SEGS : 1];
  __shared__ BvhRef seg_bvh[BATCH ? MAX_CHUNK_SEGS : 1];
  io.segs = load_seg_table<BATCH, BLOCK>(batch, chunks, chunk_cur, seg_begin, seg_bvh);
  trace_loop<true, WIDTH, LDS, BLOCK, true>(io, *count, fetch, spill, stack_cap, DGPU_REFILL_ANY);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Advances the batch's current chunk; the last node of the captured bounce-loop graph.
__global__ void k_advance(uint32_t* cur) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *cur += 1u;
}

// ----------------------------------------------------------------- path state

// thr.w holds flags: bits 0..7 bounces, bit 8 last bounce delta, bit 9 no film weight,
// the rest zero.
constexpr uint32_t F_BOUNCES = 0xFFu;
constexpr uint32_t F_DELTA = 1u << 8;

// Chunk segment holding batched slot i (binary search over the segment starts).
__device__ __forceinline__ const ChunkSeg& find_seg_slot(const ChunkDesc& cd, uint32_t i) {
  uint32_t b = 0, e = cd.n_segs;
  while (b + 1u < e) {
    uint32_t m = (b + e) / 2u;
    if (cd.segs[m].slot_begin <= i) b = m; else e = m;
  }
  return cd.segs[b];
}

template <bool BATCH>
__global__ void k_raygen(WaveParams P) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t n_chunk = BATCH ? P.chunks[*P.chunk_cur].n_chunk : P.n_chunk;
  if (i < NUM_COUNTERS) P.counters[i] = (i == CNT_QUEUE) ? n_chunk : 0u;
  if (i >= n_chunk) return;
  uint32_t sample, pixel, scene = 0, xr, yr, k0, k1;
  const DevScene* S;
  if (BATCH) {
    const ChunkDesc& cd = P.chunks[*P.chunk_cur];
    const ChunkSeg& sg = find_seg_slot(cd, i);
    scene = sg.scene;
    const BatchEntry& E = P.batch[scene];
    order_sample_pixel(E.ord, sg.s_begin, sg.q_begin + (i - sg.slot_begin), sample, pixel);
    xr = E.xr; yr = E.yr; k0 = E.k0; k1 = E.k1;
    S = &E.S;
  } else {
    order_sample_pixel(P.ord, P.s0, P.q0 + i, sample, pixel);
    xr = P.xr; yr = P.yr; k0 = P.k0; k1 = P.k1;
    S = &P.S;
  }
  uint32_t x = pixel % xr, y = pixel / xr;
  float jx = 0.5f, jy = 0.5f;
  if (sample != 0) {  // iteration 0 samples pixel centres (path_renderer.cpp:17)
    float r[4];
    dgpu_rng_block(k0, k1, pixel, sample, 0, 2, r);
    jx = r[2]; jy = r[3];
  }
  V3 o, d;
  camera_ray(*S, float(x) + jx, float(y) + jy, xr, yr, o, d);
  // only the queue: bounce-0 shade starts the path state (throughput 1, radiance 0,
  // pending NEE) without reading it, and finalize recomputes the jitter's edge test
  P.qo[0][i] = make_float4(o.x, o.y, o.z, 0.f);
  P.qd[0][i] = make_float4(d.x, d.y, d.z, INFINITY);
  P.qpath[0][i] = i;
}

// A jitter of exactly 0 puts iteration `sample` (> 0) on its pixel's lower edge; dort's
// film drops such samples (the jitter is k_raygen's: block 2 of bounce 0, words 2, 3).
__device__ __forceinline__ bool jitter_on_edge(uint32_t k0, uint32_t k1, uint32_t pixel, uint32_t sample) {
  if (sample == 0u) return false;
  float r[4];
  dgpu_rng_block(k0, k1, pixel, sample, 0, 2, r);
  return r[2] == 0.f || r[3] == 0.f;
}

struct NeeOut {
  bool emit;
  V3 o, d;
  float tmin, tmax;
  uint32_t ignore;
  V3 contrib;
};

// Light::sample_pivot_radiance for the three light kinds (diffuse_light.cpp:40-65 +
// shape.cpp:4-12 + triangle.cpp:62-76; environment_light.cpp:90-102; infinite_light.cpp:25-36).
__device__ __forceinline__ V3 light_sample(const DevScene& S, const DevLight& L, V3 p, float eps,
                                           float u, float v, V3& wi, float& pdf, NeeOut& sh) {
  V3 zero = v3(0.f, 0.f, 0.f);
  if (L.kind == LK_AREA) {
    float4 a = S.tris[3 * L.tri], b = S.tris[3 * L.tri + 1], c = S.tris[3 * L.tri + 2];
    V3 p0 = xyz(a), e1 = xyz(b) - p0, e2 = xyz(c) - p0;
    float s = sqrtf(u);
    float b1 = 1.f - s, b2 = v * s;
    V3 n = normalize(cross(e2, e1));
    V3 pt = p0 + e1 * b1 + e2 * b2;
    pdf = (1.f / L.area) * length_sq(p - pt) / fabsf(dot(n, normalize(p - pt)));
    V3 sw = normalize(pt - p);
    wi = sw;
    float dist = length(pt - p);
    sh.o = p; sh.d = (pt - p) / dist; sh.tmin = eps; sh.tmax = dist - 1e-3f;
    sh.ignore = 0xFFFFFFFFu;
    return dot(sw, n) < 0.f ? xyz(L.radiance) : zero;
  } else if (L.kind == LK_ENV) {
    V3 rad;
    wi = env_sample(S, u, v, pdf, rad);
    sh.o = p; sh.d = normalize(wi); sh.tmin = eps; sh.tmax = INFINITY; sh.ignore = 0xFFFFFFFFu;
    return rad;
  } else {
    float z = 1.f - 2.f * u, r = sqrtf(fmaxf(0.f, 1.f - z * z)), phi = TWO_PI_F * v;
    float sphi, cphi;
    sincos_small(phi, sphi, cphi);
    wi = v3(cphi * r, sphi * r, z);
    pdf = INV_FOUR_PI_F;
    sh.o = p; sh.d = normalize(wi); sh.tmin = eps; sh.tmax = INFINITY; sh.ignore = 0xFFFFFFFFu;
    return xyz(L.radiance);
  }
}

// estimate_direct_from_bsdf (path_renderer.cpp:149-182) for light L: samples the BSDF
// with (u, v, c) and, if the direction reaches L, returns the shadow ray and the
// contribution it adds when unoccluded.  use_light: MIS against light sampling.
template <bool ROUGH = true>
__device__ __forceinline__ NeeOut nee_from_bsdf(const DevScene& S, const Surface& g, V3 wo, const Bsdf& bsdf,
                                                const DevLight& L, float scale, V3 thr, bool use_light,
                                                float u, float v, float c) {
  const uint32_t NON_DELTA = BSDF_ALL & ~BSDF_DELTA;
  NeeOut nb;
  nb.emit = false;
  V3 wi; float pdf; uint32_t fl;
  V3 f = bsdf_sample<true, ROUGH>(bsdf, wo, NON_DELTA, u, v, c, wi, pdf, fl);
  if (pdf == 0.f || is_black(f)) return nb;
  V3 rad = v3(0.f, 0.f, 0.f);
  float tmax = INFINITY, lp = 0.f;
  uint32_t ignore = 0xFFFFFFFFu;
  if (L.kind == LK_AREA) {
    float tl;
    if (light_tri_hit(S, L.tri, g.p, wi, tl) && tl >= g.eps) {
      V3 nl_ = light_tri_normal(S, L.tri);
      if (dot(nl_, -wi) >= 0.f) rad = xyz(L.radiance);  // eval_radiance, diffuse_light.cpp:83-93
      lp = (1.f / L.area) * (tl * tl * length_sq(wi)) / fabsf(dot(wi, nl_));
      tmax = tl;
      ignore = L.tri;
    }
  } else if (L.kind == LK_ENV) {
    rad = env_eval(S, wi, lp);
  } else {
    rad = xyz(L.radiance);
    lp = INV_FOUR_PI_F;
  }
  if (is_black(rad)) return nb;
  float w = use_light ? pdf / (pdf + lp) : 1.f;
  nb.emit = true;
  nb.o = g.p; nb.d = wi; nb.tmin = g.eps; nb.tmax = tmax; nb.ignore = ignore;
  nb.contrib = thr * f * rad * (w * fabsf(dot(g.nn, wi)) / pdf * scale);
  return nb;
}

// estimate_direct_from_light (path_renderer.cpp:184-206) for light L with the light
// sample (u, v).  use_bsdf: MIS against BSDF sampling.
template <bool ROUGH = true>
__device__ __forceinline__ NeeOut nee_from_light(const DevScene& S, const Surface& g, V3 wo, const Bsdf& bsdf,
                                                 const DevLight& L, float scale, V3 thr, bool use_bsdf,
                                                 float u, float v) {
  const uint32_t NON_DELTA = BSDF_ALL & ~BSDF_DELTA;
  NeeOut nl;
  nl.emit = false;
  V3 wi; float pdf;
  V3 rad = light_sample(S, L, g.p, g.eps, u, v, wi, pdf, nl);
  if (pdf == 0.f || is_black(rad)) return nl;
  V3 f = bsdf_eval<ROUGH>(bsdf, wi, wo, NON_DELTA);
  if (is_black(f)) return nl;
  float w = 1.f;
  if (use_bsdf) {
    float bp = bsdf_pdf(bsdf, wi, wo, NON_DELTA);
    w = pdf / (pdf + bp);
  }
  nl.emit = true;
  nl.contrib = thr * f * rad * (w * fabsf(dot(g.nn, wi)) / pdf * scale);
  return nl;
}

// estimate_direct (path_renderer.cpp:135-206) for light li: up to two shadow rays.
__device__ __forceinline__ void estimate_direct(const DevScene& S, const WaveParams& P, const Surface& g,
                                                V3 wo, const Bsdf& bsdf, uint32_t li, float scale, V3 thr,
                                                const float rb0[4], const float rb1[4],
                                                NeeOut& nb, NeeOut& nl) {
  const DevLight L = S.lights[li];
  bool use_bsdf = P.strategy != DGPU_DIRECT_LIGHT;
  bool use_light = P.strategy != DGPU_DIRECT_BSDF;
  nb.emit = false;
  nl.emit = false;
  if (use_bsdf) nb = nee_from_bsdf(S, g, wo, bsdf, L, scale, thr, use_light, rb0[1], rb0[2], rb0[3]);
  if (use_light) nl = nee_from_light(S, g, wo, bsdf, L, scale, thr, use_bsdf, rb1[0], rb1[1]);
}

__device__ __forceinline__ void emit_shadow(const WaveParams& P, const NeeOut& s, uint32_t slot_word,
                                            uint32_t qi) {
  P.so[qi] = make_float4(s.o.x, s.o.y, s.o.z, s.tmin);
  P.sd[qi] = make_float4(s.d.x, s.d.y, s.d.z, s.tmax);
  P.smeta[qi] = make_uint2(slot_word, s.ignore);
}

// Batched queues: the slot word holds the chunk path, and the NEE slot k (0: BSDF
// strategy, 1: light strategy) rides in bit 31 of the ignored-triangle word.
__device__ __forceinline__ void emit_shadow_batch(const WaveParams& P, const NeeOut& s, uint32_t path_word,
                                                  uint32_t k, uint32_t qi) {
  P.so[qi] = make_float4(s.o.x, s.o.y, s.o.z, s.tmin);
  P.sd[qi] = make_float4(s.d.x, s.d.y, s.d.z, s.tmax);
  P.smeta[qi] = make_uint2(path_word, (s.ignore & 0x7FFFFFFFu) | (k << 31));
}

// ----------------------------------------------------------------- adjoint payloads

// Gradient deposits of the adjoint replay (dgpu_adjoint), per lobe where a path scatters:
//   d loss / d theta += dldi . (d ln f / d theta) * (radiance carried by f)
// for the albedo texels of texture 0 (Lambert: d ln f_c / d albedo_c = 1 / albedo_c) and
// a GGX conductor's roughness.  A payload is a 16-byte record (x, y, z, word) with word =
// kind << 30 | index: ADJ_TEXEL: xyz = 1 / albedo per channel (0 where the albedo is 0),
// index the texel; ADJ_ROUGH: x = d ln f / d alpha at (wi, wo), index the material.  The
// radiance a lobe carries is known only later (a shadow ray's visibility, the rest of the
// path), so payloads wait in the path's records until it is.
constexpr uint32_t ADJ_NONE = 0u, ADJ_TEXEL = 1u, ADJ_ROUGH = 2u;
constexpr uint32_t ADJ_LDS_MATS = 64;  // roughness partial sums per workgroup in LDS

__device__ __forceinline__ float4 adj_payload(const WaveParams& P, const Bsdf& b, int32_t texel, uint32_t mat,
                                              V3 wi, V3 wo) {
  if (b.kind == DGPU_MAT_LAMBERT && texel >= 0 && P.adj_g_tex) {
    V3 a = b.albedo();
    return make_float4(a.x > 0.f ? 1.f / a.x : 0.f, a.y > 0.f ? 1.f / a.y : 0.f, a.z > 0.f ? 1.f / a.z : 0.f,
                       __uint_as_float((ADJ_TEXEL << 30) | (uint32_t)texel));
  }
  if (b.kind == DGPU_MAT_GGX_CONDUCTOR && P.adj_g_rough) {
    float r = conductor_dlnf_dalpha(b, b.to_local(wi), b.to_local(wo));
    return make_float4(r, 0.f, 0.f, __uint_as_float((ADJ_ROUGH << 30) | mat));
  }
  return make_float4(0.f, 0.f, 0.f, __uint_as_float(ADJ_NONE << 30));
}

// Deposits payload pl for carried radiance c at pixel weight dl (dL/dI of the path's pixel
// divided by the sample count).  Roughness sums of the first ADJ_LDS_MATS materials go to
// the workgroup's LDS partials s_rough.
__device__ __forceinline__ void adj_apply(const WaveParams& P, float4 pl, V3 c, const float dl[3], float* s_rough) {
  const uint32_t word = __float_as_uint(pl.w), kind = word >> 30, idx = word & 0x3FFFFFFFu;
  if (kind == ADJ_TEXEL) {
    float* g = P.adj_g_tex + 3 * (size_t)idx;
    if (pl.x != 0.f) atomicAdd(g + 0, dl[0] * c.x * pl.x);
    if (pl.y != 0.f) atomicAdd(g + 1, dl[1] * c.y * pl.y);
    if (pl.z != 0.f) atomicAdd(g + 2, dl[2] * c.z * pl.z);
  } else if (kind == ADJ_ROUGH) {
    float v = pl.x * (dl[0] * c.x + dl[1] * c.y + dl[2] * c.z);
    if (idx < ADJ_LDS_MATS) atomicAdd(s_rough + idx, v);
    else atomicAdd(P.adj_g_rough + idx, v);
  }
}

// One bounce of PathRenderer::sample for every path in the current queue.
// Every lane of a wave runs the same number of queue appends (inactive lanes
// append nothing), so the ballot-based compaction sees the whole wave.
// sh: the workgroup's LDS words for block_append.
// ADJ: the adjoint replay's second pass (non-batched): the same path, plus the gradient
// payloads of its lobes and their deposits once the radiance they carry is known.
template <bool BATCH, bool ALL, bool ROUGH, bool ADJ = false>
__device__ __forceinline__ void shade_body(const WaveParams& P, const DevScene& S, uint32_t bounce,
                                           bool active, uint32_t i, uint32_t p, uint32_t scene,
                                           const PathOrder& ord, uint32_t s0, uint32_t t, uint32_t k0,
                                           uint32_t k1, uint32_t* sh, float* s_rough = nullptr) {
  const uint32_t cur = bounce & 1u, nxt = cur ^ 1u;
  const uint32_t SL = P.slots;
  uint32_t flags = 0, bounces = 0;
  V3 thr = v3(0.f, 0.f, 0.f), rad = v3(0.f, 0.f, 0.f), wo = v3(0.f, 0.f, 1.f);
  Surface g;
  Bsdf bsdf;
  bsdf.kind = DGPU_MAT_NONE;
  bool shading = false;
  // random blocks are drawn where they are first needed (they are pure functions of
  // the path, bounce and block), so fewer of them are live across the NEE code
  float rb0[4] = {0.f, 0.f, 0.f, 0.f}, rb1[4] = {0.f, 0.f, 0.f, 0.f}, rb2[4] = {0.f, 0.f, 0.f, 0.f};
  // the path's sample and pixel (position t from the start of sample s0), needed only by
  // a path that scatters here (and by the adjoint's deposits)
  uint32_t sample = 0, pixel = 0;
  // adjoint: the path's first-pass radiance, whether its sample counts, and its pixel weight
  float4 adj_l = make_float4(0.f, 0.f, 0.f, 0.f);
  float dl[3] = {0.f, 0.f, 0.f};
  int32_t texel = -1;
  if (ADJ && active) {
    order_sample_pixel(ord, s0, t, sample, pixel);
    adj_l = P.adj_L[p];
    dl[0] = P.adj_dldi[3 * (size_t)pixel] * P.adj_inv_spp;
    dl[1] = P.adj_dldi[3 * (size_t)pixel + 1] * P.adj_inv_spp;
    dl[2] = P.adj_dldi[3 * (size_t)pixel + 2] * P.adj_inv_spp;
  }
  const bool adj_on = ADJ && adj_l.w != 0.f;
  // With nee_lag 2 the NEE of bounce b is added at bounce b + 2 (its visibility test runs
  // beside bounce b + 1), from the set of pending records of b's parity, which P.pend is
  // for this launch; the first two bounces start their sets.
  const uint32_t lag = (ADJ || P.nee_lag < 2u) ? 1u : 2u;
  if (active) {
    if (bounce == 0u) {
      // the path starts here (k_raygen writes only the queue)
      thr = v3(1.f, 1.f, 1.f);
      flags = 0u;
    } else {
      float4 t4 = P.thr[p], r4 = P.rad[p];
      thr = xyz(t4);
      rad = xyz(r4);
      flags = __float_as_uint(t4.w);
    }
    if (bounce < lag) {
      for (uint32_t k = 2u * (ALL ? P.S.num_lights : 1u); k < SL; ++k)  // slots no light writes
        P.pend[(size_t)p * SL + k] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      // NEE of the vertex lag bounces back, now that its visibility is known
      for (uint32_t k = 0; k < SL; ++k) {
        size_t s = (size_t)p * SL + k;
        float4 pe = P.pend[s];
        if (pe.w != 0.f) {
          rad = rad + xyz(pe);
          P.pend[s].w = 0.f;
          if (adj_on) adj_apply(P, P.adj_pend[s], xyz(pe), dl, s_rough);
        }
      }
    }
    // adjoint: the previous vertex's continuation carried everything the path gathers
    // from here on: its first-pass radiance less what it has gathered so far
    if (adj_on && bounce > 0u) adj_apply(P, P.adj_cont[p], xyz(adj_l) - rad, dl, s_rough);
    bounces = flags & F_BOUNCES;
    bool last_delta = (flags & F_DELTA) != 0u;
    float4 d4 = P.qd[cur][i], h4 = P.hit[i];
    V3 rd = xyz(d4);
    uint32_t leaf = __float_as_uint(h4.w);
    bool is = leaf != 0xFFFFFFFFu;
    if (is) g = surface_at(S, leaf, h4.x, h4.y, h4.z);
    if ((bounces == 0u || last_delta) && bounces >= P.min_depth && bounces <= P.max_depth) {
      if (is) {
        int32_t li = g.light;
        // Intersection::eval_radiance -> DiffuseLight::eval_radiance (diffuse_light.cpp:83-93);
        // the ray's origin is read only for a hit on an emitter
        if (li >= 0 && dot(g.nn, xyz(P.qo[cur][i]) - g.p) >= 0.f) rad = rad + thr * xyz(S.lights[li].radiance);
      } else {
        rad = rad + thr * background_radiance(S, rd);
      }
    }
    bounces += 1u;
    if (is && bounces <= P.max_depth) {
      shading = true;
      wo = normalize(-rd);
      bsdf = make_bsdf(S, g, ADJ ? &texel : nullptr);
      if (!ADJ) order_sample_pixel(ord, s0, t, sample, pixel);
      dgpu_rng_block(k0, k1, pixel, sample, bounce, 0, rb0);
    }
  }
  // direct lighting (path_renderer.cpp:68-76, 113-132): up to 2 shadow rays per light
  bool do_nee = shading && bounces >= P.min_depth && bsdf.matches(BSDF_ALL & ~BSDF_DELTA) &&
                S.num_lights > 0u;
  uint32_t lights_loop = ALL ? P.S.num_lights : 1u;
  const bool use_bsdf = P.strategy != DGPU_DIRECT_LIGHT, use_light = P.strategy != DGPU_DIRECT_BSDF;
  for (uint32_t li = 0; li < lights_loop; ++li) {
    // one light: all of them in turn (sample_all_lights), or one picked by power with
    // DiscreteDistrib1d::sample / pdf (discrete_distrib_1d.cpp:22-42)
    uint32_t light = li;
    float scale = 0.f;
    if (do_nee) {
      scale = 1.f;
      if (!ALL) {
        float u = rb0[0];
        uint32_t b = 0, e = S.num_lights;
        while (b + 1u < e) {
          uint32_t m = b + (e - b) / 2u;
          if (S.light_cdf[m] < u) b = m; else e = m;
        }
        float pp = S.light_cdf[b + 1] - S.light_cdf[b];
        light = b;
        scale = pp != 0.f ? 1.f / pp : 0.f;
      }
    }
    const bool nee = scale != 0.f;
    DevLight L;
    if (nee) L = S.lights[light];
    // the two strategies one after the other, each shadow ray queued as soon as it exists
    uint32_t slot_b = (uint32_t)((size_t)p * SL + 2 * li), slot_l = slot_b + 1;
    {
      NeeOut nb;
      nb.emit = false;
      if (nee && use_bsdf) nb = nee_from_bsdf<ROUGH>(S, g, wo, bsdf, L, scale, thr, use_light, rb0[1], rb0[2], rb0[3]);
      uint32_t qb = block_append(&P.counters[CNT_SHADOW + bounce], nb.emit, sh);
      if (nb.emit) {
        P.pend[slot_b] = make_float4(nb.contrib.x, nb.contrib.y, nb.contrib.z, 1.f);
        if (ADJ) P.adj_pend[slot_b] = adj_payload(P, bsdf, texel, g.mat, nb.d, wo);
        if (BATCH) emit_shadow_batch(P, nb, p, 0u, qb);
        else emit_shadow(P, nb, slot_b, qb);
      } else if (active && bounce < lag) {
        P.pend[slot_b] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    {
      NeeOut nl;
      nl.emit = false;
      if (nee && use_light) {
        dgpu_rng_block(k0, k1, pixel, sample, bounce, 1, rb1);
        nl = nee_from_light<ROUGH>(S, g, wo, bsdf, L, scale, thr, use_bsdf, rb1[0], rb1[1]);
      }
      uint32_t ql = block_append(&P.counters[CNT_SHADOW + bounce], nl.emit, sh);
      if (nl.emit) {
        P.pend[slot_l] = make_float4(nl.contrib.x, nl.contrib.y, nl.contrib.z, 1.f);
        if (ADJ) P.adj_pend[slot_l] = adj_payload(P, bsdf, texel, g.mat, nl.d, wo);
        if (BATCH) emit_shadow_batch(P, nl, p, 1u, ql);
        else emit_shadow(P, nl, slot_l, ql);
      } else if (active && bounce < lag) {
        P.pend[slot_l] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  }
  // continuation: Bsdf::sample_light_f, throughput, Russian roulette (path_renderer.cpp:78-107)
  bool cont = false;
  V3 next_o = v3(0.f, 0.f, 0.f), next_d = v3(0.f, 0.f, 1.f);
  float next_eps = 0.f;
  if (shading) {
    uint32_t req = (P.only_direct || bounces == P.max_depth) ? (BSDF_MODES | BSDF_DELTA) : BSDF_ALL;
    V3 wi; float pdf; uint32_t fl;
    dgpu_rng_block(k0, k1, pixel, sample, bounce, 1, rb1);
    dgpu_rng_block(k0, k1, pixel, sample, bounce, 2, rb2);
    V3 f = bsdf_sample<false, ROUGH>(bsdf, wo, req, rb1[2], rb1[3], rb2[0], wi, pdf, fl);
    if (!is_black(f) && pdf != 0.f) {
      V3 contrib = f * (fabsf(dot(wi, g.nn)) / pdf);
      bool delta = (fl & BSDF_DELTA) != 0u;
      thr = thr * contrib;
      bool go = true;
      if (bounces == P.max_depth && !delta) {
        go = false;
      } else if (bounces >= P.rr_depth && bounces > P.min_depth) {
        float q = clampf(average(contrib), 0.1f, 0.99f);
        if (rb2[1] > q) go = false;
        else thr = thr / q;
      }
      if (go) {
        cont = true;
        next_o = g.p; next_d = wi; next_eps = g.eps;
        flags = (flags & ~(F_BOUNCES | F_DELTA)) | bounces | (delta ? F_DELTA : 0u);
        // adjoint: a delta lobe has no parameter here; a path that ends here carries
        // nothing further, so only continuing paths keep a payload
        if (ADJ) P.adj_cont[p] = delta ? make_float4(0.f, 0.f, 0.f, __uint_as_float(ADJ_NONE << 30))
                                       : adj_payload(P, bsdf, texel, g.mat, wi, wo);
      }
    }
  }
  if (active) {
    P.thr[p] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(flags));
    P.rad[p] = make_float4(rad.x, rad.y, rad.z, bounce > 0u ? 1.f : 0.f);  // w: the odd set is started
  }
  uint32_t qn = block_append(&P.counters[CNT_QUEUE + bounce + 1], cont, sh);
  if (cont) {
    P.qo[nxt][qn] = make_float4(next_o.x, next_o.y, next_o.z, next_eps);
    P.qd[nxt][qn] = make_float4(next_d.x, next_d.y, next_d.z, INFINITY);
    P.qpath[nxt][qn] = p;
  }
}

// A fixed grid strides over the queue, whose length is known only on the device;
// the loop bound is uniform per workgroup, so every wave runs the same number of
// queue appends.
#ifndef DGPU_SHADE_WAVES
#define DGPU_SHADE_WAVES 4
#endif
// ALL: sample_all_lights (one NEE per light and vertex) instead of one light picked by power.
// ROUGH: with the rough dielectric's lobe (scenes that have one).
template <bool BATCH, bool ALL, bool ROUGH, bool ADJ = false>
__global__ __launch_bounds__(SHADE_BLOCK) __attribute__((amdgpu_waves_per_eu(DGPU_SHADE_WAVES)))
void k_shade(WaveParams P, uint32_t bounce) {
  __shared__ uint32_t sh[SHADE_BLOCK / 64 + 1];
  __shared__ float s_rough[ADJ ? ADJ_LDS_MATS : 1];
  if (ADJ) {
    for (uint32_t k = threadIdx.x; k < ADJ_LDS_MATS; k += SHADE_BLOCK) s_rough[k] = 0.f;
    __syncthreads();
  }
  __shared__ uint32_t seg_begin[BATCH ? MAX_CHUNK_SEGS : 1], seg_scene[BATCH ? MAX_CHUNK_SEGS : 1];
  __shared__ uint32_t seg_q[BATCH ? MAX_CHUNK_SEGS : 1], seg_s[BATCH ? MAX_CHUNK_SEGS : 1];
  const uint32_t cur = bounce & 1u;
  const uint32_t n = P.counters[CNT_QUEUE + bounce];
  uint32_t nsegs = 0;
  if (BATCH) {
    const ChunkDesc& cd = P.chunks[*P.chunk_cur];
    nsegs = cd.n_segs;
    for (uint32_t k = threadIdx.x; k < nsegs; k += SHADE_BLOCK) {
      seg_begin[k] = cd.segs[k].slot_begin;
      seg_scene[k] = cd.segs[k].scene;
      seg_q[k] = cd.segs[k].q_begin;
      seg_s[k] = cd.segs[k].s_begin;
    }
    __syncthreads();
  }
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool active = i < n;
    if (BATCH) {
      uint32_t p = 0, scene = 0, sg = 0;
      if (active) {
        p = P.qpath[cur][i];
        sg = seg_of(seg_begin, nsegs, p);
        scene = seg_scene[sg];
      }
      // the batch table as constant memory: the compiler then knows the pointers read
      // from it address global memory (else every material, light and texture load is
      // a flat load) and that the shade's stores do not change it
      using CEntry = const __attribute__((address_space(4))) BatchEntry;
      // Queues keep the chunk's scene-major order, so nearly every wave holds paths of
      // one segment.  Such a wave reads its scene's record with scalar loads (a
      // wave-uniform index into constant memory); a wave that straddles segments reads
      // each lane's record with vector loads.  Both copies run the same queue appends,
      // so every wave of the workgroup meets the same barriers.
      const uint32_t sg0 = __builtin_amdgcn_readfirstlane(sg);
      if (__ballot(active && sg != sg0) == 0ull) {
        const uint32_t sc0 = __builtin_amdgcn_readfirstlane(seg_scene[sg0]);
        const CEntry& E = ((CEntry*)P.batch)[sc0];
        const uint32_t s_begin = __builtin_amdgcn_readfirstlane(seg_s[sg0]);
        const uint32_t q_first = __builtin_amdgcn_readfirstlane(seg_q[sg0] - seg_begin[sg0]);
        shade_body<true, false, ROUGH>(P, *(const DevScene*)&E.S, bounce, active, i, p, sc0,
                                       *(const PathOrder*)&E.ord, s_begin, q_first + p, E.k0, E.k1, sh);
      } else {
        const CEntry& E = ((CEntry*)P.batch)[scene];
        // the path's pixel and sample, as k_raygen derived them from its segment
        shade_body<true, false, ROUGH>(P, *(const DevScene*)&E.S, bounce, active, i, p, scene,
                                       *(const PathOrder*)&E.ord, seg_s[sg], seg_q[sg] + (p - seg_begin[sg]), E.k0,
                                       E.k1, sh);
      }
    } else {
      uint32_t p = active ? P.qpath[cur][i] : 0u;
      shade_body<false, ALL, ROUGH, ADJ>(P, P.S, bounce, active, i, p, 0u, P.ord, P.s0, P.q0 + p, P.k0, P.k1, sh,
                                         s_rough);
    }
  }
  if (ADJ) {
    __syncthreads();
    if (P.adj_g_rough)
      for (uint32_t k = threadIdx.x; k < ADJ_LDS_MATS && k < P.S.num_mats; k += SHADE_BLOCK)
        if (s_rough[k] != 0.f) atomicAdd(P.adj_g_rough + k, s_rough[k]);
  }
}

// Adjoint, first pass: each path's radiance as the film would receive it (rad plus its
// visible pending NEE), with w = 1 when the sample counts (finite, non-negative, not on a
// pixel's lower edge); per pixel in sample order, which also gives the optional image
// (the per-pixel mean) without atomics.
__global__ __launch_bounds__(256) void k_adj_total(WaveParams P) {
  uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= P.npix) return;
  uint64_t g_end = P.g0 + P.n_chunk;
  uint64_t g = P.g0 + (((uint64_t)tile_pos(px, P.xr, P.yr) + P.npix - (uint32_t)(P.g0 % P.npix)) % P.npix);
  V3 sum = v3(0.f, 0.f, 0.f);
  for (; g < g_end; g += P.npix) {
    uint32_t slot = (uint32_t)(g - P.g0);
    V3 rad = xyz(P.rad[slot]);
    for (uint32_t k = 0; k < P.slots; ++k) {
      float4 pe = P.pend[(size_t)slot * P.slots + k];
      if (pe.w != 0.f) rad = rad + xyz(pe);
    }
    bool ok = isfinite(rad.x) && isfinite(rad.y) && isfinite(rad.z) && rad.x >= 0.f && rad.y >= 0.f &&
              rad.z >= 0.f;
    bool counts = ok && !jitter_on_edge(P.k0, P.k1, px, P.sample_begin + (uint32_t)(g / P.npix));
    P.adj_L[slot] = make_float4(rad.x, rad.y, rad.z, counts ? 1.f : 0.f);
    if (counts) sum = sum + rad;
  }
  if (P.adj_image) {
    P.adj_image[3 * (size_t)px] += sum.x * P.adj_inv_spp;
    P.adj_image[3 * (size_t)px + 1] += sum.y * P.adj_inv_spp;
    P.adj_image[3 * (size_t)px + 2] += sum.z * P.adj_inv_spp;
  }
}

// Adjoint, second pass: the deposits of the NEE still pending after the last bounce.
__global__ __launch_bounds__(256) void k_adj_final(WaveParams P) {
  __shared__ float s_rough[ADJ_LDS_MATS];
  for (uint32_t k = threadIdx.x; k < ADJ_LDS_MATS; k += 256) s_rough[k] = 0.f;
  __syncthreads();
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P.n_chunk && P.adj_L[i].w != 0.f) {
    uint32_t sample, pixel;
    order_sample_pixel(P.ord, P.s0, P.q0 + i, sample, pixel);
    float dl[3] = {P.adj_dldi[3 * (size_t)pixel] * P.adj_inv_spp, P.adj_dldi[3 * (size_t)pixel + 1] * P.adj_inv_spp,
                   P.adj_dldi[3 * (size_t)pixel + 2] * P.adj_inv_spp};
    for (uint32_t k = 0; k < P.slots; ++k) {
      size_t s = (size_t)i * P.slots + k;
      float4 pe = P.pend[s];
      if (pe.w != 0.f) adj_apply(P, P.adj_pend[s], xyz(pe), dl, s_rough);
    }
  }
  __syncthreads();
  if (P.adj_g_rough)
    for (uint32_t k = threadIdx.x; k < ADJ_LDS_MATS && k < P.S.num_mats; k += 256)
      if (s_rough[k] != 0.f) atomicAdd(P.adj_g_rough + k, s_rough[k]);
}

// Film::add_sample with the r = 0.5 box filter: each sample lands on one pixel with
// weight 1; non-finite or negative samples are dropped (renderer.cpp:36-41).  Samples
// are added to the film value one at a time in sample order, so the film does not
// depend on how paths were split into chunks or batched with other scenes.
__device__ __forceinline__ void film_accumulate(const WaveParams& P, float* film_rgb, float* film_w,
                                                uint32_t px, uint32_t slot, uint32_t npix,
                                                uint64_t local, uint64_t local_end, uint32_t sample_begin,
                                                uint32_t k0, uint32_t k1) {
  float fr = film_rgb[3 * (size_t)px], fg = film_rgb[3 * (size_t)px + 1], fb = film_rgb[3 * (size_t)px + 2];
  float fw = film_w[px];
  uint32_t dropped = 0;
  bool touched = false;
  for (; local < local_end; local += npix, slot += npix) {
    float4 r4 = P.rad[slot];
    V3 rad = xyz(r4);
    for (uint32_t k = 0; k < P.slots; ++k) {
      size_t s = (size_t)slot * P.slots + k;
      float4 pe = P.pend[s];
      if (pe.w != 0.f) rad = rad + xyz(pe);
    }
    // nee_lag 2: the odd bounces' set too, once the path has reached bounce 1
    if (P.nee_lag >= 2u && r4.w != 0.f)
      for (uint32_t k = 0; k < P.slots; ++k) {
        float4 pe = P.pend_b[(size_t)slot * P.slots + k];
        if (pe.w != 0.f) rad = rad + xyz(pe);
      }
    bool ok = isfinite(rad.x) && isfinite(rad.y) && isfinite(rad.z) && rad.x >= 0.f &&
              rad.y >= 0.f && rad.z >= 0.f;
    if (!ok) { ++dropped; continue; }
    if (jitter_on_edge(k0, k1, px, sample_begin + (uint32_t)(local / npix))) continue;
    fr += rad.x; fg += rad.y; fb += rad.z;
    fw += 1.f;
    touched = true;
  }
  if (touched) {
    film_rgb[3 * (size_t)px] = fr;
    film_rgb[3 * (size_t)px + 1] = fg;
    film_rgb[3 * (size_t)px + 2] = fb;
    film_w[px] = fw;
  }
  if (dropped) atomicAdd((unsigned long long*)&P.stats[3], (unsigned long long)dropped);
}

__device__ __forceinline__ void add_chunk_stats(const WaveParams& P, uint32_t n_chunk) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint64_t ext = 0, shd = 0;
    for (uint32_t b = 0; b < MAX_BOUNCES; ++b) {
      ext += P.counters[CNT_QUEUE + b];
      shd += P.counters[CNT_SHADOW + b];
      P.stats[4 + (b < 15u ? b : 15u)] += P.counters[CNT_QUEUE + b];
      P.stats[20 + (b < 15u ? b : 15u)] += P.counters[CNT_SHADOW + b];
    }
    P.stats[0] += n_chunk;
    P.stats[1] += ext;
    P.stats[2] += shd;
  }
}

// Resolves the last pending NEE of every path and adds the chunk's samples to the film.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_finalize(WaveParams P) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (BATCH) {
    const ChunkDesc& cd = P.chunks[*P.chunk_cur];
    add_chunk_stats(P, cd.n_chunk);
    if (t >= cd.n_pairs) return;
    uint32_t b = 0, e = cd.n_segs;
    while (b + 1u < e) {
      uint32_t m = (b + e) / 2u;
      if (cd.segs[m].pair_begin <= t) b = m; else e = m;
    }
    const ChunkSeg& sg = cd.segs[b];
    using CEntry = const __attribute__((address_space(4))) BatchEntry;
    const CEntry& E = ((CEntry*)P.batch)[sg.scene];  // constant: the film pointers are global
    uint32_t q = t - sg.pair_begin;
    uint64_t local = sg.local_begin + q;
    uint32_t px = tile_pixel((uint32_t)(local % E.npix), E.xr, E.yr);
    film_accumulate(P, E.film_rgb, E.film_w, px, sg.slot_begin + q, E.npix, local, sg.local_end, E.sample_begin,
                    E.k0, E.k1);
  } else {
    add_chunk_stats(P, P.n_chunk);
    uint32_t px = t;
    if (px >= P.npix) return;
    uint64_t g_end = P.g0 + P.n_chunk;
    // the chunk's first path on pixel px (paths of a sample visit pixels in tile order)
    uint32_t q = tile_pos(px, P.xr, P.yr);
    uint64_t g = P.g0 + (((uint64_t)q + P.npix - (uint32_t)(P.g0 % P.npix)) % P.npix);
    if (g >= g_end) return;
    film_accumulate(P, P.film_rgb, P.film_w, px, (uint32_t)(g - P.g0), P.npix, g, g_end, P.sample_begin, P.k0,
                    P.k1);
  }
}

// ----------------------------------------------------------------- adjoint

// Path-replay adjoint.  Each thread owns one path: pass 1 traces it to get its
// total radiance, pass 2 re-traces it from the same counter-based random numbers
// (memory stays O(pixels): nothing is taped) and deposits
//   d loss / d theta = dldi . sum_j (d ln f_j / d theta) * (radiance carried by f_j)
// for the albedo texels of texture 0 and every material's GGX roughness.  Sampling
// pdfs, MIS weights and Russian-roulette decisions are held constant (detached),
// which keeps the estimator unbiased for the gradient of the expected image.
// The arithmetic per vertex is the same device code as the wavefront kernels
// (surface_at, make_bsdf, estimate_direct, bsdf_sample).

struct AdjSink {
  float dl[3];          // dL/dI of the path's pixel divided by the sample count
  float* g_tex;         // texture-0 gradient (RGB per texel) or null
  float* g_rough;       // per-material roughness gradient or null
  float* s_rough;       // workgroup partial sums for the first ADJ_LDS_MATS materials
  uint32_t num_mats;
};

constexpr uint32_t ADJ_BLOCK = 64;

__device__ __forceinline__ void adj_deposit(const AdjSink& k, const Bsdf& b, int32_t texel, uint32_t mat,
                                            V3 wi, V3 wo, V3 carried) {
  if (b.kind == DGPU_MAT_LAMBERT && texel >= 0 && k.g_tex) {
    // f = albedo / pi: d ln f_c / d albedo_c = 1 / albedo_c
    float* g = k.g_tex + 3 * (size_t)texel;
    if (b.albedo().x > 0.f) atomicAdd(g + 0, k.dl[0] * carried.x / b.albedo().x);
    if (b.albedo().y > 0.f) atomicAdd(g + 1, k.dl[1] * carried.y / b.albedo().y);
    if (b.albedo().z > 0.f) atomicAdd(g + 2, k.dl[2] * carried.z / b.albedo().z);
  } else if (b.kind == DGPU_MAT_GGX_CONDUCTOR && k.g_rough) {
    float r = conductor_dlnf_dalpha(b, b.to_local(wi), b.to_local(wo));
    float v = r * (k.dl[0] * carried.x + k.dl[1] * carried.y + k.dl[2] * carried.z);
    if (mat < ADJ_LDS_MATS) atomicAdd(k.s_rough + mat, v);
    else atomicAdd(k.g_rough + mat, v);
  }
}

template <bool ADJ, int WIDTH>
__device__ V3 replay_path(const WaveParams& P, uint32_t pixel, uint32_t sample, LaneStack& st,
                          V3 total, const AdjSink& sink, bool& film) {
  const DevScene& S = P.S;
  uint32_t x = pixel % P.xr, y = pixel / P.xr;
  float jx = 0.5f, jy = 0.5f;
  if (sample != 0) {
    float r[4];
    dgpu_rng_block(P.k0, P.k1, pixel, sample, 0, 2, r);
    jx = r[2]; jy = r[3];
  }
  film = !(jx == 0.f || jy == 0.f);
  TraceRay ray;
  camera_ray(S, float(x) + jx, float(y) + jy, P.xr, P.yr, ray.o, ray.d);
  ray.tmin = 0.f;
  ray.tmax = INFINITY;
  ray.ignore = 0xFFFFFFFFu;
  V3 thr = v3(1.f, 1.f, 1.f), rad = v3(0.f, 0.f, 0.f);
  uint32_t bounces = 0;
  bool last_delta = false;
  const uint32_t NON_DELTA = BSDF_ALL & ~BSDF_DELTA;
  for (uint32_t bounce = 0; bounce < MAX_BOUNCES; ++bounce) {
    TraceHit h;
    bool is = trace_one<false, WIDTH, true>(BvhRef{S.nodes, S.tris}, ray, st, h);
    Surface g;
    if (is) g = surface_at(S, h.tri, h.t, h.b1, h.b2);
    if ((bounces == 0u || last_delta) && bounces >= P.min_depth && bounces <= P.max_depth) {
      if (is) {
        int32_t li = g.light;
        if (li >= 0 && dot(g.nn, ray.o - g.p) >= 0.f) rad = rad + thr * xyz(S.lights[li].radiance);
      } else {
        rad = rad + thr * background_radiance(S, ray.d);
      }
    }
    bounces += 1u;
    if (!is || bounces > P.max_depth) break;
    V3 wo = normalize(-ray.d);
    int32_t texel = -1;
    Bsdf bsdf = make_bsdf(S, g, &texel);
    uint32_t mat = g.mat;
    float rb0[4], rb1[4], rb2[4];
    dgpu_rng_block(P.k0, P.k1, pixel, sample, bounce, 0, rb0);
    dgpu_rng_block(P.k0, P.k1, pixel, sample, bounce, 1, rb1);
    dgpu_rng_block(P.k0, P.k1, pixel, sample, bounce, 2, rb2);
    if (bounces >= P.min_depth && bsdf.matches(NON_DELTA) && S.num_lights > 0u) {
      uint32_t nl_loop = P.sample_all ? S.num_lights : 1u;
      for (uint32_t li = 0; li < nl_loop; ++li) {
        NeeOut nb, nl;
        nb.emit = false;
        nl.emit = false;
        if (P.sample_all) {
          estimate_direct(S, P, g, wo, bsdf, li, 1.f, thr, rb0, rb1, nb, nl);
        } else {
          float u = rb0[0];
          uint32_t b = 0, e = S.num_lights;
          while (b + 1u < e) {
            uint32_t m = b + (e - b) / 2u;
            if (S.light_cdf[m] < u) b = m; else e = m;
          }
          float pp = S.light_cdf[b + 1] - S.light_cdf[b];
          if (pp != 0.f) estimate_direct(S, P, g, wo, bsdf, b, 1.f / pp, thr, rb0, rb1, nb, nl);
        }
        auto shadow = [&](const NeeOut& n) {
          if (!n.emit) return;
          TraceRay sr;
          sr.o = n.o; sr.d = n.d; sr.tmin = n.tmin; sr.tmax = n.tmax; sr.ignore = n.ignore;
          TraceHit sh;
          if (trace_one<true, WIDTH, true>(BvhRef{S.nodes, S.tris}, sr, st, sh)) return;
          rad = rad + n.contrib;
          if (ADJ) adj_deposit(sink, bsdf, texel, mat, n.d, wo, n.contrib);
        };
        shadow(nb);
        shadow(nl);
      }
    }
    uint32_t req = (P.only_direct || bounces == P.max_depth) ? (BSDF_MODES | BSDF_DELTA) : BSDF_ALL;
    V3 wi; float pdf; uint32_t fl;
    V3 f = bsdf_sample(bsdf, wo, req, rb1[2], rb1[3], rb2[0], wi, pdf, fl);
    if (is_black(f) || pdf == 0.f) break;
    V3 contrib = f * (fabsf(dot(wi, g.nn)) / pdf);
    bool delta = (fl & BSDF_DELTA) != 0u;
    // radiance still to come through this continuation = total - accumulated so far
    if (ADJ && !delta) adj_deposit(sink, bsdf, texel, mat, wi, wo, total - rad);
    thr = thr * contrib;
    last_delta = delta;
    if (bounces == P.max_depth && !delta) break;
    if (bounces >= P.rr_depth && bounces > P.min_depth) {
      float q = clampf(average(contrib), 0.1f, 0.99f);
      if (rb2[1] > q) break;
      thr = thr / q;
    }
    ray.o = g.p; ray.d = wi; ray.tmin = g.eps; ray.tmax = INFINITY; ray.ignore = 0xFFFFFFFFu;
  }
  return rad;
}

template <int WIDTH>
__global__ __launch_bounds__(ADJ_BLOCK) void k_adjoint(WaveParams P, AdjArgs A) {
  __shared__ uint32_t stk[16 * ADJ_BLOCK];
  __shared__ float s_rough[ADJ_LDS_MATS];
  for (uint32_t k = threadIdx.x; k < ADJ_LDS_MATS; k += ADJ_BLOCK) s_rough[k] = 0.f;
  __syncthreads();
  LaneStack st;
  st.lds = stk + threadIdx.x;
  st.lds_stride = ADJ_BLOCK;
  st.lds_cap = 16;
  st.spill = P.spill + (size_t)blockIdx.x * ADJ_BLOCK + threadIdx.x;
  st.spill_stride = (size_t)gridDim.x * ADJ_BLOCK;
  st.cap = P.S.stack_cap;
  uint32_t i = blockIdx.x * ADJ_BLOCK + threadIdx.x;
  if (i < P.n_chunk) {
    uint64_t g = P.g0 + i;
    uint32_t sample = P.sample_begin + (uint32_t)(g / P.npix);
    uint32_t pixel = (uint32_t)(g % P.npix);
    AdjSink sink;
    sink.dl[0] = A.dldi[3 * (size_t)pixel] * A.inv_spp;
    sink.dl[1] = A.dldi[3 * (size_t)pixel + 1] * A.inv_spp;
    sink.dl[2] = A.dldi[3 * (size_t)pixel + 2] * A.inv_spp;
    sink.g_tex = A.g_tex;
    sink.g_rough = A.g_rough;
    sink.s_rough = s_rough;
    sink.num_mats = P.S.num_mats;
    bool film = true;
    V3 total = replay_path<false, WIDTH>(P, pixel, sample, st, v3(0.f, 0.f, 0.f), sink, film);
    bool ok = isfinite(total.x) && isfinite(total.y) && isfinite(total.z) && total.x >= 0.f &&
              total.y >= 0.f && total.z >= 0.f;
    if (ok && film) {
      replay_path<true, WIDTH>(P, pixel, sample, st, total, sink, film);
      if (A.image) {
        atomicAdd(A.image + 3 * (size_t)pixel, total.x * A.inv_spp);
        atomicAdd(A.image + 3 * (size_t)pixel + 1, total.y * A.inv_spp);
        atomicAdd(A.image + 3 * (size_t)pixel + 2, total.z * A.inv_spp);
      }
    }
  }
  __syncthreads();
  if (A.g_rough) {
    uint32_t nm = P.S.num_mats < ADJ_LDS_MATS ? P.S.num_mats : ADJ_LDS_MATS;
    for (uint32_t k = threadIdx.x; k < nm; k += ADJ_BLOCK)
      if (s_rough[k] != 0.f) atomicAdd(A.g_rough + k, s_rough[k]);
  }
}

// ----------------------------------------------------------------- launchers

// Compiled (BVH width, LDS stack entries, workgroup size) variants of the trace kernels.
#define DG_VARIANTS(X)                                                              \
  X(2, 8, 64) X(2, 16, 64) X(2, 32, 64) X(2, 8, 128) X(2, 16, 128) X(2, 32, 128)      \
  X(2, 8, 256) X(2, 16, 256) X(2, 32, 256) X(4, 8, 64) X(4, 16, 64) X(4, 32, 64)     \
  X(4, 8, 128) X(4, 16, 128) X(4, 32, 128) X(4, 8, 256) X(4, 16, 256) X(4, 32, 256)  \
  X(8, 8, 64) X(8, 16, 64) X(8, 32, 64) X(8, 8, 128) X(8, 16, 128) X(8, 32, 128)      \
  X(8, 8, 256) X(8, 16, 256) X(8, 32, 256)

bool variant_supported(uint32_t width, uint32_t lds, uint32_t block) {
#define X(W, L, B) if (width == W && lds == L && block == B) return true;
  DG_VARIANTS(X)
#undef X
  return false;
}

const void* trace_kernel_ptr(int which, uint32_t width, uint32_t lds, uint32_t block) {
#define X(W, L, B)                                                                         \
  if (width == W && lds == L && block == B) {                                              \
    if (which == 0) return (const void*)&k_trace_api<false, W, L, B, false>;               \
    if (which == 1) return (const void*)&k_trace_api<true, W, L, B, false>;                \
    if (which == 2) return (const void*)&k_extend<W, L, B, false>;                         \
    if (which == 3) return (const void*)&k_connect<W, L, B, false>;                        \
    if (which == 4) return (const void*)&k_extend<W, L, B, true>;                          \
    return (const void*)&k_connect<W, L, B, true>;                                         \
  }
  DG_VARIANTS(X)
#undef X
  return nullptr;
}

// ----------------------------------------------------------------- light tracing
//
// dort's LightRenderer (src/dort/light_renderer.cpp:13-181) as a wavefront.  k_lt_start
// picks a light per path, connects a point on it to the camera (sample_immediate, a
// path of 2 vertices) and samples the first ray (Light::sample_ray_radiance);
// k_lt_shade connects each hit to the camera (connect_to_camera) and scatters
// (Bsdf::sample_camera_f, Russian roulette); k_connect tests the connections and
// k_lt_splat adds the visible ones to the film (Film::add_splat) with float atomics.
// Stage s of the shadow queue holds the connections of paths of s + 2 vertices; a
// connection's pending record (contribution, visible flag, film position) sits at its
// own queue slot.  Path state: thr = (throughput, bounces), rad = (normal of the last
// vertex, pdf of the direction sampled there).

// Four uniforms of block `blk` of stage `stage` of light path (pixel, sample): Philox
// with a counter word of its own, independent of the path tracer's streams.
__device__ __forceinline__ void lt_rng(const WaveParams& P, uint32_t pixel, uint32_t sample, uint32_t stage,
                                       uint32_t blk, float out[4]) {
  dgpu_u32x4 r = dgpu_philox4x32_10(pixel, sample, (stage << 2) | blk, 0x6C747261u, P.k0, P.k1);
  for (int k = 0; k < 4; ++k) out[k] = dgpu_u01(r.v[k]);
}

__device__ __forceinline__ void lt_emit_connection(const WaveParams& P, bool emit, V3 from, float eps, V3 to,
                                                   V3 contrib, float2 fp, uint32_t stage, uint32_t* sh) {
  uint32_t q = block_append(&P.counters[CNT_SHADOW + stage], emit, sh);
  if (emit) {
    // ShadowTest::init_point_point(from, eps, to, 0) (light.cpp:40-49)
    V3 d = to - from;
    float dist = length(d);
    d = d / dist;
    P.so[q] = make_float4(from.x, from.y, from.z, eps);
    P.sd[q] = make_float4(d.x, d.y, d.z, dist);
    P.smeta[q] = make_uint2(q, 0xFFFFFFFFu);
    P.pend[q] = make_float4(contrib.x, contrib.y, contrib.z, 1.f);
    P.fpos[q] = fp;
  }
}

__global__ __launch_bounds__(SHADE_BLOCK) void k_lt_start(WaveParams P) {
  __shared__ uint32_t sh[SHADE_BLOCK / 64 + 1];
  const DevScene& S = P.S;
  const V3 cam = v3(S.c2w[3], S.c2w[7], S.c2w[11]);
  for (uint32_t base = blockIdx.x * blockDim.x; base < P.n_chunk; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool imm = false, go = false;
    V3 lp = v3(0.f, 0.f, 0.f), contrib = lp, thr = lp;
    float2 fp = make_float2(0.f, 0.f);
    LtRay lr;
    lr.o = lr.d = lr.nn = lr.rad = lp;
    lr.tmin = lr.pos_pdf = lr.dir_pdf = 0.f;
    if (i < P.n_chunk && S.num_lights > 0u) {
      uint64_t g = P.g0 + i;
      uint32_t sample = P.sample_begin + (uint32_t)(g / P.npix), pixel = (uint32_t)(g % P.npix);
      float r0[4], r1[4];
      lt_rng(P, pixel, sample, 0, 0, r0);
      float pick;
      uint32_t li = lt_pick_light(S, r0[0], pick);
      if (pick != 0.f) {
        const DevLight L = S.lights[li];
        if (P.min_len <= 2u && 2u <= P.max_len && L.kind == LK_AREA) {
          // sample_immediate (light_renderer.cpp:107-144); only area lights have points
          V3 ln;
          lp = lt_area_point(S, L, r0[1], r0[2], ln);
          float pos_pdf = 1.f / L.area;
          float imp;
          if (pos_pdf != 0.f && !(lp.x == cam.x && lp.y == cam.y && lp.z == cam.z) &&
              lt_importance(P, lt_cam_vector(S, normalize(lp - cam)), fp, imp) && dot(ln, cam - lp) >= 0.f) {
            V3 rad = xyz(L.radiance);
            contrib = rad * imp / (length_sq(lp - cam) * (pick * pos_pdf));
            contrib = contrib * fabsf(dot(ln, normalize(lp - cam)));
            imm = !is_black(contrib);
          }
        }
        if (P.max_len >= 3u) {
          lt_rng(P, pixel, sample, 0, 1, r1);
          lr = lt_light_ray(S, L, r1);
          float light_pdf = pick * lr.pos_pdf * lr.dir_pdf;
          if (!is_black(lr.rad) && light_pdf != 0.f) {
            thr = lr.rad / (lr.pos_pdf * pick);
            go = true;
          }
        }
      }
    }
    lt_emit_connection(P, imm, lp, 1e-3f, cam, contrib, fp, 0, sh);
    uint32_t qn = block_append(&P.counters[CNT_QUEUE + 0], go, sh);
    if (go) {
      P.qo[0][qn] = make_float4(lr.o.x, lr.o.y, lr.o.z, lr.tmin);
      P.qd[0][qn] = make_float4(lr.d.x, lr.d.y, lr.d.z, INFINITY);
      P.qpath[0][qn] = i;
      P.thr[i] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(0u));
      P.rad[i] = make_float4(lr.nn.x, lr.nn.y, lr.nn.z, lr.dir_pdf);
    }
  }
}

// One vertex of LightRenderer::sample_path (light_renderer.cpp:66-104) for every hit of
// extension stage `bounce`.  Bsdf::sample_camera_f calls each lobe's sample function
// with the same arguments as sample_light_f (bsdf.cpp:95-110; dielectric_material.cpp:10-22,
// rough_dielectric_material.cpp:39-52 pass fix_is_light = false to both), and every
// material here is one lobe, so it is bsdf_sample with the fixed direction wi.
__global__ __launch_bounds__(SHADE_BLOCK) void k_lt_shade(WaveParams P, uint32_t bounce) {
  __shared__ uint32_t sh[SHADE_BLOCK / 64 + 1];
  const DevScene& S = P.S;
  const V3 cam = v3(S.c2w[3], S.c2w[7], S.c2w[11]);
  const uint32_t cur = bounce & 1u, nxt = cur ^ 1u;
  const uint32_t n = P.counters[CNT_QUEUE + bounce];
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool conn = false, go = false;
    uint32_t p = 0, bounces = 0;
    V3 gp = v3(0.f, 0.f, 0.f), contrib = gp, thr = gp, gn = gp, wo = gp;
    float eps = 0.f, pdf = 0.f;
    float2 fp = make_float2(0.f, 0.f);
    if (i < n) {
      p = P.qpath[cur][i];
      float4 h4 = P.hit[i];
      uint32_t leaf = __float_as_uint(h4.w);
      if (leaf != 0xFFFFFFFFu) {
        V3 wi = normalize(-xyz(P.qd[cur][i]));
        Surface g = surface_at(S, leaf, h4.x, h4.y, h4.z);
        Bsdf bsdf = make_bsdf(S, g);
        float4 t4 = P.thr[p], n4 = P.rad[p];
        thr = xyz(t4);
        bounces = __float_as_uint(t4.w);
        float geom = fabsf(dot(xyz(n4), wi)) / n4.w;
        uint64_t gi = P.g0 + p;
        uint32_t sample = P.sample_begin + (uint32_t)(gi / P.npix), pixel = (uint32_t)(gi % P.npix);
        float r[4];
        lt_rng(P, pixel, sample, bounces + 1u, 0, r);
        gp = g.p; gn = g.nn; eps = g.eps;
        if (bounces + 3u >= P.min_len && bsdf.matches(BSDF_ALL & ~BSDF_DELTA)) {
          // connect_to_camera (light_renderer.cpp:146-180)
          float imp;
          if (lt_importance(P, lt_cam_point(S, g.p), fp, imp)) {
            V3 cw = normalize(cam - g.p);
            V3 f = bsdf_eval(bsdf, wi, cw, BSDF_ALL);
            if (!is_black(f)) {
              contrib = thr * geom * imp * f * (fabsf(dot(g.nn, cw)) / length_sq(cam - g.p));
              conn = !is_black(contrib);
            }
          }
        }
        uint32_t fl;
        V3 f = bsdf_sample(bsdf, wi, BSDF_ALL, r[0], r[1], r[2], wo, pdf, fl);
        V3 bc = f * geom;
        thr = thr * bc;
        go = !is_black(thr);
        if (go && bounces + 3u > P.min_len && bounces >= 2u) {
          float q = clampf(average(bc), 0.1f, 0.9f);
          if (r[3] > q) go = false;
          else thr = thr / q;
        }
        go = go && bounces + 4u <= P.max_len && pdf != 0.f;
      }
    }
    lt_emit_connection(P, conn, gp, eps, cam, contrib, fp, bounce + 1u, sh);
    uint32_t qn = block_append(&P.counters[CNT_QUEUE + bounce + 1u], go, sh);
    if (go) {
      P.qo[nxt][qn] = make_float4(gp.x, gp.y, gp.z, eps);
      P.qd[nxt][qn] = make_float4(wo.x, wo.y, wo.z, INFINITY);
      P.qpath[nxt][qn] = p;
      P.thr[p] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(bounces + 1u));
      P.rad[p] = make_float4(gn.x, gn.y, gn.z, pdf);
    }
  }
}

// Film::add_splat (film.cpp:32-67) with the 1 px box filter: the pixels whose centres
// lie within 0.5 of the position (one, or two or four on pixel edges) share the
// contribution equally.  Non-finite contributions are dropped and counted.
__global__ __launch_bounds__(256) void k_lt_splat(WaveParams P, uint32_t stage) {
  const uint32_t n = P.counters[CNT_SHADOW + stage];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 pe = P.pend[i];
    if (pe.w == 0.f) continue;
    if (!(isfinite(pe.x) && isfinite(pe.y) && isfinite(pe.z))) {
      atomicAdd((unsigned long long*)&P.stats[3], 1ull);
      continue;
    }
    float2 fp = P.fpos[i];
    int x0 = max(0, (int)ceilf(fp.x - 1.f)), x1 = min((int)P.xr - 1, (int)floorf(fp.x));
    int y0 = max(0, (int)ceilf(fp.y - 1.f)), y1 = min((int)P.yr - 1, (int)floorf(fp.y));
    if (x0 > x1 || y0 > y1) continue;
    float w = 1.f / float((x1 - x0 + 1) * (y1 - y0 + 1));
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) {
        float* f = P.film_rgb + 3 * ((size_t)y * P.xr + x);
        atomicAdd(f, pe.x * w);
        atomicAdd(f + 1, pe.y * w);
        atomicAdd(f + 2, pe.z * w);
      }
  }
}

__global__ void k_lt_stats(WaveParams P) { add_chunk_stats(P, P.n_chunk); }

hipError_t launch_trace_api(const DevScene& S, const KernelCfg& cfg, const dgpu_rays& rays,
                            const dgpu_hits* hits, uint32_t* occluded, uint32_t n, uint32_t* fetch,
                            uint32_t* spill, bool bounded, hipStream_t st) {
  ApiIO io;
  io.rays = rays;
  if (hits) io.hits = *hits; else io.hits = dgpu_hits{nullptr, nullptr, nullptr, nullptr};
  io.occluded = occluded;
  io.bvh = BvhRef{S.nodes, S.tris};
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, fetch, 1u);
  bool any = occluded != nullptr;
#define X(W, L, B)                                                                                 \
  if (cfg.width == W && cfg.lds == L && cfg.block == B) {                                          \
    if (any && bounded)                                                                            \
      hipLaunchKernelGGL((k_trace_api<true, W, L, B, true>), dim3(cfg.grid), dim3(B), 0, st, io, n,  \
                         fetch, spill, S.stack_cap);                                               \
    else if (any)                                                                                  \
      hipLaunchKernelGGL((k_trace_api<true, W, L, B, false>), dim3(cfg.grid), dim3(B), 0, st, io, n, \
                         fetch, spill, S.stack_cap);                                               \
    else if (bounded)                                                                              \
      hipLaunchKernelGGL((k_trace_api<false, W, L, B, true>), dim3(cfg.grid), dim3(B), 0, st, io, n, \
                         fetch, spill, S.stack_cap);                                               \
    else                                                                                           \
      hipLaunchKernelGGL((k_trace_api<false, W, L, B, false>), dim3(cfg.grid), dim3(B), 0, st, io, n, \
                         fetch, spill, S.stack_cap);                                               \
    return hipGetLastError();                                                                      \
  }
  DG_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_extend(const WaveParams& P, const KernelCfg& cfg, uint32_t bounce, bool batch,
                         hipStream_t st) {
  const uint32_t* count = P.counters + CNT_QUEUE + bounce;
  uint32_t* fetch = P.counters + CNT_FETCH_EXT + bounce;
  ExtendIO<false> io{P.qo[bounce & 1u], P.qd[bounce & 1u], P.qpath[bounce & 1u], P.hit,
                     BvhRef{P.S.nodes, P.S.tris}, SegTable{nullptr, nullptr, 0u}};
  ExtendIO<true> iob{P.qo[bounce & 1u], P.qd[bounce & 1u], P.qpath[bounce & 1u], P.hit,
                     BvhRef{nullptr, nullptr}, SegTable{nullptr, nullptr, 0u}};
  // camera rays (bounce 0) are traced in 64-ray parts, scattered rays with refills
  const uint32_t refill = bounce == 0u ? 0u : (uint32_t)DGPU_REFILL_LATE;
#define X(W, L, B)                                                                               \
  if (cfg.width == W && cfg.lds == L && cfg.block == B) {                                        \
    if (batch) hipLaunchKernelGGL((k_extend<W, L, B, true>), dim3(cfg.grid), dim3(B), 0, st, iob, \
                                  count, fetch, P.spill, P.stack_cap, P.batch, P.chunks, P.chunk_cur, refill); \
    else hipLaunchKernelGGL((k_extend<W, L, B, false>), dim3(cfg.grid), dim3(B), 0, st, io, count, \
                            fetch, P.spill, P.stack_cap, P.batch, P.chunks, P.chunk_cur, refill); \
    return hipGetLastError();                                                                    \
  }
  DG_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_connect(const WaveParams& P, const KernelCfg& cfg, uint32_t bounce, bool batch,
                          hipStream_t st) {
  const uint32_t* count = P.counters + CNT_SHADOW + bounce;
  uint32_t* fetch = P.counters + CNT_FETCH_SHD + bounce;
  float* pend_w = (float*)P.pend + 3;
  ConnectIO<false> io{P.so, P.sd, P.smeta, pend_w, BvhRef{P.S.nodes, P.S.tris}, SegTable{nullptr, nullptr, 0u}};
  ConnectIO<true> iob{P.so, P.sd, P.smeta, pend_w, BvhRef{nullptr, nullptr}, SegTable{nullptr, nullptr, 0u}};
#define X(W, L, B)                                                                                \
  if (cfg.width == W && cfg.lds == L && cfg.block == B) {                                         \
    if (batch) hipLaunchKernelGGL((k_connect<W, L, B, true>), dim3(cfg.grid), dim3(B), 0, st, iob, \
                                  count, fetch, P.spill, P.stack_cap, P.batch, P.chunks, P.chunk_cur); \
    else hipLaunchKernelGGL((k_connect<W, L, B, false>), dim3(cfg.grid), dim3(B), 0, st, io, count, \
                            fetch, P.spill, P.stack_cap, P.batch, P.chunks, P.chunk_cur); \
    return hipGetLastError();                                                                     \
  }
  DG_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_raygen(const WaveParams& P, uint32_t capacity, bool batch, hipStream_t st) {
  uint32_t blocks = (std::max(capacity, (uint32_t)NUM_COUNTERS) + 255u) / 256u;
  if (batch) hipLaunchKernelGGL(k_raygen<true>, dim3(blocks), dim3(256), 0, st, P);
  else hipLaunchKernelGGL(k_raygen<false>, dim3(blocks), dim3(256), 0, st, P);
  return hipGetLastError();
}

const void* shade_kernel_ptr(bool batch, bool all, bool rough, bool adj) {
  if (adj) {
    if (all) return rough ? (const void*)&k_shade<false, true, true, true> : (const void*)&k_shade<false, true, false, true>;
    return rough ? (const void*)&k_shade<false, false, true, true> : (const void*)&k_shade<false, false, false, true>;
  }
  if (batch) return rough ? (const void*)&k_shade<true, false, true> : (const void*)&k_shade<true, false, false>;
  if (all) return rough ? (const void*)&k_shade<false, true, true> : (const void*)&k_shade<false, true, false>;
  return rough ? (const void*)&k_shade<false, false, true> : (const void*)&k_shade<false, false, false>;
}

hipError_t launch_shade(const WaveParams& P, uint32_t capacity, uint32_t bounce, bool batch, hipStream_t st,
                        uint32_t max_blocks) {
  uint32_t blocks = std::min((capacity + SHADE_BLOCK - 1u) / SHADE_BLOCK, std::max(1u, max_blocks));
#define DG_SHADE(B, A, R) hipLaunchKernelGGL((k_shade<B, A, R>), dim3(blocks), dim3(SHADE_BLOCK), 0, st, P, bounce)
  const bool rough = P.rough != 0u;
  if (batch) { if (rough) DG_SHADE(true, false, true); else DG_SHADE(true, false, false); }
  else if (P.sample_all) { if (rough) DG_SHADE(false, true, true); else DG_SHADE(false, true, false); }
  else { if (rough) DG_SHADE(false, false, true); else DG_SHADE(false, false, false); }
#undef DG_SHADE
  return hipGetLastError();
}

hipError_t launch_shade_adjoint(const WaveParams& P, uint32_t capacity, uint32_t bounce, hipStream_t st,
                                uint32_t max_blocks) {
  uint32_t blocks = std::min((capacity + SHADE_BLOCK - 1u) / SHADE_BLOCK, std::max(1u, max_blocks));
#define DG_SHADE(A, R) hipLaunchKernelGGL((k_shade<false, A, R, true>), dim3(blocks), dim3(SHADE_BLOCK), 0, st, P, bounce)
  const bool rough = P.rough != 0u;
  if (P.sample_all) { if (rough) DG_SHADE(true, true); else DG_SHADE(true, false); }
  else { if (rough) DG_SHADE(false, true); else DG_SHADE(false, false); }
#undef DG_SHADE
  return hipGetLastError();
}

hipError_t launch_adj_total(const WaveParams& P, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_total, dim3((P.npix + 255u) / 256u), dim3(256), 0, st, P);
  return hipGetLastError();
}

hipError_t launch_adj_final(const WaveParams& P, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_final, dim3((P.n_chunk + 255u) / 256u), dim3(256), 0, st, P);
  return hipGetLastError();
}

hipError_t launch_finalize(const WaveParams& P, uint32_t capacity, bool batch, hipStream_t st) {
  if (batch) {
    uint32_t blocks = (capacity + 255u) / 256u;
    hipLaunchKernelGGL(k_finalize<true>, dim3(blocks), dim3(256), 0, st, P);
  } else {
    uint32_t blocks = (P.npix + 255u) / 256u;
    hipLaunchKernelGGL(k_finalize<false>, dim3(blocks), dim3(256), 0, st, P);
  }
  return hipGetLastError();
}

hipError_t launch_advance(uint32_t* cur, hipStream_t st) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, cur);
  return hipGetLastError();
}

hipError_t launch_adjoint(const WaveParams& P, const AdjArgs& A, uint32_t width, hipStream_t st) {
  uint32_t blocks = (P.n_chunk + ADJ_BLOCK - 1u) / ADJ_BLOCK;
  if (width == 8) hipLaunchKernelGGL(k_adjoint<8>, dim3(blocks), dim3(ADJ_BLOCK), 0, st, P, A);
  else if (width == 4) hipLaunchKernelGGL(k_adjoint<4>, dim3(blocks), dim3(ADJ_BLOCK), 0, st, P, A);
  else hipLaunchKernelGGL(k_adjoint<2>, dim3(blocks), dim3(ADJ_BLOCK), 0, st, P, A);
  return hipGetLastError();
}

hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u32, dim3((n + 63u) / 64u), dim3(64), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_lt_start(const WaveParams& P, hipStream_t st, uint32_t grid) {
  hipLaunchKernelGGL(k_lt_start, dim3(grid), dim3(SHADE_BLOCK), 0, st, P);
  return hipGetLastError();
}

hipError_t launch_lt_shade(const WaveParams& P, uint32_t bounce, hipStream_t st, uint32_t grid) {
  hipLaunchKernelGGL(k_lt_shade, dim3(grid), dim3(SHADE_BLOCK), 0, st, P, bounce);
  return hipGetLastError();
}

hipError_t launch_lt_splat(const WaveParams& P, uint32_t stage, hipStream_t st, uint32_t grid) {
  hipLaunchKernelGGL(k_lt_splat, dim3(grid), dim3(256), 0, st, P, stage);
  return hipGetLastError();
}

hipError_t launch_lt_stats(const WaveParams& P, hipStream_t st) {
  hipLaunchKernelGGL(k_lt_stats, dim3(1), dim3(64), 0, st, P);
  return hipGetLastError();
}

}  // namespace dgpu

