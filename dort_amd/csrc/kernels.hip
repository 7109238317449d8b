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
#include "dgpu_internal.hpp"
#include "traverse.hpp"

namespace dgpu {

// ----------------------------------------------------------------- wave helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Appends `pred` lanes to a queue: one atomic per wave, ranks by ballot + mbcnt.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return 0u;
  uint32_t lane = lane_id();
  uint32_t leader = (uint32_t)__ffsll((long long)mask) - 1u;
  uint32_t base = 0u;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  unsigned long long below = lane == 0 ? 0ull : (mask & ((1ull << lane) - 1ull));
  return base + (uint32_t)__popcll(below);
}

// ----------------------------------------------------------------- trace kernels

template <bool ANY, int WIDTH>
__device__ __forceinline__ bool trace_one(const DevScene& S, const TraceRay& r, LaneStack& st,
                                          TraceHit& h) {
  if (WIDTH == 4) return traverse_bvh4<ANY>(S.nodes, S.tris, r, st, h);
  return traverse_bvh2<ANY>(S.nodes, S.tris, r, st, h);
}

// Shared persistent loop: `io.load(i, ray)` / `io.store(i, hit, found)`.
template <bool ANY, int WIDTH, int LDS, int BLOCK, class IO>
__device__ __forceinline__ void trace_loop(const DevScene& S, IO& io, uint32_t n, uint32_t* fetch,
                                           uint32_t* spill) {
  __shared__ uint32_t stk[LDS * BLOCK];
  LaneStack st;
  st.lds = stk + threadIdx.x;
  st.lds_stride = BLOCK;
  st.lds_cap = LDS;
  size_t gthreads = (size_t)gridDim.x * BLOCK;
  st.spill = spill + (size_t)blockIdx.x * BLOCK + threadIdx.x;
  st.spill_stride = gthreads;
  st.cap = S.stack_cap;
  for (;;) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(fetch, 64u);
    base = __shfl(base, 0);
    if (base >= n) break;
    uint32_t i = base + lane_id();
    if (i < n) {
      TraceRay r;
      io.load(i, r);
      TraceHit h;
      bool found = trace_one<ANY, WIDTH>(S, r, st, h);
      io.store(i, h, found);
    }
  }
}

struct ApiIO {
  dgpu_rays rays;
  dgpu_hits hits;
  uint32_t* occluded;
  const float4* tris;
  __device__ __forceinline__ void load(uint32_t i, TraceRay& r) const {
    r.o = v3(rays.ox[i], rays.oy[i], rays.oz[i]);
    r.d = v3(rays.dx[i], rays.dy[i], rays.dz[i]);
    r.tmin = rays.tmin[i];
    r.tmax = rays.tmax[i];
    r.ignore = 0xFFFFFFFFu;
  }
  __device__ __forceinline__ void store(uint32_t i, const TraceHit& h, bool found) const {
    if (occluded) { occluded[i] = found ? 1u : 0u; return; }
    hits.t[i] = found ? h.t : INFINITY;
    hits.prim[i] = found ? (int32_t)__float_as_uint(tris[3 * h.tri].w) : -1;
    hits.b1[i] = found ? h.b1 : 0.f;
    hits.b2[i] = found ? h.b2 : 0.f;
  }
};

template <bool ANY, int WIDTH, int LDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_api(DevScene S, ApiIO io, uint32_t n,
                                                     uint32_t* fetch, uint32_t* spill) {
  trace_loop<ANY, WIDTH, LDS, BLOCK>(S, io, n, fetch, spill);
}

struct ExtendIO {
  const float4* qo;
  const float4* qd;
  float4* hit;
  __device__ __forceinline__ void load(uint32_t i, TraceRay& r) const {
    float4 o = qo[i], d = qd[i];
    r.o = xyz(o); r.tmin = o.w;
    r.d = xyz(d); r.tmax = d.w;
    r.ignore = 0xFFFFFFFFu;
  }
  __device__ __forceinline__ void store(uint32_t i, const TraceHit& h, bool found) const {
    hit[i] = make_float4(found ? h.t : INFINITY, h.b1, h.b2, __uint_as_float(found ? h.tri : 0xFFFFFFFFu));
  }
};

struct ConnectIO {
  const float4* so;
  const float4* sd;
  const uint2* meta;   // (pending slot, ignored triangle)
  uint8_t* vis;
  __device__ __forceinline__ void load(uint32_t i, TraceRay& r) const {
    float4 o = so[i], d = sd[i];
    r.o = xyz(o); r.tmin = o.w;
    r.d = xyz(d); r.tmax = d.w;
    r.ignore = meta[i].y;
  }
  __device__ __forceinline__ void store(uint32_t i, const TraceHit&, bool found) const {
    if (found) vis[meta[i].x] = 0;
  }
};

template <int WIDTH, int LDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_extend(DevScene S, ExtendIO io, const uint32_t* count,
                                                  uint32_t* fetch, uint32_t* spill) {
  trace_loop<false, WIDTH, LDS, BLOCK>(S, io, *count, fetch, spill);
}

template <int WIDTH, int LDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_connect(DevScene S, ConnectIO io, const uint32_t* count,
                                                   uint32_t* fetch, uint32_t* spill) {
  trace_loop<true, WIDTH, LDS, BLOCK>(S, io, *count, fetch, spill);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ----------------------------------------------------------------- path state

// thr.w holds flags: bits 0..15 bounces, bit 16 last bounce delta, bit 17 no film weight.
constexpr uint32_t F_DELTA = 1u << 16;
constexpr uint32_t F_NOFILM = 1u << 17;

__global__ void k_raygen(WaveParams P) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NUM_COUNTERS) P.counters[i] = (i == CNT_QUEUE) ? P.n_chunk : 0u;
  if (i >= P.n_chunk) return;
  uint64_t g = P.g0 + i;
  uint32_t sample = P.sample_begin + (uint32_t)(g / P.npix);
  uint32_t pixel = (uint32_t)(g % P.npix);
  uint32_t x = pixel % P.xr, y = pixel / P.xr;
  float jx = 0.5f, jy = 0.5f;
  if (sample != 0) {  // iteration 0 samples pixel centres (path_renderer.cpp:17)
    float r[4];
    dgpu_rng_block(P.k0, P.k1, pixel, sample, 0, 2, r);
    jx = r[2]; jy = r[3];
  }
  uint32_t flags = (jx == 0.f || jy == 0.f) ? F_NOFILM : 0u;
  V3 o, d;
  camera_ray(P.S, float(x) + jx, float(y) + jy, P.xr, P.yr, o, d);
  P.qo[0][i] = make_float4(o.x, o.y, o.z, 0.f);
  P.qd[0][i] = make_float4(d.x, d.y, d.z, INFINITY);
  P.qpath[0][i] = i;
  P.thr[i] = make_float4(1.f, 1.f, 1.f, __uint_as_float(flags));
  P.rad[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t k = 0; k < P.slots; ++k) P.vis[(size_t)i * P.slots + k] = 0;
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
    float eu, ev;
    wi = env_sample(S, u, v, eu, ev, pdf);
    sh.o = p; sh.d = normalize(wi); sh.tmin = eps; sh.tmax = INFINITY; sh.ignore = 0xFFFFFFFFu;
    return env_radiance(S, eu, ev);
  } else {
    float z = 1.f - 2.f * u, r = sqrtf(fmaxf(0.f, 1.f - z * z)), phi = TWO_PI_F * v;
    wi = v3(cosf(phi) * r, sinf(phi) * r, z);
    pdf = INV_FOUR_PI_F;
    sh.o = p; sh.d = normalize(wi); sh.tmin = eps; sh.tmax = INFINITY; sh.ignore = 0xFFFFFFFFu;
    return xyz(L.radiance);
  }
}

// Intersects the ray (p, w) with an emitter triangle, t in [0, inf) (shape.cpp:14-26).
__device__ __forceinline__ bool light_tri_hit(const DevScene& S, uint32_t tri, V3 p, V3 w, float& t) {
  WtRay wr = wt_setup(p, w);
  float b1, b2;
  return wt_hit(wr, S.tris[3 * tri], S.tris[3 * tri + 1], S.tris[3 * tri + 2], 0.f, INFINITY, t, b1, b2);
}

__device__ __forceinline__ V3 light_tri_normal(const DevScene& S, uint32_t tri) {
  float4 a = S.tris[3 * tri], b = S.tris[3 * tri + 1], c = S.tris[3 * tri + 2];
  V3 p0 = xyz(a);
  return normalize(cross(xyz(c) - p0, xyz(b) - p0));
}

// estimate_direct (path_renderer.cpp:135-206) for one light; writes up to two shadow rays.
__device__ __forceinline__ void estimate_direct(const WaveParams& P, const Surface& g, V3 wo,
                                                const Bsdf& bsdf, uint32_t li, float scale, V3 thr,
                                                const float rb0[4], const float rb1[4],
                                                NeeOut& nb, NeeOut& nl) {
  const DevScene& S = P.S;
  const DevLight L = S.lights[li];
  bool use_bsdf = P.strategy != DGPU_DIRECT_LIGHT;
  bool use_light = P.strategy != DGPU_DIRECT_BSDF;
  nb.emit = false;
  nl.emit = false;
  const uint32_t NON_DELTA = BSDF_ALL & ~BSDF_DELTA;
  if (use_bsdf) {  // estimate_direct_from_bsdf, path_renderer.cpp:149-182
    V3 wi; float pdf; uint32_t fl;
    V3 f = bsdf_sample(bsdf, wo, NON_DELTA, rb0[1], rb0[2], rb0[3], wi, pdf, fl);
    if (pdf != 0.f && !is_black(f)) {
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
        rad = env_background(S, wi);
        lp = env_dir_pdf(S, wi);
      } else {
        rad = xyz(L.radiance);
        lp = INV_FOUR_PI_F;
      }
      if (!is_black(rad)) {
        float w = use_light ? pdf / (pdf + lp) : 1.f;
        nb.emit = true;
        nb.o = g.p; nb.d = wi; nb.tmin = g.eps; nb.tmax = tmax; nb.ignore = ignore;
        nb.contrib = thr * f * rad * (w * fabsf(dot(g.nn, wi)) / pdf * scale);
      }
    }
  }
  if (use_light) {  // estimate_direct_from_light, path_renderer.cpp:184-206
    V3 wi; float pdf;
    V3 rad = light_sample(S, L, g.p, g.eps, rb1[0], rb1[1], wi, pdf, nl);
    if (pdf != 0.f && !is_black(rad)) {
      V3 f = bsdf_eval(bsdf, wi, wo, NON_DELTA);
      if (!is_black(f)) {
        float w = 1.f;
        if (use_bsdf) {
          float bp = bsdf_pdf(bsdf, wi, wo, NON_DELTA);
          w = pdf / (pdf + bp);
        }
        nl.emit = true;
        nl.contrib = thr * f * rad * (w * fabsf(dot(g.nn, wi)) / pdf * scale);
      }
    }
  }
}

__device__ __forceinline__ void emit_shadow(const WaveParams& P, const NeeOut& s, uint32_t slot,
                                            uint32_t qi) {
  P.so[qi] = make_float4(s.o.x, s.o.y, s.o.z, s.tmin);
  P.sd[qi] = make_float4(s.d.x, s.d.y, s.d.z, s.tmax);
  P.smeta[qi] = make_uint2(slot, s.ignore);
}

// One bounce of PathRenderer::sample for every path in the current queue.
// Every lane of a wave runs the same number of queue appends (inactive lanes
// append nothing), so the ballot-based compaction sees the whole wave.
__global__ __launch_bounds__(256) void k_shade(WaveParams P, uint32_t bounce) {
  const uint32_t cur = bounce & 1u, nxt = cur ^ 1u;
  const DevScene& S = P.S;
  const uint32_t SL = P.slots;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t n = P.counters[CNT_QUEUE + bounce];
  bool active = i < n;
  uint32_t p = 0, flags = 0, bounces = 0;
  V3 thr = v3(0.f, 0.f, 0.f), rad = v3(0.f, 0.f, 0.f), wo = v3(0.f, 0.f, 1.f);
  Surface g;
  Bsdf bsdf;
  bsdf.kind = DGPU_MAT_NONE;
  bool shading = false;
  float rb0[4] = {0.f, 0.f, 0.f, 0.f}, rb1[4] = {0.f, 0.f, 0.f, 0.f}, rb2[4] = {0.f, 0.f, 0.f, 0.f};
  if (active) {
    p = P.qpath[cur][i];
    float4 t4 = P.thr[p], r4 = P.rad[p];
    thr = xyz(t4);
    rad = xyz(r4);
    flags = __float_as_uint(t4.w);
    // NEE of the previous vertex, now that its visibility is known
    for (uint32_t k = 0; k < SL; ++k) {
      size_t s = (size_t)p * SL + k;
      if (P.vis[s]) { rad = rad + xyz(P.pend[s]); P.vis[s] = 0; }
    }
    bounces = flags & 0xFFFFu;
    bool last_delta = (flags & F_DELTA) != 0u;
    float4 o4 = P.qo[cur][i], d4 = P.qd[cur][i], h4 = P.hit[i];
    V3 ro = xyz(o4), rd = xyz(d4);
    uint32_t leaf = __float_as_uint(h4.w);
    bool is = leaf != 0xFFFFFFFFu;
    if (is) g = surface_at(S, leaf, h4.x, h4.y, h4.z);
    if ((bounces == 0u || last_delta) && bounces >= P.min_depth && bounces <= P.max_depth) {
      if (is) {
        int32_t li = S.tri_light[g.tri];
        // Intersection::eval_radiance -> DiffuseLight::eval_radiance (diffuse_light.cpp:83-93)
        if (li >= 0 && dot(g.nn, ro - g.p) >= 0.f) rad = rad + thr * xyz(S.lights[li].radiance);
      } else {
        rad = rad + thr * background_radiance(S, rd);
      }
    }
    bounces += 1u;
    if (is && bounces <= P.max_depth) {
      shading = true;
      wo = normalize(-rd);
      bsdf = make_bsdf(S, g);
      uint64_t gidx = P.g0 + p;
      uint32_t sample = P.sample_begin + (uint32_t)(gidx / P.npix);
      uint32_t pixel = (uint32_t)(gidx % P.npix);
      dgpu_rng_block(P.k0, P.k1, pixel, sample, bounce, 0, rb0);
      dgpu_rng_block(P.k0, P.k1, pixel, sample, bounce, 1, rb1);
      dgpu_rng_block(P.k0, P.k1, pixel, sample, bounce, 2, rb2);
    }
  }
  // direct lighting (path_renderer.cpp:68-76, 113-132): up to 2 shadow rays per light
  bool do_nee = shading && bounces >= P.min_depth && bsdf.matches(BSDF_ALL & ~BSDF_DELTA) &&
                S.num_lights > 0u;
  uint32_t lights_loop = P.sample_all ? S.num_lights : 1u;
  for (uint32_t li = 0; li < lights_loop; ++li) {
    NeeOut nb, nl;
    nb.emit = false;
    nl.emit = false;
    if (do_nee) {
      if (P.sample_all) {
        estimate_direct(P, g, wo, bsdf, li, 1.f, thr, rb0, rb1, nb, nl);
      } else {
        // DiscreteDistrib1d::sample / pdf (discrete_distrib_1d.cpp:22-42)
        float u = rb0[0];
        uint32_t b = 0, e = S.num_lights;
        while (b + 1u < e) {
          uint32_t m = b + (e - b) / 2u;
          if (S.light_cdf[m] < u) b = m; else e = m;
        }
        float pp = S.light_cdf[b + 1] - S.light_cdf[b];
        if (pp != 0.f) estimate_direct(P, g, wo, bsdf, b, 1.f / pp, thr, rb0, rb1, nb, nl);
      }
    }
    uint32_t slot_b = (uint32_t)((size_t)p * SL + 2 * li), slot_l = slot_b + 1;
    uint32_t qb = wave_append(&P.counters[CNT_SHADOW + bounce], nb.emit);
    if (nb.emit) {
      P.pend[slot_b] = make_float4(nb.contrib.x, nb.contrib.y, nb.contrib.z, 0.f);
      P.vis[slot_b] = 1;
      emit_shadow(P, nb, slot_b, qb);
    }
    uint32_t ql = wave_append(&P.counters[CNT_SHADOW + bounce], nl.emit);
    if (nl.emit) {
      P.pend[slot_l] = make_float4(nl.contrib.x, nl.contrib.y, nl.contrib.z, 0.f);
      P.vis[slot_l] = 1;
      emit_shadow(P, nl, slot_l, ql);
    }
  }
  // continuation: Bsdf::sample_light_f, throughput, Russian roulette (path_renderer.cpp:78-107)
  bool cont = false;
  V3 next_o = v3(0.f, 0.f, 0.f), next_d = v3(0.f, 0.f, 1.f);
  float next_eps = 0.f;
  if (shading) {
    uint32_t req = (P.only_direct || bounces == P.max_depth) ? (BSDF_MODES | BSDF_DELTA) : BSDF_ALL;
    V3 wi; float pdf; uint32_t fl;
    V3 f = bsdf_sample(bsdf, wo, req, rb1[2], rb1[3], rb2[0], wi, pdf, fl);
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
        flags = (flags & ~(0xFFFFu | F_DELTA)) | bounces | (delta ? F_DELTA : 0u);
      }
    }
  }
  if (active) {
    P.thr[p] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(flags));
    P.rad[p] = make_float4(rad.x, rad.y, rad.z, 0.f);
  }
  uint32_t qn = wave_append(&P.counters[CNT_QUEUE + bounce + 1], cont);
  if (cont) {
    P.qo[nxt][qn] = make_float4(next_o.x, next_o.y, next_o.z, next_eps);
    P.qd[nxt][qn] = make_float4(next_d.x, next_d.y, next_d.z, INFINITY);
    P.qpath[nxt][qn] = p;
  }
}

// Resolves the last pending NEE of every path and adds the chunk's samples to the
// film pixel by pixel in sample order (Film::add_sample with the r = 0.5 box filter;
// non-finite or negative samples are dropped, renderer.cpp:36-41).
__global__ __launch_bounds__(256) void k_finalize(WaveParams P) {
  uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint64_t ext = 0, shd = 0;
    for (uint32_t b = 0; b < MAX_BOUNCES; ++b) {
      ext += P.counters[CNT_QUEUE + b];
      shd += P.counters[CNT_SHADOW + b];
    }
    P.stats[0] += P.n_chunk;
    P.stats[1] += ext;
    P.stats[2] += shd;
  }
  if (px >= P.npix) return;
  uint64_t g_end = P.g0 + P.n_chunk;
  uint64_t g = P.g0 + ((px + P.npix - (uint32_t)(P.g0 % P.npix)) % P.npix);
  V3 sum = v3(0.f, 0.f, 0.f);
  float wsum = 0.f;
  uint32_t dropped = 0;
  for (; g < g_end; g += P.npix) {
    uint32_t i = (uint32_t)(g - P.g0);
    float4 r4 = P.rad[i];
    V3 rad = xyz(r4);
    for (uint32_t k = 0; k < P.slots; ++k) {
      size_t s = (size_t)i * P.slots + k;
      if (P.vis[s]) rad = rad + xyz(P.pend[s]);
    }
    bool ok = isfinite(rad.x) && isfinite(rad.y) && isfinite(rad.z) && rad.x >= 0.f &&
              rad.y >= 0.f && rad.z >= 0.f;
    if (!ok) { ++dropped; continue; }
    uint32_t flags = __float_as_uint(P.thr[i].w);
    if (flags & F_NOFILM) continue;
    sum = sum + rad;
    wsum += 1.f;
  }
  if (wsum > 0.f) {
    P.film_rgb[3 * (size_t)px] += sum.x;
    P.film_rgb[3 * (size_t)px + 1] += sum.y;
    P.film_rgb[3 * (size_t)px + 2] += sum.z;
    P.film_w[px] += wsum;
  }
  if (dropped) atomicAdd((unsigned long long*)&P.stats[3], (unsigned long long)dropped);
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

constexpr uint32_t ADJ_LDS_MATS = 64;
constexpr uint32_t ADJ_BLOCK = 64;

__device__ __forceinline__ void adj_deposit(const AdjSink& k, const Bsdf& b, int32_t texel, uint32_t mat,
                                            V3 wi, V3 wo, V3 carried) {
  if (b.kind == DGPU_MAT_LAMBERT && texel >= 0 && k.g_tex) {
    // f = albedo / pi: d ln f_c / d albedo_c = 1 / albedo_c
    float* g = k.g_tex + 3 * (size_t)texel;
    if (b.albedo.x > 0.f) atomicAdd(g + 0, k.dl[0] * carried.x / b.albedo.x);
    if (b.albedo.y > 0.f) atomicAdd(g + 1, k.dl[1] * carried.y / b.albedo.y);
    if (b.albedo.z > 0.f) atomicAdd(g + 2, k.dl[2] * carried.z / b.albedo.z);
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
    bool is = trace_one<false, WIDTH>(S, ray, st, h);
    Surface g;
    if (is) g = surface_at(S, h.tri, h.t, h.b1, h.b2);
    if ((bounces == 0u || last_delta) && bounces >= P.min_depth && bounces <= P.max_depth) {
      if (is) {
        int32_t li = S.tri_light[g.tri];
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
    uint32_t mat = S.tri_mat[g.tri];
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
          estimate_direct(P, g, wo, bsdf, li, 1.f, thr, rb0, rb1, nb, nl);
        } else {
          float u = rb0[0];
          uint32_t b = 0, e = S.num_lights;
          while (b + 1u < e) {
            uint32_t m = b + (e - b) / 2u;
            if (S.light_cdf[m] < u) b = m; else e = m;
          }
          float pp = S.light_cdf[b + 1] - S.light_cdf[b];
          if (pp != 0.f) estimate_direct(P, g, wo, bsdf, b, 1.f / pp, thr, rb0, rb1, nb, nl);
        }
        NeeOut* outs[2] = {&nb, &nl};
        for (int k = 0; k < 2; ++k) {
          const NeeOut& n = *outs[k];
          if (!n.emit) continue;
          TraceRay sr;
          sr.o = n.o; sr.d = n.d; sr.tmin = n.tmin; sr.tmax = n.tmax; sr.ignore = n.ignore;
          TraceHit sh;
          if (trace_one<true, WIDTH>(S, sr, st, sh)) continue;
          rad = rad + n.contrib;
          if (ADJ) adj_deposit(sink, bsdf, texel, mat, n.d, wo, n.contrib);
        }
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

template <bool ANY, int WIDTH, int LDS, int BLOCK>
static hipError_t launch_api_t(const DevScene& S, const ApiIO& io, uint32_t n, uint32_t* fetch,
                               uint32_t* spill, uint32_t grid, hipStream_t st) {
  hipLaunchKernelGGL((k_trace_api<ANY, WIDTH, LDS, BLOCK>), dim3(grid), dim3(BLOCK), 0, st, S, io, n,
                     fetch, spill);
  return hipGetLastError();
}

template <int WIDTH, int LDS, int BLOCK>
static hipError_t launch_ext_t(const DevScene& S, const ExtendIO& io, const uint32_t* count,
                               uint32_t* fetch, uint32_t* spill, uint32_t grid, hipStream_t st) {
  hipLaunchKernelGGL((k_extend<WIDTH, LDS, BLOCK>), dim3(grid), dim3(BLOCK), 0, st, S, io, count, fetch,
                     spill);
  return hipGetLastError();
}

template <int WIDTH, int LDS, int BLOCK>
static hipError_t launch_con_t(const DevScene& S, const ConnectIO& io, const uint32_t* count,
                               uint32_t* fetch, uint32_t* spill, uint32_t grid, hipStream_t st) {
  hipLaunchKernelGGL((k_connect<WIDTH, LDS, BLOCK>), dim3(grid), dim3(BLOCK), 0, st, S, io, count, fetch,
                     spill);
  return hipGetLastError();
}

// Dispatch over the compiled (width, LDS stack, block) variants.
#define DG_VARIANTS(X)          \
  X(2, 8, 128) X(2, 16, 128) X(2, 32, 128) X(2, 16, 64) X(2, 16, 256) \
  X(4, 8, 128) X(4, 16, 128) X(4, 32, 128) X(4, 16, 64) X(4, 16, 256)

bool variant_supported(uint32_t width, uint32_t lds, uint32_t block) {
#define X(W, L, B) if (width == W && lds == L && block == B) return true;
  DG_VARIANTS(X)
#undef X
  return false;
}

const void* trace_kernel_ptr(int which, uint32_t width, uint32_t lds, uint32_t block) {
#define X(W, L, B)                                                                         \
  if (width == W && lds == L && block == B) {                                              \
    if (which == 0) return (const void*)&k_trace_api<false, W, L, B>;                      \
    if (which == 1) return (const void*)&k_trace_api<true, W, L, B>;                       \
    if (which == 2) return (const void*)&k_extend<W, L, B>;                                \
    return (const void*)&k_connect<W, L, B>;                                               \
  }
  DG_VARIANTS(X)
#undef X
  return nullptr;
}

hipError_t launch_trace_api(const DevScene& S, const KernelCfg& cfg, const dgpu_rays& rays,
                            const dgpu_hits* hits, uint32_t* occluded, uint32_t n, uint32_t* fetch,
                            uint32_t* spill, hipStream_t st) {
  ApiIO io;
  io.rays = rays;
  if (hits) io.hits = *hits; else io.hits = dgpu_hits{nullptr, nullptr, nullptr, nullptr};
  io.occluded = occluded;
  io.tris = S.tris;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, fetch, 1u);
  bool any = occluded != nullptr;
#define X(W, L, B)                                                                          \
  if (cfg.width == W && cfg.lds == L && cfg.block == B)                                    \
    return any ? launch_api_t<true, W, L, B>(S, io, n, fetch, spill, cfg.grid, st)           \
               : launch_api_t<false, W, L, B>(S, io, n, fetch, spill, cfg.grid, st);
  DG_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_extend(const WaveParams& P, const KernelCfg& cfg, uint32_t bounce, hipStream_t st) {
  ExtendIO io{P.qo[bounce & 1u], P.qd[bounce & 1u], P.hit};
  const uint32_t* count = P.counters + CNT_QUEUE + bounce;
  uint32_t* fetch = P.counters + CNT_FETCH_EXT + bounce;
#define X(W, L, B)                                           \
  if (cfg.width == W && cfg.lds == L && cfg.block == B)     \
    return launch_ext_t<W, L, B>(P.S, io, count, fetch, P.spill, cfg.grid, st);
  DG_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_connect(const WaveParams& P, const KernelCfg& cfg, uint32_t bounce, hipStream_t st) {
  ConnectIO io{P.so, P.sd, P.smeta, P.vis};
  const uint32_t* count = P.counters + CNT_SHADOW + bounce;
  uint32_t* fetch = P.counters + CNT_FETCH_SHD + bounce;
#define X(W, L, B)                                           \
  if (cfg.width == W && cfg.lds == L && cfg.block == B)     \
    return launch_con_t<W, L, B>(P.S, io, count, fetch, P.spill, cfg.grid, st);
  DG_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_raygen(const WaveParams& P, hipStream_t st) {
  uint32_t blocks = (std::max(P.n_chunk, (uint32_t)NUM_COUNTERS) + 255u) / 256u;
  hipLaunchKernelGGL(k_raygen, dim3(blocks), dim3(256), 0, st, P);
  return hipGetLastError();
}

hipError_t launch_shade(const WaveParams& P, uint32_t bounce, hipStream_t st) {
  uint32_t blocks = (P.n_chunk + 255u) / 256u;
  hipLaunchKernelGGL(k_shade, dim3(blocks), dim3(256), 0, st, P, bounce);
  return hipGetLastError();
}

hipError_t launch_finalize(const WaveParams& P, hipStream_t st) {
  uint32_t blocks = (P.npix + 255u) / 256u;
  hipLaunchKernelGGL(k_finalize, dim3(blocks), dim3(256), 0, st, P);
  return hipGetLastError();
}

hipError_t launch_adjoint(const WaveParams& P, const AdjArgs& A, uint32_t width, hipStream_t st) {
  uint32_t blocks = (P.n_chunk + ADJ_BLOCK - 1u) / ADJ_BLOCK;
  if (width == 4) hipLaunchKernelGGL(k_adjoint<4>, dim3(blocks), dim3(ADJ_BLOCK), 0, st, P, A);
  else hipLaunchKernelGGL(k_adjoint<2>, dim3(blocks), dim3(ADJ_BLOCK), 0, st, P, A);
  return hipGetLastError();
}

}  // namespace dgpu
