// bdpt.hip -- dort's bidirectional path tracer (BdptRenderer, src/dort/bdpt_renderer.cpp)
// on the GPU.
//
// One thread per path sample (pixel, iteration).  It draws the light subpath
// (random_light_walk, bdpt_renderer.cpp:118-181) and the camera subpath
// (random_camera_walk, :183-273) into vertex arrays in HBM, then forms every strategy of
// sample_path (:43-116): s light vertices joined to t camera vertices, the contribution
// of path_contrib (:275-466) times the weight of path_weight (:468-667), dort's balance
// heuristic computed from the ratios of the vertices' forward and backward area pdfs.
// Strategies with t >= 2 add to the pixel's sample; the per-path sums are staged and
// added to the film in sample order (k_bdpt_film), so the film does not depend on the
// chunking.  Strategies with t = 1 reach the camera from the light subpath and are
// splatted with float atomics (Film::add_splat).  Rays are traced inline with the
// traversal of the wavefront kernels (traverse.hpp, per-lane LDS + spill stack).
//
// Random numbers: Philox blocks keyed by (pixel, sample, stage, block) in a domain of
// their own.  Stage 0: block 0 film jitter x, y and the light pick; block 1 the light
// ray.  Stage 1 + b: block 0 light bounce b (BSDF uvc, Russian roulette), block 1 camera
// bounce b (Russian roulette, BSDF uvc), block 2 its background-light pick, block 3 the
// fresh light vertex of the strategies s = 1 at camera vertex b (pick, position).
#include <algorithm>

#include "dgpu_internal.hpp"
#include "light_paths.hpp"
#include "traverse.hpp"
#include "wave_ops.hpp"

namespace dgpu {

constexpr uint32_t BD_BLOCK = 64;
constexpr uint32_t BD_DOMAIN = 0x62647074u;

// vertex kinds (bits 1-2 of BdV::info)
enum : uint32_t { BV_LIGHT = 0, BV_SURFACE = 1, BV_CAMERA = 2, BV_BACKGROUND = 3 };

// BdptRenderer::Vertex (bdpt_renderer.hpp:44-81).  A surface vertex keeps its hit
// (leaf-order triangle, barycentrics) instead of a Bsdf; the Bsdf is rebuilt from it
// where a connection needs it.
struct BdV {
  V3 p;
  float eps;
  V3 nn;
  float fwd;     // fwd_pdf
  V3 alpha;
  float bwd;     // bwd_pdf
  uint32_t leaf;
  float b1, b2;
  uint32_t info;  // bit 0 delta, bits 1-2 kind, bits 8-31: BV_SURFACE area light + 1 (0: none),
                  // BV_BACKGROUND the background light's index
};

DG_INLINE uint32_t bv_kind(const BdV& v) { return (v.info >> 1) & 3u; }
DG_INLINE bool bv_delta(const BdV& v) { return (v.info & 1u) != 0u; }

// Vertex k of path i: component c at [(4 k + c) n + i], so a wave's lanes read and write
// consecutive 16-byte records.
DG_INLINE void bd_store(float4* base, uint32_t n, uint32_t i, uint32_t k, const BdV& v) {
  size_t o = (size_t)4 * k * n + i;
  base[o] = make_float4(v.p.x, v.p.y, v.p.z, v.eps);
  base[o + n] = make_float4(v.nn.x, v.nn.y, v.nn.z, v.fwd);
  base[o + 2 * (size_t)n] = make_float4(v.alpha.x, v.alpha.y, v.alpha.z, v.bwd);
  base[o + 3 * (size_t)n] = make_float4(__uint_as_float(v.leaf), v.b1, v.b2, __uint_as_float(v.info));
}

DG_INLINE BdV bd_load(const float4* base, uint32_t n, uint32_t i, uint32_t k) {
  size_t o = (size_t)4 * k * n + i;
  float4 a = base[o], b = base[o + n], c = base[o + 2 * (size_t)n], d = base[o + 3 * (size_t)n];
  BdV v;
  v.p = xyz(a); v.eps = a.w;
  v.nn = xyz(b); v.fwd = b.w;
  v.alpha = xyz(c); v.bwd = c.w;
  v.leaf = __float_as_uint(d.x); v.b1 = d.y; v.b2 = d.z; v.info = __float_as_uint(d.w);
  return v;
}

__device__ __forceinline__ void bd_rng(const WaveParams& P, uint32_t pixel, uint32_t sample, uint32_t stage,
                                       uint32_t blk, float out[4]) {
  dgpu_u32x4 r = dgpu_philox4x32_10(pixel, sample, (stage << 2) | blk, BD_DOMAIN, P.k0, P.k1);
  for (int k = 0; k < 4; ++k) out[k] = dgpu_u01(r.v[k]);
}

DG_INLINE Bsdf bd_bsdf(const DevScene& S, const BdV& v) {
  return make_bsdf(S, surface_at(S, v.leaf, 0.f, v.b1, v.b2));
}

// A Bsdf with no lobe (a vertex that is not on a surface): zero f and pdf.
DG_INLINE Bsdf bd_no_bsdf() {
  Bsdf b;
  b.nn = b.sn = b.c0 = b.c1 = b.c2 = v3(0.f, 0.f, 0.f);
  b.kind = DGPU_MAT_NONE;
  b.a0 = b.a1 = 0.f;
  return b;
}

// ---------------------------------------------------------------- lights and camera

// Light::pivot_radiance_pdf (diffuse_light.cpp:104-110 with Shape::point_pivot_pdf,
// shape.cpp:14-26; environment_light.cpp:85-87; infinite_light.cpp:49-51): the solid-angle
// pdf of light sampling direction wi from pivot.
__device__ __forceinline__ float bd_pivot_pdf(const DevScene& S, const DevLight& L, V3 wi, V3 pivot) {
  if (L.kind == LK_AREA) {
    float t;
    if (!light_tri_hit(S, L.tri, pivot, wi, t)) return 0.f;
    V3 pt = pivot + wi * t;
    return (1.f / L.area) * length_sq(pt - pivot) / fabsf(dot(wi, light_tri_normal(S, L.tri)));
  }
  if (L.kind == LK_ENV) {
    float pdf;
    env_eval(S, wi, pdf);
    return pdf;
  }
  return INV_FOUR_PI_F;
}

// Light::ray_radiance_pdf (diffuse_light.cpp:95-102, environment_light.cpp:79-83,
// infinite_light.cpp:43-47): pdf of sample_ray_radiance giving the ray (origin, dir) from
// a light point with normal nn, w.r.t. area x solid angle.
__device__ __forceinline__ float bd_ray_pdf(const DevScene& S, const DevLight& L, V3 dir, V3 nn) {
  if (L.kind == LK_AREA) return dot(dir, nn) <= 0.f ? 0.f : (1.f / L.area) * INV_TWO_PI_F;
  if (L.kind == LK_ENV) {
    float pdf;
    env_eval(S, -dir, pdf);
    return pdf * INV_PI_F / sqf(S.radius);
  }
  return INV_PI_F * INV_FOUR_PI_F / sqf(S.radius);
}

struct BdPivot {
  V3 rad, wi, p, nn;
  float eps, pdf;
  TraceRay shadow;
};

// Light::sample_pivot_radiance with the point sampled (diffuse_light.cpp:40-65 with
// Shape::sample_point_pivot, shape.cpp:4-12; environment_light.cpp:54-77;
// infinite_light.cpp:25-36) and its shadow test (ShadowTest::init_point_point /
// init_point_dir, light.cpp:40-58).
__device__ __forceinline__ BdPivot bd_sample_pivot(const DevScene& S, const DevLight& L, V3 pivot, float eps,
                                                   float u, float v) {
  BdPivot o;
  o.rad = v3(0.f, 0.f, 0.f);
  o.p = o.nn = o.rad;
  o.eps = 0.f;
  o.shadow.ignore = 0xFFFFFFFFu;
  o.shadow.o = pivot;
  o.shadow.tmin = eps;
  if (L.kind == LK_AREA) {
    V3 n;
    V3 pt = lt_area_point(S, L, u, v, n);
    o.pdf = (1.f / L.area) * length_sq(pivot - pt) / fabsf(dot(n, normalize(pivot - pt)));
    o.wi = normalize(pt - pivot);
    o.p = pt;
    o.nn = n;
    o.eps = 1e-3f;  // Triangle::sample_point (triangle.cpp:74)
    float dist = length(pt - pivot);
    o.shadow.d = (pt - pivot) / dist;
    o.shadow.tmax = dist - o.eps;
    if (dot(o.wi, n) < 0.f) o.rad = xyz(L.radiance);
  } else {
    if (L.kind == LK_ENV) {
      o.wi = env_sample(S, u, v, o.pdf, o.rad);
    } else {
      float z = 1.f - 2.f * u, r = sqrtf(fmaxf(0.f, 1.f - z * z)), sp, cp;  // uniform_sphere_sample
      sincos_small(TWO_PI_F * v, sp, cp);
      o.wi = v3(cp * r, sp * r, z);
      o.pdf = INV_FOUR_PI_F;
      o.rad = xyz(L.radiance);
    }
    o.shadow.d = normalize(o.wi);
    o.shadow.tmax = INFINITY;
  }
  return o;
}

// PinholeCamera::sample_ray_importance's direction pdf (and importance) at film position
// (fx, fy): 1 / (A cos^3 theta) (pinhole_camera.cpp:13-31, 128-131).
__device__ __forceinline__ float bd_camera_pdf_at(const WaveParams& P, float fx, float fy) {
  float rx = float(P.xr), ry = float(P.yr);
  float cx = fx / rx - 0.5f, cy = fy / ry - 0.5f;
  float nx, ny;
  if (rx > ry) { nx = cx; ny = cy * ry / rx; } else { nx = cx * rx / ry; ny = cy; }
  float px = nx * P.S.project_dim, py = ny * P.S.project_dim;
  float ct = 1.f / sqrtf(px * px + py * py + 1.f);
  float area = fminf(rx, ry) / fmaxf(rx, ry) * sqf(P.S.project_dim);
  return 1.f / (area * (ct * ct * ct));
}

// PinholeCamera::ray_importance_pdf (pinhole_camera.cpp:84-98) of world direction wi.
__device__ __forceinline__ float bd_camera_ray_pdf(const WaveParams& P, V3 wi) {
  float2 fp;
  float imp;
  return lt_importance(P, lt_cam_vector(P.S, wi), fp, imp) ? imp : 0.f;
}

// The background lights' DiscreteDistrib1d (environment, then infinite light).
__device__ __forceinline__ float bd_pick_background(const DevScene& S, float u, uint32_t& light) {
  if (S.env_light >= 0 && (S.inf_light < 0 || u <= S.bg_env_prob)) {
    light = (uint32_t)S.env_light;
    return S.bg_env_prob;
  }
  if (S.inf_light < 0) return 0.f;
  light = (uint32_t)S.inf_light;
  return S.bg_inf_prob;
}

template <bool ANY, int WIDTH, bool FAST>
__device__ __forceinline__ bool bd_trace(const DevScene& S, const TraceRay& r, LaneStack& st, TraceHit& h) {
  return traverse<ANY, WIDTH, FAST>(S.nodes, S.tris, r, st, h);
}

// ---------------------------------------------------------------- subpaths

struct BdptArgs {
  float4* lv;        // light subpaths, 4 float4 per vertex (bd_store)
  float4* cv;        // camera subpaths
  uint32_t n;        // the arrays' path stride (paths per chunk)
  uint32_t nv;       // vertex slots per subpath
  uint32_t use_t1;   // BdptRenderer::use_t1_paths
  float* splat;      // splat sums, npix * 3
};

// random_light_walk (bdpt_renderer.cpp:118-181): returns the number of vertices.
template <int WIDTH>
__device__ uint32_t bd_light_walk(const WaveParams& P, const BdptArgs& A, uint32_t i, uint32_t pixel,
                                  uint32_t sample, float pick_u, LaneStack& st, uint32_t& light, uint32_t& rays) {
  const DevScene& S = P.S;
  float pick;
  light = lt_pick_light(S, pick_u, pick);
  if (pick == 0.f) return 0;
  const DevLight L = S.lights[light];
  const bool distant = L.kind != LK_AREA;
  float r[4];
  bd_rng(P, pixel, sample, 0, 1, r);
  LtRay lr = lt_light_ray(S, L, r);
  if (lr.pos_pdf == 0.f || lr.dir_pdf == 0.f) return 0;
  BdV prev;
  prev.p = lr.o; prev.eps = lr.tmin; prev.nn = lr.nn;
  prev.fwd = lr.pos_pdf; prev.bwd = 0.f;
  prev.alpha = lr.rad / (lr.pos_pdf * pick);
  prev.leaf = 0xFFFFFFFFu; prev.b1 = prev.b2 = 0.f;
  prev.info = BV_LIGHT << 1;
  uint32_t nvert = 1;
  V3 prev_f = v3(1.f, 1.f, 1.f);
  float fwd_dir_pdf = lr.dir_pdf;
  TraceRay ray;
  ray.o = lr.o; ray.d = lr.d; ray.tmin = lr.tmin; ray.tmax = INFINITY; ray.ignore = 0xFFFFFFFFu;
  for (uint32_t bounces = 0; bounces <= P.max_depth && nvert < A.nv; ++bounces) {
    TraceHit h;
    ++rays;
    // a distant light's ray starts on the scene's bounding disk, outside the range the
    // fast box test's padding covers: the first ray takes the exact test
    bool hit = bounces == 0u ? bd_trace<false, WIDTH, false>(S, ray, st, h) : bd_trace<false, WIDTH, true>(S, ray, st, h);
    if (!hit) break;
    Surface g = surface_at(S, h.tri, h.t, h.b1, h.b2);
    Bsdf bsdf = make_bsdf(S, g);
    V3 wi = normalize(-ray.d);
    float rb[4];
    bd_rng(P, pixel, sample, 1u + bounces, 0, rb);
    V3 wo;
    float wo_pdf;
    uint32_t fl;
    // Bsdf::sample_camera_f: every lobe here samples alike with either direction fixed
    V3 f = bsdf_sample(bsdf, wi, BSDF_ALL, rb[0], rb[1], rb[2], wo, wo_pdf, fl);
    if (wo_pdf == 0.f) break;
    float bwd_dir_pdf = bsdf_pdf(bsdf, wi, wo, BSDF_ALL);  // light_f_pdf(wi, wo)
    float d2 = length_sq(g.p - prev.p);
    if (d2 < 1e-9f) break;
    V3 scale = prev_f * (fabsf(dot(prev.nn, wi)) / fwd_dir_pdf);
    if (bounces >= 2u) {
      float rr = fminf(average(scale), 0.9f);
      if (rb[3] > rr) break;
      scale = scale / rr;
    }
    BdV y;
    y.p = g.p; y.eps = g.eps; y.nn = g.nn;
    y.fwd = (bounces == 0u && distant) ? lr.dir_pdf : fwd_dir_pdf * fabsf(dot(g.nn, wi)) / d2;
    y.bwd = 0.f;
    y.alpha = prev.alpha * scale;
    y.leaf = h.tri; y.b1 = h.b1; y.b2 = h.b2;
    y.info = (BV_SURFACE << 1) | ((fl & BSDF_DELTA) ? 1u : 0u);
    if (is_black(y.alpha)) break;
    prev.bwd = bwd_dir_pdf * fabsf(dot(prev.nn, wi)) / d2;
    bd_store(A.lv, A.n, i, nvert - 1u, prev);
    prev = y;
    ++nvert;
    prev_f = f;
    fwd_dir_pdf = wo_pdf;
    ray.o = g.p; ray.d = wo; ray.tmin = g.eps; ray.tmax = INFINITY;
  }
  bd_store(A.lv, A.n, i, nvert - 1u, prev);
  return nvert;
}

// random_camera_walk (bdpt_renderer.cpp:183-273): returns the number of vertices.
template <int WIDTH>
__device__ uint32_t bd_camera_walk(const WaveParams& P, const BdptArgs& A, uint32_t i, uint32_t pixel,
                                   uint32_t sample, float fx, float fy, LaneStack& st, uint32_t& rays) {
  const DevScene& S = P.S;
  TraceRay ray;
  camera_ray(S, fx, fy, P.xr, P.yr, ray.o, ray.d);
  ray.tmin = 0.f; ray.tmax = INFINITY; ray.ignore = 0xFFFFFFFFu;
  float dir_pdf = bd_camera_pdf_at(P, fx, fy);
  BdV prev;
  prev.p = ray.o; prev.eps = 0.f; prev.nn = ray.d;
  prev.fwd = 1.f; prev.bwd = 0.f;
  prev.alpha = v3(dir_pdf, dir_pdf, dir_pdf);  // importance / pos_pdf (pos_pdf = 1)
  prev.leaf = 0xFFFFFFFFu; prev.b1 = prev.b2 = 0.f;
  prev.info = BV_CAMERA << 1;
  uint32_t nvert = 1;
  V3 prev_f = v3(1.f, 1.f, 1.f);
  float fwd_dir_pdf = dir_pdf;
  for (uint32_t bounces = 0; bounces <= P.max_depth && nvert < A.nv; ++bounces) {
    if (is_black(prev.alpha)) break;
    float rb[4];
    bd_rng(P, pixel, sample, 1u + bounces, 1, rb);
    V3 scale = prev_f * (fabsf(dot(prev.nn, ray.d)) / fwd_dir_pdf);
    if (bounces >= 2u) {
      float rr = fminf(average(scale), 0.9f);
      if (rb[0] > rr) break;
      scale = scale / rr;
    }
    TraceHit h;
    ++rays;
    if (!bd_trace<false, WIDTH, true>(S, ray, st, h)) {
      float r2[4];
      bd_rng(P, pixel, sample, 1u + bounces, 2, r2);
      uint32_t bg = 0;
      float bgp = bd_pick_background(S, r2[0], bg);
      if (bgp == 0.f) break;
      // the "virtual vertex" of the escaped ray (bdpt_renderer.cpp:223-239)
      BdV z;
      z.p = prev.p + ray.d; z.eps = 0.f; z.nn = v3(0.f, 0.f, 0.f);
      z.fwd = fwd_dir_pdf; z.bwd = 0.f;
      z.alpha = prev.alpha * scale / bgp;
      z.leaf = 0xFFFFFFFFu; z.b1 = z.b2 = 0.f;
      z.info = (BV_BACKGROUND << 1) | (bg << 8);
      bd_store(A.cv, A.n, i, nvert - 1u, prev);
      prev = z;
      ++nvert;
      break;
    }
    Surface g = surface_at(S, h.tri, h.t, h.b1, h.b2);
    Bsdf bsdf = make_bsdf(S, g);
    V3 wo = -ray.d;
    V3 wi;
    float wi_pdf;
    uint32_t fl;
    V3 f = bsdf_sample(bsdf, wo, BSDF_ALL, rb[1], rb[2], rb[3], wi, wi_pdf, fl);  // sample_light_f
    float d2 = length_sq(g.p - prev.p);
    if (d2 < 1e-9f) break;
    uint32_t emitter = g.light >= 0 ? (uint32_t)g.light + 1u : 0u;
    BdV z;
    z.p = g.p; z.eps = g.eps; z.nn = g.nn;
    z.fwd = fwd_dir_pdf * fabsf(dot(g.nn, wo)) / d2;
    z.bwd = 0.f;
    z.alpha = prev.alpha * scale;
    z.leaf = h.tri; z.b1 = h.b1; z.b2 = h.b2;
    z.info = (BV_SURFACE << 1) | ((fl & BSDF_DELTA) ? 1u : 0u) | (emitter << 8);
    if (is_black(z.alpha) && emitter == 0u) break;
    if (wi_pdf == 0.f) {
      prev.bwd = 0.f;
      bd_store(A.cv, A.n, i, nvert - 1u, prev);
      prev = z;
      ++nvert;
      break;
    }
    float bwd_dir_pdf = bsdf_pdf(bsdf, wo, wi, BSDF_ALL);  // camera_f_pdf(wo, wi)
    prev.bwd = bwd_dir_pdf * fabsf(dot(prev.nn, wo)) / d2;
    bd_store(A.cv, A.n, i, nvert - 1u, prev);
    prev = z;
    ++nvert;
    prev_f = f;
    fwd_dir_pdf = wi_pdf;
    ray.o = g.p; ray.d = wi; ray.tmin = g.eps; ray.tmax = INFINITY;
  }
  bd_store(A.cv, A.n, i, nvert - 1u, prev);
  return nvert;
}

// ---------------------------------------------------------------- strategies

// The vertices of strategy (s, t) in path order x_0 (light end) .. x_{s+t-1} (camera):
// light_at / camera_at / vertex_at of path_weight (bdpt_renderer.cpp:476-492), with
// the fresh first light vertex of s = 1 and the fresh camera vertex of t = 1.
struct BdPath {
  const BdptArgs* A;
  uint32_t i, s, t;
  BdV first_light, first_camera;
  __device__ BdV at(uint32_t k) const {
    if (k < s) return (k == 0u && s == 1u) ? first_light : bd_load(A->lv, A->n, i, k);
    uint32_t j = s + t - 1u - k;
    return (j == 0u && t == 1u) ? first_camera : bd_load(A->cv, A->n, i, j);
  }
  // one word of vertex k's record c (1: normal + forward pdf, 2: throughput + backward
  // pdf, 3: triangle + barycentrics + info), without loading the rest of the vertex
  __device__ float word(uint32_t k, uint32_t c) const {
    if (k < s) {
      if (k == 0u && s == 1u) return c == 1u ? first_light.fwd : c == 2u ? first_light.bwd : __uint_as_float(first_light.info);
      return A->lv[(4 * (size_t)k + c) * A->n + i].w;
    }
    uint32_t j = s + t - 1u - k;
    if (j == 0u && t == 1u) return c == 1u ? first_camera.fwd : c == 2u ? first_camera.bwd : __uint_as_float(first_camera.info);
    return A->cv[(4 * (size_t)j + c) * A->n + i].w;
  }
  __device__ float fwd(uint32_t k) const { return word(k, 1u); }
  __device__ float bwd(uint32_t k) const { return word(k, 2u); }
  __device__ bool delta(uint32_t k) const { return (__float_as_uint(word(k, 3u)) & 1u) != 0u; }
};

// path_weight (bdpt_renderer.cpp:468-667) for a strategy with s + t > 2.  L: the path's
// light; bl / bc: the Bsdfs of the last light vertex (s >= 2) and the last camera vertex
// (t >= 2), the two vertices whose pdfs toward each other are computed here.
__device__ float bd_weight(const WaveParams& P, const BdPath& X, const DevLight& L, const Bsdf& bl,
                           const Bsdf& bc) {
  const DevScene& S = P.S;
  const uint32_t s = X.s, t = X.t, n = s + t;
  const bool distant = L.kind != LK_AREA;
  float x0_pivot;
  if (s == 1u) {
    x0_pivot = X.fwd(0);
  } else {
    BdV x0 = X.at(0), x1 = X.at(1);
    V3 wi = normalize(x0.p - x1.p);
    float dp = bd_pivot_pdf(S, L, wi, x1.p);
    x0_pivot = distant ? dp : dp * fabsf(dot(x0.nn, wi)) / length_sq(x0.p - x1.p);
  }
  // pdf of sampling x_k from x_{k-1} (from the light side)
  auto pdf_light = [&](uint32_t k) -> float {
    if (k == 0u) return x0_pivot;
    if (k == 1u) {
      if (x0_pivot == 0.f) return 0.f;
      BdV x0 = X.at(0), x1 = X.at(1);
      V3 wi = normalize(x0.p - x1.p);
      float ray_pdf;
      if (s >= 2u) {
        ray_pdf = distant ? x1.fwd * (x0.fwd * fabsf(dot(wi, x1.nn))) : x0.fwd * x1.fwd;
      } else {
        float ang = bd_ray_pdf(S, L, -wi, x0.nn);
        ray_pdf = distant ? ang * fabsf(dot(x1.nn, wi)) : ang * fabsf(dot(x1.nn, wi)) / length_sq(x0.p - x1.p);
      }
      return ray_pdf / x0_pivot;
    }
    if (k < s) return X.fwd(k);
    if (k == s && t != 1u) {
      BdV xo = X.at(k), x = X.at(k - 1u), xi = X.at(k - 2u);
      V3 bwi = normalize(xi.p - x.p), bwo = normalize(xo.p - x.p);
      float wo_pdf = bsdf_pdf(bl, bwo, bwi, BSDF_ALL);  // camera_f_pdf(bwo, bwi)
      float d2 = length_sq(xo.p - x.p);
      if (d2 == 0.f) return 0.f;
      return wo_pdf * fabsf(dot(bwo, xo.nn)) / d2;
    }
    return X.bwd(k);  // s < k < s + t - 1
  };
  // pdf of sampling x_k from x_{k+1} (from the camera side)
  auto pdf_camera = [&](uint32_t k) -> float {
    if (k == n - 2u) {
      BdV z1 = X.at(k), z0 = X.at(k + 1u);
      V3 wi = normalize(z1.p - z0.p);
      float ray_area;
      if (t >= 2u) ray_area = z0.fwd * z1.fwd;
      else ray_area = bd_camera_ray_pdf(P, wi) * fabsf(dot(wi, z1.nn)) / length_sq(z1.p - z0.p);
      float pivot = t == 1u ? z0.fwd : 1.f;  // PinholeCamera::pivot_importance_pdf is 1
      if (pivot == 0.f) return 0.f;
      return ray_area / pivot;
    }
    if (k >= s) return X.fwd(k);  // s <= k < s + t - 2
    if (k + 1u == s && t >= 2u) {
      BdV xi = X.at(k), x = X.at(k + 1u), xo = X.at(k + 2u);
      V3 bwi = normalize(xi.p - x.p), bwo = normalize(xo.p - x.p);
      float wi_pdf = bsdf_pdf(bc, bwi, bwo, BSDF_ALL);  // light_f_pdf(bwi, bwo)
      if (k == 0u && distant) return wi_pdf;
      return wi_pdf * fabsf(dot(xi.nn, bwi)) / length_sq(xi.p - x.p);
    }
    return X.bwd(k);  // k < s - 1
  };
  float inv = 1.f, r = 1.f;
  for (uint32_t j = 1; j + 1u <= t; ++j) {
    float fw = pdf_light(s + j - 1u), bw = pdf_camera(s + j - 1u);
    if (bw == 0.f) break;
    r *= fw / bw;
    if (r == 0.f) break;
    if (X.delta(s + j - 1u) || X.delta(s + j)) continue;
    inv += r;
  }
  r = 1.f;
  for (uint32_t j = 1; j <= s; ++j) {
    float fw = pdf_camera(s - j), bw = pdf_light(s - j);
    if (bw == 0.f) break;
    r *= fw / bw;
    if (r == 0.f) break;
    if (X.delta(s - j) || (j < s && X.delta(s - j - 1u))) continue;
    inv += r;
  }
  return 1.f / inv;
}

// Film::add_splat (film.cpp:32-67) with the 1 px box filter; non-finite contributions are
// dropped and counted.
__device__ __forceinline__ void bd_splat(const WaveParams& P, const BdptArgs& A, float2 fp, V3 c) {
  if (!(isfinite(c.x) && isfinite(c.y) && isfinite(c.z))) {
    atomicAdd((unsigned long long*)&P.stats[3], 1ull);
    return;
  }
  if (is_black(c)) return;
  int x0 = max(0, (int)ceilf(fp.x - 1.f)), x1 = min((int)P.xr - 1, (int)floorf(fp.x));
  int y0 = max(0, (int)ceilf(fp.y - 1.f)), y1 = min((int)P.yr - 1, (int)floorf(fp.y));
  if (x0 > x1 || y0 > y1) return;
  float w = 1.f / float((x1 - x0 + 1) * (y1 - y0 + 1));
  for (int y = y0; y <= y1; ++y)
    for (int x = x0; x <= x1; ++x) {
      float* f = A.splat + 3 * ((size_t)y * P.xr + x);
      atomicAdd(f, c.x * w);
      atomicAdd(f + 1, c.y * w);
      atomicAdd(f + 2, c.z * w);
    }
}

// One strategy (s, t) of sample_path (bdpt_renderer.cpp:43-116, path_contrib :275-466)
// for path i, whose camera vertex t - 1 is zc with Bsdf bc: its unweighted contribution c
// (black when the strategy gives nothing) and, for the strategies that need one, the
// shadow ray whose visibility decides it (the caller traces or queues it).  Sets X's
// fresh vertices, the light li of the path weight, bl (the last light vertex's Bsdf,
// s >= 2) and fp (the film position of a t = 1 splat).
struct BdStrategy {
  V3 c;
  bool ray;          // c counts only if `shadow` is unoccluded
  TraceRay shadow;
  uint32_t li;
  float2 fp;
};

__device__ __forceinline__ bool bd_strategy(const WaveParams& P, const BdptArgs& A, BdPath& X, uint32_t pixel,
                                            uint32_t sample, uint32_t light, const BdV& zc, const Bsdf& bc,
                                            Bsdf& bl, BdStrategy& o) {
  const DevScene& S = P.S;
  const uint32_t i = X.i, s = X.s, t = X.t;
  const bool z_bg = bv_kind(zc) == BV_BACKGROUND;
  const bool z_surf = bv_kind(zc) == BV_SURFACE;
  const V3 zero = v3(0.f, 0.f, 0.f);
  o.c = zero;
  o.ray = false;
  o.shadow.o = o.shadow.d = zero;
  o.shadow.tmin = o.shadow.tmax = 0.f;
  o.shadow.ignore = 0xFFFFFFFFu;
  o.li = light;
  o.fp = make_float2(0.f, 0.f);
  if (t == 1u && !A.use_t1) return false;
  if (s + t < P.min_depth + 2u || s + t > P.max_depth + 2u) return false;
  if (s > 0u && z_bg) return false;  // the virtual vertex connects to nothing
  // s = 1, t = 1 joins the camera to a fresh light point; path_weight gives that
  // strategy 0 for area and background lights (bdpt_renderer.cpp:494-503)
  if (s == 1u && t == 1u) return false;
  if (s == 0u) {
    // the camera subpath ends on an emitter (bdpt_renderer.cpp:319-339)
    if (t < 2u) return false;
    BdV zp = bd_load(A.cv, A.n, i, t - 2u);
    if (z_bg) {
      o.li = zc.info >> 8;
      V3 d = zc.p - zp.p;
      o.c = (o.li == (uint32_t)S.env_light ? env_background(S, normalize(d)) : xyz(S.lights[o.li].radiance)) * zc.alpha;
    } else {
      uint32_t e = zc.info >> 8;
      if (e == 0u) return false;
      o.li = e - 1u;
      if (dot(zc.nn, zp.p - zc.p) >= 0.f) o.c = xyz(S.lights[o.li].radiance) * zc.alpha;
    }
  } else if (s == 1u) {
    // a fresh light vertex for the camera subpath (bdpt_renderer.cpp:342-388)
    if (!z_surf) return false;
    float rc[4];
    bd_rng(P, pixel, sample, t, 3, rc);
    float pick;
    o.li = lt_pick_light(S, rc[0], pick);
    if (pick == 0.f) return false;
    const DevLight L = S.lights[o.li];
    BdPivot lp = bd_sample_pivot(S, L, zc.p, zc.eps, rc[1], rc[2]);
    if (is_black(lp.rad) || lp.pdf == 0.f) return false;
    BdV zp = bd_load(A.cv, A.n, i, t - 2u);
    V3 bwo = normalize(zp.p - zc.p);
    V3 f = bsdf_eval(bc, lp.wi, bwo, BSDF_ALL);
    if (is_black(f)) return false;
    o.ray = true;
    o.shadow = lp.shadow;
    BdV fl;
    if (L.kind != LK_AREA) {
      fl.p = zc.p + lp.wi;  // the incident direction, stored as a point
      fl.nn = zero;
      fl.eps = 0.f;
      fl.fwd = lp.pdf;
    } else {
      fl.p = lp.p;
      fl.nn = lp.nn;
      fl.eps = lp.eps;
      fl.fwd = lp.pdf * fabsf(dot(lp.nn, lp.wi)) / length_sq(lp.p - zc.p);
    }
    fl.alpha = zero;
    fl.bwd = bsdf_pdf(bc, lp.wi, bwo, BSDF_ALL);  // light_f_pdf(wi, bsdf_wo)
    fl.leaf = 0xFFFFFFFFu;
    fl.b1 = fl.b2 = 0.f;
    fl.info = BV_LIGHT << 1;
    X.first_light = fl;
    o.c = zc.alpha * f * lp.rad * (fabsf(dot(zc.nn, lp.wi)) / (pick * lp.pdf));
  } else {
    const BdV yl = bd_load(A.lv, A.n, i, s - 1u);
    const BdV yp = bd_load(A.lv, A.n, i, s - 2u);
    bl = bd_bsdf(S, yl);
    if (t == 1u) {
      // the light subpath to the camera (bdpt_renderer.cpp:390-422)
      const V3 cam = v3(S.c2w[3], S.c2w[7], S.c2w[11]);
      float imp;
      if (!lt_importance(P, lt_cam_point(S, yl.p), o.fp, imp)) return false;
      V3 wi = normalize(yp.p - yl.p), wo = normalize(cam - yl.p);
      V3 f = bsdf_eval(bl, wi, wo, BSDF_ALL);
      if (is_black(f)) return false;
      float dist = length(cam - yl.p);
      o.ray = true;
      o.shadow.o = yl.p; o.shadow.d = (cam - yl.p) / dist; o.shadow.tmin = yl.eps; o.shadow.tmax = dist;
      o.shadow.ignore = 0xFFFFFFFFu;
      BdV fc;
      fc.p = cam; fc.eps = 0.f; fc.nn = zero; fc.fwd = 1.f; fc.bwd = 0.f;
      fc.alpha = v3(imp, imp, imp);
      fc.leaf = 0xFFFFFFFFu; fc.b1 = fc.b2 = 0.f;
      fc.info = BV_CAMERA << 1;
      X.first_camera = fc;
      o.c = yl.alpha * f * imp * (fabsf(dot(yl.nn, wo)) / length_sq(cam - yl.p));
    } else {
      // the two subpaths joined (bdpt_renderer.cpp:423-443)
      if (!z_surf) return false;
      BdV zp = bd_load(A.cv, A.n, i, t - 2u);
      V3 lwi = normalize(yp.p - yl.p), cwo = normalize(zp.p - zc.p), wo = normalize(zc.p - yl.p);
      V3 lf = bsdf_eval(bl, lwi, wo, BSDF_ALL);
      if (is_black(lf)) return false;
      V3 cf = bsdf_eval(bc, -wo, cwo, BSDF_ALL);
      if (is_black(cf)) return false;
      float dist = length(zc.p - yl.p);
      o.ray = true;
      o.shadow.o = yl.p; o.shadow.d = (zc.p - yl.p) / dist; o.shadow.tmin = yl.eps; o.shadow.tmax = dist - zc.eps;
      o.shadow.ignore = 0xFFFFFFFFu;
      float geom = fabsf(dot(wo, yl.nn)) * fabsf(dot(wo, zc.nn)) / length_sq(yl.p - zc.p);
      o.c = yl.alpha * zc.alpha * lf * cf * geom;
    }
  }
  return !is_black(o.c);
}

// The strategy's MIS weight (path_weight; s + t = 2 has only the s = 0 strategy).
__device__ __forceinline__ float bd_strategy_weight(const WaveParams& P, const BdPath& X, uint32_t li,
                                                    const Bsdf& bl, const Bsdf& bc) {
  return X.s + X.t == 2u ? (X.s == 0u ? 1.f : 0.f) : bd_weight(P, X, P.S.lights[li], bl, bc);
}

// sample_path (bdpt_renderer.cpp:43-116) for path sample i of the chunk: the sum of the
// strategies with t >= 2 goes to P.rad[i] (w: 1, or 0 where dort's film drops the
// sample); the t = 1 strategies are splatted.
template <int WIDTH>
__global__ __launch_bounds__(BD_BLOCK) void k_bdpt(WaveParams P, BdptArgs A) {
  __shared__ uint32_t stk[16 * BD_BLOCK];
  LaneStack st;
  st.lds = stk + threadIdx.x;
  st.lds_stride = BD_BLOCK;
  st.lds_cap = 16;
  st.spill = P.spill + (size_t)blockIdx.x * BD_BLOCK + threadIdx.x;
  st.spill_stride = (size_t)gridDim.x * BD_BLOCK;
  st.cap = P.S.stack_cap;
  const uint32_t i = blockIdx.x * BD_BLOCK + threadIdx.x;
  if (i >= P.n_chunk) return;
  const DevScene& S = P.S;
  const uint64_t g = P.g0 + i;
  const uint32_t sample = P.sample_begin + (uint32_t)(g / P.npix);
  const uint32_t pixel = tile_pixel((uint32_t)(g % P.npix), P.xr, P.yr);  // 8x8 tiles per wave
  float r0[4];
  bd_rng(P, pixel, sample, 0, 0, r0);
  // film_pos = pixel + random_2d (bdpt_renderer.cpp:24-27); a jitter of exactly 0 puts the
  // sample on a pixel edge, where dort's box filter gives it no weight
  const float fx = float(pixel % P.xr) + r0[0], fy = float(pixel / P.xr) + r0[1];
  const bool on_edge = r0[0] == 0.f || r0[1] == 0.f;
  uint32_t light = 0, ext_rays = 0;
  const uint32_t nl = bd_light_walk<WIDTH>(P, A, i, pixel, sample, r0[2], st, light, ext_rays);
  const uint32_t nc = bd_camera_walk<WIDTH>(P, A, i, pixel, sample, fx, fy, st, ext_rays);
  V3 sum = v3(0.f, 0.f, 0.f);
  uint32_t shadow_rays = 0;
  BdPath X;
  X.A = &A;
  X.i = i;
  for (uint32_t t = 1; t <= nc; ++t) {
    const BdV zc = bd_load(A.cv, A.n, i, t - 1u);
    // the last camera vertex's Bsdf (bd_weight reads it only where it is a surface)
    const Bsdf bc = bv_kind(zc) == BV_SURFACE ? bd_bsdf(S, zc) : bd_no_bsdf();
    for (uint32_t s = 0; s <= nl; ++s) {
      X.s = s;
      X.t = t;
      Bsdf bl = bd_no_bsdf();
      BdStrategy o;
      if (!bd_strategy(P, A, X, pixel, sample, light, zc, bc, bl, o)) continue;
      if (o.ray) {
        ++shadow_rays;
        TraceHit h;
        if (bd_trace<true, WIDTH, true>(S, o.shadow, st, h)) continue;
      }
      V3 wc = o.c * bd_strategy_weight(P, X, o.li, bl, bc);
      if (t == 1u) bd_splat(P, A, o.fp, wc);
      else sum = sum + wc;
    }
  }
  P.rad[i] = make_float4(sum.x, sum.y, sum.z, on_edge ? 0.f : 1.f);
  atomicAdd((unsigned long long*)&P.stats[1], (unsigned long long)ext_rays);
  atomicAdd((unsigned long long*)&P.stats[2], (unsigned long long)shadow_rays);
}

// ---------------------------------------------------------------- wavefront form
//
// The chunk of k_bdpt in stages that reuse the path tracer's trace kernels:
//   k_bd_start, then per bounce k_extend + k_bd_shade: both subpaths of every path, the
//     light and the camera rays of one bounce traced by one launch;
//   then per part of the chunk (the strategy records need far more memory per path than
//   the subpaths): k_bd_strat<kind> for the four kinds of strategy: one thread per
//     path, the workgroup in step over the kind's slots, each contribution and MIS
//     weight (bd_strategy) recorded in its slot, its shadow ray queued;
//     k_connect: their visibility (an occluded ray clears its record); k_bd_resolve: the
//     visible records, summed in strategy order (t >= 2) or splatted.
// Each stage draws the random numbers k_bdpt draws at the same point of the walk, so both
// forms trace the same paths.  Vertices are stored as soon as they exist, and a vertex's
// backward pdf, known once the next vertex is, is patched into its record then.

__device__ __forceinline__ void bd_patch_bwd(float4* base, uint32_t n, uint32_t i, uint32_t k, float bwd) {
  base[(size_t)4 * k * n + 2 * (size_t)n + i].w = bwd;
}

// Queue words of the walks: the path index, with the top bit set for a camera ray.
constexpr uint32_t BD_CAMERA_RAY = 0x80000000u;

// Appends two classes of entries for a whole workgroup of SHADE_BLOCK threads with one
// atomic: the workgroup's `a` entries first, then its `b` entries, so that the waves
// which later read the queue mostly see runs of one class.  sh holds 2 * waves + 1 words.
__device__ __forceinline__ void block_append2(uint32_t* counter, bool a, bool b, uint32_t* sh,
                                              uint32_t& qa, uint32_t& qb) {
  constexpr uint32_t WAVES = SHADE_BLOCK / 64;
  const unsigned long long ma = __ballot(a), mb = __ballot(b);
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh[wave] = (uint32_t)__popcll(ma);
    sh[WAVES + 1 + wave] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ta = 0, tb = 0;
    for (uint32_t w = 0; w < WAVES; ++w) { uint32_t c = sh[w]; sh[w] = ta; ta += c; }
    for (uint32_t w = 0; w < WAVES; ++w) { uint32_t c = sh[WAVES + 1 + w]; sh[WAVES + 1 + w] = ta + tb; tb += c; }
    sh[WAVES] = ta + tb ? atomicAdd(counter, ta + tb) : 0u;
  }
  __syncthreads();
  const unsigned long long below = lane == 0 ? 0ull : ((1ull << lane) - 1ull);
  qa = sh[WAVES] + sh[wave] + (uint32_t)__popcll(ma & below);
  qb = sh[WAVES] + sh[WAVES + 1 + wave] + (uint32_t)__popcll(mb & below);
}

// Light subpath of path i, vertex 0 and the first ray (random_light_walk,
// bdpt_renderer.cpp:118-146).  Walk state at A.w0 / A.w1[i].
__device__ __forceinline__ bool bd_light_start(const WaveParams& P, const BdptArgsHost& A, uint32_t i,
                                               uint32_t pixel, uint32_t sample, float4& qo, float4& qd) {
  const DevScene& S = P.S;
  float r0[4];
  bd_rng(P, pixel, sample, 0, 0, r0);
  float pick;
  const uint32_t light = lt_pick_light(S, r0[2], pick);
  uint32_t nvert = 0;
  bool go = false;
  if (pick != 0.f) {
    const DevLight L = S.lights[light];
    float r[4];
    bd_rng(P, pixel, sample, 0, 1, r);
    const LtRay lr = lt_light_ray(S, L, r);
    if (lr.pos_pdf != 0.f && lr.dir_pdf != 0.f) {
      BdV prev;
      prev.p = lr.o; prev.eps = lr.tmin; prev.nn = lr.nn;
      prev.fwd = lr.pos_pdf; prev.bwd = 0.f;
      prev.alpha = lr.rad / (lr.pos_pdf * pick);
      prev.leaf = 0xFFFFFFFFu; prev.b1 = prev.b2 = 0.f;
      prev.info = BV_LIGHT << 1;
      bd_store(A.lv, A.n, i, 0, prev);
      nvert = 1;
      A.w0[i] = make_float4(1.f, 1.f, 1.f, lr.dir_pdf);
      A.w1[i] = make_uint4(1u, light, L.kind != LK_AREA ? 1u : 0u, 0u);
      go = nvert < A.nv;
      qo = make_float4(lr.o.x, lr.o.y, lr.o.z, lr.tmin);
      qd = make_float4(lr.d.x, lr.d.y, lr.d.z, INFINITY);
    }
  }
  A.nl[i] = nvert | (light << 8);
  return go;
}

// Camera subpath of path i, vertex 0 and the first ray (random_camera_walk,
// bdpt_renderer.cpp:183-204).  Walk state at A.w0 / A.w1[A.n + i].
__device__ __forceinline__ bool bd_camera_start(const WaveParams& P, const BdptArgsHost& A, uint32_t i,
                                                uint32_t pixel, uint32_t sample, float4& qo, float4& qd) {
  const DevScene& S = P.S;
  float r0[4];
  bd_rng(P, pixel, sample, 0, 0, r0);
  const float fx = float(pixel % P.xr) + r0[0], fy = float(pixel / P.xr) + r0[1];
  V3 o, d;
  camera_ray(S, fx, fy, P.xr, P.yr, o, d);
  const float dir_pdf = bd_camera_pdf_at(P, fx, fy);
  BdV prev;
  prev.p = o; prev.eps = 0.f; prev.nn = d;
  prev.fwd = 1.f; prev.bwd = 0.f;
  prev.alpha = v3(dir_pdf, dir_pdf, dir_pdf);  // importance / pos_pdf (pos_pdf = 1)
  prev.leaf = 0xFFFFFFFFu; prev.b1 = prev.b2 = 0.f;
  prev.info = BV_CAMERA << 1;
  bd_store(A.cv, A.n, i, 0, prev);
  A.nc[i] = 1u;
  A.w1[A.n + i] = make_uint4(1u, 0u, 0u, 0u);
  if (is_black(prev.alpha) || A.nv <= 1u) return false;
  const V3 scale = v3(1.f, 1.f, 1.f) * (fabsf(dot(prev.nn, d)) / dir_pdf);
  A.w0[A.n + i] = make_float4(scale.x, scale.y, scale.z, dir_pdf);
  qo = make_float4(o.x, o.y, o.z, 0.f);
  qd = make_float4(d.x, d.y, d.z, INFINITY);
  return true;
}

// Both subpaths' first vertex and first ray; the workgroup's light rays are queued ahead
// of its camera rays.
__global__ __launch_bounds__(SHADE_BLOCK) void k_bd_start(WaveParams P, BdptArgsHost A) {
  __shared__ uint32_t sh[2 * (SHADE_BLOCK / 64) + 1];
  const uint32_t i = blockIdx.x * SHADE_BLOCK + threadIdx.x;
  bool gl = false, gc = false;
  float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), ld = lo, co = lo, cd = lo;
  if (i < P.n_chunk) {
    uint32_t pixel, sample;
    order_sample_pixel(P.ord, P.s0, P.q0 + i, sample, pixel);
    gl = bd_light_start(P, A, i, pixel, sample, lo, ld);
    gc = bd_camera_start(P, A, i, pixel, sample, co, cd);
  }
  uint32_t ql, qc;
  block_append2(&P.counters[CNT_QUEUE], gl, gc, sh, ql, qc);
  if (gl) {
    P.qo[0][ql] = lo;
    P.qd[0][ql] = ld;
    P.qpath[0][ql] = i;
  }
  if (gc) {
    P.qo[0][qc] = co;
    P.qd[0][qc] = cd;
    P.qpath[0][qc] = i | BD_CAMERA_RAY;
  }
}

// One bounce of random_light_walk (bdpt_renderer.cpp:146-181) for the light ray at queue
// slot i of bounce b (path p); returns whether the next ray is queued.
__device__ __forceinline__ bool bd_light_step(const WaveParams& P, const BdptArgsHost& A, uint32_t b, uint32_t i,
                                              uint32_t p, V3& go_o, V3& go_d, float& go_eps) {
  const DevScene& S = P.S;
  const uint32_t cur = b & 1u;
  const float4 h4 = P.hit[i];
  const uint32_t leaf = __float_as_uint(h4.w);
  if (leaf == 0xFFFFFFFFu) return false;
  const float4 w0 = A.w0[p];
  const uint4 w1 = A.w1[p];
  uint32_t nvert = w1.x;
  const bool distant = w1.z != 0u;
  const V3 prev_f = xyz(w0);
  const float fwd_dir_pdf = w0.w;
  uint32_t pixel, sample;
  order_sample_pixel(P.ord, P.s0, P.q0 + p, sample, pixel);
  Surface g = surface_at(S, leaf, h4.x, h4.y, h4.z);
  Bsdf bsdf = make_bsdf(S, g);
  const V3 wi = normalize(-xyz(P.qd[cur][i]));
  float rb[4];
  bd_rng(P, pixel, sample, 1u + b, 0, rb);
  V3 wo;
  float wo_pdf;
  uint32_t fl;
  V3 f = bsdf_sample(bsdf, wi, BSDF_ALL, rb[0], rb[1], rb[2], wo, wo_pdf, fl);
  if (wo_pdf == 0.f) return false;
  const float bwd_dir_pdf = bsdf_pdf(bsdf, wi, wo, BSDF_ALL);
  const BdV prev = bd_load(A.lv, A.n, p, nvert - 1u);
  const float d2 = length_sq(g.p - prev.p);
  if (d2 < 1e-9f) return false;
  V3 scale = prev_f * (fabsf(dot(prev.nn, wi)) / fwd_dir_pdf);
  if (b >= 2u) {
    float rr = fminf(average(scale), 0.9f);
    if (rb[3] > rr) return false;
    scale = scale / rr;
  }
  BdV y;
  y.p = g.p; y.eps = g.eps; y.nn = g.nn;
  y.fwd = (b == 0u && distant) ? fwd_dir_pdf : fwd_dir_pdf * fabsf(dot(g.nn, wi)) / d2;
  y.bwd = 0.f;
  y.alpha = prev.alpha * scale;
  y.leaf = leaf; y.b1 = h4.y; y.b2 = h4.z;
  y.info = (BV_SURFACE << 1) | ((fl & BSDF_DELTA) ? 1u : 0u);
  if (is_black(y.alpha)) return false;
  bd_patch_bwd(A.lv, A.n, p, nvert - 1u, bwd_dir_pdf * fabsf(dot(prev.nn, wi)) / d2);
  bd_store(A.lv, A.n, p, nvert, y);
  ++nvert;
  A.nl[p] = nvert | (w1.y << 8);
  A.w0[p] = make_float4(f.x, f.y, f.z, wo_pdf);
  A.w1[p] = make_uint4(nvert, w1.y, w1.z, 0u);
  go_o = g.p; go_d = wo; go_eps = g.eps;
  return b + 1u <= P.max_depth && nvert < A.nv;
}

// One bounce of random_camera_walk (bdpt_renderer.cpp:204-273) for the camera ray at
// queue slot i of bounce b (path p); the Russian roulette of the next bounce decides
// whether its ray is queued.
__device__ __forceinline__ bool bd_camera_step(const WaveParams& P, const BdptArgsHost& A, uint32_t b, uint32_t i,
                                               uint32_t p, V3& go_o, V3& go_d, float& go_eps) {
  const DevScene& S = P.S;
  const uint32_t cur = b & 1u;
  const float4 h4 = P.hit[i];
  const uint32_t leaf = __float_as_uint(h4.w);
  const float4 w0 = A.w0[A.n + p];
  uint32_t nvert = A.w1[A.n + p].x;
  const V3 scale = xyz(w0);
  const float fwd_dir_pdf = w0.w;
  const V3 rd = xyz(P.qd[cur][i]);
  const BdV prev = bd_load(A.cv, A.n, p, nvert - 1u);
  uint32_t pixel, sample;
  order_sample_pixel(P.ord, P.s0, P.q0 + p, sample, pixel);
  if (leaf == 0xFFFFFFFFu) {
    // the "virtual vertex" of the escaped ray (bdpt_renderer.cpp:223-239)
    float r2[4];
    bd_rng(P, pixel, sample, 1u + b, 2, r2);
    uint32_t bg = 0;
    const float bgp = bd_pick_background(S, r2[0], bg);
    if (bgp != 0.f) {
      BdV z;
      z.p = prev.p + rd; z.eps = 0.f; z.nn = v3(0.f, 0.f, 0.f);
      z.fwd = fwd_dir_pdf; z.bwd = 0.f;
      z.alpha = prev.alpha * scale / bgp;
      z.leaf = 0xFFFFFFFFu; z.b1 = z.b2 = 0.f;
      z.info = (BV_BACKGROUND << 1) | (bg << 8);
      bd_store(A.cv, A.n, p, nvert, z);
      A.nc[p] = nvert + 1u;
    }
    return false;
  }
  Surface g = surface_at(S, leaf, h4.x, h4.y, h4.z);
  Bsdf bsdf = make_bsdf(S, g);
  const V3 wo = -rd;
  float rb[4];
  bd_rng(P, pixel, sample, 1u + b, 1, rb);
  V3 wi;
  float wi_pdf;
  uint32_t fl;
  V3 f = bsdf_sample(bsdf, wo, BSDF_ALL, rb[1], rb[2], rb[3], wi, wi_pdf, fl);  // sample_light_f
  const float d2 = length_sq(g.p - prev.p);
  if (d2 < 1e-9f) return false;
  const uint32_t emitter = g.light >= 0 ? (uint32_t)g.light + 1u : 0u;
  BdV z;
  z.p = g.p; z.eps = g.eps; z.nn = g.nn;
  z.fwd = fwd_dir_pdf * fabsf(dot(g.nn, wo)) / d2;
  z.bwd = 0.f;
  z.alpha = prev.alpha * scale;
  z.leaf = leaf; z.b1 = h4.y; z.b2 = h4.z;
  z.info = (BV_SURFACE << 1) | ((fl & BSDF_DELTA) ? 1u : 0u) | (emitter << 8);
  if (is_black(z.alpha) && emitter == 0u) return false;
  if (wi_pdf != 0.f) {
    const float bwd_dir_pdf = bsdf_pdf(bsdf, wo, wi, BSDF_ALL);  // camera_f_pdf(wo, wi)
    bd_patch_bwd(A.cv, A.n, p, nvert - 1u, bwd_dir_pdf * fabsf(dot(prev.nn, wo)) / d2);
  }
  bd_store(A.cv, A.n, p, nvert, z);
  ++nvert;
  A.nc[p] = nvert;
  A.w1[A.n + p] = make_uint4(nvert, 0u, 0u, 0u);
  // the top of the next bounce: its Russian roulette, before its ray is traced
  if (wi_pdf == 0.f || b + 1u > P.max_depth || nvert >= A.nv || is_black(z.alpha)) return false;
  float rn[4];
  bd_rng(P, pixel, sample, 2u + b, 1, rn);
  V3 sc = f * (fabsf(dot(z.nn, wi)) / wi_pdf);
  if (b + 1u >= 2u) {
    float rr = fminf(average(sc), 0.9f);
    if (rn[0] > rr) return false;
    sc = sc / rr;
  }
  A.w0[A.n + p] = make_float4(sc.x, sc.y, sc.z, wi_pdf);
  go_o = g.p; go_d = wi; go_eps = g.eps;
  return true;
}

// One bounce of both walks: each traced ray of bounce b extends its own subpath.
__global__ __launch_bounds__(SHADE_BLOCK) void k_bd_shade(WaveParams P, BdptArgsHost A, uint32_t b) {
  __shared__ uint32_t sh[SHADE_BLOCK / 64 + 1];
  const uint32_t cur = b & 1u, nxt = cur ^ 1u;
  const uint32_t n = P.counters[CNT_QUEUE + b];
  for (uint32_t base = blockIdx.x * SHADE_BLOCK; base < n; base += gridDim.x * SHADE_BLOCK) {
    const uint32_t i = base + threadIdx.x;
    bool go = false;
    uint32_t qp = 0;
    V3 go_o = v3(0.f, 0.f, 0.f), go_d = go_o;
    float go_eps = 0.f;
    if (i < n) {
      qp = P.qpath[cur][i];
      const uint32_t p = qp & ~BD_CAMERA_RAY;
      go = (qp & BD_CAMERA_RAY) ? bd_camera_step(P, A, b, i, p, go_o, go_d, go_eps)
                                : bd_light_step(P, A, b, i, p, go_o, go_d, go_eps);
    }
    const uint32_t q = block_append(&P.counters[CNT_QUEUE + b + 1u], go, sh);
    if (go) {
      P.qo[nxt][q] = make_float4(go_o.x, go_o.y, go_o.z, go_eps);
      P.qd[nxt][q] = make_float4(go_d.x, go_d.y, go_d.z, INFINITY);
      P.qpath[nxt][q] = qp;
    }
  }
}

// Strategy slots: slot k of the part's record arrays holds the strategy (s, t) that is
// k-th in sample_path's order (t outer from 1, s inner from 0, s + t <= nv); the slots
// before t number sum over t' < t of (nv - t' + 1).
__device__ __forceinline__ uint32_t bd_slot_of(uint32_t s, uint32_t t, uint32_t nv) {
  return (t - 1u) * nv - (t - 1u) * (t - 2u) / 2u + s;
}

// The strategies of one kind for the workgroup's 256 paths, one path per thread: CAT 0
// the camera subpath alone (s = 0), 1 a fresh light point (s = 1), 2 the light subpath to
// the camera (t = 1, s >= 2), 3 the two subpaths joined (s, t >= 2).  Each kind compiles
// to its own kernel with only its branches of bd_strategy and bd_weight, and so its own
// register count (80, 128, 105 and 168 VGPRs): one kernel for all kinds ran at the
// largest.  The workgroup walks its kind's slots in order up to its paths' largest vertex
// counts, a thread taking part where its own subpaths reach (s, t), so the lanes of a
// wave work on one (s, t) at a time (coalesced vertex loads, one loop length in the MIS
// weight).  Paths j0 .. j0 + m - 1 of the chunk: the record of path j0 + j's slot k at
// [k * m + j] (w = 1: summed, 2: splatted, 0: none), written for every slot of the kind
// whose s and t lie within the subpaths' vertex counts; a strategy that needs a shadow
// ray appends it to the shadow queue with its record's index (one atomic per workgroup
// and slot).
template <int CAT>
__global__ __launch_bounds__(SHADE_BLOCK, CAT == 3 ? 3 : 4) void k_bd_strat(WaveParams P, BdptArgsHost H, uint32_t j0, uint32_t m) {
  __shared__ uint32_t sh[SHADE_BLOCK / 64 + 1];
  __shared__ uint32_t smax[2];
  const DevScene& S = P.S;
  const uint32_t j = blockIdx.x * SHADE_BLOCK + threadIdx.x;
  const bool in = j < m;
  const uint32_t i = j0 + j;
  uint32_t nl = 0, nc = 0, light = 0, pixel = 0, sample = 0;
  if (threadIdx.x == 0) { smax[0] = 0u; smax[1] = 0u; }
  __syncthreads();
  if (in) {
    const uint32_t lw = H.nl[i];
    nl = lw & 0xFFu;
    light = lw >> 8;
    nc = H.nc[i];
    order_sample_pixel(P.ord, P.s0, P.q0 + i, sample, pixel);
    atomicMax(&smax[0], nl);
    atomicMax(&smax[1], nc);
  }
  __syncthreads();
  const uint32_t mnl = smax[0], mnc = smax[1];
  BdptArgs A;
  A.lv = H.lv; A.cv = H.cv; A.n = H.n; A.nv = H.nv; A.use_t1 = H.use_t1; A.splat = H.splat;
  // CAT 0: s = 0; 1: s = 1; 2: t = 1, s >= 2; 3: s, t >= 2
  const uint32_t t_lo = CAT == 2 ? 1u : 2u, t_hi = CAT == 2 ? std::min(1u, mnc) : mnc;
  for (uint32_t t = t_lo; t <= t_hi; ++t) {
    const uint32_t s_lo = CAT == 0 ? 0u : CAT == 1 ? 1u : 2u;
    const uint32_t s_end = CAT <= 1 ? std::min(s_lo, std::min(mnl, H.nv - t)) : std::min(mnl, H.nv - t);
    for (uint32_t s = s_lo; s <= s_end; ++s) {
      bool go = false;
      BdStrategy o;
      size_t rk = 0;
      if (in && s <= nl && t <= nc) {
        BdPath X;
        X.A = &A;
        X.i = i;
        X.s = s;
        X.t = t;
        const BdV zc = bd_load(A.cv, A.n, i, t - 1u);
        const Bsdf bc = bv_kind(zc) == BV_SURFACE ? bd_bsdf(S, zc) : bd_no_bsdf();
        Bsdf bl = bd_no_bsdf();
        float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
        rk = (size_t)bd_slot_of(s, t, H.nv) * m + j;
        if (bd_strategy(P, A, X, pixel, sample, light, zc, bc, bl, o)) {
          const V3 wc = o.c * bd_strategy_weight(P, X, o.li, bl, bc);
          r = make_float4(wc.x, wc.y, wc.z, t == 1u ? 2.f : 1.f);
          if (t == 1u) H.rec_fp[rk] = o.fp;
          go = o.ray;
        }
        H.rec[rk] = r;
      }
      const uint32_t q = block_append(&P.counters[CNT_SHADOW], go, sh);
      if (go) {
        P.so[q] = make_float4(o.shadow.o.x, o.shadow.o.y, o.shadow.o.z, o.shadow.tmin);
        P.sd[q] = make_float4(o.shadow.d.x, o.shadow.d.y, o.shadow.d.z, o.shadow.tmax);
        P.smeta[q] = make_uint2((uint32_t)rk, 0xFFFFFFFFu);
      }
    }
  }
}

// The visible records of path j0 + j: t >= 2 summed in strategy order into P.rad (w: 1,
// or 0 where dort's film drops the sample), t = 1 splatted.
__global__ __launch_bounds__(SHADE_BLOCK) void k_bd_resolve(WaveParams P, BdptArgsHost H, uint32_t j0, uint32_t m) {
  const uint32_t j = blockIdx.x * SHADE_BLOCK + threadIdx.x;
  if (j >= m) return;
  const uint32_t i = j0 + j;
  BdptArgs A;
  A.lv = H.lv; A.cv = H.cv; A.n = H.n; A.nv = H.nv; A.use_t1 = H.use_t1; A.splat = H.splat;
  uint32_t pixel, sample;
  order_sample_pixel(P.ord, P.s0, P.q0 + i, sample, pixel);
  float r0[4];
  bd_rng(P, pixel, sample, 0, 0, r0);
  const bool on_edge = r0[0] == 0.f || r0[1] == 0.f;
  const uint32_t nl = H.nl[i] & 0xFFu, nc = H.nc[i];
  V3 sum = v3(0.f, 0.f, 0.f);
  for (uint32_t t = 1; t <= nc; ++t) {
    const uint32_t s_end = std::min(nl, H.nv - t);
    for (uint32_t s = t == 1u ? 2u : 0u; s <= s_end; ++s) {  // (0, 1) and (1, 1) give nothing
      const size_t rk = (size_t)bd_slot_of(s, t, H.nv) * m + j;
      const float4 r = H.rec[rk];
      if (r.w == 1.f) sum = sum + xyz(r);
      else if (r.w == 2.f) bd_splat(P, A, H.rec_fp[rk], xyz(r));
    }
  }
  P.rad[i] = make_float4(sum.x, sum.y, sum.z, on_edge ? 0.f : 1.f);
}

// Adds a stage's ray counts to the statistics: the extension queues (walks) or the
// shadow queue (strategies).
__global__ void k_bd_count(WaveParams P, uint32_t shadow) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (shadow) {
    P.stats[2] += P.counters[CNT_SHADOW];
  } else {
    unsigned long long e = 0;
    for (uint32_t b = 0; b < MAX_BOUNCES; ++b) e += P.counters[CNT_QUEUE + b];
    P.stats[1] += e;
  }
}

// The chunk's per-path sums added to the film pixel by pixel in sample order, as
// film_accumulate does for the path tracer; non-finite or negative sums are dropped
// and counted (renderer.cpp:36-41).
__global__ __launch_bounds__(256) void k_bdpt_film(WaveParams P) {
  uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) P.stats[0] += P.n_chunk;
  if (px >= P.npix) return;
  uint64_t g_end = P.g0 + P.n_chunk;
  uint64_t g = P.g0 + (((uint64_t)tile_pos(px, P.xr, P.yr) + P.npix - (uint32_t)(P.g0 % P.npix)) % P.npix);
  if (g >= g_end) return;
  float fr = P.film_rgb[3 * (size_t)px], fg = P.film_rgb[3 * (size_t)px + 1], fb = P.film_rgb[3 * (size_t)px + 2];
  float fw = P.film_w[px];
  uint32_t dropped = 0;
  for (; g < g_end; g += P.npix) {
    float4 r = P.rad[g - P.g0];
    bool ok = isfinite(r.x) && isfinite(r.y) && isfinite(r.z) && r.x >= 0.f && r.y >= 0.f && r.z >= 0.f;
    if (!ok) { ++dropped; continue; }
    if (r.w == 0.f) continue;
    fr += r.x; fg += r.y; fb += r.z;
    fw += 1.f;
  }
  P.film_rgb[3 * (size_t)px] = fr;
  P.film_rgb[3 * (size_t)px + 1] = fg;
  P.film_rgb[3 * (size_t)px + 2] = fb;
  P.film_w[px] = fw;
  if (dropped) atomicAdd((unsigned long long*)&P.stats[3], (unsigned long long)dropped);
}

hipError_t launch_bdpt_wave(const WaveParams& P, const BdptArgsHost& A, const KernelCfg& ce,
                            const KernelCfg& cc, uint32_t shade_grid, uint32_t part, hipStream_t st,
                            const hipStream_t side[2], const hipEvent_t ev[3]) {
  const uint32_t blocks = (P.n_chunk + SHADE_BLOCK - 1u) / SHADE_BLOCK;
  const uint32_t grid = std::min(blocks, std::max(1u, shade_grid));
  hipError_t e;
  // both subpaths: bounce b's light and camera rays in one queue
  if ((e = launch_zero(P.counters, NUM_COUNTERS, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_bd_start, dim3(blocks), dim3(SHADE_BLOCK), 0, st, P, A);
  for (uint32_t b = 0; b <= P.max_depth; ++b) {
    if ((e = launch_extend(P, ce, b, false, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bd_shade, dim3(grid), dim3(SHADE_BLOCK), 0, st, P, A, b);
  }
  hipLaunchKernelGGL(k_bd_count, dim3(1), dim3(64), 0, st, P, 0u);
  // strategies, their shadow rays, the visible contributions: `part` paths at a time
  for (uint32_t j0 = 0; j0 < P.n_chunk; j0 += part) {
    const uint32_t m = std::min(part, P.n_chunk - j0);
    const uint32_t mb = (m + SHADE_BLOCK - 1u) / SHADE_BLOCK;
    if ((e = launch_zero(P.counters, NUM_COUNTERS, st)) != hipSuccess) return e;
    // the four kinds write disjoint records and append to one queue: the joined subpaths
    // (the largest kernel, 3 waves per SIMD) on st, the others on the side streams, so
    // the small kernels' waves fill what the large one leaves
    if ((e = hipEventRecord(ev[0], st)) != hipSuccess) return e;
    for (int k = 0; k < 2; ++k)
      if ((e = hipStreamWaitEvent(side[k], ev[0], 0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bd_strat<3>, dim3(mb), dim3(SHADE_BLOCK), 0, st, P, A, j0, m);
    hipLaunchKernelGGL(k_bd_strat<1>, dim3(mb), dim3(SHADE_BLOCK), 0, side[0], P, A, j0, m);
    hipLaunchKernelGGL(k_bd_strat<2>, dim3(mb), dim3(SHADE_BLOCK), 0, side[1], P, A, j0, m);
    hipLaunchKernelGGL(k_bd_strat<0>, dim3(mb), dim3(SHADE_BLOCK), 0, side[1], P, A, j0, m);
    for (int k = 0; k < 2; ++k) {
      if ((e = hipEventRecord(ev[1 + k], side[k])) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(st, ev[1 + k], 0)) != hipSuccess) return e;
    }
    if ((e = launch_connect(P, cc, 0, false, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bd_resolve, dim3(mb), dim3(SHADE_BLOCK), 0, st, P, A, j0, m);
    hipLaunchKernelGGL(k_bd_count, dim3(1), dim3(64), 0, st, P, 1u);
  }
  hipLaunchKernelGGL(k_bdpt_film, dim3((P.npix + 255u) / 256u), dim3(256), 0, st, P);
  return hipGetLastError();
}

hipError_t launch_bdpt(const WaveParams& P, const BdptArgsHost& H, uint32_t width, hipStream_t st) {
  BdptArgs A;
  A.lv = H.lv; A.cv = H.cv; A.n = H.n; A.nv = H.nv; A.use_t1 = H.use_t1; A.splat = H.splat;
  uint32_t blocks = (P.n_chunk + BD_BLOCK - 1u) / BD_BLOCK;
  if (width == 8) hipLaunchKernelGGL(k_bdpt<8>, dim3(blocks), dim3(BD_BLOCK), 0, st, P, A);
  else if (width == 4) hipLaunchKernelGGL(k_bdpt<4>, dim3(blocks), dim3(BD_BLOCK), 0, st, P, A);
  else hipLaunchKernelGGL(k_bdpt<2>, dim3(blocks), dim3(BD_BLOCK), 0, st, P, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bdpt_film, dim3((P.npix + 255u) / 256u), dim3(256), 0, st, P);
  return hipGetLastError();
}

}  // namespace dgpu
