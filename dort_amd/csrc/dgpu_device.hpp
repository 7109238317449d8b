// dgpu_device.hpp -- device-side data layout and shading math of the MI355X path tracer.
//
// Layout in HBM (all structure-of-arrays or 16-byte aligned records):
//  * BVH2 nodes: 64 B each (Aila & Laine 2009 layout: both child boxes in the parent),
//      n0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//      n1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//      n2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//      n3 = (child0, child1, -, -)      child >= 0: inner node; child < 0: leaf
//  * BVH4 nodes: 64 B each, child boxes quantized to 8 bits per plane in the node's frame:
//      (origin.xyz, exponents), (lo.x[4] hi.x[4] lo.y[4] hi.y[4] bytes),
//      (lo.z[4] hi.z[4] bytes, child0, child1), (child2, child3, -, -); an empty slot has
//      an inverted box (lo 255, hi 0) and the code of a leaf holding one degenerate
//      triangle, appended after the scene's triangles, that no ray hits
//  * Triangles in leaf order: 48 B each (3 x float4 vertices; v0.w = scene triangle id).
//  * Leaf child encoding: bit 31 set, bits 27..30 = count-1, bits 0..26 = first triangle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dgpu_rng.h"
#include "../../include/dortgpu.h"
#include "path_order.hpp"

#define DG_INLINE __host__ __device__ __forceinline__

namespace dgpu {

constexpr float PI_F = 3.14159265358979323846f;
constexpr float INV_PI_F = 1.0f / PI_F;
constexpr float TWO_PI_F = 2.0f * PI_F;
constexpr float INV_TWO_PI_F = 1.0f / TWO_PI_F;
constexpr float INV_FOUR_PI_F = 1.0f / (4.0f * PI_F);

constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr int MAX_STACK = 64;                   // BVH depth bound enforced by the builder

DG_INLINE uint32_t leaf_encode(uint32_t first, uint32_t count) {
  return LEAF_BIT | ((count - 1u) << 27) | first;
}
DG_INLINE uint32_t leaf_first(uint32_t c) { return c & 0x07FFFFFFu; }
DG_INLINE uint32_t leaf_count(uint32_t c) { return ((c >> 27) & 15u) + 1u; }

struct V3 {
  float x, y, z;
};
DG_INLINE V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
DG_INLINE V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
DG_INLINE V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
DG_INLINE V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
DG_INLINE V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
DG_INLINE V3 operator*(float s, V3 a) { return v3(a.x * s, a.y * s, a.z * s); }
DG_INLINE V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
DG_INLINE V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
DG_INLINE float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DG_INLINE V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
DG_INLINE float length_sq(V3 a) { return dot(a, a); }
DG_INLINE float length(V3 a) { return sqrtf(dot(a, a)); }
DG_INLINE V3 normalize(V3 a) { return a / length(a); }
DG_INLINE float sqf(float a) { return a * a; }
DG_INLINE float clampf(float a, float lo, float hi) { return a < lo ? lo : (a > hi ? hi : a); }
DG_INLINE bool is_black(V3 c) { return c.x == 0.f && c.y == 0.f && c.z == 0.f; }
DG_INLINE float average(V3 c) { return (c.x + c.y + c.z) / 3.0f; }
DG_INLINE float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
DG_INLINE V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

// sin and cos of x for |x| <= 8; every caller passes an angle in [-2 pi, 2 pi].
// Cody-Waite reduction by pi/2 (two-part constant, fused multiply-adds) to
// r in [-pi/4, pi/4], then the Taylor series of sin to r^9 and of cos to r^10, whose
// truncation errors (< 2e-9 and < 1e-10) are below half an ulp of the result.  The
// library's sinf/cosf carry a large-argument reduction that this range never needs
// and that costs registers in every kernel that inlines it.
DG_INLINE void sincos_small(float x, float& s, float& c) {
  float k = rintf(x * 0.63661977236758134f);           // x * 2/pi
  float r = fmaf(-k, 1.57079637050628662f, x);          // pi/2 rounded to float ...
  r = fmaf(-k, -4.37113900018624283e-08f, r);           // ... and the rest of pi/2
  float r2 = r * r;
  float sp = r + r * r2 * (-1.f / 6.f + r2 * (1.f / 120.f + r2 * (-1.f / 5040.f + r2 * (1.f / 362880.f))));
  float cp = 1.f + r2 * (-0.5f + r2 * (1.f / 24.f + r2 * (-1.f / 720.f +
                                                     r2 * (1.f / 40320.f + r2 * (-1.f / 3628800.f)))));
  int q = (int)k & 3;
  float ss = (q & 1) ? cp : sp, cc = (q & 1) ? sp : cp;
  s = (q & 2) ? -ss : ss;
  c = ((q + 1) & 2) ? -cc : cc;
}

// geometry.cpp:4-11
DG_INLINE void coordinate_system(V3 v0, V3& v1, V3& v2) {
  if (fabsf(v0.x) > fabsf(v0.y)) v1 = normalize(v3(-v0.z, 0.f, v0.x));
  else v1 = normalize(v3(0.f, -v0.z, v0.y));
  v2 = cross(v0, v1);
}

// ----------------------------------------------------------------- scene records

struct DevMaterial {  // 96 B
  uint32_t type;
  int32_t tex;
  float roughness;
  float ior_in;
  float4 albedo;   // w = ior_out
  float4 eta;
  float4 k;
  float4 refl;
  float4 trans;
};

enum { LK_AREA = 0, LK_ENV = 1, LK_INF = 2 };

struct DevLight {  // 32 B
  uint32_t kind;
  uint32_t tri;      // leaf-order triangle index of the emitter
  float area;
  float pad;
  float4 radiance;
};

struct DevScene {
  // acceleration structure
  const float4* nodes;
  const float4* tris;        // leaf order, 3 per triangle
  uint32_t width;            // 2, 4 or 8
  uint32_t num_tris;
  uint32_t stack_cap;        // traversal stack entries per lane (LDS + spill)
  // per-triangle attributes in leaf order (the leaf-order triangle record holds the
  // scene id in v0.w, the material in v1.w and the light index or -1 in v2.w)
  const float2* tri_uv;      // 3 per triangle, or null (default uvs)
  const DevMaterial* mats;
  uint32_t num_mats;
  // lights
  const DevLight* lights;
  const float* light_cdf;    // num_lights + 1
  uint32_t num_lights;
  int32_t env_light;         // index of the environment light or -1
  // environment map (environment_light.cpp, piecewise_distrib_2d.cpp)
  uint32_t env_w, env_h;
  const float4* env_rgb;     // w*h: rgb premultiplied by scale, w = the texel's probability
  const float* env_marg;     // h+1
  const float* env_cond;     // (w+1)*h
  const uint32_t* env_marg_guide;  // h: start row of the search for u in [g/h, (g+1)/h)
  const uint32_t* env_cond_guide;  // w*h: start column per row, likewise
  V3 env_up, env_s, env_t;
  int32_t has_inf;
  V3 inf_rad;
  // background lights (environment, infinite) picked by power among themselves, as
  // compute_light_distrib(scene, true) does (bdpt_renderer.cpp:13): the infinite light's
  // index (or -1) and the two lights' probabilities in that pick
  int32_t inf_light;
  float bg_env_prob, bg_inf_prob;
  // textures
  const float4* tex_data;
  const uint4* tex_info;     // (w, h, offset, mapping DGPU_TEXMAP_*)
  const float4* tex_xform;   // 2 per texture: rows 0, 1 of world_to_texture (DGPU_TEXMAP_XY)
  uint32_t has_rough;        // some material is a rough dielectric (kernels with that lobe)
  uint32_t num_tex;
  // camera (pinhole_camera.cpp:6-11)
  float c2w[12];
  float w2c[12];             // inverse of c2w (Transform::apply_inv)
  float project_dim;
  // bounding sphere of the scene (Scene::centroid, radius; lua_builder.cpp:124-128)
  V3 centroid;
  float radius;
};

// ----------------------------------------------------------------- RNG

struct PathRng {
  uint32_t k0, k1, pixel, sample;
  DG_INLINE void block(uint32_t bounce, uint32_t blk, float out[4]) const {
    dgpu_rng_block(k0, k1, pixel, sample, bounce, blk, out);
  }
};

// ----------------------------------------------------------------- BSDFs

enum { BSDF_REFLECTION = 1, BSDF_TRANSMISSION = 2, BSDF_DIFFUSE = 4, BSDF_GLOSSY = 8,
       BSDF_DELTA = 16, BSDF_MODES = 3, BSDF_LOBES = 28, BSDF_ALL = 31 };

// The lobe parameters share three colour slots and two scalars; which kind owns them:
//   Lambert:          c0 = albedo
//   GGX conductor:    c0 = tint, c1 = eta, c2 = k, a0 = alpha
//   dielectric:       c0 = reflect tint, c1 = transmit tint, a0 = ior inside, a1 = ior outside
//   rough dielectric: as the dielectric, and c2.x = alpha
// Every field is always written, so the record lives in registers, never on the stack.
struct Bsdf {
  V3 nn, sn;       // geometric == shading normal (no shading normals on the GPU path); tn = nn x sn
  uint32_t kind;   // DGPU_MAT_*; NONE = no bxdf
  V3 c0, c1, c2;
  float a0, a1;
  // lobe flags by kind: NONE -, LAMBERT reflection|diffuse, GGX reflection|glossy,
  // DIELECTRIC reflection|transmission|delta, ROUGH_DIELECTRIC reflection|transmission|glossy
  DG_INLINE uint32_t flags() const {
    constexpr uint64_t T = ((uint64_t)(BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_GLOSSY) << 32) |
                           ((uint64_t)(BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_DELTA) << 24) |
                           ((uint64_t)(BSDF_REFLECTION | BSDF_GLOSSY) << 16) |
                           ((uint64_t)(BSDF_REFLECTION | BSDF_DIFFUSE) << 8);
    return (uint32_t)(T >> (8u * kind)) & 0xFFu;
  }
  DG_INLINE V3 tn() const { return cross(nn, sn); }
  DG_INLINE V3 albedo() const { return c0; }
  DG_INLINE float alpha() const { return a0; }
  DG_INLINE V3 eta() const { return c1; }
  DG_INLINE V3 k() const { return c2; }
  DG_INLINE V3 refl() const { return c0; }
  DG_INLINE V3 trans() const { return c1; }
  DG_INLINE float ior_in() const { return a0; }
  DG_INLINE float ior_out() const { return a1; }
  DG_INLINE bool matches(uint32_t req) const {
    uint32_t f = flags();
    return kind != DGPU_MAT_NONE && (f & (req & BSDF_MODES)) && (f & (req & BSDF_LOBES));
  }
  DG_INLINE V3 to_local(V3 v) const { return v3(dot(sn, v), dot(tn(), v), dot(nn, v)); }
  DG_INLINE V3 to_world(V3 v) const { return sn * v.x + tn() * v.y + nn * v.z; }
};

// microfacet.cpp:55-100 (GGX)
DG_INLINE float ggx_d(float alpha, V3 m) {
  float c2 = sqf(m.z), s2 = 1.f - c2, t2 = s2 / c2;
  return INV_PI_F * sqf(alpha) / sqf(c2 * (sqf(alpha) + t2));
}
DG_INLINE float ggx_g1(float alpha, V3 v, V3 m) {
  if (dot(v, m) * v.z <= 0.f) return 0.f;
  float c2 = sqf(v.z), s2 = 1.f - c2, t2 = s2 / c2;
  return 2.f / (1.f + sqrtf(1.f + sqf(alpha) * t2));
}
DG_INLINE float ggx_dlnd(float alpha, V3 m) {
  float c2 = sqf(m.z), t2 = (1.f - c2) / c2;
  return 2.f / alpha - 4.f * alpha / (sqf(alpha) + t2);
}
DG_INLINE float ggx_dlng1(float alpha, V3 v) {
  float c2 = sqf(v.z), t2 = (1.f - c2) / c2;
  float s = sqrtf(1.f + sqf(alpha) * t2);
  return -(alpha * t2 / s) / (1.f + s);
}

// Exact unpolarised Fresnel reflectance of a conductor with complex IOR eta + i k.
DG_INLINE float fresnel_conductor1(float cos_i, float eta, float k) {
  cos_i = clampf(cos_i, 0.f, 1.f);
  float c2 = cos_i * cos_i, s2 = 1.f - c2;
  float e2 = eta * eta, k2 = k * k;
  float t0 = e2 - k2 - s2;
  float a2b2 = sqrtf(fmaxf(0.f, t0 * t0 + 4.f * e2 * k2));
  float a = sqrtf(fmaxf(0.f, 0.5f * (a2b2 + t0)));
  float t1 = a2b2 + c2, t2 = 2.f * cos_i * a;
  float rs = (t1 - t2) / (t1 + t2);
  float t3 = c2 * a2b2 + s2 * s2, t4 = t2 * s2;
  float rp = rs * (t3 - t4) / (t3 + t4);
  return 0.5f * (rp + rs);
}

// fresnel.cpp:5-47
DG_INLINE bool refract_dielectric(V3 w, V3& out, float ir, float it) {
  float eta = ir / it;
  float xt = w.x * eta, yt = w.y * eta;
  float st2 = sqf(xt) + sqf(yt);
  if (st2 >= 1.f) return false;
  float zt = sqrtf(1.f - st2);
  out = v3(-xt, -yt, -copysignf(zt, w.z));
  return true;
}
DG_INLINE float fresnel_dielectric(float cr, float ct, float ir, float it) {
  float tr = it * cr, rt = ir * ct, rr = ir * cr, tt = it * ct;
  float para = (tr - rt) / (tr + rt), perp = (rr - tt) / (rr + tt);
  return 0.5f * (sqf(para) + sqf(perp));
}

// Heitz 2018, "Sampling the GGX distribution of visible normals" (JCGT 7(4)).
DG_INLINE V3 ggx_sample_vndf(float alpha, V3 wo, float u1, float u2) {
  V3 vh = normalize(v3(alpha * wo.x, alpha * wo.y, wo.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  V3 t1 = lensq > 0.f ? v3(-vh.y, vh.x, 0.f) / sqrtf(lensq) : v3(1.f, 0.f, 0.f);
  V3 t2 = cross(vh, t1);
  float r = sqrtf(u1), phi = TWO_PI_F * u2;
  float sphi, cphi;
  sincos_small(phi, sphi, cphi);
  float p1 = r * cphi, p2 = r * sphi;
  float s = 0.5f * (1.f + vh.z);
  p2 = (1.f - s) * sqrtf(fmaxf(0.f, 1.f - p1 * p1)) + s * p2;
  V3 nh = t1 * p1 + t2 * p2 + vh * sqrtf(fmaxf(0.f, 1.f - p1 * p1 - p2 * p2));
  return normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(0.f, nh.z)));
}

DG_INLINE V3 conductor_f(const Bsdf& b, V3 wi, V3 wo) {
  if (wi.z * wo.z <= 0.f) return v3(0.f, 0.f, 0.f);
  if (wo.z < 0.f) { wi.z = -wi.z; wo.z = -wo.z; }
  V3 h = wi + wo;
  if (length_sq(h) == 0.f) return v3(0.f, 0.f, 0.f);
  h = normalize(h);
  float g = ggx_g1(b.alpha(), wi, h) * ggx_g1(b.alpha(), wo, h);
  float d = ggx_d(b.alpha(), h);
  float c = fabsf(dot(wi, h));
  V3 eta = b.eta(), k = b.k();
  V3 F = v3(fresnel_conductor1(c, eta.x, k.x), fresnel_conductor1(c, eta.y, k.y),
            fresnel_conductor1(c, eta.z, k.z));
  return b.albedo() * F * (d * g / (4.f * wi.z * wo.z));
}
DG_INLINE float conductor_pdf(const Bsdf& b, V3 wi, V3 wo) {
  if (wi.z * wo.z <= 0.f) return 0.f;
  if (wo.z < 0.f) { wi.z = -wi.z; wo.z = -wo.z; }
  V3 h = wi + wo;
  if (length_sq(h) == 0.f) return 0.f;
  h = normalize(h);
  if (dot(wi, h) <= 0.f) return 0.f;
  return ggx_g1(b.alpha(), wo, h) * ggx_d(b.alpha(), h) / (4.f * wo.z);
}
DG_INLINE float conductor_dlnf_dalpha(const Bsdf& b, V3 wi, V3 wo) {
  if (wo.z < 0.f) { wi.z = -wi.z; wo.z = -wo.z; }
  V3 h = normalize(wi + wo);
  return ggx_dlnd(b.alpha(), h) + ggx_dlng1(b.alpha(), wi) + ggx_dlng1(b.alpha(), wo);
}

// dort's RoughDielectricBxdf as dort runs it (rough_dielectric_material.cpp:7-185), GGX
// distribution (every MicrofacetType falls through to the GGX formulas, microfacet.cpp).
// get_iors picks the indices by the side of its argument (:169-179).
DG_INLINE void rough_iors(const Bsdf& b, V3 w, float& ir, float& it) {
  if (w.z >= 0.f) { ir = b.ior_out(); it = b.ior_in(); } else { ir = b.ior_in(); it = b.ior_out(); }
}
// eval_f (:7-37): reflection or transmission by the request's mode (the Bsdf layer chose
// it from the hemispheres); the Fresnel term takes both directions' dot products with h.
DG_INLINE V3 rough_dielectric_f(const Bsdf& b, V3 wi, V3 wo, bool reflect) {
  float ir, it;
  rough_iors(b, wi, ir, it);
  const float alpha = b.c2.x;
  V3 h = reflect ? normalize(wi + wo) : -normalize(wi * ir + wo * it);
  float cin = fabsf(wi.z), con = fabsf(wo.z);
  float cih = fabsf(dot(wi, h)), coh = fabsf(dot(wo, h));
  float fr = fresnel_dielectric(cih, coh, ir, it);
  float g = ggx_g1(alpha, wi, h) * ggx_g1(alpha, wo, h);
  float d = ggx_d(alpha, h);
  if (reflect) return b.refl() * (fr * g * d * 0.25f / (cin * con));
  return b.trans() * (cih * coh * sqf(it) * (1.f - fr) * g * d / (cin * con * sqf(ir * cih - it * coh)));
}
// sample_f (:69-147): a normal from D(m) cos(m) (MicrofacetDistrib::sample_m, GGX,
// microfacet.cpp:67-86: x = sin(theta) sin(phi), y = sin(theta) cos(phi)) turned to the
// fixed direction's side; reflection is picked with fresnel_dielectric_refl's probability
// (fresnel.cpp:38-47); a transmission pick returns pdf 0 (dort's transmission sampling is
// commented out, :114-128), as does a reflection whose pdf is below 1e-5.
DG_INLINE V3 rough_dielectric_sample(const Bsdf& b, V3 wf, uint32_t req, float u, float v, float c, V3& wg,
                                     float& pdf, uint32_t& fl) {
  V3 zero = v3(0.f, 0.f, 0.f);
  bool s_refl = req & BSDF_REFLECTION, s_trans = req & BSDF_TRANSMISSION;
  float ir, it;
  rough_iors(b, wf, ir, it);
  const float alpha = b.c2.x;
  float ct = 1.f / sqrtf((sqf(alpha) * u) / (1.f - u) + 1.f);
  float st = sqrtf(1.f - sqf(ct));
  float sp, cp;
  sincos_small(v * TWO_PI_F, sp, cp);
  V3 m = v3(st * sp, st * cp, ct);
  float m_pdf = ggx_d(alpha, m) * ct;
  m.z = copysignf(m.z, wf.z);
  float wfm = dot(wf, m);
  float cr = fabsf(wfm);
  float ct2 = 1.f - sqf(ir / it) * (1.f - sqf(cr));
  float fr = ct2 <= 0.f ? 1.f : fresnel_dielectric(cr, sqrtf(ct2), ir, it);
  bool reflect;
  float pick;
  if (s_refl && s_trans) {
    reflect = c <= fr;
    pick = reflect ? fr : 1.f - fr;
  } else if (s_refl) {
    reflect = true; pick = 1.f;
  } else if (s_trans && fr < 1.f) {
    reflect = false; pick = 1.f;
  } else {
    return zero;
  }
  if (!reflect) return zero;
  V3 w = m * (2.f * wfm) - wf;
  float base_pdf = m_pdf * 0.25f / fabsf(wfm);
  if (base_pdf < 1e-5f) return zero;
  float g = ggx_g1(alpha, wf, m) * ggx_g1(alpha, w, m);
  float d = ggx_d(alpha, m);
  wg = w;
  pdf = pick * base_pdf;
  fl = BSDF_GLOSSY | BSDF_REFLECTION;
  return b.refl() * (0.25f * fr) * (g * d / fabsf(wf.z * w.z));
}

// Bsdf::eval_f (bsdf.cpp:72-93); shading frame == geometric frame, so shading_to_geom = 1
// except where |n.wi| == 0.  ROUGH: the rough dielectric's branch is compiled in (kernels
// for scenes without one leave it out: it raised the shade kernel's register spill).
template <bool ROUGH = true>
DG_INLINE V3 bsdf_eval(const Bsdf& b, V3 wi, V3 wo, uint32_t req) {
  V3 wil = b.to_local(wi), wol = b.to_local(wo);
  bool refl = wol.z * wil.z > 0.f;
  uint32_t ereq = req & ((BSDF_LOBES & ~BSDF_DELTA) | (refl ? BSDF_REFLECTION : BSDF_TRANSMISSION));
  if (!(ereq & BSDF_MODES)) return v3(0.f, 0.f, 0.f);
  if (!b.matches(ereq)) return v3(0.f, 0.f, 0.f);
  float ds = fabsf(dot(b.nn, wi));
  if (ds * ds == 0.f) return v3(0.f, 0.f, 0.f);  // shading_to_geom == 0
  if (b.kind == DGPU_MAT_LAMBERT) return b.albedo() * INV_PI_F;
  if (b.kind == DGPU_MAT_GGX_CONDUCTOR) return conductor_f(b, wil, wol);
  if (ROUGH && b.kind == DGPU_MAT_ROUGH_DIELECTRIC)
    return rough_dielectric_f(b, wil, wol, (ereq & BSDF_REFLECTION) != 0u);
  return v3(0.f, 0.f, 0.f);
}

DG_INLINE float bsdf_pdf(const Bsdf& b, V3 wi, V3 wo, uint32_t req) {
  if (!b.matches(req)) return 0.f;
  V3 wil = b.to_local(wi), wol = b.to_local(wo);
  if (b.kind == DGPU_MAT_LAMBERT) return fabsf(wil.z * INV_PI_F);
  if (b.kind == DGPU_MAT_GGX_CONDUCTOR) return conductor_pdf(b, wil, wol);
  return 0.f;  // dielectric; rough dielectric: f_pdf is 0 (rough_dielectric_material.cpp:149-155)
}

// Bsdf::sample_light_f with one lobe (bsdf.cpp:95-192) and the lobes'
// sample functions (lambert_material.cpp:10-19, dielectric_material.cpp:50-104).
// NON_DELTA: the caller requests no delta lobe (NEE), so the dielectric branch is not compiled.
// ROUGH: the rough dielectric's branch is compiled in (see bsdf_eval).
template <bool NON_DELTA = false, bool ROUGH = true>
DG_INLINE V3 bsdf_sample(const Bsdf& b, V3 wo, uint32_t req, float u, float v, float c,
                         V3& wi, float& pdf, uint32_t& flags) {
  pdf = 0.f; flags = 0;
  V3 zero = v3(0.f, 0.f, 0.f);
  if (!b.matches(req)) return zero;
  V3 wol = b.to_local(wo), wil;
  V3 f = zero;
  float p = 0.f;
  uint32_t fl = 0;
  if (b.kind == DGPU_MAT_LAMBERT) {
    float r = sqrtf(u), th = TWO_PI_F * v;
    float sth, cth;
    sincos_small(th, sth, cth);
    float x = cth * r, y = sth * r;
    float z = sqrtf(fmaxf(0.f, 1.f - x * x - y * y));
    wil = v3(x, y, z);
    p = fabsf(z * INV_PI_F);
    fl = BSDF_REFLECTION | BSDF_DIFFUSE;
    wil.z = copysignf(wil.z, wol.z);
    f = b.albedo() * INV_PI_F;
  } else if (b.kind == DGPU_MAT_GGX_CONDUCTOR) {
    if (wol.z == 0.f) return zero;
    float sgn = wol.z < 0.f ? -1.f : 1.f;
    V3 wu = v3(wol.x, wol.y, wol.z * sgn);
    V3 m = ggx_sample_vndf(b.alpha(), wu, u, v);
    float wdm = dot(wu, m);
    V3 w = m * (2.f * wdm) - wu;
    if (w.z <= 0.f || wdm <= 0.f) return zero;
    wil = v3(w.x, w.y, w.z * sgn);
    p = conductor_pdf(b, wil, wol);
    if (!(p > 0.f)) return zero;
    fl = BSDF_REFLECTION | BSDF_GLOSSY;
    f = conductor_f(b, wil, wol);
  } else if (ROUGH && b.kind == DGPU_MAT_ROUGH_DIELECTRIC) {
    f = rough_dielectric_sample(b, wol, req, u, v, c, wil, p, fl);
    if (p == 0.f) return zero;
  } else if (!NON_DELTA && b.kind == DGPU_MAT_DIELECTRIC) {
    bool s_refl = req & BSDF_REFLECTION, s_trans = req & BSDF_TRANSMISSION;
    float ir, it;
    if (wol.z >= 0.f) { ir = b.ior_out(); it = b.ior_in(); } else { ir = b.ior_in(); it = b.ior_out(); }
    float cos_refl = fabsf(wol.z);
    V3 w_reflect = v3(-wol.x, -wol.y, wol.z);
    V3 w_trans;
    if (!refract_dielectric(wol, w_trans, ir, it)) {
      if (!s_refl) return zero;
      fl = BSDF_DELTA | BSDF_REFLECTION; p = 1.f; wil = w_reflect;
      f = b.refl() / cos_refl;
    } else {
      float cos_trans = fabsf(w_trans.z);
      float fr = fresnel_dielectric(cos_refl, cos_trans, ir, it);
      float r = fr, t = 1.f - r;
      bool reflect = c <= r && s_refl;
      bool transmit = c > r && s_trans;
      if (reflect) {
        fl = BSDF_DELTA | BSDF_REFLECTION; p = s_trans ? r : 1.f; wil = w_reflect;
        f = b.refl() * (r / cos_refl);
      } else if (transmit) {
        fl = BSDF_DELTA | BSDF_TRANSMISSION; p = s_refl ? t : 1.f; wil = w_trans;
        f = b.trans() * (t / cos_trans);
      } else {
        return zero;
      }
    }
  } else {
    return zero;
  }
  if (p == 0.f) return zero;
  flags = fl;
  wi = b.to_world(wil);
  pdf = p;
  float ds = fabsf(dot(b.nn, wi));
  if (ds * ds == 0.f) return zero;  // shading_to_geom == 0 (bsdf.cpp:237-242)
  return f;
}

// ----------------------------------------------------------------- geometry

// TriangleUv::hit differential geometry (triangle.cpp:137-171) for hit (b1, b2).
struct Surface {
  V3 p, nn, dpdu;
  float u, v, eps;
  uint32_t tri;    // scene triangle id
  uint32_t mat;    // material index
  int32_t light;   // light index or -1
};

DG_INLINE Surface surface_at(const DevScene& S, uint32_t leaf_idx, float t, float b1, float b2) {
  float4 a = S.tris[3 * leaf_idx], bb = S.tris[3 * leaf_idx + 1], cc = S.tris[3 * leaf_idx + 2];
  V3 p0 = xyz(a), p1 = xyz(bb), p2 = xyz(cc);
  uint32_t tri = __float_as_uint(a.w);
  V3 e1 = p1 - p0, e2 = p2 - p0;
  float2 uv0 = make_float2(0.f, 0.f), uv1 = make_float2(1.f, 0.f), uv2 = make_float2(0.f, 1.f);
  if (S.tri_uv) {
    uv0 = S.tri_uv[3 * leaf_idx]; uv1 = S.tri_uv[3 * leaf_idx + 1]; uv2 = S.tri_uv[3 * leaf_idx + 2];
  }
  float du1 = uv1.x - uv0.x, du2 = uv2.x - uv0.x, dv1 = uv1.y - uv0.y, dv2 = uv2.y - uv0.y;
  float det = du1 * dv2 - dv1 * du2;
  Surface s;
  if (det == 0.f) { V3 dv; coordinate_system(cross(e2, e1), s.dpdu, dv); }
  else s.dpdu = (e1 * dv2 - e2 * dv1) * (1.f / det);
  float b0 = 1.f - b1 - b2;
  s.u = b0 * uv0.x + b1 * uv1.x + b2 * uv2.x;
  s.v = b0 * uv0.y + b1 * uv1.y + b2 * uv2.y;
  s.p = p0 * b0 + p1 * b1 + p2 * b2;
  s.nn = normalize(cross(e2, e1));
  s.eps = fabsf(1e-3f * t);
  s.tri = tri;
  s.mat = __float_as_uint(bb.w);
  s.light = (int32_t)__float_as_uint(cc.w);
  return s;
}

DG_INLINE int32_t texel_of(uint4 info, float u, float v) {
  float wx = u - floorf(u), wy = v - floorf(v);
  int32_t px = (int32_t)floorf(float(info.x) * wx), py = (int32_t)floorf(float(info.y) * wy);
  px = px >= (int32_t)info.x ? (int32_t)info.x - 1 : px;
  py = py >= (int32_t)info.y ? (int32_t)info.y - 1 : py;
  return py * (int32_t)info.x + px;
}

// The texture coordinates of surface point g: its uv (uv_texture_map, texture_maps.cpp:4-8)
// or x, y of world_to_texture * p (xy_texture_map, texture_maps.cpp:10-15), summed in
// the order of dort's mul_mat_1 (mat.cpp:58-63) without fused multiply-adds, as the
// CPU computes it.
DG_INLINE void tex_coords(const DevScene& S, uint32_t tex, uint4 info, const Surface& g, float& u, float& v) {
  u = g.u;
  v = g.v;
  if (info.w == DGPU_TEXMAP_XY) {
#pragma clang fp contract(off)
    float4 r0 = S.tex_xform[2 * tex], r1 = S.tex_xform[2 * tex + 1];
    u = r0.x * g.p.x + r0.y * g.p.y + r0.z * g.p.z + r0.w;
    v = r1.x * g.p.x + r1.y * g.p.y + r1.z * g.p.z + r1.w;
  }
}

// Material::get_bsdf -> Bsdf frame (bsdf.cpp:50-66) + the material's bxdf.
DG_INLINE Bsdf make_bsdf(const DevScene& S, const Surface& g, int32_t* texel_out = nullptr) {
  Bsdf b;
  b.nn = g.nn;
  float len = length(g.dpdu);
  if (len < 1e-9f) { V3 t; coordinate_system(b.nn, b.sn, t); }  // its t is cross(nn, sn) too
  else b.sn = g.dpdu / len;
  const DevMaterial m = S.mats[g.mat];
  V3 zero = v3(0.f, 0.f, 0.f);
  b.kind = DGPU_MAT_NONE;
  b.c0 = zero; b.c1 = zero; b.c2 = zero;
  b.a0 = 0.f; b.a1 = 0.f;
  if (texel_out) *texel_out = -1;
  if (m.type == DGPU_MAT_LAMBERT) {
    V3 a = xyz(m.albedo);
    if (m.tex >= 0) {
      uint4 info = S.tex_info[m.tex];
      float tu, tv;
      tex_coords(S, (uint32_t)m.tex, info, g, tu, tv);
      int32_t ti = texel_of(info, tu, tv);
      a = xyz(S.tex_data[info.z + (uint32_t)ti]);
      if (texel_out && m.tex == 0) *texel_out = ti;
    }
    if (!is_black(a)) { b.kind = DGPU_MAT_LAMBERT; b.c0 = a; }
  } else if (m.type == DGPU_MAT_GGX_CONDUCTOR) {
    V3 a = xyz(m.albedo);
    if (!is_black(a)) {
      b.kind = DGPU_MAT_GGX_CONDUCTOR;
      b.c0 = a; b.c1 = xyz(m.eta); b.c2 = xyz(m.k); b.a0 = m.roughness;
    }
  } else if (m.type == DGPU_MAT_DIELECTRIC) {
    V3 r = xyz(m.refl), t = xyz(m.trans);
    if (m.ior_in > 0.f && m.albedo.w > 0.f && !(is_black(r) && is_black(t))) {
      b.kind = DGPU_MAT_DIELECTRIC;
      b.c0 = r; b.c1 = t; b.a0 = m.ior_in; b.a1 = m.albedo.w;
    }
  } else if (m.type == DGPU_MAT_ROUGH_DIELECTRIC) {
    // RoughDielectricMaterial::add_bxdfs (rough_dielectric_material.cpp:166-185): no lobe for
    // black tints, the smooth dielectric below roughness 0.001
    V3 r = xyz(m.refl), t = xyz(m.trans);
    if (!(is_black(r) && is_black(t))) {
      b.kind = (double)m.roughness < 0.001 ? DGPU_MAT_DIELECTRIC : DGPU_MAT_ROUGH_DIELECTRIC;
      b.c0 = r; b.c1 = t; b.a0 = m.ior_in; b.a1 = m.albedo.w;
      b.c2 = v3(m.roughness, 0.f, 0.f);
    }
  }
  return b;
}

// ----------------------------------------------------------------- environment

DG_INLINE void env_dir_to_uv(const DevScene& S, V3 d, float& u, float& v, float& sin_theta) {
  float cu = dot(d, S.env_up), cs = dot(d, S.env_s), ct = dot(d, S.env_t);
  float phi = atan2f(ct, cs);
  float theta = acosf(clampf(cu, -1.f, 1.f));
  sin_theta = sqrtf(fmaxf(0.f, 1.f - sqf(cu)));
  u = phi * INV_TWO_PI_F + 0.5f;
  v = theta * INV_PI_F;
}

// The texel of (u, v) (nearest, environment_light.cpp:153-159): rgb and probability;
// zero outside the map.
DG_INLINE float4 env_texel(const DevScene& S, float u, float v) {
  int32_t x = (int32_t)floorf(u * float(S.env_w)), y = (int32_t)floorf(v * float(S.env_h));
  if (x < 0 || y < 0 || x >= (int32_t)S.env_w || y >= (int32_t)S.env_h) return make_float4(0.f, 0.f, 0.f, 0.f);
  return S.env_rgb[(uint32_t)y * S.env_w + (uint32_t)x];
}

// Solid-angle pdf of a direction whose texel has probability `prob` (environment_light.cpp:99-118).
DG_INLINE float env_solid_angle_pdf(const DevScene& S, float prob, float sin_theta) {
  return sin_theta == 0.f ? 0.f : prob * (float(S.env_w) * float(S.env_h)) / (2.f * sqf(PI_F) * sin_theta);
}

// Radiance toward d and the pdf of sampling d, from one texel fetch.
DG_INLINE V3 env_eval(const DevScene& S, V3 d, float& pdf) {
  float u, v, st;
  env_dir_to_uv(S, d, u, v, st);
  float4 t = env_texel(S, u, v);
  pdf = env_solid_angle_pdf(S, t.w, st);
  return xyz(t);
}

DG_INLINE V3 env_background(const DevScene& S, V3 d) {
  float u, v, st;
  env_dir_to_uv(S, d, u, v, st);
  return xyz(env_texel(S, u, v));
}

// PiecewiseDistrib2d::sample_cdf_slice (piecewise_distrib_2d.cpp:58-80): the bin
// b = the last index below size-1 with cdf[b] < u (0 if none), plus the linear
// remainder.  dort bisects; here a guide table (Chen & Asau 1974, "On generating
// random variates from an empirical distribution") gives a start index near b and
// short scans in both directions reach exactly the bisection's b, replacing ~log2(size)
// dependent loads by one or two.
#ifndef DGPU_ENV_GUIDE
#define DGPU_ENV_GUIDE 1
#endif
DG_INLINE float sample_slice(const float* cdf, uint32_t size, const uint32_t* guide, uint32_t G, float u,
                             uint32_t& out_idx) {
#if DGPU_ENV_GUIDE
  uint32_t g = (uint32_t)(u * float(G));
  if (g >= G) g = G - 1u;
  uint32_t b = guide[g];
  while (b > 0u && cdf[b] >= u) --b;
  while (b + 2u < size && cdf[b + 1u] < u) ++b;
#else
  (void)guide; (void)G;
  uint32_t b = 0, e = size - 1;
  while (b + 1 < e) {
    uint32_t m = b + (e - b) / 2;
    if (cdf[m] < u) b = m; else e = m;
  }
#endif
  float c0 = cdf[b], c1 = cdf[b + 1];
  out_idx = b;
  return float(b) + (u - c0) / (c1 - c0);
}

// EnvironmentLight::sample_dir (environment_light.cpp:134-146); also returns the
// radiance of the sampled texel.
DG_INLINE V3 env_sample(const DevScene& S, float su, float sv, float& pdf, V3& radiance) {
  uint32_t yi, xi;
  float y = sample_slice(S.env_marg, S.env_h + 1, S.env_marg_guide, S.env_h, sv, yi);
  float x = sample_slice(S.env_cond + (size_t)(S.env_w + 1) * yi, S.env_w + 1,
                         S.env_cond_guide + (size_t)S.env_w * yi, S.env_w, su, xi);
  float u = x / float(S.env_w);
  float v = y / float(S.env_h);
  float theta = v * PI_F, phi = (u - 0.5f) * TWO_PI_F;
  float ct, st, cp, sp;
  sincos_small(theta, st, ct);
  sincos_small(phi, sp, cp);
  float4 t = env_texel(S, u, v);
  pdf = env_solid_angle_pdf(S, t.w, clampf(st, 0.f, 1.f));
  radiance = xyz(t);
  return S.env_up * ct + S.env_s * (st * cp) + S.env_t * (st * sp);
}

DG_INLINE V3 background_radiance(const DevScene& S, V3 d) {
  V3 r = v3(0.f, 0.f, 0.f);
  if (S.env_light >= 0) r = r + env_background(S, d);
  if (S.has_inf) r = r + S.inf_rad;
  return r;
}

// ----------------------------------------------------------------- camera

// PinholeCamera::sample_ray_importance (pinhole_camera.cpp:13-31)
DG_INLINE void camera_ray(const DevScene& S, float fx, float fy, uint32_t xr, uint32_t yr,
                          V3& o, V3& d) {
  float rx = float(xr), ry = float(yr);
  float cx = fx / rx - 0.5f, cy = fy / ry - 0.5f;
  float nx, ny;
  if (rx > ry) { nx = cx; ny = cy * ry / rx; } else { nx = cx * rx / ry; ny = cy; }
  V3 cd = v3(nx * S.project_dim, ny * S.project_dim, 1.f);
  const float* m = S.c2w;
  V3 wd = v3(m[0] * cd.x + m[1] * cd.y + m[2] * cd.z, m[4] * cd.x + m[5] * cd.y + m[6] * cd.z,
             m[8] * cd.x + m[9] * cd.y + m[10] * cd.z);
  o = v3(m[3], m[7], m[11]);
  d = normalize(wd);
}

}  // namespace dgpu
