// light_paths.hpp -- device helpers shared by the kernels that start paths on lights or
// connect to the camera: the light tracer and the bidirectional path tracer (kernels.hip,
// bdpt.hip), and the emitter-triangle tests of the path tracer's NEE.
#pragma once
#include "dgpu_internal.hpp"
#include "traverse.hpp"

namespace dgpu {

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

// DiscreteDistrib1d::sample / pdf over the lights' power (discrete_distrib_1d.cpp:22-42,
// scene.cpp:20-28): the light for u and its probability.
__device__ __forceinline__ uint32_t lt_pick_light(const DevScene& S, float u, float& prob) {
  uint32_t b = 0, e = S.num_lights;
  while (b + 1u < e) {
    uint32_t m = b + (e - b) / 2u;
    if (S.light_cdf[m] < u) b = m; else e = m;
  }
  prob = S.light_cdf[b + 1] - S.light_cdf[b];
  return b;
}

// Triangle::sample_point (triangle.cpp:62-76) on an area light's triangle.
__device__ __forceinline__ V3 lt_area_point(const DevScene& S, const DevLight& L, float u, float v, V3& n) {
  float4 a = S.tris[3 * L.tri], b = S.tris[3 * L.tri + 1], c = S.tris[3 * L.tri + 2];
  V3 p0 = xyz(a), e1 = xyz(b) - p0, e2 = xyz(c) - p0;
  float s = sqrtf(u);
  n = normalize(cross(e2, e1));
  return p0 + e1 * (1.f - s) + e2 * (v * s);
}

// World point and vector in the camera frame (Transform::apply_inv).
__device__ __forceinline__ V3 lt_cam_vector(const DevScene& S, V3 w) {
  const float* m = S.w2c;
  return v3(m[0] * w.x + m[1] * w.y + m[2] * w.z, m[4] * w.x + m[5] * w.y + m[6] * w.z,
            m[8] * w.x + m[9] * w.y + m[10] * w.z);
}
__device__ __forceinline__ V3 lt_cam_point(const DevScene& S, V3 p) {
  const float* m = S.w2c;
  return lt_cam_vector(S, p) + v3(m[3], m[7], m[11]);
}

// PinholeCamera::eval_importance / sample_pivot_importance with get_film_pos
// (pinhole_camera.cpp:33-82, 104-126): film position of the camera-frame direction c
// and the importance 1 / (A cos^3 theta); false when c is behind or outside the image.
__device__ __forceinline__ bool lt_importance(const WaveParams& P, V3 c, float2& fp, float& imp) {
  if (!(c.z > 0.f)) return false;
  float nx = c.x / c.z / P.S.project_dim, ny = c.y / c.z / P.S.project_dim;
  if (fabsf(nx) > 0.5f || fabsf(ny) > 0.5f) return false;
  float rx = float(P.xr), ry = float(P.yr);
  float cx = nx, cy = ny;
  if (rx > ry) cy = ny * rx / ry; else cx = nx * ry / rx;
  fp = make_float2((cx + 0.5f) * rx, (cy + 0.5f) * ry);
  float ct = c.z / length(c);
  float area = fminf(rx, ry) / fmaxf(rx, ry) * sqf(P.S.project_dim);
  imp = 1.f / (area * (ct * ct * ct));
  return true;
}

struct LtRay {
  V3 o, d, nn, rad;
  float tmin, pos_pdf, dir_pdf;
};

// Light::sample_ray_radiance (diffuse_light.cpp:15-38, environment_light.cpp:34-52,
// infinite_light.cpp:5-23): r[0..1] the position sample, r[2..3] the direction sample.
__device__ __forceinline__ LtRay lt_light_ray(const DevScene& S, const DevLight& L, const float r[4]) {
  LtRay out;
  if (L.kind == LK_AREA) {
    V3 n;
    V3 p = lt_area_point(S, L, r[0], r[1], n);
    float z = r[2], rr = sqrtf(fmaxf(0.f, 1.f - z * z)), sp, cp;  // uniform_hemisphere_sample
    sincos_small(TWO_PI_F * r[3], sp, cp);
    V3 w = v3(cp * rr, sp * rr, z);
    if (dot(w, n) < 0.f) w = -w;
    out.o = p; out.d = w; out.nn = n; out.tmin = 1e-3f;
    out.pos_pdf = 1.f / L.area;
    out.dir_pdf = 1.f / TWO_PI_F;
    out.rad = xyz(L.radiance);
    return out;
  }
  V3 dir;
  if (L.kind == LK_ENV) {
    dir = env_sample(S, r[2], r[3], out.dir_pdf, out.rad);
  } else {
    float z = 1.f - 2.f * r[2], rr = sqrtf(fmaxf(0.f, 1.f - z * z)), sp, cp;  // uniform_sphere_sample
    sincos_small(TWO_PI_F * r[3], sp, cp);
    dir = v3(cp * rr, sp * rr, z);
    out.dir_pdf = INV_FOUR_PI_F;
    out.rad = xyz(L.radiance);  // the light table holds the infinite light's radiance
  }
  float dr = sqrtf(r[0]), ds, dc;  // uniform_disk_sample
  sincos_small(TWO_PI_F * r[1], ds, dc);
  V3 su, sv;
  coordinate_system(dir, su, sv);
  V3 off = su * (dc * dr) + sv * (ds * dr);
  if (L.kind == LK_ENV) {  // rays come in from the direction sampled
    out.o = S.centroid + (dir + off) * S.radius;
    out.d = -dir;
  } else {
    out.o = S.centroid + (-dir + off) * S.radius;
    out.d = dir;
  }
  // The disk lies outside the scene's bounding sphere except at its centre; advance the
  // origin along the ray to the sphere (|off| = dr): no geometry lies before it, and from
  // there the trace kernels' fast box test holds, whose padding covers origins within
  // the sphere but not the whole disk.
  out.o = out.o + out.d * (S.radius - S.radius * sqrtf(fmaxf(0.f, 1.f - dr * dr)));
  out.nn = out.d;
  out.tmin = 0.f;
  out.pos_pdf = INV_PI_F / sqf(S.radius);
  return out;
}

}  // namespace dgpu
