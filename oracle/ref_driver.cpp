// ref_driver.cpp -- a C ABI over dort's own pt-path classes (TEST INFRASTRUCTURE ONLY).
//
// Compiled together with dort's sources where they lie (oracle/ref.mk) into
// oracle/_ref/libdortref.so.  It lets the tests run the real reference on the
// same dgpu_scene_desc the oracle and the GPU path take, to pin the oracle:
//   ref_scene_create   Scene assembled the way lua_builder.cpp:108-142 and
//                      lua_make_aggregate (lua_builder.cpp:700-712) do it
//   ref_trace          Scene::intersect / intersect_p (scene.cpp)
//   ref_render         PathRenderer::render (path_renderer.cpp:11-21) into a
//                      Film with the default 1 px box filter (lua_render.cpp:108-115)
//   ref_render_light   LightRenderer::render (light_renderer.cpp:13-45), the splat sums
//   ref_render_bdpt    BdptRenderer::render (bdpt_renderer.cpp:9-41): colour, weight and splat sums
//
// Albedo textures become dort's ImageTexture (8-bit texels) composed with uv_texture_map or
// xy_texture_map, as lua_texture.cpp:346-388 composes them; float texels must be exact
// 8-bit values k / 256.  Scenes with a GGX conductor, or textures that are not 8-bit, are
// refused (dort has no rough conductor; its ImageTexture is Image<PixelRgb8>).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "dort/camera.hpp"
#include "dort/ctx.hpp"
#include "dort/dielectric_material.hpp"
#include "dort/diffuse_light.hpp"
#include "dort/environment_light.hpp"
#include "dort/film.hpp"
#include "dort/filter.hpp"
#include "dort/image.hpp"
#include "dort/image_texture.hpp"
#include "dort/infinite_light.hpp"
#include "dort/lambert_material.hpp"
#include "dort/list_primitive.hpp"
#include "dort/bvh_primitive.hpp"
#include "dort/bdpt_renderer.hpp"
#include "dort/mat.hpp"
#include "dort/mesh.hpp"
#include "dort/light_renderer.hpp"
#include "dort/path_renderer.hpp"
#include "dort/pinhole_camera.hpp"
#include "dort/primitive.hpp"
#include "dort/rough_dielectric_material.hpp"
#include "dort/random_sampler.hpp"
#include "dort/rng.hpp"
#include "dort/scene.hpp"
#include "dort/texture.hpp"
#include "dort/texture_maps.hpp"
#include "dort/thread_pool.hpp"
#include "dort/transform.hpp"
#include "dort/triangle_shape.hpp"
#include "dort/triangle_shape_primitive.hpp"

#include "../include/dortgpu.h"

using namespace dort;

namespace {

struct RefScene {
  std::shared_ptr<Scene> scene;
  std::shared_ptr<Mesh> mesh;
  std::shared_ptr<Camera> camera;
  std::map<const Primitive*, int32_t> prim_index;  // leaf primitive -> triangle
};

struct NoProgress final : Progress {
  bool is_cancelled() const override { return false; }
  void set_percent_done(float) override {}
};

Spectrum spec3(const float* v) { return Spectrum(v[0], v[1], v[2]); }

// 4x4 inverse by cofactors, in double.
void invert4(const double m[4][4], double out[4][4]) {
  double inv[16], a[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) a[c * 4 + r] = m[c][r];
  inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
  inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
  inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
  inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
  inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
  inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
  inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
  inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
  inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
  inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
  inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
  inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
  inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
  inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
  inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
  inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
  double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) out[c][r] = inv[c * 4 + r] / det;
}

// camera_to_world is row-major 3x4; dort's Mat4x4 stores cols[col][row].
Transform camera_transform(const dgpu_camera& cam) {
  double m[4][4], mi[4][4];
  for (int c = 0; c < 4; ++c) {
    for (int r = 0; r < 3; ++r) m[c][r] = cam.camera_to_world[r * 4 + c];
    m[c][3] = c == 3 ? 1.0 : 0.0;
  }
  invert4(m, mi);
  float f[4][4], fi[4][4];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) { f[c][r] = (float)m[c][r]; fi[c][r] = (float)mi[c][r]; }
  return Transform(Mat4x4(f), Mat4x4(fi));
}

// dort's 8-bit image of texture t, or null when a float texel is not some k / 256.
std::shared_ptr<Image<PixelRgb8>> texture_image(const dgpu_texture& t) {
  auto img = std::make_shared<Image<PixelRgb8>>(t.width, t.height);
  for (uint32_t y = 0; y < t.height; ++y)
    for (uint32_t x = 0; x < t.width; ++x) {
      size_t i = 3 * (size_t(y) * t.width + x);
      uint8_t c[3];
      for (int k = 0; k < 3; ++k) {
        if (t.rgb8) { c[k] = t.rgb8[i + k]; continue; }
        float v = t.rgb[i + k] * 256.f;
        if (!(v >= 0.f && v <= 255.f && v == std::floor(v))) return nullptr;
        c[k] = (uint8_t)v;
      }
      img->set_pixel(x, y, PixelRgb8(c[0], c[1], c[2]));
    }
  return img;
}

// ImageTexture composed with the texture's mapping (lua_texture.cpp:346-388).
std::shared_ptr<TextureGeom<Spectrum>> albedo_texture(const dgpu_texture& t) {
  auto img = texture_image(t);
  if (!img) return nullptr;
  std::shared_ptr<Texture<Spectrum, Vec2>> tex = std::make_shared<ImageTexture>(img);
  std::shared_ptr<TextureGeom<Vec2>> map;
  if (t.mapping == DGPU_TEXMAP_XY) {
    // xy_texture_map applies tex_to_world's inverse matrix, which is world_to_texture as given
    double m[4][4], mi[4][4];
    for (int c = 0; c < 4; ++c) {
      for (int r = 0; r < 3; ++r) m[c][r] = t.world_to_texture[r * 4 + c];
      m[c][3] = c == 3 ? 1.0 : 0.0;
    }
    invert4(m, mi);
    float f[4][4], fi[4][4];
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 4; ++r) { f[c][r] = (float)mi[c][r]; fi[c][r] = (float)m[c][r]; }
    map = xy_texture_map(Transform(Mat4x4(f), Mat4x4(fi)));
  } else {
    map = uv_texture_map();
  }
  return compose_texture(tex, map);
}

std::shared_ptr<Material> make_material(const dgpu_material& m, const dgpu_scene_desc& d) {
  switch (m.type) {
    case DGPU_MAT_NONE:
      return std::make_shared<LambertMaterial>(const_texture(Spectrum(0.f)));
    case DGPU_MAT_LAMBERT: {
      if (m.albedo_tex < 0) return std::make_shared<LambertMaterial>(const_texture(spec3(m.albedo)));
      if (m.albedo_tex >= (int32_t)d.num_textures) return nullptr;
      auto tex = albedo_texture(d.textures[m.albedo_tex]);
      if (!tex) return nullptr;
      return std::make_shared<LambertMaterial>(tex);
    }
    case DGPU_MAT_ROUGH_DIELECTRIC:
      return std::make_shared<RoughDielectricMaterial>(const_texture(spec3(m.reflect_tint)),
                                                       const_texture(spec3(m.transmit_tint)),
                                                       const_texture(m.roughness), m.ior_inside, m.ior_outside,
                                                       MicrofacetType::GGX);
    case DGPU_MAT_DIELECTRIC:
      return std::make_shared<DielectricMaterial>(const_texture(spec3(m.reflect_tint)),
                                                  const_texture(spec3(m.transmit_tint)),
                                                  m.ior_inside, m.ior_outside, false);
    default:
      return nullptr;
  }
}

std::shared_ptr<ThreadPool> make_pool(uint32_t threads) {
  auto pool = std::make_shared<ThreadPool>();
  pool->start(threads < 1 ? 1 : threads);
  return pool;
}

}  // namespace

extern "C" {

void* ref_scene_create(const dgpu_scene_desc* d, uint32_t split_method, uint32_t leaf_size,
                       uint32_t max_leaf_size, uint32_t threads) {
  if (!d || d->num_triangles == 0) return nullptr;
  std::vector<std::shared_ptr<Material>> mats;
  for (uint32_t i = 0; i < d->num_materials; ++i) {
    mats.push_back(make_material(d->materials[i], *d));
    if (!mats.back()) return nullptr;
  }
  auto rs = new RefScene();
  rs->mesh = std::make_shared<Mesh>();
  Mesh& mesh = *rs->mesh;
  for (uint32_t i = 0; i < d->num_vertices; ++i)
    mesh.points.push_back(Point(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2]));
  if (d->uvs) {
    for (uint32_t i = 0; i < d->num_vertices; ++i) mesh.uvs.push_back(Vec2(d->uvs[2 * i], d->uvs[2 * i + 1]));
  }
  mesh.vertices.assign(d->indices, d->indices + 3 * size_t(d->num_triangles));

  std::vector<const dgpu_area_light*> tri_light(d->num_triangles, nullptr);
  for (uint32_t i = 0; i < d->num_area_lights; ++i) tri_light[d->area_lights[i].triangle] = &d->area_lights[i];

  auto scene = std::make_shared<Scene>();
  std::vector<std::shared_ptr<Light>> area_lights;
  std::vector<std::unique_ptr<Primitive>> prims;
  // A triangle is named by the offset of its first vertex in mesh.vertices
  // (lua_builder.cpp:522 passes t * 3).
  for (uint32_t t = 0; t < d->num_triangles; ++t) {
    std::unique_ptr<Primitive> p;
    const auto& mat = mats.at(d->tri_material[t]);
    if (tri_light[t]) {
      // add_shape with an area light (lua_builder.cpp:327-348)
      auto shape = std::make_shared<TriangleShape>(rs->mesh.get(), 3 * t);
      auto light = std::make_shared<DiffuseLight>(shape, Transform(), spec3(tri_light[t]->radiance));
      area_lights.push_back(light);
      p = std::make_unique<ShapePrimitive>(shape, mat, light, Transform());
    } else {
      // add_triangle (lua_builder.cpp:375-390)
      p = std::make_unique<TriangleShapePrimitive>(rs->mesh.get(), 3 * t, mat);
    }
    rs->prim_index[p.get()] = (int32_t)t;
    prims.push_back(std::move(p));
  }
  // Lights in the order the oracle and the GPU path use: area lights by
  // description order, then the environment map, then the infinite light.
  for (uint32_t i = 0; i < d->num_area_lights; ++i) {
    uint32_t t = d->area_lights[i].triangle;
    uint32_t k = 0;
    for (uint32_t u = 0; u < t; ++u) k += tri_light[u] != nullptr;
    scene->lights.push_back(area_lights.at(k));
  }
  if (d->env) {
    auto img = std::make_shared<Image<PixelRgbFloat>>(d->env->width, d->env->height);
    for (uint32_t y = 0; y < d->env->height; ++y)
      for (uint32_t x = 0; x < d->env->width; ++x) {
        const float* c = d->env->rgb + 3 * (size_t(y) * d->env->width + x);
        img->set_pixel(x, y, PixelRgbFloat(c[0], c[1], c[2]));
      }
    scene->lights.push_back(std::make_shared<EnvironmentLight>(
        img, spec3(d->env->scale), Vector(d->env->up[0], d->env->up[1], d->env->up[2]),
        Vector(d->env->fwd[0], d->env->fwd[1], d->env->fwd[2])));
  }
  if (d->infinite_radiance) scene->lights.push_back(std::make_shared<InfiniteLight>(spec3(d->infinite_radiance)));

  BvhOpts opts;
  opts.split_method = split_method == 0 ? BvhSplitMethod::Middle : BvhSplitMethod::Sah;
  opts.leaf_size = leaf_size;
  opts.max_leaf_size = max_leaf_size;
  auto pool = make_pool(threads);
  if (prims.size() == 1) {
    scene->primitive = std::move(prims.at(0));
  } else if (prims.size() <= opts.leaf_size) {
    scene->primitive = std::make_unique<ListPrimitive>(std::move(prims));
  } else {
    scene->primitive = std::make_unique<BvhPrimitive>(std::move(prims), opts, *pool);
  }
  pool->stop();
  scene->bounds = scene->primitive->bounds();
  scene->centroid = scene->bounds.centroid();
  scene->radius = scene->bounds.radius();
  for (const auto& light : scene->lights)
    if (light->flags & LIGHT_BACKGROUND) scene->background_lights.push_back(light);
  scene->meshes.push_back(rs->mesh);
  rs->camera = std::make_shared<PinholeCamera>(camera_transform(d->camera), d->camera.fov);
  scene->default_camera = rs->camera;
  rs->scene = scene;
  return rs;
}

void ref_scene_destroy(void* p) { delete (RefScene*)p; }

}  // extern "C"

// For other test drivers linked with this one (oracle/ref_gpu_glue.cpp).
std::shared_ptr<Scene> ref_scene_scene(void* p) { return ((RefScene*)p)->scene; }
std::shared_ptr<Camera> ref_scene_camera(void* p) { return ((RefScene*)p)->camera; }

extern "C" {

// rays: 8 pointers (ox, oy, oz, dx, dy, dz, tmin, tmax).  Closest hits also give
// the surface uv and point of DiffGeom (uv3 = u, v, ray epsilon; pt = p), or NULL.
void ref_trace(void* p, uint32_t n, const float* const* rays, int any, float* t, int32_t* prim,
               uint32_t* occluded, float* uv3, float* pt) {
  const RefScene& rs = *(const RefScene*)p;
  for (uint32_t i = 0; i < n; ++i) {
    Ray ray(Point(rays[0][i], rays[1][i], rays[2][i]), Vector(rays[3][i], rays[4][i], rays[5][i]),
            rays[6][i], rays[7][i]);
    if (any) {
      occluded[i] = rs.scene->intersect_p(ray) ? 1u : 0u;
      continue;
    }
    Intersection isect;
    if (rs.scene->intersect(ray, isect)) {
      t[i] = ray.t_max;
      auto it = rs.prim_index.find(isect.primitive);
      prim[i] = it == rs.prim_index.end() ? -2 : it->second;
      const DiffGeom& g = isect.world_diff_geom;
      if (uv3) { uv3[3 * i] = g.uv.x; uv3[3 * i + 1] = g.uv.y; uv3[3 * i + 2] = isect.ray_epsilon; }
      if (pt) { pt[3 * i] = g.p.v.x; pt[3 * i + 1] = g.p.v.y; pt[3 * i + 2] = g.p.v.z; }
    } else {
      t[i] = INFINITY;
      prim[i] = -1;
    }
  }
}

// Renders opts->sample_end iterations (sample_begin must be 0) with dort's
// RandomSampler seeded by opts->seed.  Outputs the film's per-pixel filtered
// colour sums and weights.  The pool runs `threads` threads; with 1 thread
// the result is deterministic.
// LightRenderer::render (light_renderer.cpp:13-45) with opts->sample_end iterations and
// lengths min_depth + 2 .. max_depth + 2 (lua_render.cpp:146-150).  Outputs the film's
// per-pixel splat sums (Film::Pixel::splat; the image is the sum / iterations).
int ref_render_light(void* p, const dgpu_render_opts* o, uint32_t threads, float* rgb) {
  const RefScene& rs = *(const RefScene*)p;
  if (o->sample_begin != 0 || o->sample_end == 0) return 1;
  CtxG ctx;
  ctx.pool = make_pool(threads);
  auto film = std::make_shared<Film>(o->x_res, o->y_res, std::make_shared<BoxFilter>(Vec2(0.5f, 0.5f)));
  auto sampler = std::make_shared<RandomSampler>(1, Rng((uint32_t)o->seed));
  LightRenderer r(rs.scene, film, sampler, rs.camera, o->sample_end, o->min_depth + 2, o->max_depth + 2);
  NoProgress progress;
  r.render(ctx, progress);
  ctx.pool->stop();
  for (size_t i = 0; i < film->pixels.size(); ++i) {
    Spectrum sp = film->pixels[i].splat.load_relaxed();
    rgb[3 * i] = sp.red();
    rgb[3 * i + 1] = sp.green();
    rgb[3 * i + 2] = sp.blue();
  }
  return 0;
}

// BdptRenderer::render (bdpt_renderer.cpp:9-41) with opts->sample_end iterations, depths
// min_depth .. max_depth and use_t1_paths (lua_render.cpp:151-160).  Outputs the film's
// per-pixel colour sums, weights and splat sums (the image is colour / weight +
// splat / iterations).
int ref_render_bdpt(void* p, const dgpu_render_opts* o, uint32_t threads, uint32_t use_t1, float* rgb, float* w,
                    float* splat) {
  const RefScene& rs = *(const RefScene*)p;
  if (o->sample_begin != 0 || o->sample_end == 0) return 1;
  CtxG ctx;
  ctx.pool = make_pool(threads);
  auto film = std::make_shared<Film>(o->x_res, o->y_res, std::make_shared<BoxFilter>(Vec2(0.5f, 0.5f)));
  auto sampler = std::make_shared<RandomSampler>(1, Rng((uint32_t)o->seed));
  BdptRenderer r(rs.scene, film, sampler, rs.camera, o->sample_end, o->min_depth, o->max_depth, use_t1 != 0, "");
  NoProgress progress;
  r.render(ctx, progress);
  ctx.pool->stop();
  for (size_t i = 0; i < film->pixels.size(); ++i) {
    const Film::Pixel& px = film->pixels[i];
    Spectrum sp = px.splat.load_relaxed();
    rgb[3 * i] = px.color.red();
    rgb[3 * i + 1] = px.color.green();
    rgb[3 * i + 2] = px.color.blue();
    w[i] = px.weight;
    splat[3 * i] = sp.red();
    splat[3 * i + 1] = sp.green();
    splat[3 * i + 2] = sp.blue();
  }
  return 0;
}

int ref_render(void* p, const dgpu_render_opts* o, uint32_t threads, float* rgb, float* w) {
  const RefScene& rs = *(const RefScene*)p;
  if (o->sample_begin != 0 || o->sample_end == 0) return 1;
  CtxG ctx;
  ctx.pool = make_pool(threads);
  auto film = std::make_shared<Film>(o->x_res, o->y_res, std::make_shared<BoxFilter>(Vec2(0.5f, 0.5f)));
  auto sampler = std::make_shared<RandomSampler>(1, Rng((uint32_t)o->seed));
  PathRenderer::DirectStrategy strat = o->direct_strategy == DGPU_DIRECT_BSDF
                                           ? PathRenderer::DirectStrategy::SAMPLE_BSDF
                                       : o->direct_strategy == DGPU_DIRECT_LIGHT
                                           ? PathRenderer::DirectStrategy::SAMPLE_LIGHT
                                           : PathRenderer::DirectStrategy::MIS;
  PathRenderer r(rs.scene, film, sampler, rs.camera, o->sample_end, o->min_depth, o->max_depth,
                 o->only_direct != 0, o->sample_all_lights != 0, strat);
  NoProgress progress;
  r.render(ctx, progress);
  ctx.pool->stop();
  for (size_t i = 0; i < film->pixels.size(); ++i) {
    const auto& px = film->pixels[i];
    rgb[3 * i] = px.color.red();
    rgb[3 * i + 1] = px.color.green();
    rgb[3 * i + 2] = px.color.blue();
    w[i] = px.weight;
  }
  return 0;
}

}  // extern "C"
