// ref_gpu_glue.cpp -- drives dort_amd::GpuPathRenderer, GpuLightRenderer and GpuBdptRenderer inside dort
// (TEST INFRASTRUCTURE ONLY).
//
// Linked with dort's own objects and ref_driver.cpp into oracle/_ref/libdortref_gpu.so
// (oracle/ref.mk).  It renders a scene assembled by ref_scene_create through the
// Renderer seam exactly as dort's Lua front-end would (a Film with the default box
// filter, a RandomSampler, the scene's camera, Renderer::render with a Progress),
// with the GPU renderer in place of PathRenderer.
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <vector>

#include "dort/camera.hpp"
#include "dort/ctx.hpp"
#include "dort/film.hpp"
#include "dort/filter.hpp"
#include "dort/random_sampler.hpp"
#include "dort/rng.hpp"
#include "dort/scene.hpp"
#include "dort/thread_pool.hpp"
#include "../dort_amd/host/gpu_path_renderer.hpp"

std::shared_ptr<dort::Scene> ref_scene_scene(void* p);
std::shared_ptr<dort::Camera> ref_scene_camera(void* p);

namespace {

// Cancels once `after` chunks have been reported (0: never).
struct CountingProgress final : dort::Progress {
  uint32_t after = 0, reports = 0;
  float last = 0.f;
  bool is_cancelled() const override { return after != 0 && reports >= after; }
  void set_percent_done(float done) override { ++reports; last = done; }
};

std::shared_ptr<dort_amd::GpuSceneDesc> typed_scene(const dgpu_scene_desc* d) {
  auto g = std::make_shared<dort_amd::GpuSceneDesc>();
  for (uint32_t i = 0; i < d->num_materials; ++i) g->add_material(d->materials[i]);
  for (uint32_t i = 0; i < d->num_textures; ++i) {
    const dgpu_texture& t = d->textures[i];
    g->add_texture(t.width, t.height, std::vector<float>(t.rgb, t.rgb + size_t(t.width) * t.height * 3));
  }
  std::vector<float> pts(d->positions, d->positions + 3 * size_t(d->num_vertices));
  std::vector<uint32_t> tris(d->indices, d->indices + 3 * size_t(d->num_triangles));
  std::vector<uint32_t> mats(d->tri_material, d->tri_material + d->num_triangles);
  std::vector<float> uvs;
  if (d->uvs) uvs.assign(d->uvs, d->uvs + 2 * size_t(d->num_vertices));
  g->add_mesh(pts, tris, mats, d->uvs ? &uvs : nullptr);
  for (uint32_t i = 0; i < d->num_area_lights; ++i) g->add_area_light(d->area_lights[i].triangle, d->area_lights[i].radiance);
  if (d->env) {
    const dgpu_env_light& e = *d->env;
    g->set_environment(e.width, e.height, std::vector<float>(e.rgb, e.rgb + size_t(e.width) * e.height * 3),
                       e.up, e.fwd, e.scale);
  }
  if (d->infinite_radiance) g->set_infinite_light(d->infinite_radiance);
  g->set_camera(d->camera.camera_to_world, d->camera.fov);
  return g;
}

}  // namespace

extern "C" {

// Renders o->sample_end iterations (sample_begin must be 0) with the GPU renderer.
// filter_kind 0: dort's box filter of radius 0.5; 1: a triangle filter (rejected).
// Returns 0, or 1 with the exception's message in `err`.
int refgpu_render(void* ref_scene, const dgpu_scene_desc* d, const dgpu_render_opts* o,
                  uint32_t sampler_seed, uint32_t cancel_after, uint32_t chunk, int filter_kind,
                  float* rgb, float* w, uint64_t* out_seed, uint64_t* out_paths, float* out_percent,
                  char* err, size_t err_len) {
  try {
    std::shared_ptr<dort::Filter> filter;
    if (filter_kind == 0) filter = std::make_shared<dort::BoxFilter>(dort::Vec2(0.5f, 0.5f));
    else filter = std::make_shared<dort::TriangleFilter>(dort::Vec2(0.5f, 0.5f));
    auto film = std::make_shared<dort::Film>(o->x_res, o->y_res, filter);
    auto sampler = std::make_shared<dort::RandomSampler>(1, dort::Rng(sampler_seed));
    dort_amd::GpuPathOptions opts;
    opts.iterations = o->sample_end;
    opts.min_depth = o->min_depth;
    opts.max_depth = o->max_depth;
    opts.rr_depth = o->rr_depth;
    opts.only_direct = o->only_direct != 0;
    opts.sample_all_lights = o->sample_all_lights != 0;
    opts.direct_strategy = o->direct_strategy == DGPU_DIRECT_BSDF ? dort::PathRenderer::DirectStrategy::SAMPLE_BSDF
                           : o->direct_strategy == DGPU_DIRECT_LIGHT ? dort::PathRenderer::DirectStrategy::SAMPLE_LIGHT
                                                                     : dort::PathRenderer::DirectStrategy::MIS;
    opts.iterations_per_chunk = chunk;
    dort_amd::GpuPathRenderer r(ref_scene_scene(ref_scene), film, sampler, ref_scene_camera(ref_scene),
                                typed_scene(d), opts);
    dort::CtxG ctx;
    ctx.pool = std::make_shared<dort::ThreadPool>();
    ctx.pool->start(1);
    CountingProgress progress;
    progress.after = cancel_after;
    r.render(ctx, progress);
    ctx.pool->stop();
    for (size_t i = 0; i < film->pixels.size(); ++i) {
      rgb[3 * i] = film->pixels[i].color.red();
      rgb[3 * i + 1] = film->pixels[i].color.green();
      rgb[3 * i + 2] = film->pixels[i].color.blue();
      w[i] = film->pixels[i].weight;
    }
    *out_seed = r.last_seed();
    *out_paths = r.last_stats().paths;
    *out_percent = progress.last;
    return 0;
  } catch (const std::exception& e) {
    std::snprintf(err, err_len, "%s", e.what());
    return 1;
  }
}

// GpuLightRenderer (dort's "lt") in the same seam: the film's splat sums and splat_scale.
int refgpu_render_light(void* ref_scene, const dgpu_scene_desc* d, const dgpu_render_opts* o,
                        uint32_t sampler_seed, uint32_t chunk, float* rgb, float* out_scale, uint64_t* out_seed,
                        uint64_t* out_paths, char* err, size_t err_len) {
  try {
    auto film = std::make_shared<dort::Film>(o->x_res, o->y_res,
                                             std::make_shared<dort::BoxFilter>(dort::Vec2(0.5f, 0.5f)));
    auto sampler = std::make_shared<dort::RandomSampler>(1, dort::Rng(sampler_seed));
    dort_amd::GpuLightOptions opts;
    opts.iterations = o->sample_end;
    opts.min_depth = o->min_depth;
    opts.max_depth = o->max_depth;
    opts.iterations_per_chunk = chunk;
    dort_amd::GpuLightRenderer r(ref_scene_scene(ref_scene), film, sampler, ref_scene_camera(ref_scene),
                                 typed_scene(d), opts);
    dort::CtxG ctx;
    ctx.pool = std::make_shared<dort::ThreadPool>();
    ctx.pool->start(1);
    CountingProgress progress;
    r.render(ctx, progress);
    ctx.pool->stop();
    for (size_t i = 0; i < film->pixels.size(); ++i) {
      dort::Spectrum sp = film->pixels[i].splat.load_relaxed();
      rgb[3 * i] = sp.red();
      rgb[3 * i + 1] = sp.green();
      rgb[3 * i + 2] = sp.blue();
    }
    *out_scale = film->splat_scale;
    *out_seed = r.last_seed();
    *out_paths = r.last_stats().paths;
    return 0;
  } catch (const std::exception& e) {
    std::snprintf(err, err_len, "%s", e.what());
    return 1;
  }
}

// GpuBdptRenderer (dort's "bdpt") in the same seam: colour sums, weights, splat sums and
// splat_scale of the film.
int refgpu_render_bdpt(void* ref_scene, const dgpu_scene_desc* d, const dgpu_render_opts* o,
                       uint32_t sampler_seed, uint32_t chunk, uint32_t use_t1, float* rgb, float* w, float* splat,
                       float* out_scale, uint64_t* out_seed, uint64_t* out_paths, char* err, size_t err_len) {
  try {
    auto film = std::make_shared<dort::Film>(o->x_res, o->y_res,
                                             std::make_shared<dort::BoxFilter>(dort::Vec2(0.5f, 0.5f)));
    auto sampler = std::make_shared<dort::RandomSampler>(1, dort::Rng(sampler_seed));
    dort_amd::GpuBdptOptions opts;
    opts.iterations = o->sample_end;
    opts.min_depth = o->min_depth;
    opts.max_depth = o->max_depth;
    opts.use_t1_paths = use_t1 != 0;
    opts.iterations_per_chunk = chunk;
    dort_amd::GpuBdptRenderer r(ref_scene_scene(ref_scene), film, sampler, ref_scene_camera(ref_scene),
                                typed_scene(d), opts);
    dort::CtxG ctx;
    ctx.pool = std::make_shared<dort::ThreadPool>();
    ctx.pool->start(1);
    CountingProgress progress;
    r.render(ctx, progress);
    ctx.pool->stop();
    for (size_t i = 0; i < film->pixels.size(); ++i) {
      const dort::Film::Pixel& px = film->pixels[i];
      dort::Spectrum sp = px.splat.load_relaxed();
      rgb[3 * i] = px.color.red();
      rgb[3 * i + 1] = px.color.green();
      rgb[3 * i + 2] = px.color.blue();
      w[i] = px.weight;
      splat[3 * i] = sp.red();
      splat[3 * i + 1] = sp.green();
      splat[3 * i + 2] = sp.blue();
    }
    *out_scale = film->splat_scale;
    *out_seed = r.last_seed();
    *out_paths = r.last_stats().paths;
    return 0;
  } catch (const std::exception& e) {
    std::snprintf(err, err_len, "%s", e.what());
    return 1;
  }
}

}  // extern "C"
