// gpu_path_renderer.hpp -- dort's Renderer plugin seam backed by the MI355X path tracer.
//
// GpuPathRenderer is a dort::Renderer (include/dort/renderer.hpp:12-34) that renders
// through the C-ABI of libdortgpu.so instead of PathRenderer::render
// (src/dort/path_renderer.cpp:11-21).  It takes the same options as dort's "pt"
// renderer (lua_render.cpp:125-145: iterations, min_depth, max_depth, only_direct,
// sample_all_lights, direct_strategy) plus rr_depth and the BVH / kernel knobs, and
// adds its samples to the shared Film the way Renderer::iteration_tiled_per_job
// does (Σ colour and Σ weight per pixel, renderer.cpp:78-84).
//
// dort's primitives, materials and textures are closures that cannot be inspected,
// so the scene is handed over a second time in typed form (GpuSceneDesc), recorded
// by the code that builds the dort::Scene (the Lua builder, lua_builder.cpp).
//
// Errors: the Renderer interface has no error channel; like the rest of dort, this
// class throws std::runtime_error, which dort's main turns into exit status 2
// (main.cpp:129-134).  Progress::is_cancelled() is polled between iteration
// chunks; a cancelled render leaves the iterations done so far in the film.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "dort/path_renderer.hpp"
#include "dort/renderer.hpp"
#include "dortgpu.h"

namespace dort_amd {

// The scene in the form the GPU path takes: triangles, per-triangle materials,
// triangle area lights, an optional environment map / infinite light, a pinhole
// camera.  Owns every array that desc() points into.
class GpuSceneDesc {
 public:
  uint32_t add_material(const dgpu_material& m);
  // Texture texels: width * height RGB fp32, row-major; returns its index.  mapping:
  // DGPU_TEXMAP_UV (uv_texture_map) or DGPU_TEXMAP_XY with world_to_texture (row-major
  // 3x4, the inverse of xy_texture_map's tex_to_world).
  uint32_t add_texture(uint32_t width, uint32_t height, std::vector<float> rgb,
                       uint32_t mapping = DGPU_TEXMAP_UV, const float* world_to_texture = nullptr);
  // The same with dort's 8-bit texels (Image<PixelRgb8>, k reads as k / 256).
  uint32_t add_texture_rgb8(uint32_t width, uint32_t height, std::vector<uint8_t> rgb,
                            uint32_t mapping = DGPU_TEXMAP_UV, const float* world_to_texture = nullptr);
  // A mesh: points (3 floats each), triangles (3 vertex indices each, into points),
  // optional uvs (2 floats per point), one material.  Returns the index of its
  // first triangle in the scene.
  uint32_t add_mesh(const std::vector<float>& points, const std::vector<uint32_t>& triangles,
                    uint32_t material, const std::vector<float>* uvs = nullptr);
  // The same with one material per triangle.
  uint32_t add_mesh(const std::vector<float>& points, const std::vector<uint32_t>& triangles,
                    const std::vector<uint32_t>& tri_materials, const std::vector<float>* uvs = nullptr);
  void add_area_light(uint32_t triangle, const float radiance[3]);
  void set_environment(uint32_t width, uint32_t height, std::vector<float> rgb, const float up[3],
                       const float fwd[3], const float scale[3]);
  void set_infinite_light(const float radiance[3]);
  void set_camera(const float camera_to_world[12], float fov);
  uint32_t num_triangles() const { return (uint32_t)tri_material_.size(); }

  // Valid until the next add_* / set_* call.
  const dgpu_scene_desc& desc();

 private:
  std::vector<float> positions_, uvs_;
  std::vector<uint32_t> indices_, tri_material_;
  bool any_uv_ = false;
  std::vector<dgpu_material> materials_;
  std::vector<dgpu_area_light> lights_;
  std::vector<std::vector<float>> texels_;
  std::vector<std::vector<uint8_t>> texels8_;
  std::vector<dgpu_texture> textures_;
  std::vector<float> env_rgb_;
  dgpu_env_light env_{};
  bool has_env_ = false;
  float inf_[3] = {0, 0, 0};
  bool has_inf_ = false;
  dgpu_camera camera_{};
  dgpu_scene_desc desc_{};
};

struct GpuPathOptions {
  uint32_t iterations = 1;  // samples per pixel; iteration 0 samples pixel centres
  uint32_t min_depth = 0;
  uint32_t max_depth = 5;
  uint32_t rr_depth = 2;    // dort's rule: RR from bounce 2 (path_renderer.cpp:101)
  bool only_direct = false;
  bool sample_all_lights = false;
  dort::PathRenderer::DirectStrategy direct_strategy = dort::PathRenderer::DirectStrategy::MIS;
  dgpu_bvh_opts bvh = {};   // zero fields take the library defaults
  int device = 0;
  uint32_t iterations_per_chunk = 0;  // progress / cancellation granularity (0: about 20 chunks)
};

class GpuPathRenderer final : public dort::Renderer {
 public:
  GpuPathRenderer(std::shared_ptr<dort::Scene> scene, std::shared_ptr<dort::Film> film,
                  std::shared_ptr<dort::Sampler> sampler, std::shared_ptr<dort::Camera> camera,
                  std::shared_ptr<GpuSceneDesc> gpu_scene, GpuPathOptions opts);
  ~GpuPathRenderer() override;
  GpuPathRenderer(const GpuPathRenderer&) = delete;
  GpuPathRenderer& operator=(const GpuPathRenderer&) = delete;

  void render(dort::CtxG& ctx, dort::Progress& progress) override;

  // Statistics and seed of the last render().
  const dgpu_stats& last_stats() const { return stats_; }
  uint64_t last_seed() const { return seed_; }

 private:
  void ensure_scene();
  [[noreturn]] void fail(const char* what, int status) const;

  std::shared_ptr<GpuSceneDesc> gpu_scene_;
  GpuPathOptions opts_;
  dgpu_ctx* ctx_ = nullptr;
  dgpu_scene* dev_scene_ = nullptr;
  dgpu_stats stats_{};
  uint64_t seed_ = 0;
};

// dort's "lt" renderer (LightRenderer, include/dort/light_renderer.hpp:7-33) backed by
// dgpu_render_light.  Options as lua_render.cpp:146-150 takes them (iterations,
// min_depth, max_depth; paths of min_depth + 2 .. max_depth + 2 vertices).  The splats
// are added to the film's splat sums and the film's splat_scale is set to
// 1 / iterations, as LightRenderer::render does (light_renderer.cpp:38-44).
struct GpuLightOptions {
  uint32_t iterations = 1;
  uint32_t min_depth = 0;
  uint32_t max_depth = 5;
  dgpu_bvh_opts bvh = {};
  int device = 0;
  uint32_t iterations_per_chunk = 0;  // progress / cancellation granularity (0: about 20 chunks)
};

class GpuLightRenderer final : public dort::Renderer {
 public:
  GpuLightRenderer(std::shared_ptr<dort::Scene> scene, std::shared_ptr<dort::Film> film,
                   std::shared_ptr<dort::Sampler> sampler, std::shared_ptr<dort::Camera> camera,
                   std::shared_ptr<GpuSceneDesc> gpu_scene, GpuLightOptions opts);
  ~GpuLightRenderer() override;
  GpuLightRenderer(const GpuLightRenderer&) = delete;
  GpuLightRenderer& operator=(const GpuLightRenderer&) = delete;

  void render(dort::CtxG& ctx, dort::Progress& progress) override;

  const dgpu_stats& last_stats() const { return stats_; }
  uint64_t last_seed() const { return seed_; }

 private:
  [[noreturn]] void fail(const char* what, int status) const;

  std::shared_ptr<GpuSceneDesc> gpu_scene_;
  GpuLightOptions opts_;
  dgpu_ctx* ctx_ = nullptr;
  dgpu_scene* dev_scene_ = nullptr;
  dgpu_stats stats_{};
  uint64_t seed_ = 0;
};

// dort's "bdpt" renderer (BdptRenderer, include/dort/bdpt_renderer.hpp:14-40) backed by
// dgpu_render_bdpt.  Options as lua_render.cpp:151-160 takes them (iterations, min_depth,
// max_depth, use_t1_paths; debug_image_dir is not supported).  Camera-subpath samples go
// to the film's colour sums and weights, light-subpath samples that reach the camera to
// its splat sums; the film's splat_scale is 1 / iterations, as BdptRenderer::render
// leaves it (bdpt_renderer.cpp:22-33).
struct GpuBdptOptions {
  uint32_t iterations = 1;
  uint32_t min_depth = 0;
  uint32_t max_depth = 5;
  bool use_t1_paths = true;
  dgpu_bvh_opts bvh = {};
  int device = 0;
  uint32_t iterations_per_chunk = 0;  // progress / cancellation granularity (0: about 20 chunks)
};

class GpuBdptRenderer final : public dort::Renderer {
 public:
  GpuBdptRenderer(std::shared_ptr<dort::Scene> scene, std::shared_ptr<dort::Film> film,
                  std::shared_ptr<dort::Sampler> sampler, std::shared_ptr<dort::Camera> camera,
                  std::shared_ptr<GpuSceneDesc> gpu_scene, GpuBdptOptions opts);
  ~GpuBdptRenderer() override;
  GpuBdptRenderer(const GpuBdptRenderer&) = delete;
  GpuBdptRenderer& operator=(const GpuBdptRenderer&) = delete;

  void render(dort::CtxG& ctx, dort::Progress& progress) override;

  const dgpu_stats& last_stats() const { return stats_; }
  uint64_t last_seed() const { return seed_; }

 private:
  [[noreturn]] void fail(const char* what, int status) const;

  std::shared_ptr<GpuSceneDesc> gpu_scene_;
  GpuBdptOptions opts_;
  dgpu_ctx* ctx_ = nullptr;
  dgpu_scene* dev_scene_ = nullptr;
  dgpu_stats stats_{};
  uint64_t seed_ = 0;
};

}  // namespace dort_amd
