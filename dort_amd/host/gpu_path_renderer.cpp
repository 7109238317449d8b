// gpu_path_renderer.cpp -- GpuPathRenderer: dort's Renderer seam over libdortgpu.so.
#include "gpu_path_renderer.hpp"

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

#include "dort/film.hpp"
#include "dort/sampler.hpp"
#include "dort/spectrum.hpp"
#include "dort/vec_2.hpp"

namespace dort_amd {

// ------------------------------------------------------------------ GpuSceneDesc

uint32_t GpuSceneDesc::add_material(const dgpu_material& m) {
  materials_.push_back(m);
  return (uint32_t)materials_.size() - 1;
}

namespace {
dgpu_texture texture_record(uint32_t width, uint32_t height, uint32_t mapping, const float* world_to_texture) {
  if (mapping > DGPU_TEXMAP_XY) throw std::invalid_argument("unknown texture mapping");
  if (mapping == DGPU_TEXMAP_XY && !world_to_texture) throw std::invalid_argument("xy mapping without a transform");
  dgpu_texture t{};
  t.width = width;
  t.height = height;
  t.mapping = mapping;
  if (world_to_texture) std::copy(world_to_texture, world_to_texture + 12, t.world_to_texture);
  return t;
}
}  // namespace

uint32_t GpuSceneDesc::add_texture(uint32_t width, uint32_t height, std::vector<float> rgb, uint32_t mapping,
                                   const float* world_to_texture) {
  if (rgb.size() != size_t(width) * height * 3) throw std::invalid_argument("texture size mismatch");
  textures_.push_back(texture_record(width, height, mapping, world_to_texture));
  texels_.push_back(std::move(rgb));
  texels8_.emplace_back();
  return (uint32_t)textures_.size() - 1;
}

uint32_t GpuSceneDesc::add_texture_rgb8(uint32_t width, uint32_t height, std::vector<uint8_t> rgb,
                                        uint32_t mapping, const float* world_to_texture) {
  if (rgb.size() != size_t(width) * height * 3) throw std::invalid_argument("texture size mismatch");
  textures_.push_back(texture_record(width, height, mapping, world_to_texture));
  texels_.emplace_back();
  texels8_.push_back(std::move(rgb));
  return (uint32_t)textures_.size() - 1;
}

uint32_t GpuSceneDesc::add_mesh(const std::vector<float>& points, const std::vector<uint32_t>& triangles,
                                uint32_t material, const std::vector<float>* uvs) {
  return add_mesh(points, triangles, std::vector<uint32_t>(triangles.size() / 3, material), uvs);
}

uint32_t GpuSceneDesc::add_mesh(const std::vector<float>& points, const std::vector<uint32_t>& triangles,
                                const std::vector<uint32_t>& tri_materials, const std::vector<float>* uvs) {
  if (points.size() % 3 || triangles.size() % 3) throw std::invalid_argument("mesh arrays not multiples of 3");
  if (tri_materials.size() != triangles.size() / 3) throw std::invalid_argument("one material per triangle");
  uint32_t nv = (uint32_t)(points.size() / 3), base = (uint32_t)(positions_.size() / 3);
  for (uint32_t v : triangles)
    if (v >= nv) throw std::invalid_argument("mesh triangle references a missing point");
  if (uvs && uvs->size() != size_t(nv) * 2) throw std::invalid_argument("mesh uvs size mismatch");
  uint32_t first = (uint32_t)tri_material_.size();
  positions_.insert(positions_.end(), points.begin(), points.end());
  for (uint32_t v : triangles) indices_.push_back(base + v);
  tri_material_.insert(tri_material_.end(), tri_materials.begin(), tri_materials.end());
  if (uvs) {
    any_uv_ = true;
    uvs_.insert(uvs_.end(), uvs->begin(), uvs->end());
  } else {
    uvs_.insert(uvs_.end(), size_t(nv) * 2, 0.f);
  }
  return first;
}

void GpuSceneDesc::add_area_light(uint32_t triangle, const float radiance[3]) {
  lights_.push_back(dgpu_area_light{triangle, {radiance[0], radiance[1], radiance[2]}});
}

void GpuSceneDesc::set_environment(uint32_t width, uint32_t height, std::vector<float> rgb,
                                   const float up[3], const float fwd[3], const float scale[3]) {
  if (rgb.size() != size_t(width) * height * 3) throw std::invalid_argument("environment size mismatch");
  env_rgb_ = std::move(rgb);
  env_.width = width;
  env_.height = height;
  for (int i = 0; i < 3; ++i) { env_.up[i] = up[i]; env_.fwd[i] = fwd[i]; env_.scale[i] = scale[i]; }
  has_env_ = true;
}

void GpuSceneDesc::set_infinite_light(const float radiance[3]) {
  for (int i = 0; i < 3; ++i) inf_[i] = radiance[i];
  has_inf_ = true;
}

void GpuSceneDesc::set_camera(const float camera_to_world[12], float fov) {
  std::copy(camera_to_world, camera_to_world + 12, camera_.camera_to_world);
  camera_.fov = fov;
}

const dgpu_scene_desc& GpuSceneDesc::desc() {
  for (size_t i = 0; i < textures_.size(); ++i) {
    textures_[i].rgb = texels8_[i].empty() ? texels_[i].data() : nullptr;
    textures_[i].rgb8 = texels8_[i].empty() ? nullptr : texels8_[i].data();
  }
  env_.rgb = env_rgb_.data();
  desc_.num_vertices = (uint32_t)(positions_.size() / 3);
  desc_.num_triangles = (uint32_t)tri_material_.size();
  desc_.positions = positions_.data();
  desc_.indices = indices_.data();
  // Without uvs dort uses (0,0), (1,0), (0,1) per triangle (triangle.cpp:89-93); with
  // per-vertex uvs on some meshes, the others get (0,0) (they only matter for textures).
  desc_.uvs = any_uv_ ? uvs_.data() : nullptr;
  desc_.tri_material = tri_material_.data();
  desc_.num_materials = (uint32_t)materials_.size();
  desc_.materials = materials_.data();
  desc_.num_area_lights = (uint32_t)lights_.size();
  desc_.area_lights = lights_.data();
  desc_.env = has_env_ ? &env_ : nullptr;
  desc_.infinite_radiance = has_inf_ ? inf_ : nullptr;
  desc_.num_textures = (uint32_t)textures_.size();
  desc_.textures = textures_.data();
  desc_.camera = camera_;
  return desc_;
}

// ------------------------------------------------------------------ GpuPathRenderer

GpuPathRenderer::GpuPathRenderer(std::shared_ptr<dort::Scene> scene, std::shared_ptr<dort::Film> film,
                                 std::shared_ptr<dort::Sampler> sampler,
                                 std::shared_ptr<dort::Camera> camera,
                                 std::shared_ptr<GpuSceneDesc> gpu_scene, GpuPathOptions opts)
    : dort::Renderer(std::move(scene), std::move(film), std::move(sampler), std::move(camera)),
      gpu_scene_(std::move(gpu_scene)),
      opts_(opts) {
  // the structs of dortgpu.h this file was compiled with must be the library's
  if (dgpu_abi_version() != DGPU_ABI_VERSION)
    throw std::runtime_error("GpuPathRenderer: libdortgpu ABI version " + std::to_string(dgpu_abi_version()) +
                             ", compiled against " + std::to_string(DGPU_ABI_VERSION));
  if (!gpu_scene_) throw std::invalid_argument("GpuPathRenderer: no GPU scene description");
  if (opts_.iterations == 0) throw std::invalid_argument("GpuPathRenderer: iterations must be positive");
  // The GPU film is dort's default 1-pixel box filter (lua_render.cpp:108-109):
  // every sample lands in its own pixel with weight 1.
  const dort::SampledFilter& f = this->film->filter;
  if (f.radius.x != 0.5f || f.radius.y != 0.5f || f.evaluate(dort::Vec2(0.f, 0.f)) != 1.f ||
      f.evaluate(dort::Vec2(0.49f, 0.49f)) != 1.f)
    throw std::invalid_argument("GpuPathRenderer: only the 1-pixel box filter is supported");
  int st = dgpu_create(opts_.device, &ctx_);
  if (st != DGPU_OK) fail("dgpu_create", st);
}

GpuPathRenderer::~GpuPathRenderer() {
  if (dev_scene_) dgpu_scene_destroy(dev_scene_);
  if (ctx_) dgpu_destroy(ctx_);
}

void GpuPathRenderer::fail(const char* what, int status) const {
  throw std::runtime_error(std::string("GpuPathRenderer: ") + what + " failed (status " +
                           std::to_string(status) + "): " + dgpu_last_error());
}

void GpuPathRenderer::ensure_scene() {
  if (dev_scene_) return;
  const dgpu_scene_desc& d = gpu_scene_->desc();
  int st = dgpu_scene_create(ctx_, &d, &opts_.bvh, &dev_scene_);
  if (st != DGPU_OK) fail("dgpu_scene_create", st);
}

void GpuPathRenderer::render(dort::CtxG&, dort::Progress& progress) {
  ensure_scene();
  // The path seed comes from the renderer's sampler, as dort's iteration seeds do
  // (renderer.cpp:55-57), so a fixed sampler seed gives a fixed image.
  {
    std::unique_lock<std::mutex> lock(this->sampler_mutex);
    uint64_t hi = this->sampler->rng.uniform_uint32(0xFFFFFFFFu);
    uint64_t lo = this->sampler->rng.uniform_uint32(0xFFFFFFFFu);
    seed_ = (hi << 32) | lo;
  }
  const uint32_t xr = (uint32_t)this->film->res.x, yr = (uint32_t)this->film->res.y;
  const size_t npix = size_t(xr) * yr;
  std::vector<float> rgb(npix * 3, 0.f), w(npix, 0.f);

  dgpu_render_opts o{};
  o.x_res = xr;
  o.y_res = yr;
  o.seed = seed_;
  o.min_depth = opts_.min_depth;
  o.max_depth = opts_.max_depth;
  o.rr_depth = opts_.rr_depth;
  o.only_direct = opts_.only_direct ? 1u : 0u;
  o.sample_all_lights = opts_.sample_all_lights ? 1u : 0u;
  switch (opts_.direct_strategy) {
    case dort::PathRenderer::DirectStrategy::SAMPLE_BSDF: o.direct_strategy = DGPU_DIRECT_BSDF; break;
    case dort::PathRenderer::DirectStrategy::SAMPLE_LIGHT: o.direct_strategy = DGPU_DIRECT_LIGHT; break;
    default: o.direct_strategy = DGPU_DIRECT_MIS; break;
  }
  const uint32_t chunk = opts_.iterations_per_chunk
                             ? opts_.iterations_per_chunk
                             : std::max(1u, (opts_.iterations + 19u) / 20u);
  stats_ = dgpu_stats{};
  for (uint32_t begin = 0; begin < opts_.iterations; begin += chunk) {
    if (progress.is_cancelled()) break;
    o.sample_begin = begin;
    o.sample_end = std::min(opts_.iterations, begin + chunk);
    dgpu_stats st{};
    int rc = dgpu_render(dev_scene_, &o, rgb.data(), w.data(), &st, nullptr);
    if (rc != DGPU_OK) fail("dgpu_render", rc);
    stats_.paths += st.paths;
    stats_.extension_rays += st.extension_rays;
    stats_.shadow_rays += st.shadow_rays;
    stats_.dropped_samples += st.dropped_samples;
    stats_.kernel_ms += st.kernel_ms;
    progress.set_percent_done(100.f * float(o.sample_end) / float(opts_.iterations));
  }
  // Film::add_tile (film.cpp:69-88) for the whole film at once.
  std::unique_lock<std::mutex> film_lock(this->film_mutex);
  for (size_t i = 0; i < npix; ++i) {
    dort::Film::Pixel& px = this->film->pixels[i];
    px.color = px.color + dort::Spectrum(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    px.weight += w[i];
  }
}

GpuLightRenderer::GpuLightRenderer(std::shared_ptr<dort::Scene> scene, std::shared_ptr<dort::Film> film,
                                   std::shared_ptr<dort::Sampler> sampler, std::shared_ptr<dort::Camera> camera,
                                   std::shared_ptr<GpuSceneDesc> gpu_scene, GpuLightOptions opts)
    : dort::Renderer(std::move(scene), std::move(film), std::move(sampler), std::move(camera)),
      gpu_scene_(std::move(gpu_scene)),
      opts_(opts) {
  if (!gpu_scene_) throw std::invalid_argument("GpuLightRenderer: no GPU scene description");
  if (dgpu_abi_version() != DGPU_ABI_VERSION)
    throw std::runtime_error("GpuLightRenderer: libdortgpu.so has another ABI version than dortgpu.h");
  // the splats are spread with the 1-pixel box filter only
  const dort::SampledFilter& f = this->film->filter;
  if (f.radius.x != 0.5f || f.radius.y != 0.5f || f.evaluate(dort::Vec2(0.f, 0.f)) != 1.f ||
      f.evaluate(dort::Vec2(0.49f, 0.49f)) != 1.f)
    throw std::invalid_argument("GpuLightRenderer: only the 1-pixel box filter is supported");
  int st = dgpu_create(opts_.device, &ctx_);
  if (st != DGPU_OK) fail("dgpu_create", st);
}

GpuLightRenderer::~GpuLightRenderer() {
  if (dev_scene_) dgpu_scene_destroy(dev_scene_);
  if (ctx_) dgpu_destroy(ctx_);
}

void GpuLightRenderer::fail(const char* what, int status) const {
  throw std::runtime_error(std::string("GpuLightRenderer: ") + what + " failed (status " +
                           std::to_string(status) + "): " + dgpu_last_error());
}

void GpuLightRenderer::render(dort::CtxG&, dort::Progress& progress) {
  if (!dev_scene_) {
    const dgpu_scene_desc& d = gpu_scene_->desc();
    int st = dgpu_scene_create(ctx_, &d, &opts_.bvh, &dev_scene_);
    if (st != DGPU_OK) fail("dgpu_scene_create", st);
  }
  {
    std::unique_lock<std::mutex> lock(this->sampler_mutex);
    uint64_t hi = this->sampler->rng.uniform_uint32(0xFFFFFFFFu);
    uint64_t lo = this->sampler->rng.uniform_uint32(0xFFFFFFFFu);
    seed_ = (hi << 32) | lo;
  }
  const uint32_t xr = (uint32_t)this->film->res.x, yr = (uint32_t)this->film->res.y;
  const size_t npix = size_t(xr) * yr;
  std::vector<float> rgb(npix * 3, 0.f);
  dgpu_render_opts o{};
  o.x_res = xr;
  o.y_res = yr;
  o.seed = seed_;
  o.min_depth = opts_.min_depth;
  o.max_depth = opts_.max_depth;
  const uint32_t chunk = opts_.iterations_per_chunk
                             ? opts_.iterations_per_chunk
                             : std::max(1u, (opts_.iterations + 19u) / 20u);
  stats_ = dgpu_stats{};
  for (uint32_t begin = 0; begin < opts_.iterations; begin += chunk) {
    if (progress.is_cancelled()) break;
    o.sample_begin = begin;
    o.sample_end = std::min(opts_.iterations, begin + chunk);
    dgpu_stats st{};
    int rc = dgpu_render_light(dev_scene_, &o, rgb.data(), &st, nullptr);
    if (rc != DGPU_OK) fail("dgpu_render_light", rc);
    stats_.paths += st.paths;
    stats_.extension_rays += st.extension_rays;
    stats_.shadow_rays += st.shadow_rays;
    stats_.dropped_samples += st.dropped_samples;
    stats_.kernel_ms += st.kernel_ms;
    progress.set_percent_done(100.f * float(o.sample_end) / float(opts_.iterations));
  }
  std::unique_lock<std::mutex> film_lock(this->film_mutex);
  for (size_t i = 0; i < npix; ++i)
    this->film->pixels[i].splat.add_relaxed(dort::Spectrum(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]));
  this->film->splat_scale = 1.f / float(std::max(1u, opts_.iterations));
}

GpuBdptRenderer::GpuBdptRenderer(std::shared_ptr<dort::Scene> scene, std::shared_ptr<dort::Film> film,
                                 std::shared_ptr<dort::Sampler> sampler, std::shared_ptr<dort::Camera> camera,
                                 std::shared_ptr<GpuSceneDesc> gpu_scene, GpuBdptOptions opts)
    : dort::Renderer(std::move(scene), std::move(film), std::move(sampler), std::move(camera)),
      gpu_scene_(std::move(gpu_scene)),
      opts_(opts) {
  if (!gpu_scene_) throw std::invalid_argument("GpuBdptRenderer: no GPU scene description");
  if (dgpu_abi_version() != DGPU_ABI_VERSION)
    throw std::runtime_error("GpuBdptRenderer: libdortgpu.so has another ABI version than dortgpu.h");
  if (opts_.iterations == 0) throw std::invalid_argument("GpuBdptRenderer: iterations must be positive");
  // samples and splats land with the 1-pixel box filter only
  const dort::SampledFilter& f = this->film->filter;
  if (f.radius.x != 0.5f || f.radius.y != 0.5f || f.evaluate(dort::Vec2(0.f, 0.f)) != 1.f ||
      f.evaluate(dort::Vec2(0.49f, 0.49f)) != 1.f)
    throw std::invalid_argument("GpuBdptRenderer: only the 1-pixel box filter is supported");
  int st = dgpu_create(opts_.device, &ctx_);
  if (st != DGPU_OK) fail("dgpu_create", st);
}

GpuBdptRenderer::~GpuBdptRenderer() {
  if (dev_scene_) dgpu_scene_destroy(dev_scene_);
  if (ctx_) dgpu_destroy(ctx_);
}

void GpuBdptRenderer::fail(const char* what, int status) const {
  throw std::runtime_error(std::string("GpuBdptRenderer: ") + what + " failed (status " +
                           std::to_string(status) + "): " + dgpu_last_error());
}

void GpuBdptRenderer::render(dort::CtxG&, dort::Progress& progress) {
  if (!dev_scene_) {
    const dgpu_scene_desc& d = gpu_scene_->desc();
    int st = dgpu_scene_create(ctx_, &d, &opts_.bvh, &dev_scene_);
    if (st != DGPU_OK) fail("dgpu_scene_create", st);
  }
  {
    std::unique_lock<std::mutex> lock(this->sampler_mutex);
    uint64_t hi = this->sampler->rng.uniform_uint32(0xFFFFFFFFu);
    uint64_t lo = this->sampler->rng.uniform_uint32(0xFFFFFFFFu);
    seed_ = (hi << 32) | lo;
  }
  const uint32_t xr = (uint32_t)this->film->res.x, yr = (uint32_t)this->film->res.y;
  const size_t npix = size_t(xr) * yr;
  std::vector<float> rgb(npix * 3, 0.f), w(npix, 0.f), splat(npix * 3, 0.f);
  dgpu_render_opts o{};
  o.x_res = xr;
  o.y_res = yr;
  o.seed = seed_;
  o.min_depth = opts_.min_depth;
  o.max_depth = opts_.max_depth;
  const uint32_t chunk = opts_.iterations_per_chunk
                             ? opts_.iterations_per_chunk
                             : std::max(1u, (opts_.iterations + 19u) / 20u);
  stats_ = dgpu_stats{};
  for (uint32_t begin = 0; begin < opts_.iterations; begin += chunk) {
    if (progress.is_cancelled()) break;
    o.sample_begin = begin;
    o.sample_end = std::min(opts_.iterations, begin + chunk);
    dgpu_stats st{};
    int rc = dgpu_render_bdpt(dev_scene_, &o, opts_.use_t1_paths ? 1u : 0u, rgb.data(), w.data(), splat.data(),
                              &st, nullptr);
    if (rc != DGPU_OK) fail("dgpu_render_bdpt", rc);
    stats_.paths += st.paths;
    stats_.extension_rays += st.extension_rays;
    stats_.shadow_rays += st.shadow_rays;
    stats_.dropped_samples += st.dropped_samples;
    stats_.kernel_ms += st.kernel_ms;
    progress.set_percent_done(100.f * float(o.sample_end) / float(opts_.iterations));
  }
  std::unique_lock<std::mutex> film_lock(this->film_mutex);
  for (size_t i = 0; i < npix; ++i) {
    dort::Film::Pixel& px = this->film->pixels[i];
    px.color = px.color + dort::Spectrum(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    px.weight += w[i];
    px.splat.add_relaxed(dort::Spectrum(splat[3 * i], splat[3 * i + 1], splat[3 * i + 2]));
  }
  this->film->splat_scale = 1.f / float(std::max(1u, opts_.iterations));
}

}  // namespace dort_amd
