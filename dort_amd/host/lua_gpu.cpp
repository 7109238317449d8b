// lua_gpu.cpp -- the Lua module "dort_gpu" (declared in lua_gpu.hpp).
#include "lua_gpu.hpp"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "dort/camera.hpp"
#include "dort/film.hpp"
#include "dort/filter.hpp"
#include "dort/image.hpp"
#include "dort/lua.hpp"
#include "dort/lua_builder.hpp"
#include "dort/lua_geometry.hpp"
#include "dort/lua_helpers.hpp"
#include "dort/lua_image.hpp"
#include "dort/lua_params.hpp"
#include "dort/lua_render.hpp"
#include "dort/lua_shape.hpp"
#include "dort/lua_spectrum.hpp"
#include "dort/mesh.hpp"
#include "dort/ply_mesh.hpp"
#include "dort/random_sampler.hpp"
#include "dort/rng.hpp"
#include "dort/scene.hpp"
#include "dort/transform.hpp"
#include "gpu_path_renderer.hpp"

namespace dort_amd {
namespace {

constexpr const char RECORDER_TNAME[] = "dort_gpu.Recorder";

// The scene as the script builds it: a triangle soup (three points per triangle, in
// the order dort's mesh lists them), each triangle with its uvs (dort's per-triangle
// default (0,0), (1,0), (0,1) when its mesh has none, triangle.cpp:89-93) and material.
struct Recorder {
  std::vector<float> points;   // 9 per triangle
  std::vector<float> uvs;      // 6 per triangle
  std::vector<uint32_t> tri_material;
  bool any_uv = false;
  std::vector<dgpu_material> materials;
  std::vector<dgpu_area_light> lights;
  bool has_infinite = false;
  float infinite[3] = {0, 0, 0};
  bool has_camera = false;
  dgpu_camera camera{};

  uint32_t num_triangles() const { return (uint32_t)tri_material.size(); }

  void add(const dort::Point p[3], const dort::Vec2* uv, uint32_t material) {
    for (int k = 0; k < 3; ++k) {
      points.push_back(p[k].v.x);
      points.push_back(p[k].v.y);
      points.push_back(p[k].v.z);
    }
    static const float dflt[6] = {0.f, 0.f, 1.f, 0.f, 0.f, 1.f};
    for (int k = 0; k < 3; ++k) {
      uvs.push_back(uv ? uv[k].x : dflt[2 * k]);
      uvs.push_back(uv ? uv[k].y : dflt[2 * k + 1]);
    }
    any_uv = any_uv || uv != nullptr;
    tri_material.push_back(material);
  }

  std::shared_ptr<GpuSceneDesc> desc() const {
    auto g = std::make_shared<GpuSceneDesc>();
    for (const dgpu_material& m : materials) g->add_material(m);
    std::vector<uint32_t> tris(points.size() / 3);
    for (size_t i = 0; i < tris.size(); ++i) tris[i] = (uint32_t)i;
    // without any mesh uvs the library's per-triangle default is dort's
    if (!tri_material.empty()) g->add_mesh(points, tris, tri_material, any_uv ? &uvs : nullptr);
    for (const dgpu_area_light& a : lights) g->add_area_light(a.triangle, a.radiance);
    if (has_infinite) g->set_infinite_light(infinite);
    g->set_camera(camera.camera_to_world, camera.fov);
    return g;
  }
};

Recorder& check_recorder(lua_State* l, int idx) {
  return dort::lua_check_managed_obj<Recorder, RECORDER_TNAME>(l, idx);
}

// Row-major 3x4 of a dort Transform (Mat4x4::cols is column-major, mat.cpp:58-63).
void matrix34(const dort::Transform& t, float out[12]) {
  const dort::Mat4x4& m = t.get_mat(false);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) out[4 * r + c] = m.cols[c][r];
}

float field_float(lua_State* l, int t, const char* name, float def) {
  lua_getfield(l, t, name);
  float v = lua_isnil(l, -1) ? def : (float)luaL_checknumber(l, -1);
  lua_pop(l, 1);
  return v;
}

void field_rgb(lua_State* l, int t, const char* name, float out[3], float def) {
  lua_getfield(l, t, name);
  if (lua_isnil(l, -1)) {
    out[0] = out[1] = out[2] = def;
  } else {
    luaL_checktype(l, -1, LUA_TTABLE);
    for (int i = 0; i < 3; ++i) {
      lua_rawgeti(l, -1, i + 1);
      out[i] = (float)luaL_checknumber(l, -1);
      lua_pop(l, 1);
    }
  }
  lua_pop(l, 1);
}

// ---------------------------------------------------------------- recorder methods

int rec_new(lua_State* l) {
  dort::lua_push_managed_obj<Recorder, RECORDER_TNAME>(l, Recorder());
  return 1;
}

// add_material{type = "none"|"lambert"|"dielectric"|"rough_dielectric", albedo = {r,g,b},
//   ior_inside, ior_outside, reflect_tint, transmit_tint, roughness} -> index
int rec_add_material(lua_State* l) {
  Recorder& r = check_recorder(l, 1);
  luaL_checktype(l, 2, LUA_TTABLE);
  lua_getfield(l, 2, "type");
  std::string type = luaL_checkstring(l, -1);
  lua_pop(l, 1);
  dgpu_material m{};
  m.albedo_tex = -1;
  if (type == "none") {
    m.type = DGPU_MAT_NONE;
  } else if (type == "lambert") {
    m.type = DGPU_MAT_LAMBERT;
    field_rgb(l, 2, "albedo", m.albedo, 1.f);
  } else if (type == "dielectric" || type == "rough_dielectric") {
    m.type = type == "dielectric" ? DGPU_MAT_DIELECTRIC : DGPU_MAT_ROUGH_DIELECTRIC;
    // dort's defaults (lua_material.cpp:113-118, 141-148)
    m.ior_inside = field_float(l, 2, "ior_inside", 1.5f);
    m.ior_outside = field_float(l, 2, "ior_outside", 1.f);
    field_rgb(l, 2, "reflect_tint", m.reflect_tint, 1.f);
    field_rgb(l, 2, "transmit_tint", m.transmit_tint, 1.f);
    m.roughness = field_float(l, 2, "roughness", 0.1f);
  } else {
    return luaL_error(l, "dort_gpu: unknown material type '%s'", type.c_str());
  }
  r.materials.push_back(m);
  lua_pushinteger(l, (lua_Integer)r.materials.size() - 1);
  return 1;
}

uint32_t check_material(lua_State* l, const Recorder& r, int idx) {
  lua_Integer m = luaL_checkinteger(l, idx);
  if (m < 0 || (size_t)m >= r.materials.size()) luaL_error(l, "dort_gpu: material %d out of range", (int)m);
  return (uint32_t)m;
}

// add_triangle(mesh, index, material [, transform]) -> triangle index.  `index` is
// the offset of the triangle's first vertex in mesh.vertices, as in
// dort.builder.add_triangle / make_triangle (lua_builder.cpp:375-390, 659-666); the
// transform (that of a shape primitive, lua_builder.cpp:327-346) is applied to the points.
int rec_add_triangle(lua_State* l) {
  Recorder& r = check_recorder(l, 1);
  std::shared_ptr<dort::Mesh> mesh = dort::lua_check_mesh(l, 2);
  lua_Integer index = luaL_checkinteger(l, 3);
  uint32_t mat = check_material(l, r, 4);
  const bool has_t = !lua_isnoneornil(l, 5);
  dort::Transform t = has_t ? dort::lua_check_transform(l, 5) : dort::Transform();
  if (index < 0 || (size_t)index + 3 > mesh->vertices.size()) {
    mesh.reset();
    return luaL_error(l, "dort_gpu: triangle index %d out of the mesh", (int)index);
  }
  dort::Point p[3];
  dort::Vec2 uv[3];
  for (int k = 0; k < 3; ++k) {
    uint32_t v = mesh->vertices[(size_t)index + k];
    if (v >= mesh->points.size()) {
      mesh.reset();
      return luaL_error(l, "dort_gpu: mesh vertex %d references a missing point", (int)v);
    }
    p[k] = has_t ? t.apply(mesh->points[v]) : mesh->points[v];
    if (!mesh->uvs.empty()) uv[k] = mesh->uvs.at(v);
  }
  r.add(p, mesh->uvs.empty() ? nullptr : uv, mat);
  lua_pushinteger(l, (lua_Integer)r.num_triangles() - 1);
  return 1;
}

// add_ply_mesh(ply_mesh, material [, transform]) -> first triangle index: every
// triangle of a PlyMesh, its points through transform.apply as
// lua_builder_add_ply_mesh does (lua_builder.cpp:532-560).
int rec_add_ply_mesh(lua_State* l) {
  Recorder& r = check_recorder(l, 1);
  std::shared_ptr<dort::PlyMesh> ply = dort::lua_check_ply_mesh(l, 2);
  uint32_t mat = check_material(l, r, 3);
  const bool has_t = !lua_isnoneornil(l, 4);
  dort::Transform t = has_t ? dort::lua_check_transform(l, 4) : dort::Transform();
  const uint32_t first = r.num_triangles();
  for (size_t i = 0; i + 2 < ply->vertices.size(); i += 3) {
    dort::Point p[3];
    for (int k = 0; k < 3; ++k) {
      const dort::Point& q = ply->points.at(ply->vertices[i + k]);
      p[k] = has_t ? t.apply(q) : q;
    }
    r.add(p, nullptr, mat);
  }
  lua_pushinteger(l, first);
  return 1;
}

// add_cube(transform, material) -> first of 12 triangles: dort's cube() shape (the box
// (-1,-1,-1)..(1,1,1), cube_shape.cpp:99-101) under a shape primitive's transform.
// Corner index 4*ix + 2*iy + iz; two triangles per face.
int rec_add_cube(lua_State* l) {
  Recorder& r = check_recorder(l, 1);
  const dort::Transform& t = dort::lua_check_transform(l, 2);
  uint32_t mat = check_material(l, r, 3);
  dort::Point c[8];
  for (int i = 0; i < 8; ++i)
    c[i] = t.apply(dort::Point(i & 4 ? 1.f : -1.f, i & 2 ? 1.f : -1.f, i & 1 ? 1.f : -1.f));
  static const int faces[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1},
                                  {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
  const uint32_t first = r.num_triangles();
  for (const auto& f : faces) {
    const dort::Point a[3] = {c[f[0]], c[f[1]], c[f[2]]};
    const dort::Point b[3] = {c[f[0]], c[f[2]], c[f[3]]};
    r.add(a, nullptr, mat);
    r.add(b, nullptr, mat);
  }
  lua_pushinteger(l, first);
  return 1;
}

void check_rgb(lua_State* l, int idx, float out[3]) {
  const dort::Spectrum& s = dort::lua_check_spectrum(l, idx);
  out[0] = s.red();
  out[1] = s.green();
  out[2] = s.blue();
}

// add_area_light(triangle, radiance): the triangle emits radiance on its front side.
int rec_add_area_light(lua_State* l) {
  Recorder& r = check_recorder(l, 1);
  lua_Integer tri = luaL_checkinteger(l, 2);
  if (tri < 0 || tri >= (lua_Integer)r.num_triangles()) return luaL_error(l, "dort_gpu: no triangle %d", (int)tri);
  dgpu_area_light a{};
  a.triangle = (uint32_t)tri;
  check_rgb(l, 3, a.radiance);
  r.lights.push_back(a);
  return 0;
}

int rec_set_infinite(lua_State* l) {
  Recorder& r = check_recorder(l, 1);
  check_rgb(l, 2, r.infinite);
  r.has_infinite = true;
  return 0;
}

// set_camera(transform [, fov]): a pinhole camera as make_pinhole takes it (fov default
// pi/2, clamped to [0.01, pi], lua_camera.cpp:57-65).
int rec_set_camera(lua_State* l) {
  Recorder& r = check_recorder(l, 1);
  const dort::Transform& t = dort::lua_check_transform(l, 2);
  float fov = lua_isnoneornil(l, 3) ? dort::PI * 0.5f : (float)luaL_checknumber(l, 3);
  matrix34(t, r.camera.camera_to_world);
  r.camera.fov = dort::clamp(fov, 1e-2f, dort::PI);
  r.has_camera = true;
  return 0;
}

int rec_num_triangles(lua_State* l) {
  lua_pushinteger(l, check_recorder(l, 1).num_triangles());
  return 1;
}

void json_floats(FILE* f, const float* v, size_t n) {
  std::fputc('[', f);
  for (size_t i = 0; i < n; ++i) std::fprintf(f, "%s%.9g", i ? "," : "", (double)v[i]);
  std::fputc(']', f);
}

// write_json(path): the recorded scene, for tests (floats to 9 significant digits,
// which round-trip float32 exactly).
int rec_write_json(lua_State* l) {
  const Recorder& r = check_recorder(l, 1);
  const char* path = luaL_checkstring(l, 2);
  FILE* f = std::fopen(path, "w");
  if (!f) return luaL_error(l, "dort_gpu: cannot write %s", path);
  std::fprintf(f, "{\"points\":");
  json_floats(f, r.points.data(), r.points.size());
  std::fprintf(f, ",\"uvs\":");
  if (r.any_uv) json_floats(f, r.uvs.data(), r.uvs.size()); else std::fprintf(f, "null");
  std::fprintf(f, ",\"tri_material\":[");
  for (size_t i = 0; i < r.tri_material.size(); ++i) std::fprintf(f, "%s%u", i ? "," : "", r.tri_material[i]);
  std::fprintf(f, "],\"materials\":[");
  for (size_t i = 0; i < r.materials.size(); ++i) {
    const dgpu_material& m = r.materials[i];
    std::fprintf(f, "%s{\"type\":%u,\"albedo\":", i ? "," : "", m.type);
    json_floats(f, m.albedo, 3);
    std::fprintf(f, ",\"roughness\":%.9g,\"ior_inside\":%.9g,\"ior_outside\":%.9g,\"reflect_tint\":",
                 (double)m.roughness, (double)m.ior_inside, (double)m.ior_outside);
    json_floats(f, m.reflect_tint, 3);
    std::fprintf(f, ",\"transmit_tint\":");
    json_floats(f, m.transmit_tint, 3);
    std::fputc('}', f);
  }
  std::fprintf(f, "],\"area_lights\":[");
  for (size_t i = 0; i < r.lights.size(); ++i) {
    std::fprintf(f, "%s{\"triangle\":%u,\"radiance\":", i ? "," : "", r.lights[i].triangle);
    json_floats(f, r.lights[i].radiance, 3);
    std::fputc('}', f);
  }
  std::fprintf(f, "],\"infinite\":");
  if (r.has_infinite) json_floats(f, r.infinite, 3); else std::fprintf(f, "null");
  std::fprintf(f, ",\"camera\":");
  if (r.has_camera) {
    std::fprintf(f, "{\"camera_to_world\":");
    json_floats(f, r.camera.camera_to_world, 12);
    std::fprintf(f, ",\"fov\":%.9g}", (double)r.camera.fov);
  } else {
    std::fprintf(f, "null");
  }
  std::fprintf(f, "}\n");
  const bool ok = std::fclose(f) == 0;
  if (!ok) return luaL_error(l, "dort_gpu: cannot write %s", path);
  return 0;
}

// ---------------------------------------------------------------- module functions

// rgb(x) -> r, g, b of a dort Spectrum or a number; nothing for anything else.
int lua_rgb(lua_State* l) {
  if (dort::lua_test_spectrum(l, 1)) {
    const dort::Spectrum& s = dort::lua_check_spectrum(l, 1);
    lua_pushnumber(l, s.red());
    lua_pushnumber(l, s.green());
    lua_pushnumber(l, s.blue());
    return 3;
  }
  if (lua_type(l, 1) == LUA_TNUMBER) {
    float v = (float)lua_tonumber(l, 1);
    for (int i = 0; i < 3; ++i) lua_pushnumber(l, v);
    return 3;
  }
  return 0;
}

dgpu_bvh_opts read_bvh_opts(lua_State* l, int p) {
  dgpu_bvh_opts b{};
  b.width = dort::lua_param_uint32_opt(l, p, "bvh_width", 0);
  b.max_leaf_size = dort::lua_param_uint32_opt(l, p, "bvh_max_leaf_size", 0);
  b.sah_bins = dort::lua_param_uint32_opt(l, p, "bvh_sah_bins", 0);
  b.stack_lds = dort::lua_param_uint32_opt(l, p, "stack_lds", 0);
  b.block_size = dort::lua_param_uint32_opt(l, p, "block_size", 0);
  b.queue_tile = dort::lua_param_uint32_opt(l, p, "queue_tile", 0);
  return b;
}

// make_job(scene, recorder, params [, camera_transform, camera_fov]) -> RenderJob.
// params as lua_render_make reads them (lua_render.cpp:104-150, 209): x_res (800),
// y_res (600), filter (1-pixel box), sampler (random, split with seed 42), renderer
// ("pt_gpu", "lt_gpu" or "bdpt_gpu"), iterations (1), camera (the scene's); "pt_gpu":
// min_depth (0), max_depth (5), only_direct, sample_all_lights, direct_strategy ("mis"),
// and rr_depth (2, dort's rule); "lt_gpu": min_depth, max_depth; "bdpt_gpu": min_depth,
// max_depth, use_t1_paths (true, lua_render.cpp:151-160).  All: device (0),
// iterations_per_chunk, bvh_width, bvh_max_leaf_size, bvh_sah_bins, stack_lds,
// block_size, queue_tile.  Unknown params are an error, as in dort.  The GPU camera is
// the recorded one, or camera_transform / camera_fov when given (the Lua side passes the
// pinhole behind params.camera).
int lua_make_job(lua_State* l) {
  std::shared_ptr<dort::Scene> scene = dort::lua_check_scene(l, 1);
  const Recorder& rec = check_recorder(l, 2);
  luaL_checktype(l, 3, LUA_TTABLE);
  const int p = 3;
  const bool cam_override = !lua_isnoneornil(l, 4);
  dgpu_camera cam = rec.camera;
  if (cam_override) {
    matrix34(dort::lua_check_transform(l, 4), cam.camera_to_world);
    float fov = lua_isnoneornil(l, 5) ? dort::PI * 0.5f : (float)luaL_checknumber(l, 5);
    cam.fov = dort::clamp(fov, 1e-2f, dort::PI);
  } else if (!rec.has_camera) {
    scene.reset();
    return luaL_error(l, "dort_gpu: no camera is set in the scene and no camera was given");
  }

  uint32_t x_res = dort::lua_param_uint32_opt(l, p, "x_res", 800);
  uint32_t y_res = dort::lua_param_uint32_opt(l, p, "y_res", 600);
  auto filter = dort::lua_param_filter_opt(l, p, "filter",
                                           std::make_shared<dort::BoxFilter>(dort::Vec2(0.5f, 0.5f)));
  auto sampler = dort::lua_param_sampler_opt(l, p, "sampler",
                                             std::make_shared<dort::RandomSampler>(1, dort::Rng(42)))->split(42);
  std::string method = dort::lua_param_string_opt(l, p, "renderer", "pt_gpu");
  uint32_t iterations = dort::lua_param_uint32_opt(l, p, "iterations", 1);
  auto camera = dort::lua_param_camera_opt(l, p, "camera", scene->default_camera);
  const int device = (int)dort::lua_param_uint32_opt(l, p, "device", 0);
  const uint32_t per_chunk = dort::lua_param_uint32_opt(l, p, "iterations_per_chunk", 0);

  GpuPathOptions po;
  GpuLightOptions lo;
  GpuBdptOptions bo;
  if (method == "pt_gpu") {
    po.iterations = iterations;
    po.min_depth = dort::lua_param_uint32_opt(l, p, "min_depth", 0);
    po.max_depth = dort::lua_param_uint32_opt(l, p, "max_depth", 5);
    po.rr_depth = dort::lua_param_uint32_opt(l, p, "rr_depth", 2);
    po.only_direct = dort::lua_param_bool_opt(l, p, "only_direct", false);
    po.sample_all_lights = dort::lua_param_bool_opt(l, p, "sample_all_lights", false);
    std::string strategy = dort::lua_param_string_opt(l, p, "direct_strategy", "mis");
    if (strategy == "mis") {
      po.direct_strategy = dort::PathRenderer::DirectStrategy::MIS;
    } else if (strategy == "bsdf") {
      po.direct_strategy = dort::PathRenderer::DirectStrategy::SAMPLE_BSDF;
    } else if (strategy == "light") {
      po.direct_strategy = dort::PathRenderer::DirectStrategy::SAMPLE_LIGHT;
    } else {
      lua_pushfstring(l, "Unknown direct strategy '%s'", strategy.c_str());
      strategy.clear();
      return lua_error(l);
    }
    po.bvh = read_bvh_opts(l, p);
    po.device = device;
    po.iterations_per_chunk = per_chunk;
  } else if (method == "bdpt_gpu") {
    // lua_render.cpp:151-160; debug_image_dir is not supported
    bo.iterations = iterations;
    bo.min_depth = dort::lua_param_uint32_opt(l, p, "min_depth", 0);
    bo.max_depth = dort::lua_param_uint32_opt(l, p, "max_depth", 5);
    bo.use_t1_paths = dort::lua_param_bool_opt(l, p, "use_t1_paths", true);
    bo.bvh = read_bvh_opts(l, p);
    bo.device = device;
    bo.iterations_per_chunk = per_chunk;
  } else if (method == "lt_gpu") {
    lo.iterations = iterations;
    lo.min_depth = dort::lua_param_uint32_opt(l, p, "min_depth", 0);
    lo.max_depth = dort::lua_param_uint32_opt(l, p, "max_depth", 5);
    lo.bvh = read_bvh_opts(l, p);
    lo.device = device;
    lo.iterations_per_chunk = per_chunk;
  } else {
    lua_pushfstring(l, "dort_gpu: unknown renderer '%s' (pt_gpu, lt_gpu, bdpt_gpu)", method.c_str());
    method.clear();
    return lua_error(l);
  }
  dort::lua_params_check_unused(l, p);
  if (!camera) {
    scene.reset();
    filter.reset();
    sampler.reset();
    return luaL_error(l, "No camera is set in the scene and no camera was given.");
  }

  bool failed = false;
  {
    try {
      auto film = std::make_shared<dort::Film>(x_res, y_res, filter);
      std::shared_ptr<GpuSceneDesc> gpu = rec.desc();
      gpu->set_camera(cam.camera_to_world, cam.fov);
      std::shared_ptr<dort::Renderer> renderer;
      if (method == "pt_gpu")
        renderer = std::make_shared<GpuPathRenderer>(scene, film, sampler, camera, gpu, po);
      else if (method == "bdpt_gpu")
        renderer = std::make_shared<GpuBdptRenderer>(scene, film, sampler, camera, gpu, bo);
      else
        renderer = std::make_shared<GpuLightRenderer>(scene, film, sampler, camera, gpu, lo);
      // the job exactly as lua_render_make assembles it (lua_render.cpp:211-219)
      auto job = std::make_shared<dort::RenderJob>();
      job->l = l;
      job->lua_id = std::this_thread::get_id();
      job->renderer = renderer;
      job->film = film;
      job->progress = std::make_shared<dort::RenderJob::JobProgress>();
      job->render_started = false;
      job->render_finished = false;
      dort::lua_push_render_job(l, job);
    } catch (const std::exception& e) {
      lua_pushstring(l, e.what());
      failed = true;
    }
  }
  if (failed) {
    scene.reset();
    filter.reset();
    sampler.reset();
    camera.reset();
    method.clear();
    return lua_error(l);
  }
  return 1;
}

// job_seed(job) -> the 64-bit path seed of the job's last render, as 16 hex digits.
int lua_job_seed(lua_State* l) {
  std::shared_ptr<dort::RenderJob> job = dort::lua_check_render_job(l, 1);
  uint64_t seed = 0;
  if (auto g = std::dynamic_pointer_cast<GpuPathRenderer>(job->renderer)) {
    seed = g->last_seed();
  } else if (auto g2 = std::dynamic_pointer_cast<GpuLightRenderer>(job->renderer)) {
    seed = g2->last_seed();
  } else if (auto g3 = std::dynamic_pointer_cast<GpuBdptRenderer>(job->renderer)) {
    seed = g3->last_seed();
  } else {
    job.reset();
    return luaL_error(l, "dort_gpu: not a GPU render job");
  }
  char buf[24];
  std::snprintf(buf, sizeof buf, "%016" PRIx64, seed);
  lua_pushstring(l, buf);
  return 1;
}

// job_stats(job) -> {paths, extension_rays, shadow_rays, kernel_ms} of the last render.
int lua_job_stats(lua_State* l) {
  std::shared_ptr<dort::RenderJob> job = dort::lua_check_render_job(l, 1);
  dgpu_stats st{};
  if (auto g = std::dynamic_pointer_cast<GpuPathRenderer>(job->renderer)) {
    st = g->last_stats();
  } else if (auto g2 = std::dynamic_pointer_cast<GpuLightRenderer>(job->renderer)) {
    st = g2->last_stats();
  } else if (auto g3 = std::dynamic_pointer_cast<GpuBdptRenderer>(job->renderer)) {
    st = g3->last_stats();
  } else {
    job.reset();
    return luaL_error(l, "dort_gpu: not a GPU render job");
  }
  lua_createtable(l, 0, 4);
  lua_pushinteger(l, (lua_Integer)st.paths);
  lua_setfield(l, -2, "paths");
  lua_pushinteger(l, (lua_Integer)st.extension_rays);
  lua_setfield(l, -2, "extension_rays");
  lua_pushinteger(l, (lua_Integer)st.shadow_rays);
  lua_setfield(l, -2, "shadow_rays");
  lua_pushnumber(l, st.kernel_ms);
  lua_setfield(l, -2, "kernel_ms");
  return 1;
}

// write_image_npy(image, path): an Image<PixelRgbFloat> (dort.render.get_image with
// hdr = true) as a float32 .npy of shape (y_res, x_res, 3), row 0 first.
int lua_write_image_npy(lua_State* l) {
  std::shared_ptr<dort::Image<dort::PixelRgbFloat>> img = dort::lua_check_image_f(l, 1);
  const char* path = luaL_checkstring(l, 2);
  char header[128];
  int n = std::snprintf(header, sizeof header, "{'descr': '<f4', 'fortran_order': False, 'shape': (%d, %d, 3), }",
                        (int)img->res.y, (int)img->res.x);
  // magic, version 1.0, header length; the header is padded with spaces to 64 bytes - 1
  int total = 10 + n + 1;
  int pad = (64 - total % 64) % 64;
  std::string h(header, (size_t)n);
  h.append((size_t)pad, ' ');
  h.push_back('\n');
  FILE* f = std::fopen(path, "wb");
  bool ok = f != nullptr;
  if (ok) {
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    const uint16_t hl = (uint16_t)h.size();
    ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&hl, 2, 1, f) == 1 &&
         std::fwrite(h.data(), 1, h.size(), f) == h.size() &&
         std::fwrite(img->storage.data(), sizeof(float), img->storage.size(), f) == img->storage.size();
    ok = (std::fclose(f) == 0) && ok;
  }
  img.reset();
  h.clear();
  h.shrink_to_fit();
  if (!ok) return luaL_error(l, "dort_gpu: cannot write %s", path);
  return 0;
}

}  // namespace

int luaopen_dort_gpu(lua_State* l) {
  static const luaL_Reg recorder_methods[] = {
      {"add_material", rec_add_material},   {"add_triangle", rec_add_triangle},
      {"add_ply_mesh", rec_add_ply_mesh},   {"add_cube", rec_add_cube},
      {"add_area_light", rec_add_area_light}, {"set_infinite", rec_set_infinite},
      {"set_camera", rec_set_camera},       {"num_triangles", rec_num_triangles},
      {"write_json", rec_write_json},
      {"__gc", dort::lua_gc_managed_obj<Recorder, RECORDER_TNAME>},
      {nullptr, nullptr}};
  static const luaL_Reg funs[] = {
      {"recorder", rec_new},         {"rgb", lua_rgb},
      {"make_job", lua_make_job},    {"job_seed", lua_job_seed},
      {"job_stats", lua_job_stats},  {"write_image_npy", lua_write_image_npy},
      {nullptr, nullptr}};
  dort::lua_register_type(l, RECORDER_TNAME, recorder_methods);
  luaL_newlib(l, funs);
  return 1;
}

}  // namespace dort_amd
