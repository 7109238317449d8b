// lua_gpu.hpp -- the Lua module "dort_gpu": dort's Lua front-end over the MI355X path tracer.
//
// dort selects its renderers in lua_render_make (src/dort/lua_render.cpp:102-221) and
// builds scenes with the dort.builder functions (src/dort/lua_builder.cpp:104-710).
// This module is what a dort host registers next to its own modules (main.cpp:182-199)
// to render dort Lua scenes on the GPU:
//
//   * dort_gpu.recorder() -- a GpuSceneDesc being filled in while a script builds its
//     dort::Scene: meshes' triangles (with dort's own Transform::apply on the points),
//     analytic cubes as 12 triangles, materials, area lights, the camera;
//   * dort_gpu.make_job(scene, recorder, params [, camera_transform, fov]) -- a dort
//     RenderJob (lua_render.hpp) whose renderer is GpuPathRenderer ("pt_gpu"),
//     GpuLightRenderer ("lt_gpu") or GpuBdptRenderer ("bdpt_gpu"), reading the same params
//     as lua_render_make's "pt", "lt" and "bdpt" branches (lua_render.cpp:104-160) plus
//     rr_depth and the BVH / kernel knobs.
//     dort.render.render_sync / get_image then work on it unchanged.
//
// The Lua half (dort_gpu.lua, run once after dort's modules are registered) wraps the
// dort.builder / material / shape / light / camera constructors so that every scene a
// script builds is recorded, and adds "pt_gpu", "lt_gpu" and "bdpt_gpu" to dort.render.make.
#pragma once

extern "C" {
#include <lua.h>
}

namespace dort_amd {
// Pushes the module table (use with luaL_requiref).
int luaopen_dort_gpu(lua_State* l);
}  // namespace dort_amd
