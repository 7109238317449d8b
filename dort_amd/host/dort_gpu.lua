--- dort_gpu: dort Lua scenes rendered by the MI355X path tracer.
--
-- Run once after dort's modules are registered (the `dort` global of main.cpp:181-199)
-- and the C module is loaded as the global `dort_gpu` (lua_gpu.cpp), before any
-- scene script.  It
--
--  * records, beside every scene a script builds through dort.builder
--    (lua_builder.cpp:104-710), the typed scene the GPU path takes: triangles of
--    meshes, PLY meshes and cube() shapes, the materials on them, triangle area
--    lights, an infinite light and the pinhole camera;
--  * adds the renderers "pt_gpu", "lt_gpu" and "bdpt_gpu" to dort.render.make, with the
--    options of "pt", "lt" and "bdpt" (lua_render.cpp:121-160), so dort.render.render_sync,
--    get_image and the DSL's render() work on their jobs unchanged.
--
-- Recording never stops a script: what the GPU path cannot express (a phong or mirror
-- material, a sphere, a point light, textures, frames) is noted, and making a
-- GPU job for that scene fails with the reason.
--
-- The constructors are wrapped in their module tables, so this must run before
-- dort/dsl is first required (the DSL copies some of them, dsl.lua:69-163).

local C = dort_gpu
assert(C and C.recorder, "dort_gpu: the C module must be loaded first")

local info = setmetatable({}, { __mode = "k" })      -- dort object -> what it is
local builders = setmetatable({}, { __mode = "k" })  -- builder -> recording
local scenes = setmetatable({}, { __mode = "k" })    -- scene -> recording

local function rgb_table(x, default)
  if x == nil then return { default, default, default } end
  local r, g, b = C.rgb(x)
  if r == nil then return nil end
  return { r, g, b }
end

-- Replaces tbl[name] with a function that first lets `describe` read the arguments
-- (dort's constructors consume their params table) and then files the description
-- under the object the original returns.
local function wrap_constructor(tbl, name, describe)
  local orig = tbl[name]
  tbl[name] = function(...)
    local d = describe(...)
    local obj = orig(...)
    if d ~= nil then info[obj] = d end
    return obj
  end
end

-- ------------------------------------------------------------- materials
local M = dort.material

wrap_constructor(M, "make_lambert", function(p)
  local albedo = rgb_table(p and p.albedo, 1)
  if not albedo then return { unsupported = "lambert_material with a texture albedo" } end
  if albedo[1] == 0 and albedo[2] == 0 and albedo[3] == 0 then
    -- reflects nothing: the GPU path ends such paths at once
    return { type = "none" }
  end
  return { type = "lambert", albedo = albedo }
end)

local function dielectric(kind, p)
  p = p or {}
  local reflect, transmit = rgb_table(p.reflect_tint, 1), rgb_table(p.transmit_tint, 1)
  if not reflect or not transmit then return { unsupported = kind .. " with texture tints" } end
  local d = { type = kind, reflect_tint = reflect, transmit_tint = transmit,
              ior_inside = p.ior_inside, ior_outside = p.ior_outside }
  if kind == "dielectric" and p.is_thin then
    return { unsupported = "dielectric_material with is_thin" }
  end
  if kind == "rough_dielectric" then
    if (p.distribution or "beckmann") ~= "ggx" then
      return { unsupported = "rough_dielectric_material with the " ..
        tostring(p.distribution or "beckmann") .. " distribution (the GPU lobe is GGX)" }
    end
    if p.roughness ~= nil and type(p.roughness) ~= "number" then
      return { unsupported = "rough_dielectric_material with a texture roughness" }
    end
    d.roughness = p.roughness
  end
  return d
end
wrap_constructor(M, "make_dielectric", function(p) return dielectric("dielectric", p) end)
wrap_constructor(M, "make_rough_dielectric", function(p) return dielectric("rough_dielectric", p) end)
for _, name in ipairs({ "make_oren_nayar", "make_mirror", "make_phong", "make_bump" }) do
  wrap_constructor(M, name, function()
    return { unsupported = "material " .. name:sub(6) }
  end)
end

-- ------------------------------------------------------------- shapes, lights, cameras
wrap_constructor(dort.shape, "make_cube", function() return { shape = "cube" } end)
wrap_constructor(dort.builder, "make_triangle", function(_, mesh, index)
  return { shape = "triangle", mesh = mesh, index = index }
end)

wrap_constructor(dort.light, "make_diffuse", function(p)
  return { light = "diffuse", shape = p.shape, transform = p.transform,
           radiance = p.radiance }
end)
wrap_constructor(dort.light, "make_infinite", function(p)
  return { light = "infinite", radiance = p.radiance }
end)

wrap_constructor(dort.camera, "make_pinhole", function(p)
  return { camera = "pinhole", transform = p.transform, fov = p.fov }
end)

-- ------------------------------------------------------------- the builder
local B_ = dort.builder
local orig = {}
for _, name in ipairs({ "make", "push_state", "pop_state", "push_frame", "pop_frame",
    "set_material", "set_camera", "add_shape", "add_primitive", "add_triangle", "add_light",
    "add_read_ply_mesh", "add_read_ply_mesh_as_bvh", "add_ply_mesh", "add_ply_mesh_as_bvh",
    "add_voxel_grid", "build_scene" }) do
  orig[name] = B_[name]
end

local function note(r, why)
  if not r.unsupported then r.unsupported = why end
end

local function material_index(r)
  local m = r.material
  if r.mat_index[m] then return r.mat_index[m] end
  local d = info[m]
  local idx
  if d == nil or d.unsupported then
    note(r, d and d.unsupported or "a material of an unknown kind")
    idx = r.rec:add_material({ type = "none" })
  else
    idx = r.rec:add_material(d)
  end
  r.mat_index[m] = idx
  return idx
end

function B_.make(...)
  local B = orig.make(...)
  builders[B] = { rec = C.recorder(), material = nil, stack = {}, frames = 0,
                  mat_index = {}, lights = {}, light_tris = {} }
  return B
end

function B_.push_state(B, ...)
  orig.push_state(B, ...)
  local r = builders[B]
  if r then r.stack[#r.stack + 1] = r.material end
end

function B_.pop_state(B, ...)
  orig.pop_state(B, ...)
  local r = builders[B]
  if r then r.material = table.remove(r.stack) end
end

function B_.push_frame(B, ...)
  orig.push_frame(B, ...)
  local r = builders[B]
  if r then
    r.stack[#r.stack + 1] = r.material
    r.frames = r.frames + 1
  end
end

function B_.pop_frame(B, ...)
  local prim = orig.pop_frame(B, ...)
  local r = builders[B]
  if r then
    r.material = table.remove(r.stack)
    r.frames = r.frames - 1
  end
  return prim
end

function B_.set_material(B, m, ...)
  orig.set_material(B, m, ...)
  local r = builders[B]
  if r then r.material = m end
end

function B_.set_camera(B, cam, ...)
  orig.set_camera(B, cam, ...)
  local r = builders[B]
  if r then r.camera = cam end
end

-- Shapes added inside a frame become an aggregate that add_primitive may place any
-- number of times; the recorder does not expand those.
local function in_root(r, what)
  if r.frames > 0 then
    note(r, what .. " inside a frame (aggregate primitives)")
    return false
  end
  return true
end

function B_.add_triangle(B, mesh, index, ...)
  orig.add_triangle(B, mesh, index, ...)
  local r = builders[B]
  -- the mesh's own coordinates, without the builder transform (lua_builder.cpp:366-374)
  if r and in_root(r, "add_triangle") then
    r.rec:add_triangle(mesh, index, material_index(r))
  end
end

function B_.add_shape(B, shape, ...)
  -- pass the arguments on as given: dort looks at how many there are (lua_builder.cpp:335)
  orig.add_shape(B, shape, ...)
  local light = ...
  local r = builders[B]
  if not r or not in_root(r, "add_shape") then return end
  local d = info[shape]
  local t = B_.get_transform(B)
  local first, count
  if d and d.shape == "cube" then
    local md = info[r.material]
    if md and (md.type == "dielectric" or md.type == "rough_dielectric") then
      -- CubeShape's normal faces the ray on both sides (cube_shape.cpp:57), which a
      -- triangle's fixed front side cannot give a dielectric
      note(r, "cube() with a dielectric material")
    end
    first, count = r.rec:add_cube(t, material_index(r)), 12
  elseif d and d.shape == "triangle" then
    first, count = r.rec:add_triangle(d.mesh, d.index, material_index(r), t), 1
  else
    note(r, "a shape other than cube() and triangle()")
    return
  end
  if light ~= nil then
    if d.shape ~= "triangle" then note(r, "an area light on a cube()") end
    r.light_tris[light] = { first = first, count = count, transform = t, shape = shape }
  end
end

function B_.add_primitive(B, ...)
  orig.add_primitive(B, ...)
  local r = builders[B]
  if r then note(r, "add_primitive (aggregate primitives)") end
end

function B_.add_light(B, light, ...)
  orig.add_light(B, light, ...)
  local r = builders[B]
  if r then r.lights[#r.lights + 1] = light end
end

local function add_ply(r, ply, t)
  if in_root(r, "a PLY mesh") then r.rec:add_ply_mesh(ply, material_index(r), t) end
end

function B_.add_read_ply_mesh(B, file_name, ...)
  orig.add_read_ply_mesh(B, file_name, ...)
  local r = builders[B]
  if r then add_ply(r, dort.shape.read_ply_mesh(file_name), B_.get_transform(B)) end
end

function B_.add_read_ply_mesh_as_bvh(B, file_name, ...)
  orig.add_read_ply_mesh_as_bvh(B, file_name, ...)
  local r = builders[B]
  if r then add_ply(r, dort.shape.read_ply_mesh(file_name), B_.get_transform(B)) end
end

function B_.add_ply_mesh(B, ply, ...)
  orig.add_ply_mesh(B, ply, ...)
  local r = builders[B]
  if r then add_ply(r, ply, B_.get_transform(B)) end
end

function B_.add_ply_mesh_as_bvh(B, ply, ...)
  orig.add_ply_mesh_as_bvh(B, ply, ...)
  local r = builders[B]
  if r then add_ply(r, ply, B_.get_transform(B)) end
end

function B_.add_voxel_grid(B, ...)
  orig.add_voxel_grid(B, ...)
  local r = builders[B]
  if r then note(r, "a voxel grid") end
end

local identity = dort.geometry.identity()

-- Lights and the camera are resolved when the scene is built: a diffuse light must
-- have been added with add_shape(shape, light) on a triangle, under its own transform.
local function finish(r)
  for _, light in ipairs(r.lights) do
    local d = info[light]
    if d and d.light == "diffuse" then
      local lt = r.light_tris[light]
      if not lt then
        note(r, "a diffuse light whose shape is not in the scene")
      elseif lt.shape ~= d.shape or lt.transform ~= (d.transform or identity) then
        note(r, "a diffuse light whose shape or transform differs from its primitive's")
      else
        for i = 0, lt.count - 1 do r.rec:add_area_light(lt.first + i, d.radiance) end
      end
    elseif d and d.light == "infinite" then
      if r.has_infinite then note(r, "two infinite lights") end
      r.has_infinite = true
      r.rec:set_infinite(d.radiance)
    else
      note(r, "a light other than diffuse and infinite lights")
    end
  end
  if r.camera ~= nil then
    local d = info[r.camera]
    if d and d.camera == "pinhole" then
      r.rec:set_camera(d.transform, d.fov)
    else
      r.camera_unsupported = "a camera other than pinhole_camera"
    end
  end
end

function B_.build_scene(B, ...)
  local scene = orig.build_scene(B, ...)
  local r = builders[B]
  if r then
    finish(r)
    scenes[scene] = r
  end
  return scene
end

-- ------------------------------------------------------------- rendering
local render_make = dort.render.make

function dort.render.make(scene, params)
  local kind = params and params.renderer
  if kind ~= "pt_gpu" and kind ~= "lt_gpu" and kind ~= "bdpt_gpu" then
    return render_make(scene, params)
  end
  local r = scenes[scene]
  if not r then
    error(kind .. ": the scene was not built through dort.builder", 2)
  end
  if r.unsupported then
    error(kind .. ": the GPU path cannot render this scene: " .. r.unsupported, 2)
  end
  local cam_t, cam_fov
  if params.camera ~= nil then
    local d = info[params.camera]
    if not d or d.camera ~= "pinhole" then
      error(kind .. ": the GPU path renders through a pinhole_camera only", 2)
    end
    cam_t, cam_fov = d.transform, d.fov
  elseif r.camera_unsupported then
    error(kind .. ": the GPU path cannot render this scene: " .. r.camera_unsupported, 2)
  end
  return C.make_job(scene, r.rec, params, cam_t, cam_fov)
end

--- The recording of `scene`: its recorder (write_json, num_triangles) and the
-- reason the GPU path cannot render it, if any.
function C.recording(scene)
  local r = scenes[scene]
  if not r then return nil end
  return r.rec, r.unsupported or r.camera_unsupported
end

return C
