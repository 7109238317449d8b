-- The Cornell scene of dort's test suite (test/test_box.lua:1-197, geometry "box",
-- surfaces "diff", light "area", camera "pin") at scale 0.01, built with dort's C
-- modules directly (dort.builder, dort.shape, ...), so that it runs where dort's .lua
-- files are not.  The Lua front-end tests check that it records the same GPU scene as
-- test_box.lua itself does through dort/dsl.
--
--   dort_lua --prelude dort_amd/host/dort_gpu.lua tests/lua/box_area_scene.lua record OUT.json
--   dort_lua --prelude ... tests/lua/box_area_scene.lua render OUT.npy META.json RENDERER
--       X_RES Y_RES ITERATIONS MIN_DEPTH MAX_DEPTH SAMPLER_SEED
local b, g, shape, mat, light, spec = dort.builder, dort.geometry, dort.shape,
  dort.material, dort.light, dort.spectrum
local s = 0.01

local function scene()
  local B = b.make()
  local white = mat.make_lambert { albedo = spec.rgb(0.5, 0.5, 0.5) }
  local green = mat.make_lambert { albedo = spec.rgb(0, 0.5, 0) }
  local red = mat.make_lambert { albedo = spec.rgb(0.5, 0, 0) }

  b.push_state(B)
  local m = shape.make_mesh {
    transform = b.get_transform(B),
    points = {
      g.point(552.8 * s, 0, 0), g.point(0, 0, 0), g.point(0, 0, 559.2 * s),
      g.point(549.6 * s, 0, 559.2 * s), g.point(556 * s, 548.8 * s, 0),
      g.point(556 * s, 548.8 * s, 559.2 * s), g.point(0, 548.8 * s, 559.2 * s),
      g.point(0, 548.8 * s, 0),
    },
    vertices = { 0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7, 6, 5, 3, 6, 3, 2, 7, 6, 2, 7, 2, 1,
                 0, 3, 5, 0, 5, 4 },
  }
  b.set_material(B, white)
  for _, i in ipairs({ 0, 3, 6, 9, 12, 15 }) do b.add_triangle(B, m, i) end
  b.set_material(B, green)
  for _, i in ipairs({ 18, 21 }) do b.add_triangle(B, m, i) end
  b.set_material(B, red)
  for _, i in ipairs({ 24, 27 }) do b.add_triangle(B, m, i) end
  b.pop_state(B)

  for _, box in ipairs({
      { g.translate(185 * s, 82.5 * s, 169 * s), -0.29, g.scale(165 * s / 2) },
      { g.translate(368 * s, 165 * s, 351 * s), -1.27, g.scale(165 * s / 2, 330 * s / 2, 165 * s / 2) } }) do
    b.push_state(B)
    b.set_material(B, white)
    b.set_transform(B, box[1] * g.rotate_y(box[2]) * box[3])
    b.add_shape(B, shape.make_cube())
    b.pop_state(B)
  end

  b.set_camera(B, dort.camera.make_pinhole {
    transform = g.look_at(g.point(278 * s, 273 * s, -800 * s), g.point(278 * s, 273 * s, 0),
      g.vector(0, 1, 0)) * g.scale(1, -1, 1),
    fov = 0.686,
  })

  b.push_state(B)
  b.set_material(B, mat.make_lambert { albedo = spec.rgb(0) })
  b.set_transform(B, g.translate(0, -1 * s, 0))
  local lm = shape.make_mesh {
    transform = b.get_transform(B),
    points = {
      g.point(343 * s, 548.8 * s, 227 * s), g.point(343 * s, 548.8 * s, 332 * s),
      g.point(213 * s, 548.8 * s, 332 * s), g.point(213 * s, 548.8 * s, 227 * s),
    },
    vertices = { 2, 1, 0, 3, 2, 0 },
  }
  for _, index in ipairs({ 0, 3 }) do
    local tri = b.make_triangle(B, lm, index)
    local l = light.make_diffuse { radiance = spec.rgb(100), shape = tri, transform = b.get_transform(B) }
    b.add_light(B, l)
    b.add_shape(B, tri, l)
  end
  b.pop_state(B)
  return b.build_scene(B)
end

local sc = scene()
local mode = arg[1]
if mode == "record" then
  local rec, why = dort_gpu.recording(sc)
  assert(rec, "the scene was not recorded")
  assert(why == nil, why)
  rec:write_json(arg[2])
elseif mode == "render" then
  local job = dort.render.make(sc, {
    renderer = arg[4],
    x_res = math.tointeger(tonumber(arg[5])), y_res = math.tointeger(tonumber(arg[6])),
    iterations = math.tointeger(tonumber(arg[7])),
    min_depth = math.tointeger(tonumber(arg[8])), max_depth = math.tointeger(tonumber(arg[9])),
    sampler = dort.sampler.make_random { samples_per_pixel = 1, seed = math.tointeger(tonumber(arg[10])) },
  })
  dort.render.render_sync(job)
  dort_gpu.write_image_npy(dort.render.get_image(job, { hdr = true }), arg[2])
  local meta = "{}"
  if arg[4]:sub(-4) == "_gpu" then
    local st = dort_gpu.job_stats(job)
    meta = string.format('{"seed": "%s", "paths": %d, "extension_rays": %d, "shadow_rays": %d}',
      dort_gpu.job_seed(job), st.paths, st.extension_rays, st.shadow_rays)
  end
  local f = assert(io.open(arg[3], "w"))
  f:write(meta, "\n")
  f:close()
  print(string.format("rendered %s x %s, %s iterations with %s: %s", arg[5], arg[6], arg[7],
    arg[4], meta))
else
  error("mode must be record or render")
end
