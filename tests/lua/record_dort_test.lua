-- Records one scene of one of dort's own test files with the dort_gpu prelude.
-- A dort test file (test/test_box.lua and its like) returns a function that calls
-- t:test{name = ..., scene = ..., renders = ..., ref_opts = ...} once per scene
-- (test/test.lua:59-176, 207-211); this script passes it a `t` that keeps the scene
-- named SCENE and stops.  Needs dort's lua/ tree (dort_lua --lua-dir).
--
--   record_dort_test.lua TEST_FILE SCENE OUT.json [RENDERER]
--
-- Writes the recorded GPU scene to OUT.json and prints "unsupported: <reason>" when
-- the GPU path cannot render it.  With RENDERER, it also makes a job of that renderer
-- for the scene (8 x 8, one iteration) and renders it.
require "dort/std"

local test_file, wanted, out, renderer = arg[1], arg[2], arg[3], arg[4]
local STOP = {}
local scene
local ok, err = pcall(dofile(test_file), { test = function(_, def)
  if def.name == wanted then
    scene = def.scene
    error(STOP)
  end
end })
if not ok and err ~= STOP then error(err, 0) end
assert(scene, "no scene named " .. wanted)

local rec, why = dort_gpu.recording(scene)
assert(rec, "the scene was not recorded")
rec:write_json(out)
print("triangles: " .. rec:num_triangles())
if why then print("unsupported: " .. why) end

if renderer then
  local job = dort.render.make(scene, { renderer = renderer, x_res = 8, y_res = 8, iterations = 1 })
  dort.render.render_sync(job)
  print("rendered with " .. renderer)
end
