"""dort's Lua front-end over the GPU path: the dort_gpu module (dort_amd/host/lua_gpu.cpp)
and its Lua half (dort_amd/host/dort_gpu.lua) in a dort Lua host built from dort's own
sources (oracle/_ref/dort_lua, oracle/lua_host.cpp).

CPU tests (need the host; those that run dort's own .lua files also need the reference
checkout):

* dort's test/test_box.lua, run through dort/dsl, records the scene
  scenes.dort_test_box("area") / ("infi", open_box=True) describes: the same triangles
  (to float rounding of dort's transforms), materials, lights and camera;
* tests/lua/box_area_scene.lua, which states that scene with dort's C modules directly
  (and runs where dort's .lua files are not), records byte-for-byte the same scene;
* scenes the GPU path cannot express are refused with the reason when a "pt_gpu" job is
  made; "pt" still renders with dort; without a GPU, a "pt_gpu" job fails loudly.

GPU test: a "pt_gpu", "lt_gpu" or "bdpt_gpu" job made and rendered through
dort.render.make / render_sync / get_image gives the image the C-ABI renders for the same
scene and seed.
"""
import json
import os

import numpy as np
import pytest

from dort_amd import abi, scenes
from oracle import ref

HERE = os.path.dirname(os.path.abspath(__file__))
BOX_SCRIPT = os.path.join(HERE, "lua", "box_area_scene.lua")
RECORD_SCRIPT = os.path.join(HERE, "lua", "record_dort_test.lua")
REF_TEST_DIR = os.path.join(ref.REF_ROOT, "test")
REF_LUA_DIR = os.path.join(ref.REF_ROOT, "lua")

needs_host = pytest.mark.skipif(not ref.lua_available(), reason="oracle/_ref/dort_lua not built")
needs_reference = pytest.mark.skipif(not os.path.isdir(REF_LUA_DIR), reason="the reference checkout is not here")


def _record_dort_test(tmp_path, scene, renderer=None):
    out = tmp_path / f"{scene}.json"
    args = [os.path.join(REF_TEST_DIR, "test_box.lua"), scene, out]
    if renderer:
        args.append(renderer)
    p = ref.run_lua(RECORD_SCRIPT, *args, lua_dir=REF_LUA_DIR, cwd=REF_TEST_DIR)
    return p, (json.loads(out.read_text()) if out.exists() else None)


def _soup(sc):
    return np.asarray(sc.positions, np.float32)[np.asarray(sc.indices)].reshape(-1, 9)


def _assert_records(d, sc):
    pts = np.asarray(d["points"], np.float32).reshape(-1, 9)
    soup = _soup(sc)
    assert pts.shape == soup.shape
    # dort composes and applies its transforms in float32; the scene module in float64
    np.testing.assert_allclose(pts, soup, rtol=0, atol=5e-7)
    assert d["uvs"] is None and sc.uvs is None
    # each triangle's material (the scene module also lists materials no triangle uses)
    got = [(d["materials"][i]["type"], tuple(np.float32(d["materials"][i]["albedo"])))
           for i in d["tri_material"]]
    want = [(sc.materials[i].type, tuple(np.float32(list(sc.materials[i].albedo)))) for i in sc.tri_material]
    assert got == want
    assert [(a["triangle"], a["radiance"]) for a in d["area_lights"]] == \
        [(int(t), [float(x) for x in r]) for t, r in sc.area_lights]
    if sc.infinite is None:
        assert d["infinite"] is None
    else:
        assert d["infinite"] == [float(x) for x in sc.infinite]
    np.testing.assert_allclose(np.float32(d["camera"]["camera_to_world"]).reshape(3, 4),
                               np.asarray(sc.camera_to_world, np.float32), rtol=0, atol=5e-7)
    assert np.float32(d["camera"]["fov"]) == np.float32(sc.fov)


def _scene_from_record(d):
    """A scenes.Scene of a recorded GPU scene (write_json of the dort_gpu recorder)."""
    pts = np.asarray(d["points"], np.float32).reshape(-1, 3)
    mats = []
    for m in d["materials"]:
        mat = abi.Material()
        mat.type = m["type"]
        mat.albedo_tex = -1
        mat.albedo = abi.f3(*m["albedo"])
        mat.roughness = m["roughness"]
        mat.ior_inside = m["ior_inside"]
        mat.ior_outside = m["ior_outside"]
        mat.reflect_tint = abi.f3(*m["reflect_tint"])
        mat.transmit_tint = abi.f3(*m["transmit_tint"])
        mats.append(mat)
    return scenes.Scene(
        positions=pts, indices=np.arange(len(pts), dtype=np.uint32).reshape(-1, 3),
        tri_material=np.asarray(d["tri_material"], np.uint32), materials=mats,
        area_lights=[(a["triangle"], tuple(a["radiance"])) for a in d["area_lights"]],
        camera_to_world=np.float32(d["camera"]["camera_to_world"]).reshape(3, 4), fov=d["camera"]["fov"],
        infinite=tuple(d["infinite"]) if d["infinite"] else None, name="lua_record")


@needs_host
@needs_reference
@pytest.mark.parametrize("scene, expected", [
    ("box_diff_area_pin", lambda: scenes.dort_test_box("area")),
    ("openbox_diff_infi_pin", lambda: scenes.dort_test_box("infi", open_box=True))])
def test_dorts_test_box_records_the_scene_module_scene(tmp_path, scene, expected):
    p, d = _record_dort_test(tmp_path, scene)
    assert p.returncode == 0, p.stderr
    assert "unsupported" not in p.stdout
    _assert_records(d, expected())


@needs_host
@needs_reference
def test_c_module_script_records_what_dorts_test_box_records(tmp_path):
    p, _ = _record_dort_test(tmp_path, "box_diff_area_pin")
    assert p.returncode == 0, p.stderr
    out = tmp_path / "script.json"
    q = ref.run_lua(BOX_SCRIPT, "record", out)
    assert q.returncode == 0, q.stderr
    assert out.read_bytes() == (tmp_path / "box_diff_area_pin.json").read_bytes()


@needs_host
@needs_reference
@pytest.mark.parametrize("scene, reason", [
    ("box_glos_area_pin", "material phong"),
    ("box_diff_sphe_pin", "a shape other than cube() and triangle()"),
    ("box_diff_poin_pin", "a light other than diffuse and infinite lights")])
def test_scenes_the_gpu_path_cannot_express_are_refused(tmp_path, scene, reason):
    p, _ = _record_dort_test(tmp_path, scene, renderer="pt_gpu")
    assert f"unsupported: {reason}" in p.stdout
    assert p.returncode == 1
    assert f"pt_gpu: the GPU path cannot render this scene: {reason}" in p.stderr


@needs_host
def test_pt_still_renders_with_dort(tmp_path):
    npy, meta = tmp_path / "pt.npy", tmp_path / "pt.json"
    p = ref.run_lua(BOX_SCRIPT, "render", npy, meta, "pt", 16, 12, 2, 0, 2, 5)
    assert p.returncode == 0, p.stderr
    img = np.load(npy)
    assert img.shape == (12, 16, 3) and np.isfinite(img).all() and img.mean() > 0


@needs_host
def test_unknown_params_are_an_error_as_in_dort(tmp_path):
    # a params table with a key lua_render_make would not read (lua_params_check_unused)
    script2 = tmp_path / "bad.lua"
    script2.write_text(
        "local B = dort.builder.make()\n"
        "dort.builder.set_camera(B, dort.camera.make_pinhole{transform = dort.geometry.identity()})\n"
        "local sc = dort.builder.build_scene(B)\n"
        "dort.render.make(sc, {renderer = 'pt_gpu', x_res = 4, y_res = 4, bogus = 1})\n")
    p = ref.run_lua(str(script2))
    assert p.returncode == 1
    assert "bogus" in p.stderr


@needs_host
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the failure without a GPU")
def test_pt_gpu_fails_loudly_without_a_gpu(tmp_path):
    p = ref.run_lua(BOX_SCRIPT, "render", tmp_path / "g.npy", tmp_path / "g.json", "pt_gpu", 8, 8, 1, 0, 2, 5)
    assert p.returncode == 1
    assert "GpuPathRenderer: dgpu_create failed" in p.stderr


@pytest.mark.gpu
@needs_host
@pytest.mark.parametrize("renderer", ["pt_gpu", "lt_gpu", "bdpt_gpu"])
def test_lua_job_renders_what_the_c_abi_renders(gpu_ctx, tmp_path, renderer):
    """dort.render.make{renderer = "pt_gpu" | "lt_gpu" | "bdpt_gpu"} -> render_sync -> get_image{hdr}
    on the test-box scene built in Lua equals the C-ABI render of the scene it recorded
    for the seed the job drew from its sampler."""
    x, y, spp, d0, d1 = 96, 64, 8, 0, 3
    npy, meta = tmp_path / "img.npy", tmp_path / "meta.json"
    print(f"lua {renderer}: {x}x{y}, {spp} iterations, depth {d0}..{d1}", flush=True)
    p = ref.run_lua(BOX_SCRIPT, "render", npy, meta, renderer, x, y, spp, d0, d1, 11, timeout=120)
    assert p.returncode == 0, p.stderr
    img, info = np.load(npy), json.loads(meta.read_text())
    opts = abi.render_opts(x, y, spp=spp, seed=int(info["seed"], 16), min_depth=d0, max_depth=d1)
    # the scene exactly as the script recorded it (dort's float32 transforms place its points
    # up to a few ulps from scenes.dort_test_box's)
    rec = tmp_path / "scene.json"
    q = ref.run_lua(BOX_SCRIPT, "record", rec)
    assert q.returncode == 0, q.stderr
    gsc = gpu_ctx.scene(_scene_from_record(json.loads(rec.read_text())))
    if renderer == "pt_gpu":
        assert info["paths"] == x * y * spp
        rgb, w, _ = gsc.render(opts)
        want = rgb / w[..., None]
    elif renderer == "lt_gpu":
        # LightRenderer's film: splat sums times 1 / iterations (light_renderer.cpp:38-44)
        want = gsc.render_light(opts)[0] / spp
    else:
        # BdptRenderer's film: colour / weight + splat sums / iterations
        want = gsc.bdpt_image(opts)
    if renderer == "pt_gpu":
        # the job renders in chunks; per pixel the samples add up in the same order
        np.testing.assert_allclose(img, want, rtol=1e-5, atol=1e-6)
    else:
        # the same light paths; splat sums differ only in the order of float atomics
        np.testing.assert_allclose(img, want, rtol=1e-4, atol=1e-5)
    assert img.mean() > 0
