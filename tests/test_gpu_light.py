"""dgpu_render_light (dort's `lt` renderer, LightRenderer) on the GPU against dort.

* dort's own test of its light tracer (test/test_box.lua:198-330: the Cornell box
  "box_diff_area_pin" rendered with `lt` must pass test_convergence against the `pt`
  reference images test/ref/box_diff_area_pin_d{0..3}.hdr and box_diff_area_pin.hdr),
  here with the stricter golden_ok rule of the path-tracer tests;
* dort's LightRenderer itself, compiled from the reference checkout into oracle/_ref:
  the GPU's converged light-tracer image must accept dort's noisy one under dort's
  test_convergence, and their means agree, for an area light and for an open box under
  an InfiniteLight (sample_ray_radiance from the scene's bounding disk);
* the GPU path tracer on the same scenes: two estimators of one image must agree.
"""
import os

import numpy as np
import pytest

from dort_amd import abi, scenes
from oracle import convergence, ref, rgbe
from tests.test_oracle import golden_ok

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name,min_depth,max_depth", [
    ("box_diff_area_pin_d0", 0, 0), ("box_diff_area_pin_d1", 1, 1), ("box_diff_area_pin_d2", 2, 2),
    ("box_diff_area_pin_d3", 3, 3), ("box_diff_area_pin", 0, 3)])
def test_lt_dort_reference_images(gpu_ctx, name, min_depth, max_depth):
    golden = rgbe.read_hdr(os.path.join(GOLDEN, name + ".hdr"))
    sc = scenes.dort_test_box("area")
    opts = abi.render_opts(512, 512, spp=64, seed=21, min_depth=min_depth, max_depth=max_depth)
    s = gpu_ctx.scene(sc)
    img = s.light_image(opts)
    print(name, img.reshape(-1, 3).mean(0), golden.reshape(-1, 3).mean(0))
    ok, msg = golden_ok(golden, img)
    assert ok, msg
    # dort's own rule at dort's own iteration count (test_box.lua:236-239: 10)
    img10 = s.light_image(abi.render_opts(512, 512, spp=10, seed=22, min_depth=min_depth, max_depth=max_depth))
    status, msg = convergence.compare_with_ref(golden, img10)
    assert status != "err", msg


def test_lt_stats(gpu_ctx):
    sc = scenes.dort_test_box("area")
    opts = abi.render_opts(64, 48, spp=4, seed=5, max_depth=5)
    film, st = gpu_ctx.scene(sc).render_light(opts)
    assert st["paths"] == 64 * 48 * 4
    assert st["extension_rays"] > 0 and st["shadow_rays"] > 0
    # stage 0 holds the light-to-camera connections: at most one per path
    assert 0 < st["bounce_shadow_rays"][0] <= st["paths"]
    # paths of at most max_depth + 2 vertices: extension stages 0 .. max_depth - 1
    assert st["bounce_extension_rays"][5] == 0 and st["bounce_extension_rays"][4] > 0
    assert np.isfinite(film).all() and (film >= 0).all()


def _dort_lt(sc, opts):
    return ref.RefScene(sc).render_light(opts) / float(opts.sample_end - opts.sample_begin)


@pytest.mark.parametrize("kind", ["area", "infi"])
def test_lt_matches_dort_light_renderer(gpu_ctx, kind):
    if not ref.available():
        pytest.skip("oracle/_ref (dort compiled from the reference checkout) is not built")
    sc = scenes.dort_test_box(kind, open_box=(kind == "infi"))
    res = 128
    gpu = gpu_ctx.scene(sc).light_image(abi.render_opts(res, res, spp=2048, seed=31, max_depth=3))
    theirs = _dort_lt(sc, abi.render_opts(res, res, spp=32, seed=7, max_depth=3))
    a, b = gpu.reshape(-1, 3).mean(0), theirs.reshape(-1, 3).mean(0)
    print(kind, "gpu", a, "dort", b)
    assert np.allclose(a, b, rtol=0.03), (a, b)
    msg = convergence.test_convergence(gpu, theirs)
    assert msg == "", msg


@pytest.mark.parametrize("kind", ["area", "infi"])
def test_lt_agrees_with_pt(gpu_ctx, kind):
    """The light tracer and the path tracer estimate the same image (pt with max_depth
    d counts paths of up to d + 2 vertices, like lt with max_depth d).  Under the
    InfiniteLight the camera sees the sky directly, a 2-vertex path that a light path
    cannot reach (the light has no position), so both start at 3 vertices there."""
    sc = scenes.dort_test_box(kind, open_box=(kind == "infi"))
    s = gpu_ctx.scene(sc)
    res = 128
    lo = 0 if kind == "area" else 1
    pt = s.image(abi.render_opts(res, res, spp=1024, seed=41, min_depth=lo, max_depth=3))
    lt = s.light_image(abi.render_opts(res, res, spp=1024, seed=42, min_depth=lo, max_depth=3))
    a, b = pt.reshape(-1, 3).mean(0), lt.reshape(-1, 3).mean(0)
    print(kind, "pt", a, "lt", b)
    assert np.allclose(a, b, rtol=0.02), (a, b)
    msg = convergence.test_convergence(pt, lt)
    assert msg == "", msg


def test_lt_errors(gpu_ctx):
    from dort_amd import gpu
    s = gpu_ctx.scene(scenes.dort_test_box("area"))
    with pytest.raises(gpu.DgpuError):
        s.render_light(abi.render_opts(16, 16, spp=1, min_depth=3, max_depth=2))
    with pytest.raises(gpu.DgpuError):
        s.render_light(abi.render_opts(16, 16, spp=1, max_depth=100))
    with pytest.raises(gpu.DgpuError):
        s.render_light(abi.render_opts(0, 16, spp=1))
    # an empty sample range adds nothing
    film, st = s.render_light(abi.render_opts(16, 16, spp=0))
    assert st["paths"] == 0 and not film.any()
