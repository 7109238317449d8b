"""dort_amd::GpuPathRenderer and GpuLightRenderer in dort's own Renderer seam
(renderer.hpp:12-34).

oracle/_ref/libdortref_gpu.so is dort (its pt-path sources, built in place) with
the GPU renderer of dort_amd/host linked in.  The tests build a dort::Scene, a Film
with dort's box filter, a RandomSampler and the scene's camera, and call
Renderer::render on the GPU renderer, as dort's Lua front-end does for "pt":

* the film equals what the C-ABI renders for the same seed (the adapter adds
  nothing and drops nothing);
* the image agrees with dort's own PathRenderer (CPU) on the same scene;
* Progress cancellation stops between chunks, and an unsupported film filter is
  rejected with an exception, dort's only error channel;
* GpuLightRenderer ("lt") fills the film's splat sums and splat_scale as
  LightRenderer::render does, with what the C-ABI light tracer gives for the seed.
Skipped where the reference checkout was not there to build the library.
"""
import numpy as np
import pytest

from dort_amd import abi, scenes
from oracle import ref

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref.gpu_available(), reason="oracle/_ref/libdortref_gpu.so not built")]


def _opts(spp, **kw):
    o = abi.render_opts(64, 48, spp=spp, seed=0, max_depth=5, rr_depth=2)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_renderer_seam_film_equals_c_abi_render(gpu_ctx):
    sc = scenes.cornell()
    opts = _opts(32)
    rgb, w, info = ref.GpuRendererInDort(sc).render(opts, sampler_seed=42, chunk=5)
    assert info["paths"] == 64 * 48 * 32 and info["percent"] == pytest.approx(100.0)
    np.testing.assert_array_equal(w, np.full((48, 64), 32.0, np.float32))
    direct = abi.RenderOpts.from_buffer_copy(opts)
    direct.seed = info["seed"]
    rgb2, w2, _ = gpu_ctx.scene(sc).render(direct)
    np.testing.assert_array_equal(w, w2)
    # chunked calls accumulate each pixel's samples in the same order as one call
    np.testing.assert_allclose(rgb, rgb2, rtol=1e-6, atol=1e-7)
    assert rgb.sum() > 0


@pytest.mark.parametrize("name", ["cornell", "box_area"])
def test_renderer_seam_agrees_with_dorts_path_renderer(name):
    sc = scenes.cornell() if name == "cornell" else scenes.dort_test_box("area")
    opts = _opts(256)
    rgb, w, _ = ref.GpuRendererInDort(sc).render(opts, sampler_seed=7)
    crgb, cw = ref.RefScene(sc).render(opts)
    img, cimg = rgb / w[..., None], crgb / cw[..., None]
    # the same estimator with different random numbers: image means agree to the
    # noise of 786k paths, 8x8-pixel tile means to a few percent
    m, cm = img.reshape(-1, 3).mean(0), cimg.reshape(-1, 3).mean(0)
    print(f"{name}: GPU mean {m}, dort mean {cm}")
    np.testing.assert_allclose(m, cm, rtol=0.02, atol=1e-4)
    tiles = img.reshape(6, 8, 8, 8, 3).mean((1, 3)).sum(-1)
    ctiles = cimg.reshape(6, 8, 8, 8, 3).mean((1, 3)).sum(-1)
    lit = ctiles > 0.05 * ctiles.max()
    assert np.abs(tiles[lit] - ctiles[lit]).max() / ctiles[lit].max() < 0.08


def test_renderer_seam_progress_cancellation():
    sc = scenes.cornell()
    rgb, w, info = ref.GpuRendererInDort(sc).render(_opts(32), cancel_after=2, chunk=4)
    np.testing.assert_array_equal(w, np.full((48, 64), 8.0, np.float32))
    assert info["percent"] == pytest.approx(25.0)


def test_renderer_seam_rejects_other_filters():
    with pytest.raises(RuntimeError, match="box filter"):
        ref.GpuRendererInDort(scenes.cornell()).render(_opts(4), filter_kind=1)


def test_light_renderer_seam_film_equals_c_abi_render(gpu_ctx):
    sc = scenes.dort_test_box("area")
    opts = abi.render_opts(64, 48, spp=12, seed=0, max_depth=3)
    rgb, info = ref.GpuRendererInDort(sc).render_light(opts, sampler_seed=5, chunk=4)
    assert info["paths"] == 64 * 48 * 12 and info["splat_scale"] == pytest.approx(1 / 12)
    direct = abi.RenderOpts.from_buffer_copy(opts)
    direct.seed = info["seed"]
    rgb2, _ = gpu_ctx.scene(sc).render_light(direct)
    # the same light paths; splat sums differ only in the order of float atomics
    np.testing.assert_allclose(rgb, rgb2, rtol=1e-4, atol=1e-5)
    assert rgb.sum() > 0


def test_bdpt_renderer_seam_film_equals_c_abi_render(gpu_ctx):
    """GpuBdptRenderer (dort's bdpt) adds to the film's colour sums, weights and splat
    sums what the C-ABI renders for its seed, and leaves splat_scale at 1 / iterations
    as BdptRenderer::render does (bdpt_renderer.cpp:22-33)."""
    sc = scenes.dort_test_box("area")
    opts = abi.render_opts(64, 48, spp=12, seed=0, max_depth=3)
    rgb, w, splat, info = ref.GpuRendererInDort(sc).render_bdpt(opts, sampler_seed=9, chunk=4)
    assert info["paths"] == 64 * 48 * 12 and info["splat_scale"] == pytest.approx(1 / 12)
    direct = abi.RenderOpts.from_buffer_copy(opts)
    direct.seed = info["seed"]
    rgb2, w2, splat2, _ = gpu_ctx.scene(sc).render_bdpt(direct)
    np.testing.assert_array_equal(w, w2)
    np.testing.assert_allclose(rgb, rgb2, rtol=1e-6, atol=1e-7)
    # the same light subpaths; splat sums differ only in the order of float atomics
    np.testing.assert_allclose(splat, splat2, rtol=1e-4, atol=1e-5)
    assert rgb.sum() > 0 and splat.sum() > 0
