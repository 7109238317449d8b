"""The autotuner (dort_amd/autotune.py) on the GPU: it times the candidates it lists,
returns one of them, and the knobs it picks render the same samples as the defaults
(the BVH shape and the launch knobs change how rays are traced, not which triangle
they hit, except at FP ties)."""
import numpy as np
import pytest

from dort_amd import abi, autotune, scenes

pytestmark = pytest.mark.gpu


def _parity(a, b):
    return float((np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-4).all(axis=-1).mean())


def test_autotune_one_scene(gpu_ctx):
    sc = scenes.terrain_scene(quads=60, seed=4, sky_res=(128, 64))
    opts = abi.render_opts(48, 48, spp=2, seed=3, max_depth=6, rr_depth=3)
    best, table = autotune.autotune(gpu_ctx, sc, opts, shape_space=[(2, 4), (4, 2), (4, 8), (8, 4)],
                                    launch_space=[("stack_lds", (8, 32)), ("block_size", (64, 256)),
                                                  ("queue_tile", (1000,))])
    print("best", best, "of", [(k, round(ms, 3)) for k, ms in table])
    assert len(table) == 10
    assert all(np.isfinite(ms) and ms > 0 for _, ms in table)
    assert best in [k for k, _ in table]
    assert min(ms for _, ms in table) == dict((str(k), ms) for k, ms in table)[str(best)]
    ref_rgb, ref_w, _ = gpu_ctx.scene(sc).render(opts)
    rgb, w, _ = gpu_ctx.scene(sc, **best).render(opts)
    assert np.array_equal(w, ref_w)
    assert _parity(rgb / np.maximum(w, 1)[..., None], ref_rgb / np.maximum(ref_w, 1)[..., None]) >= 0.99


def test_autotune_batch_knobs_keep_batch_equal_to_per_scene(gpu_ctx):
    batch = scenes.batch_scenes(4, seed=5, max_tris=20000, sky_res=(64, 32))
    sc_list = [s for s, _ in batch]
    opts = [abi.render_opts(24, 24, spp=2, seed=40 + k, max_depth=6, rr_depth=3, scene_id=k) for k in range(4)]
    best, table = autotune.autotune_batch(gpu_ctx, sc_list, opts, shape_space=[(2, 4)],
                                          launch_space=[("stack_lds", (8,)), ("queue_tile", (700,))])
    print("best", best, "of", [(k, round(ms, 3)) for k, ms in table])
    assert len(table) == 4 and all(np.isfinite(ms) for _, ms in table)
    items = [(gpu_ctx.scene(s, **best), o) for s, o in zip(sc_list, opts)]
    films, _ = gpu_ctx.render_batch(items)
    for (g, o), (rgb, w) in zip(items, films):
        ref_rgb, ref_w, _ = g.render(o)
        assert np.array_equal(w, ref_w)
        np.testing.assert_allclose(rgb, ref_rgb, rtol=1e-5, atol=1e-6)


def test_autotune_batch_classes(gpu_ctx):
    """Per-class knobs: scenes split by triangle count, each class tuned on its own
    sample; the batch rendered with every scene's class knobs equals the per-scene renders."""
    batch = scenes.batch_scenes(6, seed=8, min_tris=1000, max_tris=30000, sky_res=(64, 32))
    sc_list = [s for s, _ in batch]
    opts = [abi.render_opts(20, 20, spp=2, seed=70 + k, max_depth=5, rr_depth=3, scene_id=k) for k in range(6)]
    cut = sorted(s.num_triangles for s in sc_list)[3]
    knobs, classes = autotune.autotune_batch_classes(
        gpu_ctx, sc_list, opts, bounds=(cut,), per_class=2, shape_space=[(8, 4)],
        launch_space=[("stack_lds", (32,))])
    print("classes", [(b, idx, best) for b, idx, best, _ in classes])
    assert sorted(i for _, idx, _, _ in classes for i in idx) == list(range(6))
    for (lo, hi), idx, best, table in classes:
        assert all(lo <= sc_list[i].num_triangles < hi for i in idx)
        assert len(table) == 3 and all(np.isfinite(ms) for _, ms in table)
        assert all(knobs[i] == best for i in idx)
    items = [(gpu_ctx.scene(s, **k), o) for s, k, o in zip(sc_list, knobs, opts)]
    films, _ = gpu_ctx.render_batch(items)
    for (g, o), (rgb, w) in zip(items, films):
        ref_rgb, ref_w, _ = g.render(o)
        assert np.array_equal(w, ref_w)
        np.testing.assert_allclose(rgb, ref_rgb, rtol=1e-5, atol=1e-6)
