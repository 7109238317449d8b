"""Batched rendering (dgpu_render_batch): many scenes of uneven size in one batched
wavefront with the bounce loops of every chunk captured in one hipGraph, scenes of
different kernel knobs in groups, from a few small scenes up to
the 64-scene batch workload itself.  Each scene's film must equal
its own dgpu_render: the same counter-based random numbers give the same paths and
the same per-pixel order; the batched and single-scene kernels are separate
instantiations whose FMA contraction may differ, so radiance agrees to a few ULP."""
import numpy as np
import pytest

from dort_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _mixed_scenes():
    out = []
    for k, (sc, res) in enumerate(scenes.batch_scenes(count=5, seed=11, min_tris=500, max_tris=20000,
                                                      sky_res=(64, 32))):
        r = [24, 40, 17, 33, 8][k]
        spp = [3, 1, 4, 2, 5][k]
        out.append((sc, abi.render_opts(r, r + 3, spp=spp, sample_begin=k, seed=100 + k, max_depth=6,
                                        rr_depth=3, scene_id=k)))
    out.append((scenes.cornell(), abi.render_opts(7, 5, spp=3, seed=9, max_depth=6, rr_depth=3, scene_id=9)))
    return out


@pytest.mark.parametrize("tile", [0, 2000])
def test_batch_equals_per_scene(gpu_ctx, tile):
    items = [(gpu_ctx.scene(sc, queue_tile=tile), o) for sc, o in _mixed_scenes()]
    films, st = gpu_ctx.render_batch(items)
    n_paths = sum(o.x_res * o.y_res * (o.sample_end - o.sample_begin) for _, o in items)
    assert st["paths"] == n_paths
    for (gsc, o), (rgb, w) in zip(items, films):
        ref_rgb, ref_w, _ = gsc.render(o)
        assert np.array_equal(w, ref_w)
        np.testing.assert_allclose(rgb, ref_rgb, rtol=1e-5, atol=1e-6)


def test_batch_graph_equals_direct_launches(gpu_ctx):
    items = [(gpu_ctx.scene(sc, queue_tile=3000), o) for sc, o in _mixed_scenes()]
    a, _ = gpu_ctx.render_batch(items, graph=True)
    b, _ = gpu_ctx.render_batch(items, graph=False)
    for (ra, wa), (rb, wb) in zip(a, b):
        assert np.array_equal(ra, rb) and np.array_equal(wa, wb)


@pytest.mark.parametrize("graph", [True, False])
def test_batch_with_mixed_knobs_equals_per_scene(gpu_ctx, graph):
    """Items whose scenes differ in BVH width, LDS stack and trace workgroup form groups
    traced by their own kernel variants (one graph launch for every chunk of every group);
    each film equals the scene's own render with the same knobs.  Small chunks make some
    groups span several chunks."""
    knobs = [dict(width=8, stack_lds=32, block_size=64, queue_tile=900),
             dict(width=4, queue_tile=900), dict(width=2, stack_lds=8, block_size=256, queue_tile=900),
             dict(width=8, stack_lds=32, block_size=64, queue_tile=900), dict(width=4, queue_tile=900),
             dict(width=8, stack_lds=16, block_size=128, queue_tile=900)]
    items = [(gpu_ctx.scene(sc, **k), o) for (sc, o), k in zip(_mixed_scenes(), knobs)]
    films, st = gpu_ctx.render_batch(items, graph=graph)
    assert st["paths"] == sum(o.x_res * o.y_res * (o.sample_end - o.sample_begin) for _, o in items)
    for (gsc, o), (rgb, w) in zip(items, films):
        ref_rgb, ref_w, _ = gsc.render(o)
        assert np.array_equal(w, ref_w)
        np.testing.assert_allclose(rgb, ref_rgb, rtol=1e-5, atol=1e-6)


def test_batch_rejects_mismatched_options(gpu_ctx):
    from dort_amd import gpu
    (s1, o1), (s2, o2) = _mixed_scenes()[:2]
    o2.max_depth = 3
    with pytest.raises(gpu.DgpuError):
        gpu_ctx.render_batch([(gpu_ctx.scene(s1), o1), (gpu_ctx.scene(s2), o2)])


@pytest.mark.parametrize("count,tile", [(1, 500), (2, 500), (130, 500), (255, 500), (130, 0), (255, 0)])
def test_batch_chunk_segments(gpu_ctx, count, tile):
    """The kernels find a path's scene by a binary search over the chunk's segment starts:
    from 1 scene to the limit of 255, with 500-path chunks that end inside scenes and
    with chunks of the most segments one chunk holds (64; 255 scenes make chunks of 64,
    64, 64 and 63), every film equals the scene's own render."""
    base = scenes.cornell()
    uniq = [gpu_ctx.scene(base, queue_tile=tile), gpu_ctx.scene(scenes.dort_test_box("area"), queue_tile=tile)]
    items = []
    for k in range(count):
        o = abi.render_opts(6 + k % 5, 4 + k % 3, spp=1 + k % 2, seed=1000 + k, max_depth=4, rr_depth=2,
                            scene_id=k)
        items.append((uniq[k % 2], o))
    films, st = gpu_ctx.render_batch(items)
    assert st["paths"] == sum(o.x_res * o.y_res * (o.sample_end - o.sample_begin) for _, o in items)
    for k in (0, count // 2, count - 1):
        gsc, o = items[k]
        ref_rgb, ref_w, _ = gsc.render(o)
        rgb, w = films[k]
        assert np.array_equal(w, ref_w)
        np.testing.assert_allclose(rgb, ref_rgb, rtol=1e-5, atol=1e-6)


def test_batch_device_films_must_not_overlap(gpu_ctx):
    """The batched finalize adds to film pixels without atomics, so device films of two
    items (or one item's colour and weight arrays) that share memory are rejected; host
    films are staged per item and may be shared (the film receives both renders)."""
    from dort_amd import gpu
    (s1, o1), (s2, o2) = _mixed_scenes()[:2]
    o2.x_res, o2.y_res = o1.x_res, o1.y_res
    g1, g2 = gpu_ctx.scene(s1), gpu_ctx.scene(s2)
    npix = o1.x_res * o1.y_res
    f = gpu.DeviceArray(gpu_ctx, npix * 3, np.float32)
    w = gpu.DeviceArray(gpu_ctx, npix, np.float32)
    f2 = gpu.DeviceArray(gpu_ctx, npix * 3, np.float32)
    w2 = gpu.DeviceArray(gpu_ctx, npix, np.float32)
    for films in ([(f, w), (f, w)],             # two items, one film
                  [(f, w), (f2, w)],            # shared weight array
                  [(f, w), (f2, f)]):           # weight array inside another item's colours
        with pytest.raises(gpu.DgpuError) as e:
            gpu_ctx.render_batch([(g1, o1), (g2, o2)], films=films, device=True)
        assert e.value.status == abi.DGPU_EINVAL
        assert "overlap" in str(e.value)
    # separate device films are accepted and equal the per-scene renders
    for a in (f, w, f2, w2):
        a.copy_from(np.zeros(a.shape, np.float32))
    gpu_ctx.render_batch([(g1, o1), (g2, o2)], films=[(f, w), (f2, w2)], device=True)
    for (g, o), (a, b) in zip([(g1, o1), (g2, o2)], [(f, w), (f2, w2)]):
        ref_rgb, ref_w, _ = g.render(o)
        assert np.array_equal(b.numpy().reshape(ref_w.shape), ref_w)
        np.testing.assert_allclose(a.numpy().reshape(ref_rgb.shape), ref_rgb, rtol=1e-5, atol=1e-6)
    # host films may be shared: the film holds the sum of both renders
    rgb = np.zeros((o1.y_res, o1.x_res, 3), np.float32)
    wh = np.zeros((o1.y_res, o1.x_res), np.float32)
    gpu_ctx.render_batch([(g1, o1), (g2, o2)], films=[(rgb, wh), (rgb, wh)])
    r1, w1, _ = g1.render(o1)
    r2, w2_, _ = g2.render(o2)
    assert np.array_equal(wh, w1 + w2_)
    np.testing.assert_allclose(rgb, r1 + r2, rtol=1e-5, atol=1e-6)
    for a in (f, w, f2, w2):
        a.free()


@pytest.mark.timeout(900)
def test_batch_64_scene_workload(gpu_ctx):
    """The batch workload at its own size: 64 scenes of 1k-500k triangles (5.2M in all),
    256^2 / 512^2 / 768^2, 16 spp, max depth 6, RR from depth 3, each under a 1024x512
    sky, rendered as one batch (331M paths in three chunks).  Every scene's film equals
    its own dgpu_render: weights bitwise, radiance to rtol 1e-5."""
    batch = scenes.batch_scenes(64, seed=3)
    items = []
    for k, (sc, res) in enumerate(batch):
        o = abi.render_opts(res, res, spp=16, seed=500 + k, max_depth=6, rr_depth=3, scene_id=k)
        items.append((gpu_ctx.scene(sc), o))
    print(f"64 scenes, {sum(sc.num_triangles for sc, _ in batch)} triangles", flush=True)
    films, st = gpu_ctx.render_batch(items)
    assert st["paths"] == sum(o.x_res * o.y_res * 16 for _, o in items)
    bad = []
    for k, ((gsc, o), (rgb, w)) in enumerate(zip(items, films)):
        ref_rgb, ref_w, _ = gsc.render(o)
        w_ok = np.array_equal(w, ref_w)
        far = ~np.isclose(rgb, ref_rgb, rtol=1e-5, atol=1e-6)
        if not w_ok or far.any():
            bad.append((k, w_ok, int(far.any(axis=-1).sum()), float(np.abs(rgb - ref_rgb).max())))
        gsc.close()
    print(f"scenes differing from their own render: {bad}", flush=True)
    assert not bad
