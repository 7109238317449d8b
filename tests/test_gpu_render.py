"""GPU path-tracing parity against the oracle and dort's own reference images.

* Counter-RNG parity: the oracle can draw the same counter-based random numbers as
  the GPU, so both trace the same paths; images agree pixel by pixel except where a
  path diverges on an FP tie (traversal edge, RR threshold).
* dort's fixtures: test/ref/box_diff_area_pin_d{0..3}.hdr and box_diff_area_pin.hdr
  (dort `pt`, 512x512, 500 iterations) judged by dort's own test_convergence.
"""
import os

import numpy as np
import pytest

from dort_amd import abi, scenes
from oracle import convergence, oracle, rgbe
from tests.test_oracle import golden_ok

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _pixel_parity(img_g, img_c, rel=1e-3, abs_=1e-4):
    close = np.abs(img_g - img_c) <= rel * np.abs(img_c) + abs_
    return float(close.all(axis=-1).mean())


def _paired_z(img_g, img_c):
    """Largest |mean / standard error| over the channels of the difference image, taken
    as zero where the pixels agree to the parity tolerance.  With the same counter-based
    random numbers the two renders trace the same paths, so the difference is zero
    except at pixels where a path diverged on an FP tie; for an unbiased GPU path those
    differences have zero mean, while an error in one branch (a delta lobe, total
    internal reflection) shifts them all one way."""
    d = (img_g.astype(np.float64) - img_c.astype(np.float64)).reshape(-1, 3)
    # pixels within rounding of each other hold the same paths: their last-bit
    # differences are not samples of the divergence and are left out
    same = np.abs(d) <= 1e-3 * np.abs(img_c.astype(np.float64).reshape(-1, 3)) + 1e-4
    d = np.where(same.all(axis=1, keepdims=True), 0.0, d)
    se = d.std(axis=0) / np.sqrt(len(d))
    z = np.where(se > 0, np.abs(d.mean(axis=0)) / np.where(se > 0, se, 1.0), 0.0)
    return float(z.max())


@pytest.mark.parametrize("strategy", [abi.DIRECT_MIS, abi.DIRECT_LIGHT, abi.DIRECT_BSDF])
def test_cornell_counter_rng_parity(gpu_ctx, strategy):
    sc = scenes.cornell()
    opts = abi.render_opts(64, 64, spp=16, seed=7, max_depth=5, direct_strategy=strategy)
    ref = oracle.OracleScene(sc).image(opts, rng="counter")
    img = gpu_ctx.scene(sc).image(opts)
    frac = _pixel_parity(img, ref)
    rel_mean = abs(img.mean() - ref.mean()) / ref.mean()
    z = _paired_z(img, ref)
    print(f"strategy {strategy}: {frac:.4f} pixels match, mean rel diff {rel_mean:.2e}, paired z {z:.2f}")
    assert frac >= 0.99
    assert z < 5.0
    assert rel_mean < 2e-3


def test_terrain_counter_rng_parity(gpu_ctx):
    sc = scenes.terrain_scene(quads=60, seed=4, sky_res=(128, 64))
    opts = abi.render_opts(48, 48, spp=8, seed=3, max_depth=8, rr_depth=3)
    ref = oracle.OracleScene(sc).image(opts, rng="counter")
    img = gpu_ctx.scene(sc).image(opts)
    frac = _pixel_parity(img, ref)
    rel_mean = abs(img.mean() - ref.mean()) / ref.mean()
    z = _paired_z(img, ref)
    print(f"terrain: {frac:.4f} pixels match, mean rel diff {rel_mean:.2e}, paired z {z:.2f}")
    assert frac >= 0.99
    assert z < 5.0
    assert rel_mean < 5e-3


# dort's pt options on its test box (two light triangles): sample_all_lights runs the
# shade kernel that estimates every light at every vertex (two NEE slots per light),
# only_direct restricts continuation to delta lobes (path_renderer.cpp:82-90, 113-132)
OPTION_CASES = [
    ("sample_all_lights", {"sample_all_lights": 1}),
    ("only_direct", {"only_direct": 1}),
    ("sample_all_lights+only_direct", {"sample_all_lights": 1, "only_direct": 1}),
    ("sample_all_lights+light_strategy", {"sample_all_lights": 1, "direct_strategy": abi.DIRECT_LIGHT}),
    ("min_depth_2", {"min_depth": 2}),
]


@pytest.mark.parametrize("name,kw", OPTION_CASES, ids=[c[0] for c in OPTION_CASES])
def test_test_box_option_parity(gpu_ctx, name, kw):
    sc = scenes.dort_test_box("area")
    opts = abi.render_opts(64, 64, spp=16, seed=21, max_depth=5, **kw)
    ref = oracle.OracleScene(sc).image(opts, rng="counter")
    img = gpu_ctx.scene(sc).image(opts)
    frac = _pixel_parity(img, ref)
    rel_mean = abs(img.mean() - ref.mean()) / ref.mean()
    z = _paired_z(img, ref)
    print(f"{name}: {frac:.4f} pixels match, mean rel diff {rel_mean:.2e}, paired z {z:.2f}")
    assert frac >= 0.99
    assert z < 5.0
    assert rel_mean < 2e-3


@pytest.mark.parametrize("max_depth,min_match", [(5, 0.96), (12, 0.93)])
def test_glass_cornell_counter_rng_parity(gpu_ctx, max_depth, min_match):
    """Dielectric-heavy Cornell box (the tall block is glass of IOR 1.5): delta
    reflection and refraction with Fresnel choice, total internal reflection and the
    emission-after-delta rule.  Chains of delta bounces amplify last-bit differences of
    the hit point (watertight test here, Moller-Trumbore in the oracle) and of the
    normals of coplanar triangles, so more paths diverge than in diffuse scenes: 97.2%
    of pixels match at depth 5 and 94.9% at depth 12 (MI355X), against 99.2-99.9% in
    the diffuse scenes; independent seeds move the image mean by 0.6%.  The paired test
    on the difference image is what pins the delta branches: diverged paths must not
    shift the image one way."""
    sc = scenes.glass_cornell()
    opts = abi.render_opts(64, 64, spp=16, seed=13, max_depth=max_depth)
    ref = oracle.OracleScene(sc).image(opts, rng="counter")
    img = gpu_ctx.scene(sc).image(opts)
    frac = _pixel_parity(img, ref)
    rel_mean = abs(img.mean() - ref.mean()) / ref.mean()
    z = _paired_z(img, ref)
    print(f"glass cornell depth {max_depth}: {frac:.4f} pixels match, mean rel diff {rel_mean:.2e}, "
          f"paired z {z:.2f}")
    assert frac >= min_match
    assert z < 4.0
    assert rel_mean < 5e-3


@pytest.mark.parametrize("name,min_depth,max_depth", [
    ("box_diff_area_pin_d0", 0, 0), ("box_diff_area_pin_d1", 1, 1), ("box_diff_area_pin_d2", 2, 2),
    ("box_diff_area_pin_d3", 3, 3), ("box_diff_area_pin", 0, 3)])
def test_dort_reference_images(gpu_ctx, name, min_depth, max_depth):
    ref = rgbe.read_hdr(os.path.join(GOLDEN, name + ".hdr"))
    sc = scenes.dort_test_box("area")
    opts = abi.render_opts(512, 512, spp=64, seed=11, min_depth=min_depth, max_depth=max_depth)
    img = gpu_ctx.scene(sc).image(opts)
    print(name, img.reshape(-1, 3).mean(0), ref.reshape(-1, 3).mean(0))
    ok, msg = golden_ok(ref, img)
    assert ok, msg


def test_deterministic_and_progressive(gpu_ctx):
    sc = scenes.cornell()
    s = gpu_ctx.scene(sc)
    opts = abi.render_opts(32, 32, spp=8, seed=5)
    a, wa, _ = s.render(opts)
    b, wb, _ = s.render(opts)
    assert np.array_equal(a, b) and np.array_equal(wa, wb)
    # [0,4) + [4,8) into the same film equals [0,8) up to summation order
    c, wc, _ = s.render(abi.render_opts(32, 32, spp=4, sample_begin=0, seed=5))
    c, wc, _ = s.render(abi.render_opts(32, 32, spp=4, sample_begin=4, seed=5), c, wc)
    assert np.array_equal(wc, wa)
    assert np.allclose(c, a, rtol=1e-5, atol=1e-6)
    # a different chunk size (queue tile) gives the same image
    s2 = gpu_ctx.scene(sc, queue_tile=1000)
    d, wd, _ = s2.render(opts)
    assert np.array_equal(wd, wa)
    assert np.allclose(d, a, rtol=1e-5, atol=1e-6)


def test_furnace_convex_sphere(gpu_ctx):
    """A convex Lambert sphere of albedo a under a uniform infinite light L = 1 reflects
    exactly a*L; the background is L."""
    p, t = scenes.displaced_sphere(24, seed=1, amp=0.0)
    b = scenes._Builder()
    a = 0.6
    m = b.material(scenes.lambert((a, a, a)))
    b.mesh(p, t, m)
    sc = b.build(scenes.look_at((0, 0, -4), (0, 0, 0), (0, 1, 0)), 0.6, infinite=(1.0, 1.0, 1.0))
    img = gpu_ctx.scene(sc).image(abi.render_opts(64, 64, spp=256, seed=2, max_depth=1))
    ys, xs = np.mgrid[0:64, 0:64]
    r = np.hypot(xs + 0.5 - 32, ys + 0.5 - 32)
    inside = r < 20
    outside = r > 29
    assert abs(img[inside].mean() - a) < 0.01 * a
    assert np.allclose(img[outside], 1.0)


def test_stats_and_errors(gpu_ctx):
    sc = scenes.cornell()
    s = gpu_ctx.scene(sc)
    _, _, st = s.render(abi.render_opts(16, 16, spp=2))
    assert st["paths"] == 16 * 16 * 2
    assert st["extension_rays"] >= st["paths"]
    assert st["dropped_samples"] == 0
    from dort_amd import gpu
    with pytest.raises(gpu.DgpuError):
        s.render(abi.render_opts(16, 16, spp=1, max_depth=100))
    cancel = np.ones(1, np.int32)
    with pytest.raises(gpu.DgpuError) as e:
        s.render(abi.render_opts(16, 16, spp=1), cancel=cancel)
    assert e.value.status == abi.DGPU_ECANCELLED


@pytest.mark.parametrize("texels", ["u8", "f32"])
def test_textured_cornell_counter_rng_parity(gpu_ctx, texels):
    """Image textures as dort composes them (ImageTexture over uv_texture_map on the floor,
    over xy_texture_map on the back wall; 8-bit texels or their float values k / 256):
    the oracle's lookup is pinned path for path against dort's own classes, and the GPU
    must pick the same texels."""
    sc = scenes.textured_cornell(texels=texels)
    opts = abi.render_opts(64, 64, spp=16, seed=17, max_depth=5)
    ref = oracle.OracleScene(sc).image(opts, rng="counter")
    img = gpu_ctx.scene(sc).image(opts)
    frac = _pixel_parity(img, ref)
    rel_mean = abs(img.mean() - ref.mean()) / ref.mean()
    z = _paired_z(img, ref)
    print(f"textured cornell ({texels}): {frac:.4f} pixels match, mean rel diff {rel_mean:.2e}, paired z {z:.2f}")
    assert frac >= 0.99
    assert z < 5.0
    assert rel_mean < 2e-3


# the Cornell cases' short block is smooth glass, whose delta chains diverge on FP ties as
# in the glass Cornell test (96.2% and 95.0% of pixels match, MI355X); the pane has none
ROUGH_CASES = [("rough_glass_cornell_0.2", lambda: scenes.rough_glass_cornell(0.2), 5, 0.95),
               ("rough_glass_cornell_0.6", lambda: scenes.rough_glass_cornell(0.6), 10, 0.94),
               ("rough_pane", scenes.rough_pane, 5, 0.99)]


@pytest.mark.parametrize("name,make,max_depth,min_match", ROUGH_CASES, ids=[c[0] for c in ROUGH_CASES])
def test_rough_dielectric_counter_rng_parity(gpu_ctx, name, make, max_depth, min_match):
    """dort's rough dielectric as dort runs it (reflection sampled with the Fresnel
    probability, transmission reached by light sampling, a zero pdf in the MIS weights):
    the oracle is pinned path for path against dort's own RoughDielectricBxdf, and the
    GPU must trace the same paths."""
    sc = make()
    opts = abi.render_opts(64, 64, spp=16, seed=23, max_depth=max_depth)
    ref = oracle.OracleScene(sc).image(opts, rng="counter")
    img = gpu_ctx.scene(sc).image(opts)
    frac = _pixel_parity(img, ref)
    rel_mean = abs(img.mean() - ref.mean()) / ref.mean()
    z = _paired_z(img, ref)
    print(f"{name}: {frac:.4f} pixels match, mean rel diff {rel_mean:.2e}, paired z {z:.2f}")
    assert frac >= min_match
    assert z < 5.0
    assert rel_mean < 5e-3
