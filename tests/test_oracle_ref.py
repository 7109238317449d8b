"""Pins the oracle against dort itself.

oracle/_ref/libdortref.so is dort's own pt-path sources (compiled where they lie
in the reference checkout by oracle/ref.mk, with dort's clang++ build) driven
through oracle/ref_driver.cpp.  Given the same scene and seed, dort's
PathRenderer with its mt19937 RandomSampler and the oracle's "mt" mode draw the
same random numbers in the same order, so every path is the same path: the
images agree pixel by pixel up to float rounding, and hit records agree bit for
bit.  The tests skip where the library has not been built (no checkout).
"""
import dataclasses
import os

import numpy as np
import pytest

from dort_amd import abi, scenes
from oracle import oracle, ref, rgbe

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (no reference checkout)")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dortable(sc):
    """dort has no rough conductor: GGX conductors become Lambert."""
    mats = [scenes.lambert((0.6, 0.5, 0.4)) if m.type == abi.MAT_GGX_CONDUCTOR else m for m in sc.materials]
    return dataclasses.replace(sc, materials=mats)


def terrain():
    return dortable(scenes.terrain_scene(quads=12, sky_res=(64, 32)))


SCENES = {
    "cornell": scenes.cornell,
    "box_area": lambda: scenes.dort_test_box("area"),
    "box_infi": lambda: scenes.dort_test_box("infi"),
    "glass_cornell": scenes.glass_cornell,
    "terrain_sky": terrain,
    "textured_cornell": scenes.textured_cornell,
    "rough_glass_cornell": scenes.rough_glass_cornell,
    "rough_glass_cornell_0.6": lambda: scenes.rough_glass_cornell(0.6),
    "rough_pane": scenes.rough_pane,
    "textured_cornell_f32": lambda: scenes.textured_cornell(texels="f32"),
}


@pytest.mark.parametrize("split", ["middle", "sah"])
@pytest.mark.parametrize("name", ["cornell", "box_area", "terrain_sky"])
def test_trace_matches_dort_bitwise(name, split):
    sc = SCENES[name]()
    rays = scenes.intersection_workload(sc, 48, 48, 20000, seed=5)
    a = ref.RefScene(sc, split=split).trace(rays)
    b = oracle.OracleScene(sc, split=split).trace(rays, threads=1)
    assert (a["prim"] >= 0).mean() > 0.1
    np.testing.assert_array_equal(a["prim"], b["prim"])
    np.testing.assert_array_equal(a["t"], b["t"])
    occ_a = ref.RefScene(sc, split=split).trace(rays, any_hit=True)
    occ_b = oracle.OracleScene(sc, split=split).trace(rays, any_hit=True, threads=1)
    np.testing.assert_array_equal(occ_a, occ_b)


RENDER_CASES = [
    ("cornell", {}),
    ("box_area", {}),
    ("box_infi", {}),
    ("glass_cornell", {}),
    ("terrain_sky", {}),
    ("terrain_sky", {"direct_strategy": abi.DIRECT_BSDF}),
    ("terrain_sky", {"direct_strategy": abi.DIRECT_LIGHT}),
    ("box_area", {"sample_all_lights": 1}),
    ("box_area", {"min_depth": 2}),
    ("box_area", {"only_direct": 1}),
    ("glass_cornell", {"max_depth": 12}),
    ("textured_cornell", {}),
    ("textured_cornell_f32", {}),
    ("rough_glass_cornell", {}),
    ("rough_glass_cornell", {"direct_strategy": abi.DIRECT_BSDF}),
    ("rough_glass_cornell", {"direct_strategy": abi.DIRECT_LIGHT}),
    ("rough_glass_cornell_0.6", {"max_depth": 10}),
    ("rough_pane", {}),
    ("rough_pane", {"direct_strategy": abi.DIRECT_LIGHT}),
]


@pytest.mark.parametrize("name,kw", RENDER_CASES, ids=[f"{n}-{'-'.join(k) or 'default'}" for n, k in RENDER_CASES])
def test_render_matches_dort_path_for_path(name, kw):
    sc = SCENES[name]()
    opts = abi.render_opts(32, 24, spp=6, seed=11, max_depth=6)
    for k, v in kw.items():
        setattr(opts, k, v)
    ra, wa = ref.RefScene(sc).render(opts, threads=1)
    rb, wb, _ = oracle.OracleScene(sc).render(opts, rng="mt", threads=1)
    np.testing.assert_array_equal(wa, wb)
    assert ra.sum() > 0
    # same paths: per-pixel agreement to float rounding (a different random
    # draw order or a dropped term moves whole pixels by O(1))
    np.testing.assert_allclose(rb, ra, rtol=2e-4, atol=1e-5)


def test_render_detects_planted_mistake():
    """The comparison above is sharp: 1% more albedo on one wall fails it."""
    sc = scenes.cornell()
    bad = dataclasses.replace(sc, materials=list(sc.materials))
    m = abi.Material.from_buffer_copy(bad.materials[1])
    m.albedo = abi.f3(*(np.array(m.albedo[:]) * 1.01))
    bad.materials[1] = m
    opts = abi.render_opts(32, 24, spp=6, seed=11, max_depth=6)
    ra, _ = ref.RefScene(sc).render(opts, threads=1)
    rb, _, _ = oracle.OracleScene(bad).render(opts, rng="mt", threads=1)
    assert not np.allclose(rb, ra, rtol=2e-4, atol=1e-5)


def test_dort_itself_passes_its_golden_render():
    """dort rendering its own test box meets the acceptance rule the oracle and the
    GPU path are held to (pins the scene restatement and the rule together)."""
    from tests.test_oracle import golden_ok
    gold = rgbe.read_hdr(os.path.join(GOLDEN, "box_diff_area_pin.hdr"))
    opts = abi.render_opts(512, 512, spp=16, seed=42, min_depth=0, max_depth=3)
    rgb, w = ref.RefScene(scenes.dort_test_box("area")).render(opts)
    ok, msg = golden_ok(gold, rgb / np.maximum(w, 1e-30)[..., None])
    assert ok, msg


def test_counter_rng_estimator_has_dorts_relmse():
    """The counter-based random numbers the GPU draws give the same error as dort's
    mt19937 streams: relMSE against dort's 65536-spp render, 48 seeds each (the
    relMSE of one render is heavy-tailed; 48 seeds put the ratio's noise near 2%)."""
    r = np.load(os.path.join(GOLDEN, "cornell_64x48_65536spp_dort.npy"))
    sc = scenes.cornell()
    o, d = oracle.OracleScene(sc), ref.RefScene(sc)
    a, b = [], []
    for s in range(48):
        opts = abi.render_opts(64, 48, spp=8, seed=300 + s, max_depth=5, rr_depth=2)
        a.append(np.mean((o.image(opts, rng="counter") - r) ** 2 / (r ** 2 + 1e-2)))
        rgb, w = d.render(opts)
        b.append(np.mean((rgb / w[..., None] - r) ** 2 / (r ** 2 + 1e-2)))
    ratio = np.mean(a) / np.mean(b)
    print(f"relMSE counter {np.mean(a):.5f}, dort {np.mean(b):.5f}, ratio {ratio:.4f}")
    assert 0.92 < ratio < 1.08


def test_textures_are_dorts_image_textures():
    """The textured case above is sharp for the texel lookup: shifting the floor texture by
    one texel (a wrong row or column offset) or flipping the xy map's rows fails it, and
    the 8-bit and float forms of the same texels render the same image."""
    sc = scenes.textured_cornell()
    opts = abi.render_opts(32, 24, spp=6, seed=11, max_depth=6)
    ra, _ = ref.RefScene(sc).render(opts, threads=1)
    f32, _, _ = oracle.OracleScene(scenes.textured_cornell(texels="f32")).render(opts, rng="mt", threads=1)
    u8, _, _ = oracle.OracleScene(sc).render(opts, rng="mt", threads=1)
    np.testing.assert_array_equal(f32, u8)
    shifted = dataclasses.replace(sc, textures=[np.roll(sc.textures[0], 1, axis=1), sc.textures[1]])
    rs, _, _ = oracle.OracleScene(shifted).render(opts, rng="mt", threads=1)
    assert not np.allclose(rs, ra, rtol=2e-4, atol=1e-5)
    back = dict(sc.textures[1])
    back["image"] = back["image"][::-1].copy()
    flipped = dataclasses.replace(sc, textures=[sc.textures[0], back])
    rf, _, _ = oracle.OracleScene(flipped).render(opts, rng="mt", threads=1)
    assert not np.allclose(rf, ra, rtol=2e-4, atol=1e-5)
