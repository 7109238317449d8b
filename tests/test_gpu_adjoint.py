"""Adjoint path replay (dgpu_adjoint): gradients of an image loss with respect to
albedo texels and GGX roughness, checked against
* the oracle's FP64 adjoint (same counter-based random numbers, so the same paths),
* central finite differences of the GPU forward renderer with common random numbers
  (h = 1e-3), and
* linearity in dL/dI.
"""
import numpy as np
import pytest

from dort_amd import abi, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

ROUGH = [0.1, 0.15, 0.2, 0.25, 0.3, 0.35, 0.4, 0.45]


def tiny_scene():
    # the config-4 scene shrunk: textured Lambert + 8 GGX patches, sky
    return scenes.terrain_scene(quads=24, seed=5, textured=True, tex_res=16, sky_res=(64, 32),
                                ggx_roughness=ROUGH)


def _dldi(res, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, (res, res, 3)).astype(np.float32)


@pytest.fixture(scope="module")
def scene():
    return tiny_scene()


def test_adjoint_matches_fp64_oracle(gpu_ctx, scene):
    opts = abi.render_opts(64, 64, spp=4, seed=9, max_depth=6, rr_depth=3)
    dldi = _dldi(64)
    g_tex, g_rough, _, _ = gpu_ctx.scene(scene).adjoint(opts, dldi)
    o = oracle.OracleScene(scene, fp64=True)
    tex = scene.textures[0]
    r_tex, r_rough, _ = o.adjoint(opts, dldi, tex.shape[0] * tex.shape[1], len(scene.materials))
    r_tex = r_tex.reshape(tex.shape)
    rel_tex = np.linalg.norm(g_tex - r_tex) / np.linalg.norm(r_tex)
    rel_rough = np.linalg.norm(g_rough - r_rough) / np.linalg.norm(r_rough)
    print(f"texel gradient rel L2 err {rel_tex:.2e}, roughness {rel_rough:.2e}")
    print("gpu", g_rough, "\nfp64", r_rough)
    assert np.linalg.norm(r_tex) > 0 and np.linalg.norm(r_rough) > 0
    assert rel_tex < 0.02
    assert rel_rough < 0.02


def _loss(sc, opts, dldi):
    img = sc.image(opts)
    return float((img.astype(np.float64) * dldi).sum())


def test_adjoint_matches_central_fd_texels(gpu_ctx, scene):
    """Without Russian roulette, the forward estimator is linear in each albedo, so central
    differences with common random numbers equal the replayed gradient up to FP error."""
    opts = abi.render_opts(48, 48, spp=8, seed=4, max_depth=6, rr_depth=100)
    dldi = _dldi(48, 1)
    sc = gpu_ctx.scene(scene)
    g_tex, _, _, _ = sc.adjoint(opts, dldi, want_roughness=False)
    tex = scene.textures[0].copy()
    flat = np.argsort(-np.abs(g_tex).ravel())[:6]
    h = 1e-3
    for k in flat:
        y, x, c = np.unravel_index(k, g_tex.shape)
        t = tex.copy(); t[y, x, c] += h
        sc.update_texture(0, t)
        lp = _loss(sc, opts, dldi)
        t = tex.copy(); t[y, x, c] -= h
        sc.update_texture(0, t)
        lm = _loss(sc, opts, dldi)
        fd = (lp - lm) / (2 * h)
        print(f"texel {(y, x, c)}: adjoint {g_tex[y, x, c]:.6f} fd {fd:.6f}")
        assert abs(fd - g_tex[y, x, c]) <= 0.02 * abs(g_tex[y, x, c]) + 1e-4
    sc.update_texture(0, tex)


def ggx_plane(rough):
    """A GGX conductor plane under a constant infinite light: nothing occludes and the
    light is constant, so the image is smooth in the roughness."""
    b = scenes._Builder()
    m = b.material(scenes.ggx_conductor(rough, "gold"))
    b.mesh([(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)], [[0, 1, 2], [0, 2, 3]], m)
    cam = scenes.look_at((0, 1.0, -1.6), (0, 0, 0), (0, 1, 0)) @ scenes.scale_m(1, -1, 1)
    return b.build(cam, 0.9, infinite=(1.0, 0.8, 0.6))


@pytest.mark.parametrize("rough", [0.1, 0.3, 0.6])
def test_adjoint_matches_central_fd_roughness(gpu_ctx, rough):
    """Roughness also moves the GGX sample directions, so central differences (common
    random numbers) and the replayed, detached-sampling gradient agree in expectation;
    on a smooth scene at many samples they agree closely."""
    opts = abi.render_opts(48, 48, spp=2048, seed=6, max_depth=2, rr_depth=100)
    dldi = np.abs(_dldi(48, 2))
    _, g, _, _ = gpu_ctx.scene(ggx_plane(rough)).adjoint(opts, dldi, want_texture=False)
    h = 1e-3
    lp = _loss(gpu_ctx.scene(ggx_plane(rough + h)), opts, dldi)
    lm = _loss(gpu_ctx.scene(ggx_plane(rough - h)), opts, dldi)
    fd = (lp - lm) / (2 * h)
    print(f"roughness {rough}: adjoint {g[0]:.4f} fd {fd:.4f}")
    assert abs(fd - g[0]) <= 0.03 * abs(fd)


def test_adjoint_linear_in_dldi(gpu_ctx, scene):
    opts = abi.render_opts(32, 32, spp=2, seed=3, max_depth=6, rr_depth=3)
    sc = gpu_ctx.scene(scene)
    d = _dldi(32, 5)
    a_t, a_r, img, _ = sc.adjoint(opts, d, want_image=True)
    b_t, b_r, _, _ = sc.adjoint(opts, 2 * d)
    np.testing.assert_allclose(b_t, 2 * a_t, rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(b_r, 2 * a_r, rtol=1e-4, atol=1e-7)
    # the replayed paths are the forward render's paths
    fwd = sc.image(opts)
    assert np.abs(img - fwd).max() <= 1e-3 * max(1.0, np.abs(fwd).max())


@pytest.mark.parametrize("sample_all", [0, 1])
def test_wavefront_adjoint_equals_megakernel(gpu_ctx, scene, sample_all):
    """The two wavefront passes and the one-thread-per-path kernel replay the same paths
    and make the same deposits (in another order, so sums agree to float rounding), and
    the wavefront's image is the per-pixel mean of the first pass."""
    opts = abi.render_opts(40, 40, spp=6, seed=21, max_depth=6, rr_depth=3)
    opts.sample_all_lights = sample_all
    sc = gpu_ctx.scene(scene)
    d = _dldi(40, 7)
    w_t, w_r, w_img, _ = sc.adjoint(opts, d, want_image=True)
    m_t, m_r, m_img, _ = sc.adjoint(opts, d, want_image=True, megakernel=True)
    print("texel rel", np.linalg.norm(w_t - m_t) / np.linalg.norm(m_t),
          "rough", np.abs(w_r - m_r).max() / np.abs(m_r).max())
    assert np.linalg.norm(m_t) > 0 and np.abs(m_r).max() > 0
    np.testing.assert_allclose(w_t, m_t, rtol=1e-3, atol=1e-6 * np.abs(m_t).max())
    np.testing.assert_allclose(w_r, m_r, rtol=1e-3, atol=1e-6 * np.abs(m_r).max())
    np.testing.assert_allclose(w_img, m_img, rtol=1e-4, atol=1e-6)
    fwd = sc.image(opts)
    np.testing.assert_allclose(w_img, fwd, rtol=1e-4, atol=1e-6)


def test_wavefront_adjoint_chunks(gpu_ctx, scene):
    """Gradients do not depend on how the paths are cut into wavefront chunks."""
    opts = abi.render_opts(24, 20, spp=5, seed=33, max_depth=6, rr_depth=3)
    d = _dldi(24, 8)[:20]
    a_t, a_r, _, _ = gpu_ctx.scene(scene).adjoint(opts, d)
    b_t, b_r, _, _ = gpu_ctx.scene(scene, queue_tile=777).adjoint(opts, d)
    np.testing.assert_allclose(b_t, a_t, rtol=1e-4, atol=1e-6 * np.abs(a_t).max())
    np.testing.assert_allclose(b_r, a_r, rtol=1e-4, atol=1e-6 * np.abs(a_r).max())
