"""GPU intersection parity: dgpu_trace_closest / dgpu_trace_any against the oracle.

The oracle traverses dort's own BVH (middle split, 6-triangle leaves) with dort's
Moller-Trumbore test; the GPU uses its SAH BVH and the watertight test.  Hit ids
must agree on >= 99.99% of rays, mismatches only at edge or t ties, and hit
distances within 1e-5 relative.
"""
import numpy as np
import pytest

from dort_amd import scenes
from oracle import oracle
from tests.helpers import classify_hits

pytestmark = pytest.mark.gpu


def _workload(scene, res=(256, 128), n_random=32768, seed=5):
    return scenes.intersection_workload(scene, res[0], res[1], n_random, seed)


@pytest.fixture(scope="module")
def sphere():
    return scenes.icosphere_scene(freq=32, seed=11)  # 20,480 triangles


@pytest.mark.parametrize("width", [2, 4])
def test_closest_hit_parity_icosphere(gpu_ctx, sphere, width):
    rays = _workload(sphere)
    ref = oracle.OracleScene(sphere).trace(rays)
    g = gpu_ctx.scene(sphere, width=width).trace_closest(rays)
    c = classify_hits(ref, g)
    print(f"width {width}: {c['agree']}/{c['n']} agree, {c['explained']} tie/edge, {c['fail']} fail,"
          f" max rel t err {c['max_rel_t']:.2e}")
    assert c["agree"] >= 0.9999 * c["n"]
    assert c["fail"] == 0
    assert c["max_rel_t"] <= 1e-5


@pytest.mark.parametrize("width", [2, 4])
def test_any_hit_parity_icosphere(gpu_ctx, sphere, width):
    rays = _workload(sphere)
    # finite segments: half of each closest-hit distance, and twice
    ref_c = oracle.OracleScene(sphere).trace(rays)
    t = np.where(np.isfinite(ref_c["t"]), ref_c["t"], 10.0).astype(np.float32)
    osc = oracle.OracleScene(sphere)
    gsc = gpu_ctx.scene(sphere, width=width)
    for scale in (0.5, 2.0):
        r = dict(rays)
        r["tmax"] = (t * scale).astype(np.float32)
        ref = osc.trace(r, any_hit=True)
        got, _ = gsc.trace_any(r)
        diff = np.nonzero(ref != got)[0]
        # disagreements only where the segment end grazes the closest hit
        assert len(diff) <= 1e-4 * len(ref), (scale, len(diff))


def test_brute_force_small(gpu_ctx):
    sc = scenes.cornell()
    rays = _workload(sc, res=(64, 64), n_random=4096)
    osc = oracle.OracleScene(sc)
    ref = osc.trace(rays, brute=True)
    g = gpu_ctx.scene(sc).trace_closest(rays)
    c = classify_hits(ref, g)
    # the Cornell box is all quads split on a diagonal: pixel-centre rays on its symmetry
    # lines hit shared edges exactly, where either triangle is a valid answer
    assert c["fail"] == 0, c
    assert c["agree"] >= 0.99 * c["n"]


def test_watertight_closed_mesh(gpu_ctx):
    """Rays from inside a closed star-shaped mesh aimed at vertices and edge midpoints
    must all hit (no leaks through shared edges)."""
    p, t = scenes.displaced_sphere(16, seed=3, amp=0.05)
    b = scenes._Builder()
    m = b.material(scenes.lambert((0.5, 0.5, 0.5)))
    b.mesh(p, t, m)
    sc = b.build(scenes.look_at((0, 0, -3), (0, 0, 0), (0, 1, 0)), 0.8)
    targets = [p.astype(np.float64)]
    tt = t.astype(np.int64)
    for a, c in ((0, 1), (1, 2), (2, 0)):
        targets.append(0.5 * (p[tt[:, a]].astype(np.float64) + p[tt[:, c]].astype(np.float64)))
    tg = np.concatenate(targets)
    o = np.zeros_like(tg) + np.array([0.01, -0.02, 0.015])
    d = tg - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = scenes.ray_batch(o.astype(np.float32), d.astype(np.float32))
    g = gpu_ctx.scene(sc).trace_closest(rays)
    assert (g["prim"] >= 0).all(), int((g["prim"] < 0).sum())
    occ, _ = gpu_ctx.scene(sc).trace_any(rays)
    assert occ.all()


def test_device_buffers(gpu_ctx, sphere):
    """Rays and hits resident in HBM (DGPU_BUF_DEVICE) give the same hits as host buffers."""
    rays = _workload(sphere, res=(64, 32), n_random=2048)
    sc = gpu_ctx.scene(sphere)
    host = sc.trace_closest(rays)
    dev_rays = {k: gpu_ctx.device_array(v) for k, v in rays.items()}
    n = len(rays["ox"])
    out = {"t": gpu_ctx.device_array(np.zeros(n, np.float32)),
           "prim": gpu_ctx.device_array(np.zeros(n, np.int32)),
           "b1": gpu_ctx.device_array(np.zeros(n, np.float32)),
           "b2": gpu_ctx.device_array(np.zeros(n, np.float32))}
    sc.trace_closest(dev_rays, out=out, device=True)
    assert np.array_equal(out["prim"].numpy(), host["prim"])
    assert np.array_equal(out["t"].numpy(), host["t"])
    occ_h, _ = sc.trace_any(rays)
    occ_d = gpu_ctx.device_array(np.zeros(n, np.uint32))
    sc.trace_any(dev_rays, out=occ_d, device=True)
    assert np.array_equal(occ_d.numpy(), occ_h)
