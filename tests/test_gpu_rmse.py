"""Relative MSE at equal samples per pixel: the GPU path against dort on the CPU, on
the Cornell box and on the terrain benchmark scene itself.

The reference is dort's own render of the Cornell scene at 65536 spp
(tests/golden/cornell_64x48_65536spp_dort.npy, written by
tests/golden/make_cornell_reference.py).  relMSE = mean over pixels and channels of
(x - r)^2 / (r^2 + 1e-2).  Both sides render 64x48 at 64 spp, max depth 5, MIS,
RR from bounce 2, over 64 seeds each; the GPU's mean relMSE must be within 5% of
the CPU's.  The CPU side is the oracle's mt19937 mode on all host cores: it
follows dort path for path (tests/test_oracle_ref.py), and dort itself runs
single-threaded only (its thread pool has a race).
"""
import os

import numpy as np
import pytest

from dort_amd import abi, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell_64x48_65536spp_dort.npy")
SEEDS = 64


def rel_mse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


def cpu_image(osc, opts):
    return osc.image(opts, rng="mt", threads=min(16, os.cpu_count() or 1))


def test_relmse_matches_cpu_at_equal_spp(gpu_ctx):
    r = np.load(REF)
    sc = scenes.cornell()
    gsc = gpu_ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    g, c, gsum = [], [], np.zeros_like(r)
    for s in range(SEEDS):
        opts = abi.render_opts(64, 48, spp=64, seed=500 + s, max_depth=5, rr_depth=2)
        img = gsc.image(opts)
        gsum += img
        g.append(rel_mse(img, r))
        c.append(rel_mse(cpu_image(osc, opts), r))
    g, c = np.array(g), np.array(c)
    ratio = g.mean() / c.mean()
    print(f"relMSE at 64 spp: GPU {g.mean():.5f} +- {g.std() / np.sqrt(SEEDS):.5f}, "
          f"CPU {c.mean():.5f} +- {c.std() / np.sqrt(SEEDS):.5f}, ratio {ratio:.4f}")
    assert ratio <= 1.05
    # no bias: the mean of the 64 GPU renders (4096 spp in all) is close to the reference
    bias = rel_mse(gsum / SEEDS, r)
    assert bias < 0.1 * g.mean()


TERRAIN_REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "terrain_64x64_65536spp_cpu.npy")
TERRAIN_SEEDS = 48


@pytest.mark.timeout(600)
def test_terrain_relmse_matches_cpu_at_equal_spp(gpu_ctx):
    """The benchmark scene (2,000,002 triangles, 1024x512 sky, Lambert / GGX conductor /
    dielectric objects) at 64x64, 64 spp, max depth 8, RR from depth 3, against its
    65536-spp CPU render (tests/golden/terrain_64x64_65536spp_cpu.npy, written by
    tests/golden/make_terrain_reference.py with the oracle in dort's mt19937 mode): over
    48 seeds the GPU's mean relMSE is within 5% of the CPU's at the same spp."""
    r = np.load(TERRAIN_REF)
    sc = scenes.terrain_scene(quads=1000, seed=2)
    gsc = gpu_ctx.scene(sc)
    osc = oracle.OracleScene(sc)
    g, c, gsum = [], [], np.zeros_like(r)
    for s in range(TERRAIN_SEEDS):
        opts = abi.render_opts(64, 64, spp=64, seed=700 + s, max_depth=8, rr_depth=3)
        img = gsc.image(opts)
        gsum += img
        g.append(rel_mse(img, r))
        c.append(rel_mse(cpu_image(osc, opts), r))
    g, c = np.array(g), np.array(c)
    ratio = g.mean() / c.mean()
    print(f"terrain relMSE at 64 spp: GPU {g.mean():.5f} +- {g.std() / np.sqrt(TERRAIN_SEEDS):.5f}, "
          f"CPU {c.mean():.5f} +- {c.std() / np.sqrt(TERRAIN_SEEDS):.5f}, ratio {ratio:.4f}")
    assert ratio <= 1.05
    bias = rel_mse(gsum / TERRAIN_SEEDS, r)
    assert bias < 0.1 * g.mean()
