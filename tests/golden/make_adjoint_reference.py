"""Writes tests/golden/adjoint_terrain200k_64x64_4spp_fp64.npz: the FP64 reference
gradient (the oracle's double-precision path-replay adjoint) of the inverse-rendering
workload's scene rendered small -- the 199,714-triangle textured terrain under a
1024 x 512 sky, albedo texture 256 x 256 at 0.5, eight GGX roughness unknowns at 0.25 --
at 64 x 64, 4 spp, max depth 6, RR from depth 3, seed 77, for the loss
L = sum(dLdI * image) with the seeded dLdI stored in the file.  bench.py measures the
GPU adjoint's and central differences' errors against it.

    python tests/golden/make_adjoint_reference.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from dort_amd import abi, scenes  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.join(HERE, "adjoint_terrain200k_64x64_4spp_fp64.npz")


def scene():
    sc = scenes.terrain_scene(quads=316, seed=8, textured=True, tex_res=256, sky_res=(1024, 512),
                              ggx_roughness=np.full(8, 0.25))
    sc.textures[0] = np.full_like(sc.textures[0], 0.5)
    return sc


def opts():
    return abi.render_opts(64, 64, spp=4, seed=77, max_depth=6, rr_depth=3)


def dldi():
    return np.random.default_rng(2024).uniform(-1.0, 1.0, (64, 64, 3)).astype(np.float32)


def main():
    sc = scene()
    o = oracle.OracleScene(sc, fp64=True)
    tex = sc.textures[0]
    g_tex, g_rough, _ = o.adjoint(opts(), dldi(), tex.shape[0] * tex.shape[1], len(sc.materials))
    np.savez_compressed(OUT, g_tex=g_tex.reshape(tex.shape).astype(np.float64),
                        g_rough=np.asarray(g_rough, np.float64), dldi=dldi())
    print(f"wrote {OUT}: {np.count_nonzero(g_tex)} nonzero texel entries, roughness {g_rough}")


if __name__ == "__main__":
    main()
