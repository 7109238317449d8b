"""Writes tests/golden/terrain_64x64_65536spp_cpu.npy: the converged CPU render of the
terrain benchmark scene (dort_amd.scenes.terrain_scene(quads=1000, seed=2): 2,000,002
triangles, 1024 x 512 sky, Lambert / GGX conductor / dielectric objects) at 64 x 64,
65536 iterations, max depth 8, RR from depth 3, MIS -- the reference the relative-MSE
test compares GPU and CPU renders of the benchmark scene against.

The renderer is the oracle in dort's mt19937 mode: dort itself cannot express the
scene's GGX conductors, and the oracle follows dort path for path on everything dort
can express (tests/test_oracle_ref.py); the conductor is pinned by closed forms
(tests/test_oracle_pins.py).  Sixteen renders of 4096 iterations with sampler seeds
2000..2015, summed.  Run from the repository root (about 3 minutes on 8 cores):

    python tests/golden/make_terrain_reference.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dort_amd import abi, scenes  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "terrain_64x64_65536spp_cpu.npy")


def main():
    osc = oracle.OracleScene(scenes.terrain_scene(quads=1000, seed=2))
    total = np.zeros((64, 64, 3), np.float64)
    weight = np.zeros((64, 64), np.float64)
    t0 = time.time()
    for k in range(16):
        opts = abi.render_opts(64, 64, spp=4096, seed=2000 + k, max_depth=8, rr_depth=3)
        rgb, w, _ = osc.render(opts, rng="mt", threads=os.cpu_count() or 1)
        total += rgb
        weight += w
        print(f"batch {k}: {time.time() - t0:.0f}s", flush=True)
    img = (total / weight[..., None]).astype(np.float32)
    np.save(OUT, img)
    print(f"wrote {OUT}: mean {img.reshape(-1, 3).mean(0)}, weight per pixel {weight[0, 0]:.0f}")


if __name__ == "__main__":
    main()
