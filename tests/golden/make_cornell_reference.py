"""Writes tests/golden/cornell_64x48_65536spp_dort.npy: dort's own `pt` render of the
Cornell scene (dort_amd.scenes.cornell) at 64x48, 65536 iterations, max depth 5, MIS,
RR from bounce 2, sampler seeds 1000..1015 -- the converged reference that the relative-MSE
test compares GPU and CPU renders against.

dort is oracle/_ref/libdortref.so (its sources compiled in place by oracle/ref.mk).
Run from the repository root:  python tests/golden/make_cornell_reference.py
"""
import multiprocessing
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dort_amd import abi, scenes  # noqa: E402
from oracle import ref  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "cornell_64x48_65536spp_dort.npy")


def batch(k):
    # dort runs single-threaded (its thread pool has a race); batches run in parallel
    # processes instead
    opts = abi.render_opts(64, 48, spp=4096, seed=1000 + k, max_depth=5, rr_depth=2)
    rgb, w = ref.RefScene(scenes.cornell()).render(opts, threads=1)
    return rgb.astype(np.float64), w.astype(np.float64)


def main():
    if not ref.available():
        ref.build()
    total = np.zeros((48, 64, 3), np.float64)
    weight = np.zeros((48, 64), np.float64)
    t0 = time.time()
    # 16 renders of 4096 iterations with different sampler seeds (dort's tiles
    # draw their streams from the seed), summed
    with multiprocessing.Pool(min(16, os.cpu_count() or 1)) as pool:
        for k, (rgb, w) in enumerate(pool.imap(batch, range(16))):
            total += rgb
            weight += w
            print(f"batch {k}: {time.time() - t0:.0f}s", flush=True)
    img = (total / weight[..., None]).astype(np.float32)
    np.save(OUT, img)
    print(f"wrote {OUT}: mean {img.reshape(-1, 3).mean(0)}, weight per pixel {weight[0, 0]:.0f}")


if __name__ == "__main__":
    main()
