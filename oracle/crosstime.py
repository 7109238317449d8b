"""Times the oracle against dort itself, one thread each (TEST INFRASTRUCTURE ONLY).

The oracle (oracle/liboracle.so) stands in for dort as the timed CPU baseline where the
reference checkout is absent.  This script renders the same scenes with the same seeds
through dort's PathRenderer (oracle/_ref/libdortref.so, dort's sources built by
oracle/ref.mk) and through the oracle in dort's mt19937 mode, which traces the same
paths, and reports both render times (scene and BVH builds excluded, median of 3).  One
thread each: dort's thread pool is only safe with one (oracle/ref.py).

    python -m oracle.crosstime
"""
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from dort_amd import abi, scenes  # noqa: E402
from oracle import oracle, ref  # noqa: E402


def dortable(sc):
    """dort has no rough conductor: GGX conductors become Lambert (as in the parity tests)."""
    mats = [scenes.lambert((0.6, 0.5, 0.4)) if m.type == abi.MAT_GGX_CONDUCTOR else m for m in sc.materials]
    return dataclasses.replace(sc, materials=mats)


CASES = [
    ("Cornell box, 36 triangles, 256x256, 16 spp, max depth 5 (the CPU workload)",
     scenes.cornell, lambda: abi.render_opts(256, 256, spp=16, seed=5, max_depth=5)),
    ("terrain reduced to 80,002 triangles + 1024x512 sky, 128x128, 16 spp, max depth 8, RR from 3",
     lambda: dortable(scenes.terrain_scene(quads=200, seed=2)),
     lambda: abi.render_opts(128, 128, spp=16, seed=1234, max_depth=8, rr_depth=3)),
]


def _median_time(fn, runs=3):
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def main():
    if not ref.available():
        print("oracle/_ref/libdortref.so is not built (no reference checkout)")
        return 1
    rows = []
    for name, make_scene, make_opts in CASES:
        sc, opts = make_scene(), make_opts()
        d, o = ref.RefScene(sc), oracle.OracleScene(sc)
        paths = opts.x_res * opts.y_res * (opts.sample_end - opts.sample_begin)
        t_dort = _median_time(lambda: d.render(opts, threads=1))
        t_orc = _median_time(lambda: o.render(opts, rng="mt", threads=1))
        rows.append({"workload": name, "paths": paths, "dort_s": round(t_dort, 3), "oracle_s": round(t_orc, 3),
                     "dort_Mpaths_s": round(paths / t_dort / 1e6, 4),
                     "oracle_Mpaths_s": round(paths / t_orc / 1e6, 4),
                     "oracle_over_dort_time": round(t_orc / t_dort, 3)})
        print(json.dumps(rows[-1]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
