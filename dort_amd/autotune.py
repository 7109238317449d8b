"""Autotuner for the BVH / traversal / wavefront knobs.

Searches BVH width, max leaf size, the number of traversal-stack entries kept in
LDS, the traversal workgroup size and the wavefront queue tile (paths per chunk)
by timing short renders of the caller's scene on the GPU.  The search is greedy
by coordinate: the BVH shape first (it needs a rebuild per candidate), then the
kernel launch knobs on the best BVH.  Results are cached per (scene name,
triangle count, resolution) in memory.

    best, table = autotune(ctx, scene, opts)        # opts: a RenderOpts, few spp
    gsc = ctx.scene(scene, **best)
"""
import time

from . import abi
from .gpu import DgpuError

WIDTHS = (2, 4)
LEAVES = (2, 4, 8)
STACKS = (8, 16, 32)
BLOCKS = (64, 128, 256)
TILES = (1 << 19, 1 << 20, 1 << 21, 1 << 22)

_cache = {}


def _time_render(gsc, opts, repeats=2):
    o = abi.RenderOpts.from_buffer_copy(opts)
    gsc.render(o)  # warm-up
    best = float("inf")
    for _ in range(repeats):
        _, _, st = gsc.render(o)
        best = min(best, st["kernel_ms"])
    return best


def autotune(ctx, scene, opts, repeats=2, log=None):
    """Returns (best knob dict for Context.scene, list of (knobs, ms) tried)."""
    key = (scene.name, scene.num_triangles, opts.x_res, opts.y_res, opts.sample_end - opts.sample_begin)
    if key in _cache:
        return _cache[key]
    table = []

    def trial(knobs):
        t0 = time.perf_counter()
        try:
            gsc = ctx.scene(scene, **knobs)
        except DgpuError as e:
            if e.status != abi.DGPU_EUNSUPPORTED:
                raise
            return float("inf")
        ms = _time_render(gsc, opts, repeats)
        gsc.close()
        table.append((dict(knobs), ms))
        if log:
            log(f"autotune {knobs}: {ms:.2f} ms ({time.perf_counter() - t0:.1f}s)")
        return ms

    best = {"width": 4, "max_leaf_size": 4, "stack_lds": 16, "block_size": 128, "queue_tile": 0}
    best_ms = trial(best)
    for w in WIDTHS:
        for leaf in LEAVES:
            k = dict(best, width=w, max_leaf_size=leaf)
            if k == best:
                continue
            ms = trial(k)
            if ms < best_ms:
                best_ms, best = ms, k
    for name, values in (("stack_lds", STACKS), ("block_size", BLOCKS), ("queue_tile", TILES)):
        base = dict(best)
        for v in values:
            k = dict(base, **{name: v})
            if k == best:
                continue
            ms = trial(k)
            if ms < best_ms:
                best_ms, best = ms, k
    _cache[key] = (best, table)
    return best, table
