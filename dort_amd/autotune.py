"""Autotuner for the BVH / traversal / wavefront knobs.

Searches BVH width, max leaf size, the number of traversal-stack entries kept in
LDS, the traversal workgroup size and the wavefront queue tile (paths per chunk)
by timing short renders on the GPU.  The search is greedy by coordinate: the BVH
shape first (it needs a rebuild per candidate), then the kernel launch knobs on the
best BVH.  Results are cached per workload in memory.

    best, table = autotune(ctx, scene, opts)                 # one scene, opts with few spp
    gsc = ctx.scene(scene, **best)
    best, table = autotune_batch(ctx, scenes, opts_list)     # knobs for a whole batch
    knobs, classes = autotune_batch_classes(ctx, scenes, opts_list)  # knobs per size class

dgpu_render_batch groups a batch's items by their scenes' knobs and traces each group
with its own kernel variant, all in one graph launch.  The batch tuner times
dgpu_render_batch over a set of scenes (or a sample of it) per candidate; the class
tuner splits the batch by triangle count and tunes each class on a sample of its own
scenes, so small and large scenes can take different BVH widths and stacks.
"""
import time

from . import abi
from .gpu import DgpuError

WIDTHS = (2, 4, 8)
LEAVES = (2, 4, 8)
STACKS = (8, 16, 32)
BLOCKS = (64, 128, 256)
TILES = (1 << 19, 1 << 20, 1 << 21, 1 << 22)
DEFAULTS = {"width": 4, "max_leaf_size": 4, "stack_lds": 16, "block_size": 128, "queue_tile": 0}

_cache = {}


def search(measure, shape_space=None, launch_space=None, log=None):
    """Greedy coordinate search: measure(knobs) -> ms (inf when the variant is missing).
    shape_space: [(width, max_leaf_size), ...]; launch_space: [(knob, values), ...].
    Returns (best knobs, [(knobs, ms), ...] in the order tried)."""
    if shape_space is None:
        shape_space = [(w, leaf) for w in WIDTHS for leaf in LEAVES]
    if launch_space is None:
        launch_space = [("stack_lds", STACKS), ("block_size", BLOCKS), ("queue_tile", TILES)]
    table = []

    def trial(knobs):
        t0 = time.perf_counter()
        try:
            ms = measure(knobs)
        except DgpuError as e:
            if e.status != abi.DGPU_EUNSUPPORTED:
                raise
            ms = float("inf")
        table.append((dict(knobs), ms))
        if log:
            log(f"autotune {knobs}: {ms:.2f} ms ({time.perf_counter() - t0:.1f}s)")
        return ms

    best = dict(DEFAULTS)
    best_ms = trial(best)
    for w, leaf in shape_space:
        k = dict(best, width=w, max_leaf_size=leaf)
        if k == best:
            continue
        ms = trial(k)
        if ms < best_ms:
            best_ms, best = ms, k
    for name, values in launch_space:
        base = dict(best)
        for v in values:
            k = dict(base, **{name: v})
            if k == best:
                continue
            ms = trial(k)
            if ms < best_ms:
                best_ms, best = ms, k
    return best, table


def _time_render(gsc, opts, repeats=2):
    o = abi.RenderOpts.from_buffer_copy(opts)
    gsc.render(o)  # warm-up
    best = float("inf")
    for _ in range(repeats):
        _, _, st = gsc.render(o)
        best = min(best, st["kernel_ms"])
    return best


def autotune(ctx, scene, opts, repeats=2, log=None, shape_space=None, launch_space=None):
    """Knobs for one scene: returns (best knob dict for Context.scene, [(knobs, ms), ...])."""
    key = ("scene", scene.name, scene.num_triangles, opts.x_res, opts.y_res, opts.sample_end - opts.sample_begin,
           str(shape_space), str(launch_space))
    if key in _cache:
        return _cache[key]

    def measure(knobs):
        gsc = ctx.scene(scene, **knobs)
        try:
            return _time_render(gsc, opts, repeats)
        finally:
            gsc.close()

    _cache[key] = search(measure, shape_space, launch_space, log)
    return _cache[key]


def autotune_batch(ctx, scene_list, opts_list, repeats=1, log=None, shape_space=None, launch_space=None):
    """Knobs shared by a batch: times dgpu_render_batch of (scene_list[i], opts_list[i]) with
    device films per candidate; returns (best knobs, [(knobs, ms), ...])."""
    import numpy as np
    from .gpu import DeviceArray
    key = ("batch", tuple((s.name, s.num_triangles) for s in scene_list),
           tuple((o.x_res, o.y_res, o.sample_end - o.sample_begin) for o in opts_list),
           str(shape_space), str(launch_space))
    if key in _cache:
        return _cache[key]
    films = []
    for o in opts_list:
        n = o.x_res * o.y_res
        f, w = DeviceArray(ctx, n * 3, np.float32), DeviceArray(ctx, n, np.float32)
        f.copy_from(np.zeros(n * 3, np.float32))
        w.copy_from(np.zeros(n, np.float32))
        films.append((f, w))

    def measure(knobs):
        gscs = [ctx.scene(s, **knobs) for s in scene_list]
        try:
            items = list(zip(gscs, opts_list))
            ctx.render_batch(items, films=films, device=True)  # warm-up
            best = float("inf")
            for _ in range(repeats):
                _, st = ctx.render_batch(items, films=films, device=True)
                best = min(best, st["kernel_ms"])
            return best
        finally:
            for g in gscs:
                g.close()

    try:
        _cache[key] = search(measure, shape_space, launch_space, log)
    finally:
        for f, w in films:
            f.free()
            w.free()
    return _cache[key]


def autotune_batch_classes(ctx, scene_list, opts_list, bounds=(20000, 150000), per_class=8, repeats=1, log=None,
                           shape_space=None, launch_space=None):
    """Knobs per scene of a batch, tuned per size class: the scenes are split by triangle
    count at `bounds` (class k holds counts in [bounds[k-1], bounds[k])), and each class is
    tuned with autotune_batch on at most `per_class` of its scenes (evenly spaced).
    Returns ([knobs of scene i], [(class bounds, scene indices, best knobs, table), ...])."""
    edges = (0,) + tuple(bounds) + (float("inf"),)
    knobs = [None] * len(scene_list)
    classes = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        idx = [i for i, s in enumerate(scene_list) if lo <= s.num_triangles < hi]
        if not idx:
            continue
        step = max(1, len(idx) // per_class)
        sample = idx[::step][:per_class]
        best, table = autotune_batch(ctx, [scene_list[i] for i in sample], [opts_list[i] for i in sample],
                                     repeats=repeats, log=log, shape_space=shape_space, launch_space=launch_space)
        for i in idx:
            knobs[i] = dict(best)
        classes.append(((lo, hi), idx, best, table))
    return knobs, classes
