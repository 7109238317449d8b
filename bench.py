"""Benchmark of the MI355X path-tracing core.

One step = one dgpu_render of the displaced-noise terrain workload: a
2,000,000-triangle heightfield plus a 2-triangle ground plane, lit by a
1024x512 FP32 procedural sky (sun lobe of peak 50, gradient sky), per-object
Lambert / GGX-conductor / dielectric materials drawn from a seed, rendered at
1024x1024 with 64 samples per pixel, max depth 8, Russian roulette from depth 3,
MIS between BSDF and environment sampling.  The scene (BVH, triangles, sky) and
the film live in HBM before the timed region starts.

Also measured in the same run (secondary fields): the ray-intersection workload,
a 1,310,720-triangle displaced icosphere with 16,777,216 rays (half pinhole rays
at 4096x2048, half random rays), closest hit and any hit, rays and hits resident
in HBM.

Printed: ONE JSON line with the metric, the roofline of the dominant kernel
(achieved algorithmic GB/s from live HIP-event timings of every launch) and the
CPU baseline (the oracle, a C++ restatement of dort's `pt`, timed on this host's
cores over a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--quads", type=int, default=1000)
    ap.add_argument("--cpu-spp", type=int, default=16,
                    help="samples per pixel of the oracle's CPU sample (16: about 4 s on 16 threads)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU workers (default: the CPUs this process may use, capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-intersect", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-light", action="store_true")
    ap.add_argument("--no-bdpt", action="store_true")
    ap.add_argument("--no-batch-tune", action="store_true",
                    help="batch leg: the default knobs (BVH4, stack 16, block 256) instead of the autotuner's")
    ap.add_argument("--calib", action="store_true",
                    help="also run a 256 MiB device-to-device copy (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--width", type=int, default=4)
    ap.add_argument("--stack", type=int, default=16)
    ap.add_argument("--block", type=int, default=128)
    ap.add_argument("--leaf", type=int, default=4)
    ap.add_argument("--tile", type=int, default=0)
    return ap.parse_args()


def committed_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary of this
    workload (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, copy-kernel calibration), when the run
    uses the profiled configuration; otherwise None."""
    path = os.path.join(ROOT, "profiles", "terrain_2M_1024x1024_64spp_pmc.md")
    default = (args.quads, args.res, args.spp, args.width, args.stack, args.block, args.leaf) == \
        (1000, 1024, 64, 4, 16, 128, 4)
    if not default or not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith(f"| `k_{kernel}"):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            return int((float(cols[3]) + float(cols[4])) * 1e6), os.path.relpath(path, ROOT)
    return None, None


def terrain_workload(args):
    from dort_amd import abi, scenes
    sc = scenes.terrain_scene(quads=args.quads, seed=2)
    opts = abi.render_opts(args.res, args.res, spp=args.spp, seed=1234, max_depth=8, rr_depth=3)
    return sc, opts


def intersection_bench(ctx, args):
    from dort_amd import abi, gpu, scenes
    t0 = time.time()
    sc = scenes.icosphere_scene(freq=256, seed=1)
    rays = scenes.intersection_workload(sc, 4096, 2048, 4096 * 2048, seed=7)
    n = len(rays["ox"])
    log(f"intersection workload: {sc.num_triangles} triangles, {n} rays ({time.time() - t0:.1f}s to generate)")
    gsc = ctx.scene(sc, width=args.width, max_leaf_size=args.leaf, stack_lds=args.stack,
                    block_size=args.block)
    info = gsc.info()
    dev = {k: ctx.device_array(v) for k, v in rays.items()}
    out = {"t": gpu.DeviceArray(ctx, n, np.float32), "prim": gpu.DeviceArray(ctx, n, np.int32),
           "b1": gpu.DeviceArray(ctx, n, np.float32), "b2": gpu.DeviceArray(ctx, n, np.float32)}
    occ = gpu.DeviceArray(ctx, n, np.uint32)
    gsc.trace_closest(dev, out=out, device=True)
    gsc.trace_any(dev, out=occ, device=True)
    cms, ams = [], []
    for _ in range(3):
        cms.append(gsc.trace_closest(dev, out=out, device=True)["kernel_ms"])
        ams.append(gsc.trace_any(dev, out=occ, device=True)[1])
    c_ms, a_ms = float(np.median(cms)), float(np.median(ams))
    hit_frac = float((out["prim"].numpy() >= 0).mean())
    # the scene once: 36 B of positions per triangle + the BVH nodes as stored
    scene_bytes = 36 * sc.num_triangles + info["node_bytes"]
    res = {"workload": f"displaced icosphere {sc.num_triangles} tris, {n} rays (4096x2048 pinhole + random)",
           "Mrays_closest": n / c_ms / 1e3, "Mrays_any": n / a_ms / 1e3,
           "closest_ms": c_ms, "any_ms": a_ms, "hit_fraction": hit_frac,
           "closest_GBps_algorithmic": (48 * n + scene_bytes) / c_ms / 1e6,
           "bvh": {"nodes": info["nodes"], "depth": info["depth"], "build_ms": info["build_ms"]}}
    for v in list(dev.values()) + [out[k] for k in ("t", "prim", "b1", "b2")] + [occ]:
        v.free()
    return res


def light_bench(gsc, film, args):
    """dort's light tracer (dgpu_render_light) on the headline terrain scene: light paths
    per second at 1024x1024 x 16 iterations (16.8M paths), depth 8; the splat film is
    resident in HBM."""
    from dort_amd import abi
    its = 16
    opts = abi.render_opts(args.res, args.res, spp=its, seed=99, max_depth=8)
    gsc.render_light(opts, film, device=True)
    ts, st = [], None
    for _ in range(3):
        t0 = time.perf_counter()
        _, st = gsc.render_light(opts, film, device=True)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    paths = args.res * args.res * its
    return {"workload": f"terrain, light tracing, {args.res}x{args.res} x {its} iterations, max depth 8",
            "Mpaths_s": paths / t / 1e6, "ms": t * 1e3,
            "extension_rays_per_path": st["extension_rays"] / st["paths"],
            "shadow_rays_per_path": st["shadow_rays"] / st["paths"]}


def bdpt_bench(ctx, gsc, film, wfilm, args):
    """dort's bidirectional path tracer (dgpu_render_bdpt) in its two forms, the wavefront
    stages and the one-thread-per-path kernel: camera-subpath samples per second on the
    headline terrain (1024x1024 x 4 iterations, depth 8) and on dort's test box with its
    area light (512x512 x 16 iterations, depth 5).  Films resident in HBM."""
    from dort_amd import abi, gpu, scenes
    npix = args.res * args.res
    splat = gpu.DeviceArray(ctx, npix * 3, np.float32)
    splat.copy_from(np.zeros(npix * 3, np.float32))
    box = ctx.scene(scenes.dort_test_box("area"))
    out = {}
    for name, sc, res, its, depth in (("terrain", gsc, args.res, 4, 8), ("test_box", box, 512, 16, 5)):
        opts = abi.render_opts(res, res, spp=its, seed=123, max_depth=depth)
        row = {"workload": f"{name}, {res}x{res} x {its} iterations, max depth {depth}"}
        for form, mega in (("wavefront", False), ("megakernel", True)):
            sc.render_bdpt(opts, films=(film, wfilm, splat), device=True, megakernel=mega)
            ts, st = [], None
            for _ in range(3):
                t0 = time.perf_counter()
                *_, st = sc.render_bdpt(opts, films=(film, wfilm, splat), device=True, megakernel=mega)
                ts.append(time.perf_counter() - t0)
            t = float(np.median(ts))
            row[form] = {"Mpaths_s": res * res * its / t / 1e6, "ms": t * 1e3,
                         "extension_rays_per_path": st["extension_rays"] / st["paths"],
                         "shadow_rays_per_path": st["shadow_rays"] / st["paths"]}
        out[name] = row
    splat.free()
    return out


def adjoint_errors(ctx, gsc, guess):
    """RMS errors of the GPU adjoint and of central finite differences (h = 1e-3, common
    random numbers) against the FP64 adjoint, on the inverse-rendering scene rendered at
    64x64, 4 spp (tests/golden/adjoint_terrain200k_64x64_4spp_fp64.npz, written by the
    oracle's double-precision path replay).  The unknowns compared: the 64 texel entries
    with the largest reference gradients and every GGX roughness; the loss is the
    fixture's seeded linear functional sum(dLdI * image)."""
    from dort_amd import abi
    fix = np.load(os.path.join(ROOT, "tests", "golden", "adjoint_terrain200k_64x64_4spp_fp64.npz"))
    r_tex, r_rough, dldi = fix["g_tex"], fix["g_rough"], fix["dldi"]
    opts = abi.render_opts(64, 64, spp=4, seed=77, max_depth=6, rr_depth=3)
    g_tex, g_rough, _, _ = gsc.adjoint(opts, dldi)
    texels = np.argsort(-np.abs(r_tex).ravel())[:64]
    mats = [m for m in range(len(guess.materials)) if guess.materials[m].type == abi.MAT_GGX_CONDUCTOR]
    ref = np.concatenate([r_tex.ravel()[texels], r_rough[mats]])
    adj = np.concatenate([g_tex.ravel()[texels], g_rough[mats]])
    h = 1e-3

    def loss():
        return float((gsc.image(opts).astype(np.float64) * dldi).sum())

    fd = []
    tex = guess.textures[0].astype(np.float32)
    for k in texels:
        vals = []
        for sgn in (1.0, -1.0):
            t = tex.copy()
            t.reshape(-1)[k] += sgn * h
            gsc.update_texture(0, t)
            vals.append(loss())
        fd.append((vals[0] - vals[1]) / (2 * h))
    gsc.update_texture(0, tex)
    for m in mats:
        vals = []
        for sgn in (1.0, -1.0):
            mm = abi.Material.from_buffer_copy(guess.materials[m])
            mm.roughness += sgn * h
            gsc.update_material(m, mm)
            vals.append(loss())
        fd.append((vals[0] - vals[1]) / (2 * h))
        gsc.update_material(m, guess.materials[m])
    rms = lambda v: float(np.sqrt(np.mean(np.square(v))))
    return rms(adj - ref) / rms(ref), rms(np.asarray(fd) - ref) / rms(ref), len(ref)


def adjoint_bench(ctx, args):
    """Differentiable inverse rendering workload: 200k-triangle terrain, 512x512, 32 spp,
    max depth 6; unknowns = 256x256 RGB albedo texture + 8 per-object GGX roughness;
    loss = pixelwise L2 against a target rendered from hidden seeded parameters.  Central
    differences need two renders per unknown for one gradient; their error and the
    adjoint's are measured against the FP64 adjoint (adjoint_errors), so the adjoint's
    speed-up holds at equal gradient error whenever central differences are no more
    accurate than the adjoint."""
    from dort_amd import abi, scenes
    hidden = scenes.terrain_scene(quads=316, seed=8, textured=True, tex_res=256, sky_res=(1024, 512),
                                  ggx_roughness=np.linspace(0.08, 0.45, 8))
    guess = scenes.terrain_scene(quads=316, seed=8, textured=True, tex_res=256, sky_res=(1024, 512),
                                 ggx_roughness=np.full(8, 0.25))
    guess.textures[0] = np.full_like(guess.textures[0], 0.5)
    opts = abi.render_opts(512, 512, spp=32, seed=77, max_depth=6, rr_depth=3)
    target = ctx.scene(hidden).image(opts)
    gsc = ctx.scene(guess, width=args.width, max_leaf_size=args.leaf, stack_lds=args.stack,
                    block_size=args.block)
    n_params = guess.textures[0].size + 8
    # warm up, then time forward + adjoint (one full gradient vector)
    img = gsc.image(opts)
    gsc.adjoint(opts, 2 * (img - target))
    fwd, adj = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        img = gsc.image(opts)
        t1 = time.perf_counter()
        g_tex, g_rough, _, st = gsc.adjoint(opts, 2 * (img - target))
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        adj.append(t2 - t1)
    t_fwd, t_adj = float(np.median(fwd)), float(np.median(adj))
    grads_per_s = 1.0 / (t_fwd + t_adj)
    fd_grads_per_s = 1.0 / (2 * n_params * t_fwd)
    err_adj, err_fd, n_cmp = adjoint_errors(ctx, gsc, guess)
    equal_error = err_fd >= err_adj
    return {"workload": f"terrain {guess.num_triangles} tris, 512x512, 32 spp, depth 6, "
                        f"{n_params} unknowns (256x256x3 albedo + 8 roughness)",
            "grads_per_s": grads_per_s, "fd_grads_per_s": fd_grads_per_s,
            "forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3, "adjoint_kernel_ms": st["kernel_ms"],
            "rms_rel_error_vs_fp64": {"adjoint": err_adj, "central_fd_h1e-3": err_fd, "unknowns": n_cmp,
                                      "scene": "the same scene at 64x64, 4 spp"},
            "speedup_vs_fd_at_equal_error": (grads_per_s / fd_grads_per_s) if equal_error else None,
            "loss": float(((img - target) ** 2).sum()),
            "grad_norm": float(np.sqrt((g_tex.astype(np.float64) ** 2).sum() + (g_rough ** 2).sum()))}


def batch_bench(ctx, args):
    """64 independent scenes, triangle counts log-uniform in [1k, 500k] (displaced
    geodesic spheres and boxes on a ground plane), resolutions uniform in {256^2, 512^2,
    768^2}, 16 spp, max depth 6, each lit by a 1024x512 sky.  Both legs use the knobs the
    autotuner picks per triangle-count class.  Batched wavefront (every chunk in one
    hipGraph launch) against one dgpu_render launch sequence per scene; each
    leg has its own films in HBM, and before timing the batched films are checked
    against the per-scene ones (weights bitwise, radiance to rtol 1e-5)."""
    from dort_amd import abi, autotune, gpu, scenes
    t0 = time.time()
    batch = scenes.batch_scenes(64, seed=3)
    # the knobs per size class (BVH width, stack entries in LDS, trace workgroup), picked
    # by the autotuner from batched renders of up to 8 scenes of each class at the
    # workload's 16 spp (at 1 spp the launches are too small for occupancy to show, and
    # the tuner picked a deeper LDS stack that made the full batch 15% slower);
    # dgpu_render_batch traces each class's chunks with its own kernel variant
    t_tune = time.time()
    all_opts = [abi.render_opts(res, res, spp=16, seed=500 + k, max_depth=6, rr_depth=3, scene_id=k)
                for k, (_, res) in enumerate(batch)]
    if args.no_batch_tune:
        fixed = dict(autotune.DEFAULTS, block_size=256)
        scene_knobs, classes = [fixed] * len(batch), [((0, float("inf")), list(range(len(batch))), fixed, [])]
    else:
        scene_knobs, classes = autotune.autotune_batch_classes(
            ctx, [sc for sc, _ in batch], all_opts, bounds=(20000, 150000), per_class=8,
            shape_space=[(4, 4), (8, 4), (2, 4)],
            launch_space=[("stack_lds", (8, 16, 32)), ("block_size", (64, 128, 256))])
    t_tune = time.time() - t_tune
    for (lo, hi), idx, best, table in classes:
        log(f"batch autotune, {len(idx)} scenes of {lo}-{hi} triangles: {best} ({len(table)} candidates)")
    items, films_p, films_b = [], [], []

    def zeroed(n):
        a = gpu.DeviceArray(ctx, n, np.float32)
        a.copy_from(np.zeros(n, np.float32))
        return a

    for k, (sc, res) in enumerate(batch):
        gsc = ctx.scene(sc, **scene_knobs[k])
        o = all_opts[k]
        items.append((gsc, o))
        films_p.append((zeroed(res * res * 3), zeroed(res * res)))
        films_b.append((zeroed(res * res * 3), zeroed(res * res)))
    tris = sum(sc.num_triangles for sc, _ in batch)
    paths = sum(o.x_res * o.y_res * 16 for _, o in items)
    log(f"batch workload: 64 scenes, {tris} triangles, {paths / 1e6:.1f} Mpaths ({time.time() - t0:.1f}s to build)")

    def per_scene():
        for (gsc, o), (f, w) in zip(items, films_p):
            gsc.render(o, f, w, device=True)

    def batched():
        ctx.render_batch(items, films=films_b, device=True)

    per_scene()
    batched()
    mismatched = []
    for k, ((fp, wp), (fb, wb)) in enumerate(zip(films_p, films_b)):
        if not (np.array_equal(wp.numpy(), wb.numpy()) and
                np.allclose(fb.numpy(), fp.numpy(), rtol=1e-5, atol=1e-6)):
            mismatched.append(k)
    log(f"batch films equal per-scene films: {not mismatched} (mismatched scenes: {mismatched})")
    tp, tb = [], []
    for _ in range(2):
        t0 = time.perf_counter(); per_scene(); tp.append(time.perf_counter() - t0)
        t0 = time.perf_counter(); batched(); tb.append(time.perf_counter() - t0)
    t_p, t_b = min(tp), min(tb)
    for f, w in films_p + films_b:
        f.free()
        w.free()
    return {"workload": f"64 scenes, {tris} tris (log-uniform 1k-500k), 256^2/512^2/768^2, 16 spp, depth 6, "
                        "1024x512 sky each",
            "paths": paths, "per_scene_Mpaths_s": paths / t_p / 1e6, "batched_Mpaths_s": paths / t_b / 1e6,
            "speedup": t_p / t_b, "per_scene_ms": t_p * 1e3, "batched_ms": t_b * 1e3,
            "films_equal_per_scene": not mismatched,
            "graph_launches": 1,
            "autotune": {"classes": [{"triangles": [lo, hi if hi != float("inf") else None], "scenes": len(idx),
                                      "knobs": best, "candidates": [[k, round(ms, 3)] for k, ms in table]}
                                     for (lo, hi), idx, best, table in classes],
                         "seconds": round(t_tune, 1),
                         "sample": "up to 8 scenes of each triangle-count class, 16 spp, batched"}}


def cpu_baseline(sc, args):
    """dort itself (oracle/_ref/libdortref.so, its sources built by oracle/ref.mk) on the
    host's CPU share: one single-threaded dort process per CPU, each rendering the whole
    1024x1024 image for one iteration with its own seed (oracle/dort_baseline.py), timed
    together; and, beside it, the oracle (this build's C++ restatement of dort's pt, which
    traces the same paths as dort, 1.5-1.9x faster single-threaded) on the exact scene.
    Where dort is not built the oracle is the baseline."""
    from dort_amd import abi
    from oracle import dort_baseline, oracle, ref
    info = dort_baseline.cpu_info()
    procs = args.cpu_threads or dort_baseline.cpu_share()
    o = oracle.OracleScene(sc)
    opts = abi.render_opts(args.res, args.res, spp=args.cpu_spp, seed=1234, max_depth=8, rr_depth=3)
    t0 = time.time()
    _, _, cnt = o.render(opts, rng="mt", threads=procs)
    dt = time.time() - t0
    port = {"value": cnt["paths"] / dt / 1e6, "unit": "Mpaths/s", "cores": procs, "kind": "port",
            "sample": f"{args.res}x{args.res} x {args.cpu_spp} spp of the same terrain scene "
                      f"({cnt['paths']} paths, {dt:.1f}s, dort's mt19937 sampler)",
            "extension_rays_per_path": cnt["extension_rays"] / cnt["paths"],
            "shadow_rays_per_path": cnt["shadow_rays"] / cnt["paths"]}
    host = {"nproc": info["nproc"], "allowed_cpus": info["allowed_cpus"], "cpu_model": info["cpu_model"],
            "smt": info["smt"], "cpu_share": procs}
    if not ref.available():
        return {**port, **host}
    d = dort_baseline.run(procs, 1, res=args.res, quads=args.quads)
    return {"value": d["Mpaths_s"], "unit": "Mpaths/s", "cores": procs, "kind": "reference",
            "sample": f"dort's PathRenderer, {procs} single-threaded processes x 1 iteration of "
                      f"{args.res}x{args.res} ({d['paths']} paths in {d['seconds']:.1f}s), the terrain "
                      "with its GGX conductors as Lambert (dort has no rough conductor) and dort's RR "
                      "from bounce 2",
            **host, "port": port}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is described in DESIGN.md, not run)"}))
        sys.exit(2)
    from dort_amd import abi, gpu
    sc, opts = terrain_workload(args)
    log(f"terrain workload: {sc.num_triangles} triangles, {args.res}x{args.res} x {args.spp} spp")
    ctx = gpu.Context(0)
    gsc = ctx.scene(sc, width=args.width, max_leaf_size=args.leaf, stack_lds=args.stack,
                    block_size=args.block, queue_tile=args.tile)
    info = gsc.info()
    log(f"BVH: {info['nodes']} nodes, depth {info['depth']}, build {info['build_ms']:.0f} ms")
    npix = args.res * args.res
    film = gpu.DeviceArray(ctx, npix * 3, np.float32)
    film.copy_from(np.zeros(npix * 3, np.float32))
    wfilm = gpu.DeviceArray(ctx, npix, np.float32)
    wfilm.copy_from(np.zeros(npix, np.float32))
    paths = npix * args.spp
    for i in range(args.warmup):
        _, _, st = gsc.render(opts, film, wfilm, device=True)
        log(f"warmup {i}: {st['kernel_ms']:.0f} ms")
    times, stats = [], []
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        _, _, st = gsc.render(opts, film, wfilm, device=True)
        times.append(time.perf_counter() - t0)
        stats.append(st)
        log(f"step {i}: {times[-1] * 1e3:.0f} ms, {paths / times[-1] / 1e6:.1f} Mpaths/s")
    total = time.perf_counter() - t_all
    ms_per_step = total / args.steps * 1e3
    value = paths * args.steps / total / 1e6

    # live per-kernel timing (HIP events around every launch, on the stream it runs on)
    popts = abi.RenderOpts.from_buffer_copy(opts)
    popts.flags |= abi.RENDER_PROFILE
    _, _, ps = gsc.render(popts, film, wfilm, device=True)
    ext_rays, shd_rays = ps["extension_rays"], ps["shadow_rays"]
    # the scene once: 36 B of positions per triangle + the BVH nodes as stored
    scene_bytes = 36 * sc.num_triangles + info["node_bytes"]
    kernels = {
        # closest-hit: 32 B ray in + 16 B hit out per ray, plus the scene once per launch
        "extend": (ps["extend_ms"], ps["extend_launches"], 48 * ext_rays / max(1, ps["extend_launches"]) + scene_bytes),
        # any-hit: 32 B ray + 8 B slot/ignore in, 1 B visibility out, plus the scene
        "connect": (ps["connect_ms"], ps["connect_launches"], 41 * shd_rays / max(1, ps["connect_launches"]) + scene_bytes),
        # shade: per extension ray 224 B (state, queue entry, hit, triangle, next entry) + 57 B per shadow ray
        "shade": (ps["shade_ms"], ps["shade_launches"],
                  (224 * ext_rays + 57 * shd_rays) / max(1, ps["shade_launches"])),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    traffic, traffic_src = committed_traffic(args, dom)
    kms, nl, bytes_per_launch = kernels[dom]
    avg_ms = kms / max(1, nl)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": dom,
                "avg_launch_ms": round(avg_ms, 4), "launches": nl,
                "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "kernel_ms": {k: round(v[0], 2) for k, v in kernels.items()},
                "other_ms": round(ps["other_ms"], 2),
                "overlap": "each bounce's connect runs beside the next bounce's extend on a second "
                           "stream; launch times include the time the two kernels share the GPU"}
    log(f"profile: {roofline['kernel_ms']} dominant {dom} avg {avg_ms:.3f} ms, {achieved:.1f} GB/s")

    if args.calib:
        nb = 256 << 20
        a = gpu.DeviceArray(ctx, nb // 4, np.float32)
        b = gpu.DeviceArray(ctx, nb // 4, np.float32)
        for _ in range(2):
            gpu.check(gpu.lib().dgpu_copy(ctx.h, b.ptr, a.ptr, nb, 2))
        a.free()
        b.free()
        log("calibration copy: 2 x 256 MiB device-to-device")

    secondary = {}
    if not args.no_light:
        secondary["light_tracing"] = light_bench(gsc, film, args)
        lt = secondary["light_tracing"]
        log(f"light tracing: {lt['Mpaths_s']:.0f} Mpaths/s ({lt['ms']:.1f} ms), "
            f"{lt['extension_rays_per_path']:.2f} extension + {lt['shadow_rays_per_path']:.2f} shadow rays per path")
    if not args.no_bdpt:
        secondary["bdpt"] = bdpt_bench(ctx, gsc, film, wfilm, args)
        for name, row in secondary["bdpt"].items():
            log(f"bdpt {name}: wavefront {row['wavefront']['Mpaths_s']:.0f} Mpaths/s "
                f"({row['wavefront']['ms']:.1f} ms), one thread per path {row['megakernel']['Mpaths_s']:.0f} "
                f"({row['megakernel']['ms']:.1f} ms)")
    if not args.no_intersect:
        film.free()
        wfilm.free()
        secondary["intersection"] = intersection_bench(ctx, args)
        log(f"intersection: {secondary['intersection']['Mrays_closest']:.0f} Mrays/s closest, "
            f"{secondary['intersection']['Mrays_any']:.0f} Mrays/s any")

    if not args.no_adjoint:
        secondary["adjoint"] = adjoint_bench(ctx, args)
        a = secondary["adjoint"]
        e = a["rms_rel_error_vs_fp64"]
        log(f"adjoint: {a['grads_per_s']:.2f} grads/s, FD {a['fd_grads_per_s']:.2e} grads/s, forward "
            f"{a['forward_ms']:.0f} ms, adjoint {a['adjoint_ms']:.0f} ms; RMS error vs FP64: adjoint "
            f"{e['adjoint']:.2e}, central FD {e['central_fd_h1e-3']:.2e}")

    if not args.no_batch:
        secondary["batch"] = batch_bench(ctx, args)
        b = secondary["batch"]
        log(f"batch: {b['batched_Mpaths_s']:.1f} Mpaths/s batched vs {b['per_scene_Mpaths_s']:.1f} per scene "
            f"({b['speedup']:.2f}x)")

    cpu = None
    if not args.no_cpu:
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(sc, args)
        log(f"cpu ({cpu['kind']}): {cpu['value']:.3f} Mpaths/s on {cpu['cores']} cores of {cpu['nproc']} "
            f"({cpu['cpu_model']}, SMT {cpu['smt']})")

    line = {
        "metric": "Mpaths/s at equal rMSE vs CPU", "value": round(value, 3), "unit": "Mpaths/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural terrain, sky and materials)",
        "config": {"workload": f"terrain {sc.num_triangles} tris + sky 1024x512, {args.res}x{args.res}, "
                               f"{args.spp} spp, max depth 8, RR from depth 3, MIS",
                   "paths_per_step": paths, "bvh_width": args.width, "stack_lds": args.stack,
                   "block": args.block, "max_leaf": args.leaf,
                   "extension_rays_per_path": round(ext_rays / paths, 3),
                   "shadow_rays_per_path": round(shd_rays / paths, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
