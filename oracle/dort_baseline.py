"""dort itself as the timed CPU baseline (TEST INFRASTRUCTURE ONLY).

oracle/_ref/libdortref.so is dort's own path tracer (its sources built by oracle/ref.mk).
Its thread pool is only safe with one thread (oracle/ref.py), so the host's cores are
used the way independent render nodes would be: P single-threaded processes, each
rendering the whole image for its share of the iterations with its own seed (dort's
PathRenderer always starts at iteration 0).  Every process builds its scene first and
reports ready; then all start rendering together, and the baseline is the paths of all
processes over the longest render time.  Scene generation and BVH builds are outside the
timed region, as they are for the GPU.

dort renders the terrain workload as dort can express it: its only GGX lobe is a
rough dielectric without a usable pdf, so the GGX conductors become Lambert surfaces
(as in the parity tests), and its Russian roulette starts at bounce 2 (rr_depth 3 is
this library's option), which makes dort's paths slightly shorter than the GPU's.

    python -m oracle.dort_baseline --procs 4 --iters 1 --res 256
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cpu_info():
    """nproc, the CPUs this process may run on, the CPU model and the SMT state."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    smt = "unknown"
    try:
        smt = "on" if open("/sys/devices/system/cpu/smt/active").read().strip() == "1" else "off"
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "allowed_cpus": allowed, "cpu_model": model, "smt": smt}


def cpu_share():
    """Worker count for this host: the CPUs this process may use, capped by the job's CPU
    share where the environment states one (OMP_NUM_THREADS)."""
    n = cpu_info()["allowed_cpus"]
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def _child(args):
    sys.path.insert(0, ROOT)
    from dort_amd import abi, scenes
    from oracle import crosstime, ref
    sc = crosstime.dortable(scenes.terrain_scene(quads=args.quads, seed=args.scene_seed))
    d = ref.RefScene(sc)
    opts = abi.render_opts(args.res, args.res, spp=args.iters, seed=args.seed, max_depth=8, rr_depth=3)
    print("ready", flush=True)
    sys.stdin.readline()
    t0 = time.perf_counter()
    d.render(opts, threads=1)
    dt = time.perf_counter() - t0
    print(json.dumps({"seconds": dt, "paths": args.res * args.res * args.iters}), flush=True)


def run(procs, iters, res=1024, quads=1000, scene_seed=2, seed=1234, timeout=900):
    """Renders with `procs` single-threaded dort processes, `iters` iterations each;
    returns {"Mpaths_s", "paths", "seconds", "procs"}."""
    cmd = [sys.executable, "-m", "oracle.dort_baseline", "--child", "--iters", str(iters), "--res", str(res),
           "--quads", str(quads), "--scene-seed", str(scene_seed)]
    kids = [subprocess.Popen(cmd + ["--seed", str(seed + k)], cwd=ROOT, stdin=subprocess.PIPE,
                             stdout=subprocess.PIPE, text=True) for k in range(procs)]
    try:
        deadline = time.time() + timeout
        for k in kids:
            line = k.stdout.readline()
            if line.strip() != "ready" or time.time() > deadline:
                raise RuntimeError(f"dort baseline process failed to start (got {line!r})")
        for k in kids:
            k.stdin.write("go\n")
            k.stdin.flush()
        results = [json.loads(k.stdout.readline()) for k in kids]
        for k in kids:
            if k.wait(timeout=max(1.0, deadline - time.time())) != 0:
                raise RuntimeError("dort baseline process failed")
    finally:
        for k in kids:
            if k.poll() is None:
                k.kill()
    paths = sum(r["paths"] for r in results)
    seconds = max(r["seconds"] for r in results)
    return {"Mpaths_s": paths / seconds / 1e6, "paths": paths, "seconds": seconds, "procs": procs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--child", action="store_true")
    ap.add_argument("--procs", type=int, default=0)
    ap.add_argument("--iters", type=int, default=1)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--quads", type=int, default=1000)
    ap.add_argument("--scene-seed", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1234)
    args = ap.parse_args()
    if args.child:
        _child(args)
        return
    print(json.dumps({**run(args.procs or cpu_share(), args.iters, args.res, args.quads), **cpu_info()}))


if __name__ == "__main__":
    main()
