"""dgpu_trace_closest_noalloc: closest-hit tracing through the per-GPU scratch array
(no device allocation) gives exactly dgpu_trace_closest's hits, call after call."""
import ctypes as C

import numpy as np
import pytest

from dort_amd import abi, gpu, scenes

pytestmark = pytest.mark.gpu

FIELDS = ("t", "prim", "b1", "b2")


def _device_hits(ctx, n):
    return {"t": gpu.DeviceArray(ctx, n, np.float32), "prim": gpu.DeviceArray(ctx, n, np.int32),
            "b1": gpu.DeviceArray(ctx, n, np.float32), "b2": gpu.DeviceArray(ctx, n, np.float32)}


def _noalloc(gsc, dev_rays, out, flags):
    """One call of the entry point through the C-ABI; returns its status."""
    n = len(dev_rays["ox"])
    r = abi.rays_struct(dev_rays)
    h = abi.Hits(*(out[k].data_ptr() for k in FIELDS))
    ms = C.c_double()
    return gpu.lib().dgpu_trace_closest_noalloc(gsc.h, C.byref(r), n, C.byref(h), flags, C.byref(ms))


def _compare_calls(ctx, gsc, rays_list, bounded):
    flags = abi.BUF_DEVICE | (abi.TRACE_BOUNDED if bounded else 0)
    for rays in rays_list:  # calls in a row; later ones are smaller
        n = len(rays["ox"])
        dev = {k: ctx.device_array(np.ascontiguousarray(v)) for k, v in rays.items()}
        ref = _device_hits(ctx, n)
        got = _device_hits(ctx, n)
        for a in got.values():  # stale values must not survive
            a.copy_from(np.full(a.shape, 7, a.dtype))
        gsc.trace_closest(dev, out=ref, device=True, bounded=bounded)
        st = _noalloc(gsc, dev, got, flags)
        assert st == abi.DGPU_OK, gpu.lib().dgpu_last_error().decode()
        for k in FIELDS:
            a, b = ref[k].numpy(), got[k].numpy()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (k, n)
        hits = int((ref["prim"].numpy() >= 0).sum())
        print(f"{n} rays, {hits} hits: identical", flush=True)
        assert hits > 0
        for a in list(dev.values()) + [ref[k] for k in FIELDS] + [got[k] for k in FIELDS]:
            a.free()


def _subsets(rays, sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for m in sizes:
        idx = np.sort(rng.choice(len(rays["ox"]), size=m, replace=False))
        out.append({k: v[idx] for k, v in rays.items()})
    return out


def test_noalloc_equals_trace_closest(gpu_ctx):
    sc = scenes.icosphere_scene(freq=48, seed=3)
    rays = scenes.intersection_workload(sc, 128, 128, 40000, seed=5)
    gsc = gpu_ctx.scene(sc)
    _compare_calls(gpu_ctx, gsc, _subsets(rays, [50000, 7000], seed=1), bounded=False)
    cam = scenes.ray_batch(*scenes.camera_rays(sc, 64, 64))
    _compare_calls(gpu_ctx, gsc, [cam, {k: v[:1000] for k, v in cam.items()}], bounded=True)


def test_noalloc_with_a_grid_cut_to_the_scratch(gpu_ctx):
    """A deep BVH with 8 LDS stack entries and 256-thread workgroups needs more spill than
    the array holds at full occupancy: the entry point traces with fewer workgroups."""
    sc = scenes.icosphere_scene(freq=256, seed=1)
    gsc = gpu_ctx.scene(sc, stack_lds=8, block_size=256)
    info = gsc.info()
    stack_cap = (info["width"] - 1) * (info["depth"] + 1) + 2
    per_block = 256 * (stack_cap - 8) * 4
    print(f"BVH depth {info['depth']}, {per_block} spill bytes per workgroup, "
          f"{((32 << 20) - 256) // per_block} workgroups fit", flush=True)
    rays = scenes.intersection_workload(sc, 256, 256, 131072, seed=9)
    _compare_calls(gpu_ctx, gsc, _subsets(rays, [196608, 20000], seed=2), bounded=False)


def test_noalloc_refuses_host_buffers(gpu_ctx):
    sc = scenes.cornell()
    gsc = gpu_ctx.scene(sc)
    rays = scenes.ray_batch(*scenes.camera_rays(sc, 8, 8))
    dev = {k: gpu_ctx.device_array(np.ascontiguousarray(v)) for k, v in rays.items()}
    out = _device_hits(gpu_ctx, 64)
    assert _noalloc(gsc, dev, out, 0) == abi.DGPU_EINVAL
    assert b"device buffers" in gpu.lib().dgpu_last_error()
    assert _noalloc(gsc, dev, out, abi.BUF_DEVICE | 0x100) == abi.DGPU_EINVAL
    assert _noalloc(gsc, dev, out, abi.BUF_DEVICE) == abi.DGPU_OK
