"""ctypes binding of oracle/_ref/libdortref.so (TEST INFRASTRUCTURE ONLY).

libdortref.so is dort itself: its pt-path sources compiled where they lie under
the reference checkout (oracle/ref.mk) plus ref_driver.cpp.  Tests use it to pin
the oracle against the real reference; it is built only where the reference
checkout exists, and tests that need it skip when it is absent.

Run dort with one thread.  Its ThreadPool (thread_pool.cpp) lets a worker that
found a loop exhausted take the pool lock and read the Loop object after the
thread that owns it (on its stack) has seen the loop finish and returned
(ThreadPool::loop :54-75 against work_on_loop :111-146): a use after free that
crashes now and then with many threads.  With one thread the pool has no workers.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from dort_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libdortref.so")
GPU_LIB_PATH = os.path.join(HERE, "_ref", "libdortref_gpu.so")
LUA_HOST_PATH = os.path.join(HERE, "_ref", "dort_lua")
LUA_PRELUDE = os.path.join(HERE, "..", "dort_amd", "host", "dort_gpu.lua")
REF_ROOT = "/root/reference"
_lib = None
_gpu_lib = None


def available():
    return os.path.exists(LIB_PATH)


def gpu_available():
    return os.path.exists(GPU_LIB_PATH)


def lua_available():
    return os.path.exists(LUA_HOST_PATH)


def run_lua(script, *args, lua_dir=None, cwd=None, timeout=300):
    """Runs `script` in oracle/_ref/dort_lua (dort's Lua front-end with the dort_gpu
    module and its prelude loaded); lua_dir: dort's lua/ tree, for scripts that require
    dort/dsl.  Returns the CompletedProcess (text output)."""
    cmd = [LUA_HOST_PATH, "--prelude", LUA_PRELUDE]
    if lua_dir:
        cmd += ["--lua-dir", lua_dir]
    cmd += ["--", script, *[str(a) for a in args]]
    return subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=timeout)


def build():
    """Compiles dort's sources into oracle/_ref (only where the checkout exists), and
    dort with dort_amd's GpuPathRenderer and its Lua front-end (dort_lua) when
    libdortgpu.so has been built."""
    if not os.path.isdir(os.path.join(REF_ROOT, "src", "dort")):
        return False
    targets = ["all"]
    if os.path.exists(os.path.join(HERE, "..", "dort_amd", "libdortgpu.so")):
        targets += ["gpu", "lua"]
    subprocess.check_call(["make", "-s", "-j8", "-f", os.path.join(HERE, "ref.mk"), f"REF={REF_ROOT}", *targets])
    return True


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.ref_scene_create.restype = C.c_void_p
        L.ref_scene_create.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.ref_scene_destroy.argtypes = [C.c_void_p]
        L.ref_trace.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.c_int,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_render.restype = C.c_int
        L.ref_render.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_uint32, C.c_void_p, C.c_void_p]
        L.ref_render_light.restype = C.c_int
        L.ref_render_light.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_uint32, C.c_void_p]
        L.ref_render_bdpt.restype = C.c_int
        L.ref_render_bdpt.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_uint32, C.c_uint32, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
        L.ref_read_ply.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                                   C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint32)]
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_read_hdr.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32)]
        L.ref_write_hdr.argtypes = [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]
        L.ref_read_png.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32)]
        L.ref_write_png.argtypes = [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]
        L.ref_test_convergence.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32, C.c_float,
                                           C.c_float, C.c_float, C.c_char_p, C.c_size_t]
        _lib = L
    return _lib


def read_ply(path):
    """dort's read_ply_to_ply_mesh: (positions (V, 3), indices (T, 3))."""
    pos, ind = C.POINTER(C.c_float)(), C.POINTER(C.c_uint32)()
    nv, nt = C.c_uint32(), C.c_uint32()
    rc = lib().ref_read_ply(os.fsencode(path), C.byref(pos), C.byref(nv), C.byref(ind), C.byref(nt))
    if rc != 0:
        raise ValueError(f"dort could not read {path} ({rc})")
    p = np.ctypeslib.as_array(pos, (3 * nv.value + 1,))[: 3 * nv.value].reshape(-1, 3).copy()
    t = np.ctypeslib.as_array(ind, (3 * nt.value + 1,))[: 3 * nt.value].reshape(-1, 3).copy()
    lib().ref_free(pos)
    lib().ref_free(ind)
    return p, t


def read_hdr(path):
    """dort's read_image_f (stb_image): (H, W, 3) float32."""
    px = C.POINTER(C.c_float)()
    w, h = C.c_uint32(), C.c_uint32()
    if lib().ref_read_hdr(os.fsencode(path), C.byref(px), C.byref(w), C.byref(h)) != 0:
        raise ValueError(f"dort could not open {path}")
    img = np.ctypeslib.as_array(px, (3 * w.value * h.value + 1,))[: 3 * w.value * h.value]
    img = img.reshape(h.value, w.value, 3).copy()
    lib().ref_free(px)
    return img


def write_hdr(path, img):
    """dort's write_image_rgbe (stb_image_write)."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    if lib().ref_write_hdr(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0]) != 0:
        raise ValueError(f"dort could not write {path}")


def read_png(path):
    """dort's read_image_8 (stb_image, 3 channels): (H, W, 3) uint8."""
    px = C.POINTER(C.c_uint8)()
    w, h = C.c_uint32(), C.c_uint32()
    if lib().ref_read_png(os.fsencode(path), C.byref(px), C.byref(w), C.byref(h)) != 0:
        raise ValueError(f"dort could not read {path}")
    img = np.ctypeslib.as_array(px, (3 * w.value * h.value + 1,))[: 3 * w.value * h.value]
    img = img.reshape(h.value, w.value, 3).copy()
    lib().ref_free(px)
    return img


def write_png(path, img):
    """dort's write_image_png (stb_image_write) of an (H, W, 3) uint8 image."""
    a = np.ascontiguousarray(img, dtype=np.uint8)
    if lib().ref_write_png(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0]) != 0:
        raise ValueError(f"dort could not write {path}")


def test_convergence(reference, tested, min_tile_size=32, variation=8.0, bias=0.0, p_value=0.01):
    """dort's test_convergence on one thread: "" or the first failing tile."""
    a = np.ascontiguousarray(reference, dtype=np.float32)
    b = np.ascontiguousarray(tested, dtype=np.float32)
    msg = C.create_string_buffer(512)
    ok = lib().ref_test_convergence(a.ctypes.data, b.ctypes.data, a.shape[1], a.shape[0], min_tile_size,
                                    variation, bias, p_value, msg, len(msg))
    return "" if ok else msg.value.decode()


class RefScene:
    """dort's Scene, assembled as its Lua builder does (middle split, leaf 6, max 32)."""

    def __init__(self, scene, split="middle", leaf_size=6, max_leaf_size=32, threads=1):
        self.packed = scene.pack()
        self.h = lib().ref_scene_create(C.addressof(self.packed.desc), 0 if split == "middle" else 1,
                                        leaf_size, max_leaf_size, threads)
        if not self.h:
            raise ValueError("scene not expressible in dort (GGX conductor or texture)")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_scene_destroy(self.h)
            self.h = None

    def trace(self, rays, any_hit=False):
        n = len(rays["ox"])
        arrs = [np.ascontiguousarray(rays[k], dtype=np.float32)
                for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]
        ptrs = (C.c_void_p * 8)(*[a.ctypes.data for a in arrs])
        if any_hit:
            occ = np.zeros(n, np.uint32)
            lib().ref_trace(self.h, n, ptrs, 1, None, None, occ.ctypes.data, None, None)
            return occ
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.int32)
        uv3 = np.zeros((n, 3), np.float32)
        pt = np.zeros((n, 3), np.float32)
        lib().ref_trace(self.h, n, ptrs, 0, t.ctypes.data, prim.ctypes.data, None, uv3.ctypes.data,
                        pt.ctypes.data)
        return {"t": t, "prim": prim, "uv": uv3[:, :2], "eps": uv3[:, 2], "p": pt}

    def render(self, opts, threads=1):
        """PathRenderer::render with RandomSampler(Rng(seed)); (rgb_sum, weight).
        threads > 1 runs into dort's thread-pool race (see the module docstring)."""
        rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        w = np.zeros((opts.y_res, opts.x_res), np.float32)
        if lib().ref_render(self.h, C.byref(opts), threads, rgb.ctypes.data, w.ctypes.data) != 0:
            raise RuntimeError("ref_render failed")
        return rgb, w

    def render_light(self, opts, threads=1):
        """LightRenderer::render (dort's `lt`) with RandomSampler(Rng(seed)): the film's
        splat sums (H, W, 3); the image is the sum / iterations."""
        rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        if lib().ref_render_light(self.h, C.byref(opts), threads, rgb.ctypes.data) != 0:
            raise RuntimeError("ref_render_light failed")
        return rgb

    def render_bdpt(self, opts, use_t1_paths=True, threads=1):
        """BdptRenderer::render (dort's `bdpt`) with RandomSampler(Rng(seed)): the film's
        colour sums (H, W, 3), weights (H, W) and splat sums (H, W, 3)."""
        rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        w = np.zeros((opts.y_res, opts.x_res), np.float32)
        splat = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        if lib().ref_render_bdpt(self.h, C.byref(opts), threads, 1 if use_t1_paths else 0, rgb.ctypes.data,
                                 w.ctypes.data, splat.ctypes.data) != 0:
            raise RuntimeError("ref_render_bdpt failed")
        return rgb, w, splat

    def bdpt_image(self, opts, use_t1_paths=True):
        rgb, w, splat = self.render_bdpt(opts, use_t1_paths)
        img = np.where(w[..., None] > 0, rgb / np.maximum(w, 1e-30)[..., None], 0.0)
        return (img + splat / float(opts.sample_end - opts.sample_begin)).astype(np.float32)


def gpu_lib():
    global _gpu_lib
    if _gpu_lib is None:
        L = C.CDLL(GPU_LIB_PATH)
        L.ref_scene_create.restype = C.c_void_p
        L.ref_scene_create.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.ref_scene_destroy.argtypes = [C.c_void_p]
        L.refgpu_render.restype = C.c_int
        L.refgpu_render.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(abi.RenderOpts), C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_float), C.c_char_p, C.c_size_t]
        L.refgpu_render_bdpt.restype = C.c_int
        L.refgpu_render_bdpt.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(abi.RenderOpts), C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.POINTER(C.c_float), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                         C.c_char_p, C.c_size_t]
        L.refgpu_render_light.restype = C.c_int
        L.refgpu_render_light.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(abi.RenderOpts), C.c_uint32,
                                          C.c_uint32, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]
        _gpu_lib = L
    return _gpu_lib


class GpuRendererInDort:
    """A dort::Scene plus dort_amd::GpuPathRenderer plugged into dort's Renderer seam."""

    def __init__(self, scene):
        self.packed = scene.pack()
        self.h = gpu_lib().ref_scene_create(C.addressof(self.packed.desc), 0, 6, 32, 1)
        if not self.h:
            raise ValueError("scene not expressible in dort (GGX conductor or texture)")

    def __del__(self):
        if getattr(self, "h", None):
            gpu_lib().ref_scene_destroy(self.h)
            self.h = None

    def render(self, opts, sampler_seed=42, cancel_after=0, chunk=0, filter_kind=0):
        """Renderer::render of the GPU renderer; returns (rgb_sum, weight, info dict).
        Raises RuntimeError with the renderer's exception message."""
        rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        w = np.zeros((opts.y_res, opts.x_res), np.float32)
        seed, paths, pct = C.c_uint64(), C.c_uint64(), C.c_float()
        err = C.create_string_buffer(512)
        rc = gpu_lib().refgpu_render(self.h, C.addressof(self.packed.desc), C.byref(opts), sampler_seed,
                                     cancel_after, chunk, filter_kind, rgb.ctypes.data, w.ctypes.data,
                                     C.byref(seed), C.byref(paths), C.byref(pct), err, len(err))
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return rgb, w, {"seed": seed.value, "paths": paths.value, "percent": pct.value}

    def render_light(self, opts, sampler_seed=42, chunk=0):
        """Renderer::render of GpuLightRenderer (dort's lt); returns (splat sums, info dict)."""
        rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        scale, seed, paths = C.c_float(), C.c_uint64(), C.c_uint64()
        err = C.create_string_buffer(512)
        rc = gpu_lib().refgpu_render_light(self.h, C.addressof(self.packed.desc), C.byref(opts), sampler_seed,
                                           chunk, rgb.ctypes.data, C.byref(scale), C.byref(seed), C.byref(paths),
                                           err, len(err))
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return rgb, {"splat_scale": scale.value, "seed": seed.value, "paths": paths.value}

    def render_bdpt(self, opts, sampler_seed=42, chunk=0, use_t1_paths=True):
        """Renderer::render of GpuBdptRenderer (dort's bdpt); returns (colour sums, weights,
        splat sums, info dict)."""
        rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        w = np.zeros((opts.y_res, opts.x_res), np.float32)
        splat = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        scale, seed, paths = C.c_float(), C.c_uint64(), C.c_uint64()
        err = C.create_string_buffer(512)
        rc = gpu_lib().refgpu_render_bdpt(self.h, C.addressof(self.packed.desc), C.byref(opts), sampler_seed, chunk,
                                          1 if use_t1_paths else 0, rgb.ctypes.data, w.ctypes.data,
                                          splat.ctypes.data, C.byref(scale), C.byref(seed), C.byref(paths), err,
                                          len(err))
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return rgb, w, splat, {"seed": seed.value, "paths": paths.value, "splat_scale": scale.value}
