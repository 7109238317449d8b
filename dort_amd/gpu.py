"""ctypes binding of libdortgpu.so (include/dortgpu.h).

This is the Python side of the drop-in boundary: the same entry points a C++
host (dort's `Renderer` seam, include/dort/renderer.hpp:12-34) calls.  It fails
loudly when the HIP library is missing; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdortgpu.so")

_lib = None

EXPORTS = ["dgpu_abi_version", "dgpu_last_error", "dgpu_create", "dgpu_destroy", "dgpu_device_alloc",
           "dgpu_device_free", "dgpu_copy", "dgpu_scene_create", "dgpu_scene_destroy",
           "dgpu_scene_get_info", "dgpu_scene_update_texture", "dgpu_scene_update_material",
           "dgpu_trace_closest", "dgpu_trace_any", "dgpu_trace_closest_noalloc", "dgpu_render", "dgpu_render_light", "dgpu_render_batch",
           "dgpu_render_bdpt", "dgpu_adjoint",
           "dgpu_read_ply", "dgpu_free", "dgpu_read_rgbe", "dgpu_write_rgbe", "dgpu_read_png", "dgpu_write_png",
           "dgpu_test_convergence", "dgpu_bias_variance"]


class DgpuError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dort_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.dgpu_abi_version.restype = C.c_uint32
        L.dgpu_last_error.restype = C.c_char_p
        L.dgpu_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.dgpu_destroy.argtypes = [C.c_void_p]
        L.dgpu_device_alloc.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]
        L.dgpu_device_free.argtypes = [C.c_void_p, C.c_void_p]
        L.dgpu_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.dgpu_scene_create.argtypes = [C.c_void_p, C.POINTER(abi.SceneDesc), C.POINTER(abi.BvhOpts),
                                        C.POINTER(C.c_void_p)]
        L.dgpu_scene_destroy.argtypes = [C.c_void_p]
        L.dgpu_scene_get_info.argtypes = [C.c_void_p, C.POINTER(abi.SceneInfo)]
        L.dgpu_trace_closest.argtypes = [C.c_void_p, C.POINTER(abi.Rays), C.c_uint32,
                                         C.POINTER(abi.Hits), C.c_uint32, C.POINTER(C.c_double)]
        L.dgpu_trace_any.argtypes = [C.c_void_p, C.POINTER(abi.Rays), C.c_uint32, C.c_void_p,
                                     C.c_uint32, C.POINTER(C.c_double)]
        L.dgpu_trace_closest_noalloc.argtypes = [C.c_void_p, C.POINTER(abi.Rays), C.c_uint32,
                                                 C.POINTER(abi.Hits), C.c_uint32, C.POINTER(C.c_double)]
        L.dgpu_render.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_void_p, C.c_void_p,
                                  C.POINTER(abi.RenderStats), C.c_void_p]
        L.dgpu_render_light.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_void_p,
                                        C.POINTER(abi.RenderStats), C.c_void_p]
        L.dgpu_render_batch.argtypes = [C.c_void_p, C.POINTER(abi.BatchItem), C.c_uint32, C.c_uint32,
                                        C.POINTER(abi.RenderStats)]
        L.dgpu_render_bdpt.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_uint32, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.POINTER(abi.RenderStats), C.c_void_p]
        L.dgpu_scene_update_texture.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.dgpu_scene_update_material.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(abi.Material)]
        L.dgpu_adjoint.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.POINTER(abi.RenderStats)]
        L.dgpu_read_ply.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                                    C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint32)]
        L.dgpu_free.argtypes = [C.c_void_p]
        L.dgpu_free.restype = None
        L.dgpu_read_rgbe.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32)]
        L.dgpu_write_rgbe.argtypes = [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]
        L.dgpu_read_png.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]
        L.dgpu_write_png.argtypes = [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]
        L.dgpu_test_convergence.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32,
                                            C.c_float, C.c_float, C.c_float, C.c_uint32, C.POINTER(C.c_int32),
                                            C.c_char_p, C.c_size_t]
        L.dgpu_bias_variance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
        for n in ("dgpu_read_ply", "dgpu_read_rgbe", "dgpu_write_rgbe", "dgpu_read_png", "dgpu_write_png",
                  "dgpu_test_convergence", "dgpu_bias_variance"):
            getattr(L, n).restype = C.c_int
        for n in ("dgpu_create", "dgpu_device_alloc", "dgpu_device_free", "dgpu_copy",
                  "dgpu_scene_create", "dgpu_scene_get_info", "dgpu_scene_update_texture",
                  "dgpu_scene_update_material", "dgpu_trace_closest", "dgpu_trace_any", "dgpu_render",
                  "dgpu_render_light", "dgpu_render_bdpt",
                  "dgpu_render_batch", "dgpu_adjoint"):
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def check(status):
    if status != abi.DGPU_OK:
        raise DgpuError(status, lib().dgpu_last_error().decode())


def read_ply(path):
    """dgpu_read_ply: (positions (V, 3) float32, indices (T, 3) uint32), faces fan-triangulated."""
    pos, ind = C.POINTER(C.c_float)(), C.POINTER(C.c_uint32)()
    nv, nt = C.c_uint32(), C.c_uint32()
    check(lib().dgpu_read_ply(os.fsencode(path), C.byref(pos), C.byref(nv), C.byref(ind), C.byref(nt)))
    try:
        p = np.ctypeslib.as_array(pos, (max(1, nv.value) * 3,))[: 3 * nv.value].reshape(-1, 3).copy()
        t = np.ctypeslib.as_array(ind, (max(1, nt.value) * 3,))[: 3 * nt.value].reshape(-1, 3).copy()
    finally:
        lib().dgpu_free(pos)
        lib().dgpu_free(ind)
    return p, t


def read_rgbe(path):
    """dgpu_read_rgbe: (H, W, 3) float32."""
    px = C.POINTER(C.c_float)()
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().dgpu_read_rgbe(os.fsencode(path), C.byref(px), C.byref(w), C.byref(h)))
    try:
        img = np.ctypeslib.as_array(px, (h.value * w.value * 3,)).reshape(h.value, w.value, 3).copy()
    finally:
        lib().dgpu_free(px)
    return img


def write_rgbe(path, img):
    """dgpu_write_rgbe of an (H, W, 3) image."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    check(lib().dgpu_write_rgbe(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0]))


def read_png(path):
    """dgpu_read_png: (H, W, 3) uint8."""
    px = C.POINTER(C.c_uint8)()
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().dgpu_read_png(os.fsencode(path), C.byref(px), C.byref(w), C.byref(h)))
    try:
        img = np.ctypeslib.as_array(px, (h.value * w.value * 3,)).reshape(h.value, w.value, 3).copy()
    finally:
        lib().dgpu_free(px)
    return img


def write_png(path, img):
    """dgpu_write_png of an (H, W, 3) uint8 image."""
    a = np.ascontiguousarray(img, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("write_png takes an (H, W, 3) image")
    check(lib().dgpu_write_png(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0]))


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().dgpu_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().dgpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def scene(self, scene, **bvh):
        return Scene(self, scene, **bvh)

    def render_batch(self, items, films=None, device=False, graph=True):
        """Renders [(Scene, RenderOpts), ...] in one batched wavefront.  Returns
        ([(film_rgb, film_w), ...], stats); films (optional) are accumulated into."""
        n = len(items)
        arr = (abi.BatchItem * n)()
        if films is None:
            films = [(np.zeros((o.y_res, o.x_res, 3), np.float32), np.zeros((o.y_res, o.x_res), np.float32))
                     for _, o in items]
        ptr = lambda a: a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data
        for i, ((sc, o), (rgb, w)) in enumerate(zip(items, films)):
            arr[i].scene = sc.h.value
            arr[i].opts = o
            arr[i].film_rgb_sum = ptr(rgb)
            arr[i].film_weight = ptr(w)
        flags = (abi.BUF_DEVICE if device else 0) | (0 if graph else abi.BATCH_NO_GRAPH)
        st = abi.RenderStats()
        check(lib().dgpu_render_batch(self.h, arr, n, flags, C.byref(st)))
        return films, abi.stats_dict(st)

    def test_convergence(self, reference, tested, min_tile_size=32, variation=8.0, bias=0.0, p_value=0.01):
        """dgpu_test_convergence on (H, W, 3) float32 images (numpy on the host, or
        DeviceArrays): "" when the test passes, else dort's description of the first
        failing tile."""
        dev = isinstance(reference, DeviceArray)
        if dev:
            h, w = tested.shape[0], tested.shape[1]
            rp, tp = reference.ptr, tested.ptr
        else:
            reference = np.ascontiguousarray(reference, dtype=np.float32)
            tested = np.ascontiguousarray(tested, dtype=np.float32)
            h, w = reference.shape[:2]
            rp, tp = reference.ctypes.data, tested.ctypes.data
        ok = C.c_int32()
        msg = C.create_string_buffer(512)
        check(lib().dgpu_test_convergence(self.h, rp, tp, w, h, min_tile_size, variation, bias, p_value,
                                          abi.BUF_DEVICE if dev else 0, C.byref(ok), msg, len(msg)))
        return "" if ok.value else msg.value.decode()

    def bias_variance(self, reference, tested, rect=None):
        """dgpu_bias_variance over rect = (x0, y0, x1, y1) (default: the whole image)."""
        reference = np.ascontiguousarray(reference, dtype=np.float32)
        tested = np.ascontiguousarray(tested, dtype=np.float32)
        h, w = reference.shape[:2]
        x0, y0, x1, y1 = rect if rect is not None else (0, 0, w, h)
        b, v = C.c_double(), C.c_double()
        check(lib().dgpu_bias_variance(self.h, reference.ctypes.data, tested.ctypes.data, w, h, x0, y0, x1, y1,
                                       0, C.byref(b), C.byref(v)))
        return b.value, v.value

    def device_array(self, host):
        """Copies a numpy array into a new DeviceArray on this context's GPU."""
        d = DeviceArray(self, host.shape, host.dtype)
        d.copy_from(host)
        return d


class DeviceArray:
    """A device allocation of the library (HBM), with numpy round trips.  `data_ptr()`
    gives the raw device address for DGPU_BUF_DEVICE calls."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(lib().dgpu_device_alloc(ctx.h, self.nbytes, C.byref(self.ptr)))

    def data_ptr(self):
        return self.ptr.value

    def __len__(self):
        return self.shape[0]

    def copy_from(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes
        check(lib().dgpu_copy(self.ctx.h, self.ptr, host.ctypes.data, self.nbytes, 0))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().dgpu_copy(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def free(self):
        if self.ptr:
            lib().dgpu_device_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scene:
    """A scene uploaded to the GPU (BVH built on the host, flattened to the GPU layout)."""

    def __init__(self, ctx, scene, width=4, max_leaf_size=4, sah_bins=16, stack_lds=16,
                 block_size=128, queue_tile=0, sah_node_cost=0.0):
        self.ctx = ctx
        self.packed = scene.pack()
        self.scene = scene
        self.opts = abi.BvhOpts(width, max_leaf_size, sah_bins, stack_lds, block_size, queue_tile, sah_node_cost)
        self.h = C.c_void_p()
        check(lib().dgpu_scene_create(ctx.h, C.byref(self.packed.desc), C.byref(self.opts),
                                      C.byref(self.h)))

    def close(self):
        if self.h:
            lib().dgpu_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def info(self):
        i = abi.SceneInfo()
        check(lib().dgpu_scene_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in abi.SceneInfo._fields_}

    def trace_closest(self, rays, out=None, device=False, bounded=False):
        """rays: dict of 8 SoA arrays (numpy on host or torch tensors on the GPU).
        bounded: the origins lie within the scene's coordinate range (DGPU_TRACE_BOUNDED)."""
        n = len(rays["ox"])
        r = abi.rays_struct(rays)
        if out is None:
            out = {"t": np.zeros(n, np.float32), "prim": np.zeros(n, np.int32),
                   "b1": np.zeros(n, np.float32), "b2": np.zeros(n, np.float32)}
        ptr = lambda a: a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data
        h = abi.Hits(ptr(out["t"]), ptr(out["prim"]), ptr(out["b1"]), ptr(out["b2"]))
        ms = C.c_double()
        flags = (abi.BUF_DEVICE if device else 0) | (abi.TRACE_BOUNDED if bounded else 0)
        check(lib().dgpu_trace_closest(self.h, C.byref(r), n, C.byref(h), flags, C.byref(ms)))
        out["kernel_ms"] = ms.value
        return out

    def trace_any(self, rays, out=None, device=False, bounded=False):
        n = len(rays["ox"])
        r = abi.rays_struct(rays)
        if out is None:
            out = np.zeros(n, np.uint32)
        ptr = out.data_ptr() if hasattr(out, "data_ptr") else out.ctypes.data
        ms = C.c_double()
        flags = (abi.BUF_DEVICE if device else 0) | (abi.TRACE_BOUNDED if bounded else 0)
        check(lib().dgpu_trace_any(self.h, C.byref(r), n, ptr, flags, C.byref(ms)))
        return out, ms.value

    def render(self, opts, film_rgb=None, film_w=None, device=False, cancel=None):
        """Adds samples to (film_rgb (H,W,3), film_w (H,W)); returns (film_rgb, film_w, stats)."""
        if film_rgb is None:
            film_rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
            film_w = np.zeros((opts.y_res, opts.x_res), np.float32)
        ptr = lambda a: a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data
        o = abi.RenderOpts.from_buffer_copy(opts)
        if device:
            o.flags |= abi.BUF_DEVICE
        st = abi.RenderStats()
        cptr = None
        if cancel is not None:
            cptr = cancel.ctypes.data
        check(lib().dgpu_render(self.h, C.byref(o), ptr(film_rgb), ptr(film_w), C.byref(st), cptr))
        stats = abi.stats_dict(st)
        return film_rgb, film_w, stats

    def image(self, opts):
        rgb, w, _ = self.render(opts)
        return rgb / np.maximum(w, 1e-30)[..., None]

    def render_light(self, opts, film_rgb=None, device=False, cancel=None):
        """dort's light tracer: adds splats to film_rgb (H,W,3); returns (film_rgb, stats)."""
        if film_rgb is None:
            film_rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        ptr = lambda a: a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data
        o = abi.RenderOpts.from_buffer_copy(opts)
        if device:
            o.flags |= abi.BUF_DEVICE
        st = abi.RenderStats()
        cptr = cancel.ctypes.data if cancel is not None else None
        check(lib().dgpu_render_light(self.h, C.byref(o), ptr(film_rgb), C.byref(st), cptr))
        return film_rgb, abi.stats_dict(st)

    def light_image(self, opts):
        """The light tracer's image: splats / iterations (Film::to_image's splat_scale)."""
        rgb, _ = self.render_light(opts)
        return rgb / float(opts.sample_end - opts.sample_begin)

    def render_bdpt(self, opts, use_t1_paths=True, films=None, device=False, cancel=None, megakernel=False):
        """dort's bidirectional path tracer (the wavefront stages, or one thread per path sample
        with megakernel=True): adds to films (rgb (H,W,3), weight (H,W), splat (H,W,3)); returns
        (rgb, weight, splat, stats)."""
        if films is None:
            films = (np.zeros((opts.y_res, opts.x_res, 3), np.float32),
                     np.zeros((opts.y_res, opts.x_res), np.float32),
                     np.zeros((opts.y_res, opts.x_res, 3), np.float32))
        rgb, w, splat = films
        ptr = lambda a: a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data
        o = abi.RenderOpts.from_buffer_copy(opts)
        if device:
            o.flags |= abi.BUF_DEVICE
        if megakernel:
            o.flags |= abi.BDPT_MEGAKERNEL
        st = abi.RenderStats()
        cptr = cancel.ctypes.data if cancel is not None else None
        check(lib().dgpu_render_bdpt(self.h, C.byref(o), 1 if use_t1_paths else 0, ptr(rgb), ptr(w), ptr(splat),
                                     C.byref(st), cptr))
        return rgb, w, splat, abi.stats_dict(st)

    def bdpt_image(self, opts, use_t1_paths=True, megakernel=False):
        """The bidirectional image as dort's film gives it: rgb / weight + splat / iterations."""
        rgb, w, splat, _ = self.render_bdpt(opts, use_t1_paths, megakernel=megakernel)
        img = np.where(w[..., None] > 0, rgb / np.maximum(w, 1e-30)[..., None], 0.0)
        return (img + splat / float(opts.sample_end - opts.sample_begin)).astype(np.float32)

    def update_texture(self, tex, rgb):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        check(lib().dgpu_scene_update_texture(self.h, tex, rgb.ctypes.data))

    def update_material(self, index, material):
        check(lib().dgpu_scene_update_material(self.h, index, C.byref(material)))

    def adjoint(self, opts, dldi, want_texture=True, want_roughness=True, want_image=False, megakernel=False):
        """Gradient of sum(dldi * image) w.r.t. texture-0 texels and material roughness, by
        the two wavefront passes (or one thread per path with megakernel=True).
        Returns (g_tex (h, w, 3) or None, g_rough (num_materials,) or None, image or None, stats)."""
        if megakernel:
            opts = abi.RenderOpts.from_buffer_copy(opts)
            opts.flags |= abi.ADJOINT_MEGAKERNEL
        dldi = np.ascontiguousarray(dldi, dtype=np.float32)
        assert dldi.shape == (opts.y_res, opts.x_res, 3)
        g_tex = None
        if want_texture and self.scene.textures:
            g_tex = np.zeros(self.scene.textures[0].shape, np.float32)
        g_rough = np.zeros(len(self.scene.materials), np.float32) if want_roughness else None
        img = np.zeros((opts.y_res, opts.x_res, 3), np.float32) if want_image else None
        ptr = lambda a: None if a is None else a.ctypes.data
        st = abi.RenderStats()
        check(lib().dgpu_adjoint(self.h, C.byref(opts), dldi.ctypes.data, ptr(g_tex), ptr(g_rough),
                                 ptr(img), C.byref(st)))
        stats = abi.stats_dict(st)
        return g_tex, g_rough, img, stats
