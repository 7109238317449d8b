"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a CPU restatement of dort's `pt` path (see dort_oracle.cpp).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module,
as the checker or the timed CPU baseline; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from dort_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# DORT_ORACLE_LIB names another build of the oracle (the sanitizer build of
# oracle/Makefile's `san` target, run by tests/test_sanitizers.py)
LIB_PATH = os.environ.get("DORT_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_scene_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_uint32)] * 3
        L.orc_trace.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.c_int, C.c_int,
                                C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.POINTER(C.c_uint64)]
        L.orc_render.restype = C.c_int
        L.orc_render.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_int, C.c_uint32,
                                 C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]
        for n in ("orc_adjoint64", "orc_adjoint32"):
            f = getattr(L, n)
            f.restype = C.c_int
            f.argtypes = [C.c_void_p, C.POINTER(abi.RenderOpts), C.c_uint32, C.c_void_p, C.c_void_p,
                          C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_bsdf_sample.argtypes = [C.POINTER(abi.Material), C.c_uint32, C.c_void_p, C.c_void_p,
                                      C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_bsdf_eval.argtypes = [C.POINTER(abi.Material), C.c_uint32, C.c_void_p, C.c_void_p,
                                    C.c_uint32, C.c_void_p, C.c_void_p]
        L.orc_env_sample.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_env_eval.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_lights.restype = C.c_uint32
        L.orc_lights.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_fresnel_conductor.argtypes = [C.c_uint32, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_void_p]
        L.orc_fresnel_dielectric.argtypes = [C.c_uint32, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_void_p]
        L.orc_philox.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def default_threads():
    return max(1, min(os.cpu_count() or 1, 16))


BSDF_ALL = 31
BSDF_NON_DELTA = 31 & ~16


class OracleScene:
    """dort's scene: BVH built by dort's rules (middle split, leaf 6, max leaf 32)."""

    def __init__(self, scene, split="middle", leaf_size=6, max_leaf_size=32, fp64=False):
        self.packed = scene.pack()
        self.scene = scene
        self.h = lib().orc_scene_create(C.addressof(self.packed.desc), 0 if split == "middle" else 1,
                                        leaf_size, max_leaf_size, int(fp64))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def info(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        lib().orc_scene_info(self.h, C.byref(a), C.byref(b), C.byref(c))
        return {"nodes": a.value, "depth": b.value, "lights": c.value}

    def trace(self, rays, any_hit=False, brute=False, threads=None):
        n = len(rays["ox"])
        arrs = [np.ascontiguousarray(rays[k], dtype=np.float32)
                for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]
        ptrs = (C.c_void_p * 8)(*[a.ctypes.data for a in arrs])
        vis = C.c_uint64()
        if any_hit:
            occ = np.zeros(n, np.uint32)
            lib().orc_trace(self.h, n, ptrs, 1, int(brute), threads or default_threads(),
                            None, None, None, None, occ.ctypes.data, C.byref(vis))
            return occ
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.int32)
        b1 = np.zeros(n, np.float32)
        b2 = np.zeros(n, np.float32)
        lib().orc_trace(self.h, n, ptrs, 0, int(brute), threads or default_threads(),
                        t.ctypes.data, prim.ctypes.data, b1.ctypes.data, b2.ctypes.data, None,
                        C.byref(vis))
        return {"t": t, "prim": prim, "b1": b1, "b2": b2, "visited": vis.value}

    def render(self, opts, rng="counter", threads=None):
        """Returns (rgb_sum (H,W,3), weight (H,W), counters dict)."""
        rgb = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        w = np.zeros((opts.y_res, opts.x_res), np.float32)
        cnt = (C.c_uint64 * 4)()
        rc = lib().orc_render(self.h, C.byref(opts), 1 if rng == "counter" else 0,
                              threads or default_threads(), rgb.ctypes.data, w.ctypes.data, cnt)
        if rc != 0:
            raise RuntimeError("orc_render failed")
        return rgb, w, {"paths": cnt[0], "extension_rays": cnt[1], "shadow_rays": cnt[2],
                        "dropped": cnt[3]}

    def image(self, opts, rng="counter", threads=None):
        rgb, w, _ = self.render(opts, rng, threads)
        return rgb / np.maximum(w, 1e-30)[..., None]

    def adjoint(self, opts, dldi, n_texels, n_mats, fp64=True, threads=None):
        dldi = np.ascontiguousarray(dldi, dtype=np.float32)
        gt = np.zeros(n_texels * 3, np.float64)
        gr = np.zeros(n_mats, np.float64)
        img = np.zeros((opts.y_res, opts.x_res, 3), np.float32)
        f = lib().orc_adjoint64 if fp64 else lib().orc_adjoint32
        rc = f(self.h, C.byref(opts), threads or default_threads(), dldi.ctypes.data, gt.ctypes.data,
               n_texels, gr.ctypes.data, n_mats, img.ctypes.data)
        if rc != 0:
            raise RuntimeError("orc_adjoint failed (fp64 scene requires fp64=True at creation)")
        return gt, gr, img

    def env_sample(self, uv):
        uv = np.ascontiguousarray(uv, dtype=np.float32)
        n = len(uv)
        d = np.zeros((n, 3), np.float32)
        p = np.zeros(n, np.float32)
        r = np.zeros((n, 3), np.float32)
        lib().orc_env_sample(self.h, n, uv.ctypes.data, d.ctypes.data, p.ctypes.data, r.ctypes.data)
        return d, p, r

    def env_eval(self, dirs):
        dirs = np.ascontiguousarray(dirs, dtype=np.float32)
        n = len(dirs)
        p = np.zeros(n, np.float32)
        r = np.zeros((n, 3), np.float32)
        lib().orc_env_eval(self.h, n, dirs.ctypes.data, p.ctypes.data, r.ctypes.data)
        return p, r

    def lights(self):
        n = lib().orc_lights(self.h, None, None)
        kinds = np.zeros(n, np.int32)
        cdf = np.zeros(n + 1, np.float32)
        lib().orc_lights(self.h, kinds.ctypes.data, cdf.ctypes.data)
        return kinds, cdf


def bsdf_sample(material, wo, uvc, request=BSDF_ALL):
    wo = np.ascontiguousarray(wo, dtype=np.float32)
    uvc = np.ascontiguousarray(uvc, dtype=np.float32)
    n = len(wo)
    wi = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    fl = np.zeros(n, np.uint32)
    lib().orc_bsdf_sample(C.byref(material), n, wo.ctypes.data, uvc.ctypes.data, request,
                          wi.ctypes.data, f.ctypes.data, pdf.ctypes.data, fl.ctypes.data)
    return wi, f, pdf, fl


def bsdf_eval(material, wi, wo, request=BSDF_NON_DELTA):
    wi = np.ascontiguousarray(wi, dtype=np.float32)
    wo = np.ascontiguousarray(wo, dtype=np.float32)
    n = len(wi)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().orc_bsdf_eval(C.byref(material), n, wi.ctypes.data, wo.ctypes.data, request,
                        f.ctypes.data, pdf.ctypes.data)
    return f, pdf


def fresnel_conductor(cos_i, eta, k, fp64=False):
    """The oracle's conductor Fresnel reflectance at each cos_i (complex IOR eta + i k)."""
    c = np.ascontiguousarray(cos_i, dtype=np.float64)
    out = np.zeros_like(c)
    lib().orc_fresnel_conductor(len(c), c.ctypes.data, float(eta), float(k), int(fp64), out.ctypes.data)
    return out


def fresnel_dielectric(cos_i, ior_refl, ior_trans, fp64=False):
    """The oracle's dielectric Fresnel reflectance (dort's fresnel.cpp) at each cos_i."""
    c = np.ascontiguousarray(cos_i, dtype=np.float64)
    out = np.zeros_like(c)
    lib().orc_fresnel_dielectric(len(c), c.ctypes.data, float(ior_refl), float(ior_trans), int(fp64),
                                 out.ctypes.data)
    return out


def philox4x32_10(ctr, key):
    """include/dgpu_rng.h's Philox4x32-10 of one (4-word counter, 2-word key)."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out
