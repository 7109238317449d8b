"""ctypes mirror of include/dortgpu.h (struct layouts and constants).

The structs are plain C data; the same descriptor feeds the HIP library and
the CPU oracle used by the tests, so both render the identical scene.
"""
import ctypes as C

import numpy as np

DGPU_OK, DGPU_EINVAL, DGPU_ENOMEM, DGPU_EDEVICE, DGPU_EUNSUPPORTED, DGPU_ECANCELLED = range(6)
STATUS_NAMES = {0: "DGPU_OK", 1: "DGPU_EINVAL", 2: "DGPU_ENOMEM", 3: "DGPU_EDEVICE",
                4: "DGPU_EUNSUPPORTED", 5: "DGPU_ECANCELLED"}

MAT_NONE, MAT_LAMBERT, MAT_GGX_CONDUCTOR, MAT_DIELECTRIC, MAT_ROUGH_DIELECTRIC = range(5)
DIRECT_MIS, DIRECT_BSDF, DIRECT_LIGHT = range(3)
BUF_DEVICE = 1
RENDER_PROFILE = 2
BATCH_NO_GRAPH = 4
ADJOINT_MEGAKERNEL = 16
BDPT_MEGAKERNEL = 32
TRACE_BOUNDED = 8

f3 = C.c_float * 3


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("albedo_tex", C.c_int32), ("albedo", f3),
                ("roughness", C.c_float), ("eta", f3), ("k", f3),
                ("ior_inside", C.c_float), ("ior_outside", C.c_float),
                ("reflect_tint", f3), ("transmit_tint", f3)]


class AreaLight(C.Structure):
    _fields_ = [("triangle", C.c_uint32), ("radiance", f3)]


class EnvLight(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_float)),
                ("up", f3), ("fwd", f3), ("scale", f3)]


TEXMAP_UV, TEXMAP_XY = 0, 1


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_float)),
                ("rgb8", C.POINTER(C.c_uint8)), ("mapping", C.c_uint32),
                ("world_to_texture", C.c_float * 12)]


class Camera(C.Structure):
    _fields_ = [("camera_to_world", C.c_float * 12), ("fov", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32),
                ("positions", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_uint32)),
                ("uvs", C.POINTER(C.c_float)), ("tri_material", C.POINTER(C.c_uint32)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(Material)),
                ("num_area_lights", C.c_uint32), ("area_lights", C.POINTER(AreaLight)),
                ("env", C.POINTER(EnvLight)), ("infinite_radiance", C.POINTER(C.c_float)),
                ("num_textures", C.c_uint32), ("textures", C.POINTER(Texture)),
                ("camera", Camera)]


class BvhOpts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("max_leaf_size", C.c_uint32), ("sah_bins", C.c_uint32),
                ("stack_lds", C.c_uint32), ("block_size", C.c_uint32), ("queue_tile", C.c_uint32),
                ("sah_node_cost", C.c_float)]


class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class Hits(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("t", "prim", "b1", "b2")]


class RenderOpts(C.Structure):
    _fields_ = [("x_res", C.c_uint32), ("y_res", C.c_uint32),
                ("sample_begin", C.c_uint32), ("sample_end", C.c_uint32),
                ("seed", C.c_uint64), ("min_depth", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("only_direct", C.c_uint32),
                ("sample_all_lights", C.c_uint32), ("direct_strategy", C.c_uint32),
                ("flags", C.c_uint32), ("scene_id", C.c_uint32)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("opts", RenderOpts), ("film_rgb_sum", C.c_void_p),
                ("film_weight", C.c_void_p)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("dropped_samples", C.c_uint64),
                ("kernel_ms", C.c_double), ("extend_ms", C.c_double), ("shade_ms", C.c_double),
                ("connect_ms", C.c_double), ("other_ms", C.c_double),
                ("extend_launches", C.c_uint32), ("shade_launches", C.c_uint32),
                ("connect_launches", C.c_uint32), ("other_launches", C.c_uint32),
                ("bounce_extension_rays", C.c_uint64 * 16), ("bounce_shadow_rays", C.c_uint64 * 16)]


def stats_dict(st):
    """A RenderStats as a dict (arrays as lists)."""
    return {f: (list(getattr(st, f)) if f.startswith("bounce_") else getattr(st, f))
            for f, _ in RenderStats._fields_}


class SceneInfo(C.Structure):
    _fields_ = [("nodes", C.c_uint32), ("leaves", C.c_uint32), ("depth", C.c_uint32),
                ("width", C.c_uint32), ("node_bytes", C.c_uint64), ("triangle_bytes", C.c_uint64),
                ("build_ms", C.c_double), ("upload_ms", C.c_double)]


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def render_opts(x_res, y_res, spp=1, sample_begin=0, seed=1, min_depth=0, max_depth=5,
                rr_depth=2, only_direct=False, sample_all_lights=False,
                direct_strategy=DIRECT_MIS, flags=0, scene_id=0):
    """Options of dort's `pt` renderer (lua_render.cpp:125-145) plus rr_depth."""
    return RenderOpts(x_res, y_res, sample_begin, sample_begin + spp, seed, min_depth, max_depth,
                      rr_depth, int(only_direct), int(sample_all_lights), direct_strategy, flags,
                      scene_id)


class PackedScene:
    """A scene packed into a SceneDesc; keeps every numpy buffer it points into alive."""

    def __init__(self, scene):
        s = scene
        self.scene = s
        self._keep = []
        pos = np.ascontiguousarray(s.positions, dtype=np.float32)
        idx = np.ascontiguousarray(s.indices, dtype=np.uint32)
        tmat = np.ascontiguousarray(s.tri_material, dtype=np.uint32)
        self._keep += [pos, idx, tmat]
        d = SceneDesc()
        d.num_vertices = pos.shape[0]
        d.num_triangles = idx.shape[0]
        d.positions = fptr(pos)
        d.indices = uptr(idx)
        if s.uvs is not None:
            uv = np.ascontiguousarray(s.uvs, dtype=np.float32)
            self._keep.append(uv)
            d.uvs = fptr(uv)
        d.tri_material = uptr(tmat)
        mats = (Material * max(1, len(s.materials)))()
        for i, m in enumerate(s.materials):
            mats[i] = m
        self._keep.append(mats)
        d.num_materials = len(s.materials)
        d.materials = C.cast(mats, C.POINTER(Material))
        lights = (AreaLight * max(1, len(s.area_lights)))()
        for i, (tri, rad) in enumerate(s.area_lights):
            lights[i] = AreaLight(int(tri), f3(*rad))
        self._keep.append(lights)
        d.num_area_lights = len(s.area_lights)
        d.area_lights = C.cast(lights, C.POINTER(AreaLight))
        if s.env is not None:
            img = np.ascontiguousarray(s.env["image"], dtype=np.float32)
            self._keep.append(img)
            env = EnvLight(img.shape[1], img.shape[0], fptr(img), f3(*s.env["up"]),
                           f3(*s.env["fwd"]), f3(*s.env["scale"]))
            self._keep.append(env)
            d.env = C.pointer(env)
        if s.infinite is not None:
            inf = np.ascontiguousarray(s.infinite, dtype=np.float32)
            self._keep.append(inf)
            d.infinite_radiance = fptr(inf)
        texs = (Texture * max(1, len(s.textures)))()
        for i, t in enumerate(s.textures):
            # an (H, W, 3) image (float32 texels, or uint8 read as k / 256) addressed by the
            # surface uv, or {"image": ..., "mapping": "xy", "world_to_texture": (3, 4)}
            spec = t if isinstance(t, dict) else {"image": t}
            img = np.asarray(spec["image"])
            tx = Texture()
            tx.width, tx.height = img.shape[1], img.shape[0]
            if img.dtype == np.uint8:
                img = np.ascontiguousarray(img)
                tx.rgb8 = img.ctypes.data_as(C.POINTER(C.c_uint8))
            else:
                img = np.ascontiguousarray(img, dtype=np.float32)
                tx.rgb = fptr(img)
            self._keep.append(img)
            if spec.get("mapping", "uv") == "xy":
                tx.mapping = TEXMAP_XY
                w2t = np.asarray(spec["world_to_texture"], dtype=np.float32).reshape(12)
                for k in range(12):
                    tx.world_to_texture[k] = float(w2t[k])
            texs[i] = tx
        self._keep.append(texs)
        d.num_textures = len(s.textures)
        d.textures = C.cast(texs, C.POINTER(Texture))
        cam = Camera()
        c2w = np.asarray(s.camera_to_world, dtype=np.float32).reshape(12)
        for i in range(12):
            cam.camera_to_world[i] = float(c2w[i])
        cam.fov = float(s.fov)
        d.camera = cam
        self.desc = d

    @property
    def ptr(self):
        return C.byref(self.desc)


def rays_struct(arrs):
    """SoA rays from a dict of 8 arrays (numpy or torch) -> Rays (raw pointers)."""
    r = Rays()
    for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax"):
        a = arrs[n]
        setattr(r, n, a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data)
    return r
