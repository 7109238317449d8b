"""Seeded scene generators.

Every scene is plain arrays (triangles, per-triangle material, material table,
area lights on triangles, optional environment map / infinite light, pinhole
camera).  The same arrays are fed to the HIP library and to the CPU oracle, so
both render the identical scene.

Transform conventions follow dort's `Transform` (src/dort/transform.cpp):
`look_at` (:146-165) builds camera_to_world with columns (right, up, dir, eye);
`rotate_y` (:104-114) maps x' = cos*x + sin*z, z' = -sin*x + cos*z.
"""
from dataclasses import dataclass, field, replace

import numpy as np

from . import abi


@dataclass
class Scene:
    positions: np.ndarray            # (V, 3) float32
    indices: np.ndarray              # (T, 3) uint32
    tri_material: np.ndarray         # (T,) uint32
    materials: list                  # [abi.Material]
    area_lights: list                # [(triangle index, (r, g, b))]
    camera_to_world: np.ndarray      # (3, 4) float32
    fov: float
    uvs: np.ndarray = None           # (V, 2) float32 or None
    env: dict = None                 # {"image": (H, W, 3), "up", "fwd", "scale"}
    infinite: tuple = None           # rgb of an InfiniteLight
    textures: list = field(default_factory=list)
    name: str = ""

    @property
    def num_triangles(self):
        return int(self.indices.shape[0])

    def pack(self):
        return abi.PackedScene(self)


# ---------------------------------------------------------------- materials

def lambert(albedo, tex=-1):
    m = abi.Material()
    m.type = abi.MAT_LAMBERT
    m.albedo_tex = tex
    m.albedo = abi.f3(*albedo)
    return m


def black():
    """The black Lambert dort attaches to area-light shapes (builder_ext.lua:38-48)."""
    m = abi.Material()
    m.type = abi.MAT_NONE
    m.albedo_tex = -1
    return m


# Complex IOR (n, k) at ~ (650, 550, 450) nm for a few conductors.
CONDUCTORS = {
    "gold": ((0.166, 0.424, 1.38), (3.15, 2.35, 1.77)),
    "copper": ((0.214, 0.958, 1.17), (3.67, 2.58, 2.40)),
    "aluminium": ((1.66, 0.880, 0.521), (9.22, 6.27, 4.84)),
    "silver": ((0.155, 0.124, 0.136), (4.83, 3.52, 2.47)),
}


def ggx_conductor(roughness, metal="aluminium", tint=(1.0, 1.0, 1.0)):
    m = abi.Material()
    m.type = abi.MAT_GGX_CONDUCTOR
    m.albedo_tex = -1
    m.albedo = abi.f3(*tint)
    m.roughness = float(roughness)
    eta, k = CONDUCTORS[metal]
    m.eta = abi.f3(*eta)
    m.k = abi.f3(*k)
    return m


def dielectric(ior_inside=1.5, ior_outside=1.0, reflect=(1, 1, 1), transmit=(1, 1, 1)):
    m = abi.Material()
    m.type = abi.MAT_DIELECTRIC
    m.albedo_tex = -1
    m.ior_inside = ior_inside
    m.ior_outside = ior_outside
    m.reflect_tint = abi.f3(*reflect)
    m.transmit_tint = abi.f3(*transmit)
    return m


def rough_dielectric(roughness, ior_inside=1.5, ior_outside=1.0, reflect=(1, 1, 1), transmit=(1, 1, 1)):
    """dort's rough dielectric (rough_dielectric_material.cpp) as dort runs it."""
    m = dielectric(ior_inside, ior_outside, reflect, transmit)
    m.type = abi.MAT_ROUGH_DIELECTRIC
    m.roughness = float(roughness)
    return m


# ---------------------------------------------------------------- transforms

def _normalize(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def look_at(eye, look, up):
    """dort `look_at` (transform.cpp:146-165) as a 4x4 camera_to_world matrix."""
    eye = np.asarray(eye, dtype=np.float64)
    d = _normalize(np.asarray(look, dtype=np.float64) - eye)
    right = _normalize(np.cross(_normalize(up), d))
    u = np.cross(d, right)
    m = np.eye(4)
    m[:3, 0] = right
    m[:3, 1] = u
    m[:3, 2] = d
    m[:3, 3] = eye
    return m


def scale_m(x, y, z):
    return np.diag([x, y, z, 1.0])


def translate_m(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def rotate_y_m(a):
    c, s = np.cos(a), np.sin(a)
    m = np.eye(4)
    m[0, 0] = c
    m[0, 2] = s
    m[2, 0] = -s
    m[2, 2] = c
    return m


def apply_points(m, p):
    p = np.asarray(p, dtype=np.float64)
    return p @ m[:3, :3].T + m[:3, 3]


def camera34(m):
    return np.asarray(m[:3, :4], dtype=np.float32)


# ---------------------------------------------------------------- builder

class _Builder:
    def __init__(self):
        self.pos = []
        self.idx = []
        self.mat = []
        self.nv = 0
        self.materials = []
        self.area_lights = []
        self.uvs = []
        self.has_uv = False

    def material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def mesh(self, points, tris, mat, uvs=None):
        points = np.asarray(points, dtype=np.float32).reshape(-1, 3)
        tris = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
        first_tri = sum(len(i) for i in self.idx)
        self.pos.append(points)
        self.idx.append(tris + self.nv)
        self.mat.append(np.full(len(tris), mat, dtype=np.uint32))
        if uvs is not None:
            self.has_uv = True
            self.uvs.append(np.asarray(uvs, dtype=np.float32).reshape(-1, 2))
        else:
            self.uvs.append(None)
        self.nv += len(points)
        return first_tri

    def build(self, camera_to_world, fov, **kw):
        pos = np.concatenate(self.pos).astype(np.float32)
        idx = np.concatenate(self.idx).astype(np.uint32)
        mat = np.concatenate(self.mat).astype(np.uint32)
        uvs = None
        if self.has_uv:
            uvs = np.concatenate([u if u is not None else np.zeros((len(p), 2), np.float32)
                                  for u, p in zip(self.uvs, self.pos)]).astype(np.float32)
        return Scene(positions=pos, indices=idx, tri_material=mat, materials=self.materials,
                     area_lights=self.area_lights, camera_to_world=camera34(camera_to_world),
                     fov=fov, uvs=uvs, **kw)


def _quad(p0, p1, p2, p3):
    return [p0, p1, p2, p3], [[0, 1, 2], [0, 2, 3]]


def _cube_tris():
    """The cube [-1,1]^3 as 8 corners and 12 triangles (cube_shape.cpp:99-101 bounds)."""
    pts = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    # corner index = 4*ix + 2*iy + iz
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, c, d in faces:
        tris += [[a, b, c], [a, c, d]]
    return pts, np.array(tris)


def _light_quad_facing(b, corners, normal, radiance, mat):
    """Adds two one-sided light triangles whose normalize(cross(e2, e1)) is `normal`."""
    p = np.asarray(corners, dtype=np.float64)
    tris = [[0, 1, 2], [0, 2, 3]]
    e1 = p[1] - p[0]
    e2 = p[2] - p[0]
    if np.dot(np.cross(e2, e1), normal) < 0:
        tris = [[0, 2, 1], [0, 3, 2]]
    first = b.mesh(p, tris, mat)
    b.area_lights.append((first, tuple(radiance)))
    b.area_lights.append((first + 1, tuple(radiance)))


# ---------------------------------------------------------------- config 0

def cornell(seed=0):
    """Procedural Cornell box in the unit cube: 10 wall tris, a short and a tall box
    (12 tris each, rotated -18 and +18 degrees), a 0.25 x 0.25 ceiling light of
    radiance (17, 12, 4) as 2 one-sided triangles.  36 triangles in total."""
    del seed  # fully determined
    b = _Builder()
    white = b.material(lambert((0.73, 0.73, 0.73)))
    red = b.material(lambert((0.65, 0.05, 0.05)))
    green = b.material(lambert((0.12, 0.45, 0.15)))
    blk = b.material(black())
    q = lambda *ps: _quad(*ps)
    for pts, mat in [
        (q((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)), white),    # floor
        (q((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)), white),    # ceiling
        (q((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)), white),    # back
        (q((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)), red),      # left
        (q((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)), green),    # right
    ]:
        b.mesh(pts[0], pts[1], mat)
    cpts, ctris = _cube_tris()
    for (cx, cz, h, ang) in [(0.33, 0.37, 0.3, np.radians(-18)), (0.66, 0.66, 0.6, np.radians(18))]:
        m = translate_m(cx, h / 2, cz) @ rotate_y_m(ang) @ scale_m(0.15, h / 2, 0.15)
        b.mesh(apply_points(m, cpts), ctris, white)
    y = 1.0 - 1e-3
    _light_quad_facing(b, [(0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625)],
                       (0, -1, 0), (17.0, 12.0, 4.0), blk)
    cam = look_at((0.5, 0.5, -1.4), (0.5, 0.5, 0.0), (0, 1, 0)) @ scale_m(1, -1, 1)
    return b.build(cam, 0.69, name="cornell")


def glass_cornell():
    """The Cornell box with its tall block made a smooth dielectric of IOR 1.5
    (dielectric_material.cpp:32-104): camera paths through it are delta bounces."""
    sc = cornell()
    mats = list(sc.materials) + [dielectric(1.5)]
    tri_mat = sc.tri_material.copy()
    tri_mat[22:34] = len(mats) - 1  # 10 wall + 12 short-box triangles come first
    return replace(sc, materials=mats, tri_material=tri_mat)


def rough_glass_cornell(roughness=0.2):
    """The Cornell box with its tall block made of dort's rough dielectric (IOR 1.5, GGX
    alpha = roughness) and its short block a smooth dielectric: glossy reflection sampled
    with the Fresnel probability, transmission reached only through light sampling
    (light paths that enter the block through its smooth-glass neighbour's caustics and
    the walls' bounces)."""
    sc = cornell()
    mats = list(sc.materials) + [rough_dielectric(roughness, 1.5), dielectric(1.5)]
    tri_mat = sc.tri_material.copy()
    tri_mat[22:34] = len(mats) - 2   # tall block
    tri_mat[10:22] = len(mats) - 1   # short block
    # the blocks' triangles are wound with normalize(cross(e2, e1)) pointing into the
    # block; the rough block is rewound so its normals point out, which puts the camera
    # and the light on the lobe's outer (z > 0) side, where its reflection is nonzero
    idx = sc.indices.copy()
    idx[22:34] = idx[22:34][:, [0, 2, 1]]
    return replace(sc, materials=mats, tri_material=tri_mat, indices=idx, name=f"rough_glass_cornell_{roughness}")


def rough_pane(roughness=0.3):
    """A square pane of dort's rough dielectric (IOR 1.5) standing between the camera
    and a square area light of radiance (20, 18, 14), over a grey Lambert floor: light
    reaches the camera through the pane by light sampling (the lobe's transmission term)
    and is reflected off it toward the floor."""
    b = _Builder()
    glass = b.material(rough_dielectric(roughness, 1.5))
    floor = b.material(lambert((0.5, 0.5, 0.5)))
    blk = b.material(black())
    b.mesh([(-3, 0, -3), (3, 0, -3), (3, 0, 3), (-3, 0, 3)], [[0, 2, 1], [0, 3, 2]], floor)
    # the pane in the plane z = 0, normal -z (toward the camera)
    b.mesh([(-1, 0.2, 0), (1, 0.2, 0), (1, 1.8, 0), (-1, 1.8, 0)], [[0, 2, 1], [0, 3, 2]], glass)
    _light_quad_facing(b, [(-0.5, 0.6, 1.5), (0.5, 0.6, 1.5), (0.5, 1.4, 1.5), (-0.5, 1.4, 1.5)],
                       (0, 0, -1), (20.0, 18.0, 14.0), blk)
    cam = look_at((0.4, 1.1, -3.0), (0, 0.9, 0), (0, 1, 0)) @ scale_m(1, -1, 1)
    return b.build(cam, 0.8, name=f"rough_pane_{roughness}")


def textured_cornell(seed=0, texels="u8"):
    """The Cornell box with image textures on two walls, as dort's ImageTexture composes
    with its texture maps (image_texture.cpp:5-11, texture_maps.cpp:4-15): the floor
    carries a 16 x 16 texture through uv_texture_map (vertex uvs span [0, 2]^2, so it
    wraps twice), the back wall an 8 x 12 texture through xy_texture_map (rotated by 30
    degrees about z, scaled by 2.5, shifted).  Texels are seeded 8-bit values, given as
    uint8 (texels="u8") or as the float32 values k / 256 they read as ("f32").  The other
    surfaces get all-zero vertex uvs, which takes the degenerate-uv branch of
    TriangleUv::hit's tangent (triangle.cpp:137-171)."""
    sc = cornell()
    rng = np.random.default_rng(seed)
    t_floor = rng.integers(0, 256, size=(16, 16, 3), dtype=np.uint8)
    t_back = rng.integers(0, 256, size=(8, 12, 3), dtype=np.uint8)
    if texels == "f32":
        t_floor, t_back = t_floor.astype(np.float32) / 256, t_back.astype(np.float32) / 256
    a = np.radians(30.0)
    rot = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
    w2t = np.zeros((3, 4))
    w2t[:, :3] = 2.5 * rot
    w2t[:, 3] = (0.3, -0.7, 0.0)
    mats = list(sc.materials) + [lambert((0, 0, 0), tex=0), lambert((0, 0, 0), tex=1)]
    tri_mat = sc.tri_material.copy()
    tri_mat[0:2] = len(mats) - 2   # floor
    tri_mat[4:6] = len(mats) - 1   # back wall
    uvs = np.zeros((len(sc.positions), 2), np.float32)
    uvs[0:4] = [(0, 0), (2, 0), (2, 2), (0, 2)]  # the floor quad's vertices come first
    return replace(sc, materials=mats, tri_material=tri_mat, uvs=uvs, name="textured_cornell",
                   textures=[t_floor, {"image": t_back, "mapping": "xy", "world_to_texture": w2t}])


# ------------------------------------------------- dort's own test box scene

def dort_test_box(light_kind="area", open_box=False):
    """The Cornell scene of dort's test suite (test/test_box.lua:1-197) at scale 0.01,
    surface kind "diff", pinhole camera, with the analytic `cube()` shapes
    triangulated exactly.  light_kind "area" (2 triangles of radiance 100) or
    "infi" (InfiniteLight of radiance 1)."""
    s = 0.01
    b = _Builder()
    white = b.material(lambert((0.5, 0.5, 0.5)))
    green = b.material(lambert((0.0, 0.5, 0.0)))
    red = b.material(lambert((0.5, 0.0, 0.0)))
    blk = b.material(black())
    pts = np.array([(552.8, 0, 0), (0, 0, 0), (0, 0, 559.2), (549.6, 0, 559.2),
                    (556, 548.8, 0), (556, 548.8, 559.2), (0, 548.8, 559.2), (0, 548.8, 0)]) * s
    verts = [0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7, 6, 5, 3, 6, 3, 2, 7, 6, 2, 7, 2, 1, 0, 3, 5, 0, 5, 4]
    tri = lambda i: [verts[i], verts[i + 1], verts[i + 2]]
    b.mesh(pts, [tri(0), tri(3)], white)
    if not open_box:
        b.mesh(pts, [tri(6), tri(9), tri(12), tri(15)], white)
        b.mesh(pts, [tri(18), tri(21)], green)
        b.mesh(pts, [tri(24), tri(27)], red)
    cpts, ctris = _cube_tris()
    left = translate_m(185 * s, 82.5 * s, 169 * s) @ rotate_y_m(-0.29) @ scale_m(*(3 * [165 * s / 2]))
    right = (translate_m(368 * s, 165 * s, 351 * s) @ rotate_y_m(-1.27)
             @ scale_m(165 * s / 2, 330 * s / 2, 165 * s / 2))
    b.mesh(apply_points(left, cpts), ctris, white)
    b.mesh(apply_points(right, cpts), ctris, white)
    cam = look_at((278 * s, 273 * s, -800 * s), (278 * s, 273 * s, 0), (0, 1, 0)) @ scale_m(1, -1, 1)
    kw = {}
    if light_kind == "area":
        lp = np.array([(343, 548.8, 227), (343, 548.8, 332), (213, 548.8, 332), (213, 548.8, 227)]) * s
        # The block's translate(0, -s, 0) moves the light twice in dort: mesh{} applies the
        # builder transform to the points (dsl.lua:148-153, lua_shape.cpp:171) and
        # add_shape(triangle, light) places the shape under it again
        # (lua_builder.cpp:327-346); DiffuseLight samples it under the same transform.
        lp[:, 1] -= 2 * s
        first = b.mesh(lp, [[2, 1, 0], [3, 2, 0]], blk)
        b.area_lights += [(first, (100.0, 100.0, 100.0)), (first + 1, (100.0, 100.0, 100.0))]
    elif light_kind == "infi":
        kw["infinite"] = (1.0, 1.0, 1.0)
    else:
        raise ValueError(light_kind)
    return b.build(cam, 0.686, name=f"dort_test_box_{light_kind}", **kw)


# ---------------------------------------------------------------- noise

class ValueNoise:
    """Seeded 3D value noise with smoothstep interpolation, fBm over octaves."""

    def __init__(self, seed):
        rng = np.random.default_rng(seed)
        self.perm = rng.permutation(256).astype(np.int64)
        self.vals = rng.uniform(-1.0, 1.0, 256)

    def _lattice(self, ix, iy, iz):
        p = self.perm
        h = p[(p[(p[ix & 255] + iy) & 255] + iz) & 255]
        return self.vals[h]

    def noise(self, x, y, z):
        fx, fy, fz = np.floor(x), np.floor(y), np.floor(z)
        ix, iy, iz = fx.astype(np.int64), fy.astype(np.int64), fz.astype(np.int64)
        tx, ty, tz = x - fx, y - fy, z - fz
        sx, sy, sz = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty), tz * tz * (3 - 2 * tz)
        out = 0.0
        for dx in (0, 1):
            wx = sx if dx else 1 - sx
            for dy in (0, 1):
                wy = sy if dy else 1 - sy
                for dz in (0, 1):
                    wz = sz if dz else 1 - sz
                    out = out + wx * wy * wz * self._lattice(ix + dx, iy + dy, iz + dz)
        return out

    def fbm(self, p, octaves=6, freq=2.0, lacunarity=2.0, gain=0.5):
        p = np.asarray(p, dtype=np.float64)
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        amp, total = 1.0, 0.0
        for _ in range(octaves):
            total = total + amp * self.noise(x * freq, y * freq, z * freq)
            freq *= lacunarity
            amp *= gain
        return total


# ---------------------------------------------------------------- meshes

_ICO_V = None


def _icosahedron():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]])
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def geodesic_sphere(freq):
    """Class-I geodesic sphere: each icosahedron face split into freq^2 triangles,
    20*freq^2 triangles and 10*freq^2+2 shared vertices (freq = 2^k gives the
    k-times subdivided icosphere)."""
    v, f = _icosahedron()
    n = freq
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    mask = ii + jj <= n
    ii, jj = ii[mask], jj[mask]
    # local index of lattice point (i, j) on a face
    loc = -np.ones((n + 1, n + 1), dtype=np.int64)
    loc[ii, jj] = np.arange(len(ii))
    # local triangles
    lt = []
    for i in range(n):
        j = np.arange(n - i)
        lt.append(np.stack([loc[i, j], loc[i + 1, j], loc[i, j + 1]], 1))
        j2 = np.arange(n - i - 1)
        lt.append(np.stack([loc[i + 1, j2], loc[i + 1, j2 + 1], loc[i, j2 + 1]], 1))
    lt = np.concatenate(lt)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    w1 = ii / n
    w2 = jj / n
    w0 = 1 - w1 - w2
    pts = (w0[None, :, None] * a[:, None, :] + w1[None, :, None] * b[:, None, :]
           + w2[None, :, None] * c[:, None, :]).reshape(-1, 3)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    npf = len(ii)
    tris = (lt[None, :, :] + (np.arange(20) * npf)[:, None, None]).reshape(-1, 3)
    # weld duplicated edge/corner vertices
    key = np.round(pts * (1 << 20)).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    pts = pts[first]
    tris = inv.reshape(-1)[tris]
    return pts, tris


def displaced_sphere(freq, seed, amp=0.1, radius=1.0, center=(0, 0, 0)):
    p, t = geodesic_sphere(freq)
    r = radius * (1.0 + amp * ValueNoise(seed).fbm(p * 1.0, octaves=6))
    return (p * r[:, None] + np.asarray(center)).astype(np.float32), t.astype(np.uint32)


def displaced_box(n, seed, amp=0.05, half=1.0, center=(0, 0, 0)):
    """Cube with each face as an n x n grid (12 n^2 triangles), radially displaced."""
    faces = []
    g = np.linspace(-1, 1, n + 1)
    uu, vv = np.meshgrid(g, g, indexing="ij")
    uu, vv = uu.reshape(-1), vv.reshape(-1)
    base = np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]
    base = base.reshape(-1)
    lt = np.concatenate([np.stack([base, base + n + 1, base + n + 2], 1),
                         np.stack([base, base + n + 2, base + 1], 1)])
    pts, tris, off = [], [], 0
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            q = np.zeros((len(uu), 3))
            q[:, axis] = sgn
            q[:, (axis + 1) % 3] = uu
            q[:, (axis + 2) % 3] = vv
            pts.append(q)
            tris.append(lt + off)
            off += len(q)
    p = np.concatenate(pts)
    t = np.concatenate(tris)
    key = np.round(p * (1 << 20)).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    p = p[first]
    t = inv.reshape(-1)[t]
    d = 1.0 + amp * ValueNoise(seed).fbm(p, octaves=6)
    return (p * half * d[:, None] + np.asarray(center)).astype(np.float32), t.astype(np.uint32)


# ---------------------------------------------------------------- config 1

def icosphere_scene(freq=256, seed=1):
    """Displaced icosphere (20*freq^2 triangles; freq 256 -> 1,310,720) with r = 1 + 0.1 fBm_6."""
    p, t = displaced_sphere(freq, seed)
    b = _Builder()
    m = b.material(lambert((0.7, 0.7, 0.7)))
    b.mesh(p, t, m)
    cam = look_at((0.0, 0.6, -3.2), (0, 0, 0), (0, 1, 0)) @ scale_m(1, -1, 1)
    return b.build(cam, 0.9, name=f"icosphere_{freq}", infinite=(1.0, 1.0, 1.0))


def camera_rays(scene, x_res, y_res):
    """Pinhole rays through pixel centres, pinhole_camera.cpp:163-181."""
    c2w = np.asarray(scene.camera_to_world, dtype=np.float64)
    proj = 2.0 * np.tan(0.5 * scene.fov)
    xs = (np.arange(x_res) + 0.5) / x_res - 0.5
    ys = (np.arange(y_res) + 0.5) / y_res - 0.5
    cx, cy = np.meshgrid(xs, ys)
    if x_res > y_res:
        cy = cy * y_res / x_res
    else:
        cx = cx * x_res / y_res
    d = np.stack([cx.ravel() * proj, cy.ravel() * proj, np.ones(cx.size)], 1)
    w = d @ c2w[:3, :3].T
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    o = np.broadcast_to(c2w[:3, 3], w.shape)
    return o.astype(np.float32), w.astype(np.float32)


def random_rays(scene, n, seed, box_scale=3.0):
    """Origins uniform in the scene bounds scaled `box_scale`x about the centre,
    directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    lo, hi = scene.positions.min(0).astype(np.float64), scene.positions.max(0).astype(np.float64)
    c, h = (lo + hi) / 2, (hi - lo) / 2 * box_scale
    o = c + (rng.random((n, 3)) * 2 - 1) * h
    z = 1 - 2 * rng.random(n)
    r = np.sqrt(np.maximum(0, 1 - z * z))
    phi = 2 * np.pi * rng.random(n)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    return o.astype(np.float32), d.astype(np.float32)


def ray_batch(o, d, tmin=0.0, tmax=np.inf):
    n = len(o)
    return {"ox": np.ascontiguousarray(o[:, 0]), "oy": np.ascontiguousarray(o[:, 1]),
            "oz": np.ascontiguousarray(o[:, 2]), "dx": np.ascontiguousarray(d[:, 0]),
            "dy": np.ascontiguousarray(d[:, 1]), "dz": np.ascontiguousarray(d[:, 2]),
            "tmin": np.full(n, tmin, np.float32), "tmax": np.full(n, tmax, np.float32)}


def intersection_workload(scene, x_res, y_res, n_random, seed):
    """Half pinhole camera rays (pixel centres), half random rays, tmax = inf."""
    o1, d1 = camera_rays(scene, x_res, y_res)
    o2, d2 = random_rays(scene, n_random, seed)
    return ray_batch(np.concatenate([o1, o2]), np.concatenate([d1, d2]))


# ---------------------------------------------------------------- sky

def procedural_sky(width=1024, height=512, sun_dir=(0.4, 0.55, 0.3), sun_intensity=50.0,
                   sun_size=0.03):
    """Lat-long FP32 HDR sky: zenith-to-horizon gradient, dark ground, and a sun lobe
    of peak radiance `sun_intensity`.  Row y <-> theta = (y + 0.5)/H * pi from `up`
    (environment_light.cpp:163-175 with up = +y, s = +z-ish frame of up/fwd)."""
    up = np.array([0.0, 1.0, 0.0])
    fwd = np.array([1.0, 0.0, 0.0])
    t_ax = np.cross(up, fwd)
    t_ax /= np.linalg.norm(t_ax)
    s_ax = np.cross(up, t_ax)
    theta = (np.arange(height) + 0.5) / height * np.pi
    phi = ((np.arange(width) + 0.5) / width - 0.5) * 2 * np.pi
    th, ph = np.meshgrid(theta, phi, indexing="ij")
    dirs = (up[None, None, :] * np.cos(th)[..., None]
            + s_ax[None, None, :] * (np.sin(th) * np.cos(ph))[..., None]
            + t_ax[None, None, :] * (np.sin(th) * np.sin(ph))[..., None])
    cos_up = np.cos(th)
    zenith = np.array([0.25, 0.45, 1.0])
    horizon = np.array([1.0, 0.95, 0.9])
    ground = np.array([0.08, 0.07, 0.06])
    k = np.clip(cos_up, 0, 1)[..., None]
    sky = horizon * (1 - k) ** 3 + zenith * (1 - (1 - k) ** 3)
    img = np.where(cos_up[..., None] > 0, sky, ground * np.ones(3))
    sd = np.asarray(sun_dir, dtype=np.float64)
    sd /= np.linalg.norm(sd)
    ang = np.arccos(np.clip(dirs @ sd, -1, 1))
    img = img + sun_intensity * np.exp(-(ang / sun_size) ** 2)[..., None] * np.array([1.0, 0.95, 0.85])
    return {"image": img.astype(np.float32), "up": tuple(up), "fwd": tuple(fwd),
            "scale": (1.0, 1.0, 1.0)}


# ---------------------------------------------------------------- configs 2 / 4

def _draw_material(rng):
    r = rng.random()
    if r < 0.45:
        return lambert(tuple(rng.uniform(0.1, 0.9, 3)))
    if r < 0.8:
        metal = ["gold", "copper", "aluminium", "silver"][rng.integers(4)]
        return ggx_conductor(rng.uniform(0.05, 0.5), metal)
    return dielectric(1.5)


def terrain_scene(quads=1000, seed=2, patches=4, sky=None, textured=False, tex_res=256,
                  sky_res=(1024, 512), ggx_roughness=None):
    """Displaced-noise heightfield of quads x quads cells (2*quads^2 triangles) over
    [-1,1]^2 split into patches x patches objects, plus a 2-triangle ground plane.
    Per-object materials come from `seed` (Lambert albedo U(0.1,0.9), GGX conductor
    roughness U(0.05,0.5), dielectric IOR 1.5).  With `textured`, even patches are
    Lambert with a tex_res^2 RGB albedo texture (terrain uv) and odd patches are GGX
    conductors with per-object roughness (`ggx_roughness`, 8 values for 4x4 patches)."""
    rng = np.random.default_rng(seed)
    noise = ValueNoise(seed)
    n = quads
    g = np.linspace(-1.0, 1.0, n + 1)
    xx, zz = np.meshgrid(g, g, indexing="ij")
    xx, zz = xx.ravel(), zz.ravel()
    h = 0.25 * noise.fbm(np.stack([xx, np.zeros_like(xx), zz], 1), octaves=6, freq=1.5)
    pts = np.stack([xx, h, zz], 1).astype(np.float32)
    uvs = np.stack([(xx + 1) / 2, (zz + 1) / 2], 1).astype(np.float32)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    v00 = (i * (n + 1) + j).ravel()
    v10 = v00 + (n + 1)
    t1 = np.stack([v00, v10, v10 + 1], 1)
    t2 = np.stack([v00, v10 + 1, v00 + 1], 1)
    tris = np.stack([t1, t2], 1).reshape(-1, 3)
    cell_i = np.repeat((i * np.ones_like(j)).ravel(), 2)
    cell_j = np.repeat((np.ones_like(i) * j).ravel(), 2)
    patch = (cell_i * patches // n) * patches + (cell_j * patches // n)
    b = _Builder()
    mats = []
    textures = []
    if textured:
        trng = np.random.default_rng(seed + 1000)
        tex = ValueNoise(seed + 7).fbm(np.stack(np.meshgrid(np.arange(tex_res), np.arange(tex_res),
                                                             [0.0], indexing="ij"), -1).reshape(-1, 3)
                                       / tex_res * 8.0, octaves=4, freq=1.0)
        tex = tex.reshape(tex_res, tex_res)
        base = trng.uniform(0.2, 0.8, 3)
        img = np.clip(base[None, None, :] + 0.3 * tex[..., None] * np.array([1.0, 0.8, 0.6]), 0.05, 0.95)
        textures.append(img.astype(np.float32))
        rough = ggx_roughness if ggx_roughness is not None else trng.uniform(0.05, 0.5, patches * patches // 2)
        k = 0
        for pidx in range(patches * patches):
            if pidx % 2 == 0:
                mats.append(lambert((1, 1, 1), tex=0))
            else:
                mats.append(ggx_conductor(float(rough[k]), ["gold", "copper", "aluminium", "silver"][k % 4]))
                k += 1
    else:
        for pidx in range(patches * patches):
            mats.append(_draw_material(rng))
    mids = [b.material(m) for m in mats]
    # terrain: one mesh, per-triangle object material
    b.pos.append(pts)
    b.idx.append(tris.astype(np.int64))
    b.mat.append(np.asarray(mids, dtype=np.uint32)[patch])
    b.uvs.append(uvs)
    b.has_uv = True
    b.nv += len(pts)
    ground_mat = b.material(lambert(tuple(rng.uniform(0.1, 0.9, 3))))
    gy = float(h.min()) - 0.05
    b.mesh([(-10, gy, -10), (10, gy, -10), (10, gy, 10), (-10, gy, 10)], [[0, 1, 2], [0, 2, 3]],
           ground_mat, uvs=[(0, 0), (1, 0), (1, 1), (0, 1)])
    if sky is None:
        sky = procedural_sky(*sky_res)
    cam = look_at((0.0, 1.1, -2.3), (0.0, -0.1, 0.1), (0, 1, 0)) @ scale_m(1, -1, 1)
    return b.build(cam, 0.75, name=f"terrain_{quads}", env=sky, textures=textures)


# ---------------------------------------------------------------- config 3

def batch_scenes(count=64, seed=3, min_tris=1000, max_tris=500000, sky_res=(1024, 512)):
    """`count` independent scenes, triangle counts log-uniform in [min_tris, max_tris]
    (displaced geodesic spheres and boxes on a ground plane), resolutions uniform in
    {256^2, 512^2, 768^2}, materials as in the terrain scene, lit by a procedural sky of
    the terrain's resolution (1024 x 512, sun lobe of peak 50) with a seeded sun."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        target = int(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        b = _Builder()
        if rng.random() < 0.5:
            freq = max(1, int(round(np.sqrt(target / 20.0))))
            p, t = displaced_sphere(freq, seed * 1000 + k, amp=0.1, radius=0.8, center=(0, 0.8, 0))
        else:
            nq = max(1, int(round(np.sqrt(target / 12.0))))
            p, t = displaced_box(nq, seed * 1000 + k, amp=0.05, half=0.6, center=(0, 0.6, 0))
        m = b.material(_draw_material(rng))
        b.mesh(p, t, m)
        gm = b.material(lambert(tuple(rng.uniform(0.1, 0.9, 3))))
        b.mesh([(-10, 0, -10), (10, 0, -10), (10, 0, 10), (-10, 0, 10)], [[0, 1, 2], [0, 2, 3]], gm)
        sun = (rng.uniform(-1, 1), rng.uniform(0.3, 1.0), rng.uniform(-1, 1))
        sky = procedural_sky(*sky_res, sun_dir=sun, sun_size=0.05)
        cam = look_at((0.0, 1.6, -3.0), (0, 0.6, 0), (0, 1, 0)) @ scale_m(1, -1, 1)
        sc = b.build(cam, 0.8, name=f"batch_{k}", env=sky)
        res = [256, 512, 768][rng.integers(3)]
        out.append((sc, res))
    return out
