"""Scenes: seeded synthetic geometry and the vgi_scene handle.

The synthetic scene follows the benchmark description: axis-aligned boxes
(12 triangles) and subdivided icospheres (level 2: 320, level 3: 1280
triangles) resting on a ground plane (Z up, as in violet's app.rs:92-103),
albedo U[0.05, 0.95] per object, metallic 0, perceptual roughness
U[0.3, 1.0] per object (the workload does not fix roughness; this is the
recorded choice).  Icospheres carry smooth per-vertex normals, boxes and the
plane flat ones (each face has its own vertices).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

from . import _lib
from .synth import uniform


class Material(C.Structure):
    _fields_ = [("base_color", C.c_float * 4), ("metallic", C.c_float), ("roughness", C.c_float),
                ("color_metalrough_texture", C.c_uint32), ("normal_texture", C.c_uint32)]


class TextureDesc(C.Structure):
    """vgi_texture: RGBA8 mip chain in host memory."""

    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("levels", C.c_uint32), ("srgb", C.c_uint32),
                ("texels", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("num_vertices", C.c_uint32),
                ("indices", C.c_void_p), ("num_triangles", C.c_uint32), ("triangle_material", C.c_void_p),
                ("materials", C.POINTER(Material)), ("num_materials", C.c_uint32),
                ("texcoords", C.c_void_p), ("tangents", C.c_void_p), ("textures", C.POINTER(TextureDesc)),
                ("num_textures", C.c_uint32)]


TEXTURE_WHITE = 0  # VGI_TEXTURE_WHITE
TEXTURE_NORMAL_UP = 1  # VGI_TEXTURE_NORMAL_UP
TEXTURE_FIRST = 2  # VGI_TEXTURE_FIRST


@dataclass
class Texture:
    """An RGBA8 texture with its mip chain: levels[l] is (max(1, h >> l),
    max(1, w >> l), 4) uint8.  srgb marks base-colour textures."""

    levels: list
    srgb: bool = False

    @property
    def width(self) -> int:
        return self.levels[0].shape[1]

    @property
    def height(self) -> int:
        return self.levels[0].shape[0]

    def packed(self) -> np.ndarray:
        return np.concatenate([np.ascontiguousarray(l, np.uint8).reshape(-1) for l in self.levels])


def mip_count(width: int, height: int) -> int:
    """violet's level count: 1 + log2(next power of two of max(w, h))
    (model.rs:853-857)."""
    m = max(width, height)
    return 1 + (m - 1).bit_length()


def box_mips(level0: np.ndarray, levels: int | None = None) -> list:
    """A mip chain by 2x2 box averaging (rounded), halving with a floor of 1
    as violet's importer does; odd sizes drop the last row/column.  A simple
    chain for synthetic scenes (violet's own importer filters with
    Lanczos3)."""
    h, w = level0.shape[:2]
    n = mip_count(w, h) if levels is None else levels
    out = [np.ascontiguousarray(level0, np.uint8)]
    for _ in range(1, n):
        p = out[-1].astype(np.uint32)
        ph, pw = p.shape[:2]
        nh, nw = max(1, ph // 2), max(1, pw // 2)
        acc = np.zeros((nh, nw, p.shape[2]), np.uint32)
        for r in ((p[0:2 * nh:2], p[1:2 * nh:2]) if ph > 1 else (p, p)):
            for c in ((r[:, 0:2 * nw:2], r[:, 1:2 * nw:2]) if pw > 1 else (r, r)):
                acc += c
        out.append(((acc + 2) // 4).astype(np.uint8))
    return out


class SceneStats(C.Structure):
    _fields_ = [("num_triangles", C.c_uint32), ("num_vertices", C.c_uint32), ("num_nodes", C.c_uint32),
                ("depth", C.c_uint32), ("device_bytes", C.c_uint64)]


_lib.register("vgi_scene_create", C.c_int32, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)])
_lib.register("vgi_scene_destroy", C.c_int32, [C.c_void_p])
_lib.register("vgi_scene_info", C.c_int32, [C.c_void_p, C.POINTER(SceneStats)])
_lib.register("vgi_trace_rays", C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p])


class GltfConfig(C.Structure):
    _fields_ = [("scale", C.c_float)]


class GltfStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("meshes", "primitives", "primitives_skipped", "instances", "materials",
                                           "textured_materials", "triangles", "vertices", "truncated_indices",
                                           "images", "images_unsupported", "textures")]


_lib.register("vgi_gltf_default_config", C.c_int32, [C.POINTER(GltfConfig)])
_lib.register("vgi_gltf_import", C.c_int32, [C.c_char_p, C.POINTER(GltfConfig), C.POINTER(C.c_void_p)])
_lib.register("vgi_gltf_model_desc", C.c_int32, [C.c_void_p, C.POINTER(SceneDesc)])
_lib.register("vgi_gltf_model_stats", C.c_int32, [C.c_void_p, C.POINTER(GltfStats)])
_lib.register("vgi_gltf_destroy", C.c_int32, [C.c_void_p])


@dataclass
class Mesh:
    positions: np.ndarray  # (V,3) f32
    normals: np.ndarray  # (V,3) f32
    indices: np.ndarray  # (T,3) u32
    tri_material: np.ndarray  # (T,) u32
    albedo: np.ndarray  # (M,3)
    metallic: np.ndarray  # (M,)
    roughness: np.ndarray  # (M,)
    texcoords: np.ndarray | None = None  # (V,2) f32
    tangents: np.ndarray | None = None  # (V,4) f32, w = bitangent sign
    textures: list = field(default_factory=list)  # Texture, numbered from TEXTURE_FIRST
    # (M,3) u32 texture indices per material: base colour, metallic-roughness,
    # normal map (None: white, white, flat normal)
    material_textures: np.ndarray | None = None

    def texture_words(self) -> np.ndarray:
        """(M,2) u32: vgi_material's color_metalrough_texture, normal_texture."""
        nm = len(self.albedo)
        t = self.material_textures
        if t is None:
            t = np.tile(np.array([TEXTURE_WHITE, TEXTURE_WHITE, TEXTURE_NORMAL_UP], np.uint32), (nm, 1))
        t = np.asarray(t, np.uint32)
        return np.stack([t[:, 0] | (t[:, 1] << 16), t[:, 2]], -1).astype(np.uint32)

    @property
    def num_triangles(self) -> int:
        return len(self.indices)

    def materials(self) -> list[tuple]:
        return [((*self.albedo[m], 1.0), float(self.metallic[m]), float(self.roughness[m]))
                for m in range(len(self.albedo))]


@lru_cache(maxsize=None)
def icosphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(level):
        cache: dict = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, dtype=np.uint32)


def _box(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [  # (4 corners CCW seen from outside, normal)
        ([(x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0)], (0, 0, -1)),
        ([(x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)], (0, 0, 1)),
        ([(x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)], (0, -1, 0)),
        ([(x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0)], (0, 1, 0)),
        ([(x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)], (-1, 0, 0)),
        ([(x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)], (1, 0, 0)),
    ]
    P, N, I = [], [], []
    for corners, n in faces:
        b = len(P)
        P += corners
        N += [n] * 4
        I += [(b, b + 1, b + 2), (b, b + 2, b + 3)]
    return np.array(P), np.array(N, dtype=np.float64), np.array(I, dtype=np.uint32)


def make_scene(target_triangles: int = 200_000, seed: int = 11, extent: float = 40.0) -> Mesh:
    """Boxes + icospheres on a plane with about `target_triangles` triangles."""
    P, N, I, TM = [], [], [], []
    nv = 0
    # ground plane: 2 triangles, object 0
    h = extent * 1.5
    pp, pn, pi = np.array([(-h, -h, 0), (h, -h, 0), (h, h, 0), (-h, h, 0)], np.float64), np.tile([0, 0, 1.0], (4, 1)), \
        np.array([(0, 1, 2), (0, 2, 3)], np.uint32)
    P.append(pp), N.append(pn), I.append(pi + nv), TM.append(np.zeros(2, np.uint32))
    nv += 4
    obj = 1
    total = 2
    k = 0
    while total < target_triangles:
        r = uniform(seed, k * 16 + 0)
        cx = float(uniform(seed, k * 16 + 1, -extent, extent))
        cy = float(uniform(seed, k * 16 + 2, -extent, extent))
        if r < 0.5:
            sx, sy, sz = (float(uniform(seed, k * 16 + 3 + a, 0.3, 2.5)) for a in range(3))
            z0 = float(uniform(seed, k * 16 + 6, 0.0, 0.5)) if uniform(seed, k * 16 + 7) < 0.3 else 0.0
            bp, bn, bi = _box((cx - sx / 2, cy - sy / 2, z0), (cx + sx / 2, cy + sy / 2, z0 + sz))
        else:
            level = 2 if uniform(seed, k * 16 + 8) < 0.5 else 3
            rad = float(uniform(seed, k * 16 + 9, 0.3, 1.8))
            lift = float(uniform(seed, k * 16 + 10, 0.0, 1.0))
            v, f = icosphere(level)
            bp = v * rad + np.array([cx, cy, rad + lift])
            bn = v.copy()
            bi = f
        P.append(bp), N.append(bn), I.append(bi + nv), TM.append(np.full(len(bi), obj, np.uint32))
        nv += len(bp)
        total += len(bi)
        obj += 1
        k += 1
    oid = np.arange(obj)
    albedo = np.stack([uniform(seed + 1, oid * 4 + c, 0.05, 0.95) for c in range(3)], -1)
    rough = uniform(seed + 2, oid, 0.3, 1.0)
    return Mesh(np.concatenate(P).astype(np.float32), np.concatenate(N).astype(np.float32),
                np.concatenate(I).astype(np.uint32), np.concatenate(TM).astype(np.uint32), albedo,
                np.zeros(obj), rough)


def vertex_tangents(P: np.ndarray, N: np.ndarray, UV: np.ndarray, I: np.ndarray) -> np.ndarray:
    """(V,4) f32 tangents for normal mapping: each triangle's dP/du summed
    per vertex, made orthogonal to the vertex normal and normalised, with w
    the sign of the bitangent against the summed dP/dv (E. Lengyel, 2001).
    A vertex without a usable direction gets any unit vector orthogonal to
    its normal.  (violet generates missing tangents with MikkTSpace,
    model.rs:507-531, which also splits vertices on tangent seams.)"""
    P = np.asarray(P, np.float64)
    N = np.asarray(N, np.float64)
    UV = np.asarray(UV, np.float64)
    I = np.asarray(I, np.int64)
    e1, e2 = P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]]
    d1, d2 = UV[I[:, 1]] - UV[I[:, 0]], UV[I[:, 2]] - UV[I[:, 0]]
    r = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
    inv = np.where(np.abs(r) > 1e-20, 1.0 / np.where(r == 0, 1.0, r), 0.0)[:, None]
    sdir = (e1 * d2[:, 1:2] - e2 * d1[:, 1:2]) * inv
    tdir = (e2 * d1[:, 0:1] - e1 * d2[:, 0:1]) * inv
    T = np.zeros_like(P)
    B = np.zeros_like(P)
    for k in range(3):
        np.add.at(T, I[:, k], sdir)
        np.add.at(B, I[:, k], tdir)
    n = N / np.maximum(np.linalg.norm(N, axis=1, keepdims=True), 1e-30)
    t = T - n * np.sum(n * T, axis=1, keepdims=True)
    ln = np.linalg.norm(t, axis=1, keepdims=True)
    alt = np.cross(n, np.where(np.abs(n[:, :1]) < 0.9, [[1.0, 0.0, 0.0]], [[0.0, 1.0, 0.0]]))
    t = np.where(ln > 1e-12, t / np.maximum(ln, 1e-30), alt / np.maximum(np.linalg.norm(alt, axis=1, keepdims=True),
                                                                         1e-30))
    w = np.where(np.sum(np.cross(n, t) * B, axis=1) < 0.0, -1.0, 1.0)
    return np.concatenate([t, w[:, None]], 1).astype(np.float32)


def _smooth_image(seed: int, w: int, h: int, channels: int = 4, cells: int = 6) -> np.ndarray:
    """Seeded RGBA noise: a coarse random lattice bilinearly upsampled plus
    fine per-texel noise, values in [0, 1]."""
    rng = np.random.default_rng(seed)
    g = rng.random((cells + 1, cells + 1, channels))
    ys, xs = np.linspace(0, cells, h, endpoint=False), np.linspace(0, cells, w, endpoint=False)
    y0, x0 = ys.astype(int), xs.astype(int)
    fy, fx = (ys - y0)[:, None, None], (xs - x0)[None, :, None]
    a = g[y0][:, x0] * (1 - fx) + g[y0][:, x0 + 1] * fx
    b = g[y0 + 1][:, x0] * (1 - fx) + g[y0 + 1][:, x0 + 1] * fx
    img = a * (1 - fy) + b * fy
    return np.clip(0.8 * img + 0.2 * rng.random((h, w, channels)), 0.0, 1.0)


def texture_scene(mesh: Mesh, seed: int = 5, uv_scale: float = 0.5) -> Mesh:
    """The same geometry with texture coordinates, tangents and a set of
    seeded textures: planar coordinates on the plane facing each vertex
    normal's dominant axis (times uv_scale, so the ground repeats), tangents
    from vertex_tangents, and materials cycling through untextured, base
    colour, base colour + metallic-roughness, all three maps, a normal map
    alone, and the built-in textures in unusual slots."""
    P = np.asarray(mesh.positions, np.float64)
    N = np.asarray(mesh.normals, np.float64)
    ax = np.argmax(np.abs(N), axis=1)
    uv = np.where((ax == 2)[:, None], P[:, [0, 1]], np.where((ax == 0)[:, None], P[:, [1, 2]], P[:, [0, 2]]))
    uv = (uv * uv_scale).astype(np.float32)
    tg = vertex_tangents(P, N, uv, mesh.indices)

    def rgba8(x):
        return np.round(np.clip(x, 0.0, 1.0) * 255.0).astype(np.uint8)

    def normal_map(w, h):
        rng = np.random.default_rng(seed + 7)
        t = np.concatenate([(_smooth_image(seed + 3, w, h, 2) - 0.5) * 0.8, np.ones((h, w, 1))], 2)
        t /= np.linalg.norm(t, axis=2, keepdims=True)
        img = np.concatenate([t * 0.5 + 0.5, np.ones((h, w, 1))], 2)
        img[..., :2] += (rng.random((h, w, 2)) - 0.5) * 0.02
        return rgba8(img)

    metal_rough = _smooth_image(seed + 2, 32, 128)
    metal_rough[..., 0] = (metal_rough[..., 0] > 0.6)  # metallic mostly 0 or 1
    textures = [
        Texture(box_mips(rgba8(_smooth_image(seed, 64, 64))), srgb=True),  # 2
        Texture(box_mips(rgba8(_smooth_image(seed + 1, 37, 53))), srgb=True),  # 3: odd sizes
        Texture(box_mips(rgba8(metal_rough))),  # 4
        Texture(box_mips(normal_map(64, 64))),  # 5
        Texture([rgba8(np.array([[[0.9, 0.5, 0.2, 1.0]]]))], srgb=True),  # 6: 1x1
        Texture(box_mips(rgba8(_smooth_image(seed + 4, 256, 16))), srgb=True),  # 7: wide
    ]
    nm = len(mesh.albedo)
    mt = np.zeros((nm, 3), np.uint32)
    for m in range(nm):
        bc = (2, 3, 6, 7)[(m // 6) % 4]
        mt[m] = [(0, 0, 1), (bc, 0, 1), (bc, 4, 1), (bc, 4, 5), (0, 0, 5), (1, bc, 0)][(m + 3) % 6]
    # a metallic-roughness texture drives metallic through a factor of 1
    metallic = np.where(mt[:, 1] >= TEXTURE_FIRST, 1.0, np.asarray(mesh.metallic, np.float64))
    return Mesh(mesh.positions, mesh.normals, mesh.indices, mesh.tri_material, mesh.albedo, metallic,
                mesh.roughness, texcoords=uv, tangents=tg, textures=textures, material_textures=mt)


class Scene:
    """Device scene (BVH + shading data), wraps vgi_scene_create."""

    def __init__(self, mesh: Mesh):
        mats = (Material * len(mesh.albedo))()
        words = mesh.texture_words()
        for m, (bc, met, rough) in enumerate(mesh.materials()):
            mats[m].base_color[:] = list(bc)
            mats[m].metallic = met
            mats[m].roughness = rough
            mats[m].color_metalrough_texture = int(words[m, 0])
            mats[m].normal_texture = int(words[m, 1])
        pos = np.ascontiguousarray(mesh.positions, np.float32)
        nrm = np.ascontiguousarray(mesh.normals, np.float32)
        idx = np.ascontiguousarray(mesh.indices, np.uint32)
        tm = np.ascontiguousarray(mesh.tri_material, np.uint32)
        uv = None if mesh.texcoords is None else np.ascontiguousarray(mesh.texcoords, np.float32)
        tg = None if mesh.tangents is None else np.ascontiguousarray(mesh.tangents, np.float32)
        packed = [t.packed() for t in mesh.textures]
        texs = (TextureDesc * max(1, len(packed)))()
        for k, (t, p) in enumerate(zip(mesh.textures, packed)):
            texs[k] = TextureDesc(t.width, t.height, len(t.levels), int(bool(t.srgb)), p.ctypes.data)
        d = SceneDesc(pos.ctypes.data, nrm.ctypes.data, len(pos), idx.ctypes.data, len(idx), tm.ctypes.data, mats,
                      len(mats), None if uv is None else uv.ctypes.data, None if tg is None else tg.ctypes.data,
                      texs, len(packed))
        h = C.c_void_p()
        _lib.call("vgi_scene_create", C.byref(d), C.byref(h))
        self._h = h
        self.mesh = mesh

    @property
    def handle(self):
        return self._h

    def stats(self) -> SceneStats:
        s = SceneStats()
        _lib.call("vgi_scene_info", self._h, C.byref(s))
        return s

    def trace(self, rays_dev, n: int, any_hit: bool, hits_dev, stream=None) -> None:
        from . import gpu

        _lib.call("vgi_trace_rays", self._h, gpu.ptr(rays_dev), n, int(any_hit), gpu.ptr(hits_dev),
                  gpu.stream() if stream is None else stream)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().vgi_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_gltf(path, scale: float = 1.0) -> tuple[Mesh, GltfStats]:
    """Import a .gltf / .glb file through the library's importer
    (vgi_gltf_import: model.rs import semantics, instances baked to world
    space, Z up) into a Mesh for Scene(); also returns the import statistics."""
    cfg = GltfConfig(scale)
    h = C.c_void_p()
    _lib.call("vgi_gltf_import", str(path).encode(), C.byref(cfg), C.byref(h))
    try:
        d = SceneDesc()
        _lib.call("vgi_gltf_model_desc", h, C.byref(d))
        st = GltfStats()
        _lib.call("vgi_gltf_model_stats", h, C.byref(st))

        def arr(ptr, n, dt, shape):
            if n == 0:
                return np.zeros(shape, dt)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,)).reshape(
                shape).copy()

        nv, nt, nm = d.num_vertices, d.num_triangles, d.num_materials
        pos = arr(d.positions, 3 * nv, np.float32, (nv, 3))
        nrm = arr(d.normals, 3 * nv, np.float32, (nv, 3))
        idx = arr(d.indices, 3 * nt, np.uint32, (nt, 3))
        tm = arr(d.triangle_material, nt, np.uint32, (nt,))
        mats = [d.materials[k] for k in range(nm)]
        uv = arr(d.texcoords, 2 * nv, np.float32, (nv, 2)) if d.texcoords else None
        tg = arr(d.tangents, 4 * nv, np.float32, (nv, 4)) if d.tangents else None
        textures = []
        for k in range(d.num_textures):
            t = d.textures[k]
            levels, off = [], 0
            total = sum(max(1, t.width >> l) * max(1, t.height >> l) for l in range(t.levels)) * 4
            raw = arr(t.texels, total, np.uint8, (total,))
            for l in range(t.levels):
                lw, lh = max(1, t.width >> l), max(1, t.height >> l)
                levels.append(raw[off:off + lw * lh * 4].reshape(lh, lw, 4))
                off += lw * lh * 4
            textures.append(Texture(levels, srgb=bool(t.srgb)))
        mt = np.array([[m.color_metalrough_texture & 0xFFFF, m.color_metalrough_texture >> 16,
                        m.normal_texture & 0xFFFF] for m in mats], np.uint32).reshape(nm, 3)
        mesh = Mesh(pos, nrm, idx, tm, np.array([list(m.base_color)[:3] for m in mats], np.float32).reshape(nm, 3),
                    np.array([m.metallic for m in mats], np.float32), np.array([m.roughness for m in mats], np.float32),
                    texcoords=uv, tangents=tg, textures=textures, material_textures=mt)
        return mesh, st
    finally:
        _lib.lib().vgi_gltf_destroy(h)
