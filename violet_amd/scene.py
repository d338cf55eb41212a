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
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from . import _lib
from .synth import uniform


class Material(C.Structure):
    _fields_ = [("base_color", C.c_float * 4), ("metallic", C.c_float), ("roughness", C.c_float),
                ("pad", C.c_float * 2)]


class SceneDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("num_vertices", C.c_uint32),
                ("indices", C.c_void_p), ("num_triangles", C.c_uint32), ("triangle_material", C.c_void_p),
                ("materials", C.POINTER(Material)), ("num_materials", C.c_uint32)]


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
                                           "textured_materials", "triangles", "vertices", "truncated_indices")]


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


class Scene:
    """Device scene (BVH + shading data), wraps vgi_scene_create."""

    def __init__(self, mesh: Mesh):
        mats = (Material * len(mesh.albedo))()
        for m, (bc, met, rough) in enumerate(mesh.materials()):
            mats[m].base_color[:] = list(bc)
            mats[m].metallic = met
            mats[m].roughness = rough
        pos = np.ascontiguousarray(mesh.positions, np.float32)
        nrm = np.ascontiguousarray(mesh.normals, np.float32)
        idx = np.ascontiguousarray(mesh.indices, np.uint32)
        tm = np.ascontiguousarray(mesh.tri_material, np.uint32)
        d = SceneDesc(pos.ctypes.data, nrm.ctypes.data, len(pos), idx.ctypes.data, len(idx), tm.ctypes.data, mats,
                      len(mats))
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
        mesh = Mesh(pos, nrm, idx, tm, np.array([list(m.base_color)[:3] for m in mats], np.float32).reshape(nm, 3),
                    np.array([m.metallic for m in mats], np.float32), np.array([m.roughness for m in mats], np.float32))
        return mesh, st
    finally:
        _lib.lib().vgi_gltf_destroy(h)
