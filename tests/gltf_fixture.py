"""Small glTF 2.0 writer for the import tests: a document plus one binary
buffer, saved as .gltf + .bin, as .gltf with a base64 data URI, or as .glb."""
from __future__ import annotations

import base64
import json
import struct
from pathlib import Path

import numpy as np

COMP = {np.dtype(np.int8): 5120, np.dtype(np.uint8): 5121, np.dtype(np.int16): 5122, np.dtype(np.uint16): 5123,
        np.dtype(np.uint32): 5125, np.dtype(np.float32): 5126}
TYPE = {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4", 16: "MAT4"}


class Doc:
    def __init__(self):
        self.j = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [], "meshes": [],
                  "nodes": [], "materials": [], "scenes": []}
        self.blob = bytearray()

    def _align(self, n=4):
        while len(self.blob) % n:
            self.blob.append(0)

    def view(self, data: bytes, stride: int | None = None) -> int:
        self._align()
        v = {"buffer": 0, "byteOffset": len(self.blob), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.blob += data
        self.j["bufferViews"].append(v)
        return len(self.j["bufferViews"]) - 1

    def accessor(self, arr: np.ndarray, view: int | None = None, offset: int = 0, count: int | None = None) -> int:
        arr = np.asarray(arr)
        ncomp = 1 if arr.ndim == 1 else arr.shape[1]
        if view is None:
            view = self.view(np.ascontiguousarray(arr).tobytes())
        a = {"bufferView": view, "componentType": COMP[arr.dtype], "count": int(count if count is not None else len(arr)),
             "type": TYPE[ncomp]}
        if offset:
            a["byteOffset"] = offset
        if arr.dtype == np.float32 and ncomp == 3:
            a["min"] = arr.min(0).tolist()
            a["max"] = arr.max(0).tolist()
        self.j["accessors"].append(a)
        return len(self.j["accessors"]) - 1

    def interleaved(self, pos: np.ndarray, nrm: np.ndarray) -> tuple[int, int]:
        """POSITION and NORMAL interleaved in one bufferView (stride 24)."""
        inter = np.concatenate([pos.astype(np.float32), nrm.astype(np.float32)], 1)
        v = self.view(inter.tobytes(), stride=24)
        a = {"bufferView": v, "componentType": 5126, "count": len(pos), "type": "VEC3",
             "min": pos.min(0).tolist(), "max": pos.max(0).tolist()}
        self.j["accessors"].append(a)
        pa = len(self.j["accessors"]) - 1
        self.j["accessors"].append({"bufferView": v, "byteOffset": 12, "componentType": 5126, "count": len(nrm),
                                    "type": "VEC3"})
        return pa, pa + 1

    def mesh(self, primitives: list[dict]) -> int:
        self.j["meshes"].append({"primitives": primitives})
        return len(self.j["meshes"]) - 1

    def node(self, **kw) -> int:
        self.j["nodes"].append(kw)
        return len(self.j["nodes"]) - 1

    def material(self, m: dict) -> int:
        self.j["materials"].append(m)
        return len(self.j["materials"]) - 1

    def image(self, data: bytes, mime: str = "image/png", in_buffer: bool = False) -> int:
        """An image as a base64 data URI, or in a bufferView of the blob."""
        if in_buffer:
            im = {"bufferView": self.view(data), "mimeType": mime}
        else:
            im = {"uri": f"data:{mime};base64," + base64.b64encode(data).decode()}
        self.j.setdefault("images", []).append(im)
        return len(self.j["images"]) - 1

    def texture(self, image: int) -> int:
        self.j.setdefault("textures", []).append({"source": image, "sampler": 0})
        self.j.setdefault("samplers", [{"magFilter": 9729, "minFilter": 9987, "wrapS": 10497, "wrapT": 10497}])
        return len(self.j["textures"]) - 1

    def _json(self, uri):
        j = json.loads(json.dumps(self.j))
        j["buffers"] = [{"byteLength": len(self.blob)}]
        if uri is not None:
            j["buffers"][0]["uri"] = uri
        return j

    def save_gltf(self, path: Path, bin_name: str = "scene data.bin") -> Path:
        path = Path(path)
        (path.parent / bin_name).write_bytes(bytes(self.blob))
        path.write_text(json.dumps(self._json(bin_name.replace(" ", "%20")), indent=1))
        return path

    def save_data_uri(self, path: Path) -> Path:
        path = Path(path)
        uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.blob)).decode()
        path.write_text(json.dumps(self._json(uri)))
        return path

    def save_glb(self, path: Path) -> Path:
        path = Path(path)
        js = json.dumps(self._json(None)).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        bn = bytes(self.blob) + b"\0" * ((4 - len(self.blob) % 4) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bn), 0x004E4942) + bn
        path.write_bytes(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)
        return path


def scene_from_mesh(mesh, jpeg: bool = False) -> Doc:
    """A Z-up world-space mesh as a glTF document (Y up: (x, y, z) -> (x, z, -y)),
    one primitive per material, one node.  A textured mesh keeps its texture
    coordinates and tangents, and its textures' base levels become PNG
    images (a metallic-roughness texture with R and G moved back to glTF's B
    and G); jpeg=True stores base-colour images as baseline JPEG."""
    from png_fixture import encode_png

    d = Doc()
    pos = np.stack([mesh.positions[:, 0], mesh.positions[:, 2], -mesh.positions[:, 1]], 1).astype(np.float32)
    nrm = np.stack([mesh.normals[:, 0], mesh.normals[:, 2], -mesh.normals[:, 1]], 1).astype(np.float32)
    pa, na = d.interleaved(pos, nrm)
    attrs = {"POSITION": pa, "NORMAL": na}
    if getattr(mesh, "texcoords", None) is not None:
        attrs["TEXCOORD_0"] = d.accessor(np.ascontiguousarray(mesh.texcoords, np.float32))
    if getattr(mesh, "tangents", None) is not None:
        t = mesh.tangents
        attrs["TANGENT"] = d.accessor(np.stack([t[:, 0], t[:, 2], -t[:, 1], t[:, 3]], 1).astype(np.float32))
    mt = getattr(mesh, "material_textures", None)
    tex_index = {}

    def tex(k, metal_rough=False):
        if (k, metal_rough) not in tex_index:
            img = np.array(mesh.textures[k - 2].levels[0])
            if metal_rough:
                img = img[..., [2, 1, 0, 3]]
            if jpeg and mesh.textures[k - 2].srgb and not metal_rough:
                import io

                from PIL import Image

                b = io.BytesIO()
                Image.fromarray(np.ascontiguousarray(img[..., :3])).save(b, "JPEG", quality=90, restart_marker_blocks=4)
                tex_index[(k, metal_rough)] = d.texture(d.image(b.getvalue(), mime="image/jpeg"))
            else:
                tex_index[(k, metal_rough)] = d.texture(d.image(encode_png(img, 6, 8)))
        return tex_index[(k, metal_rough)]

    prims = []
    for m, (alb, met, rough) in enumerate(zip(mesh.albedo, mesh.metallic, mesh.roughness)):
        pbr = {"baseColorFactor": [float(alb[0]), float(alb[1]), float(alb[2]), 1.0], "metallicFactor": float(met),
               "roughnessFactor": float(rough)}
        mat = {"pbrMetallicRoughness": pbr}
        if mt is not None:
            if mt[m, 0] >= 2:
                pbr["baseColorTexture"] = {"index": tex(int(mt[m, 0]))}
            if mt[m, 1] >= 2:
                pbr["metallicRoughnessTexture"] = {"index": tex(int(mt[m, 1]), metal_rough=True)}
            if mt[m, 2] >= 2:
                mat["normalTexture"] = {"index": tex(int(mt[m, 2]))}
        d.material(mat)
        tri = mesh.indices[mesh.tri_material == m]
        if len(tri):
            prims.append({"attributes": attrs, "indices": d.accessor(tri.reshape(-1).astype(np.uint32)),
                          "material": m})
    d.node(mesh=d.mesh(prims))
    d.j["scenes"] = [{"nodes": [0]}]
    d.j["scene"] = 0
    return d


def corrupted_documents(d: Doc, out_dir: Path, seed: int = 99) -> list[Path]:
    """200 seeded corruptions of a document: 60 truncations of the .gltf
    text, 60 .glb forms with 1-7 random bytes replaced, and 80 .gltf forms
    with one accessor / buffer view / mesh / node field deleted, set to -1,
    to a string or to 2^40."""
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    gl = d.save_gltf(out_dir / "base.gltf")
    glb = d.save_glb(out_dir / "base.glb").read_bytes()
    text = gl.read_text()
    j = json.loads(text)
    paths = []
    for k in range(60):
        p = out_dir / f"t{k}.gltf"
        p.write_text(text[: int(rng.integers(0, len(text)))])
        paths.append(p)
    for k in range(60):
        b = bytearray(glb)
        for _ in range(int(rng.integers(1, 8))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        p = out_dir / f"f{k}.glb"
        p.write_bytes(bytes(b))
        paths.append(p)
    keys = [("accessors", "count"), ("accessors", "componentType"), ("bufferViews", "byteOffset"),
            ("bufferViews", "byteLength"), ("bufferViews", "buffer"), ("meshes", "primitives"), ("nodes", "mesh"),
            ("nodes", "children"), ("nodes", "matrix"), ("accessors", "bufferView")]
    for k in range(80):
        jj = json.loads(json.dumps(j))
        arr, key = keys[int(rng.integers(0, len(keys)))]
        items = [x for x in jj[arr] if isinstance(x, dict)]
        tgt = items[int(rng.integers(0, len(items)))]
        choice = int(rng.integers(0, 4))
        if choice == 0:
            tgt.pop(key, None)
        elif choice == 1:
            tgt[key] = -1
        elif choice == 2:
            tgt[key] = "x"
        else:
            tgt[key] = 2 ** 40
        p = out_dir / f"m{k}.gltf"
        p.write_text(json.dumps(jj))
        paths.append(p)
    return paths
