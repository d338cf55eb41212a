"""Material textures through the glTF importer (CPU only).

* PNG decoding (vgi_image.cpp) against the samples a PNG writer of the
  test's own encoded (png_fixture.py): every colour type and bit depth,
  palettes with and without tRNS, tRNS colour keys, all five row filters,
  Adam7 interlacing, odd sizes, stored / fixed-Huffman / dynamic-Huffman
  zlib blocks and images split over several IDAT chunks; corrupted files
  are refused with an error, never read out of bounds.
* The mip chain: violet's level count and halving sizes (model.rs:853-870),
  constant images stay constant, a linear ramp stays linear away from the
  edges (a normalised symmetric filter reproduces linear functions), and
  the base level is the decoded image.
* Material slots (model.rs:636-790, render_scene.rs:920-935): used images
  numbered in image order, an image used only as a metallic-roughness map
  swizzled (B, G -> R, G), only base colour marked sRGB, an image used for
  two purposes neither; default slots white / white / flat normal; a
  progressive JPEG counted as unsupported with its slot left at the default.
* Baseline JPEG decoding against PIL (libjpeg): 4:4:4, 4:2:2 and 4:2:0
  chroma, greyscale, restart intervals, odd sizes and two qualities agree
  within 3 codes (mean below 0.1: decoders differ by their inverse-DCT and
  colour-conversion rounding; the chroma upsampling is the same triangle
  filter), as do violet's Sponza textures when the reference tree is
  present; damaged files are refused or decoded, never read out of bounds.
* Texture coordinates and generated tangents: a plane with u along +x
  gets tangents +x (w = +1), and w = -1 when v runs the other way.
* violet's own Duck asset (when the reference tree is present): the
  embedded PNG decodes to what PIL decodes, sRGB, with violet's mip count.
"""
from __future__ import annotations

import zlib
from pathlib import Path

import numpy as np
import pytest

from gltf_fixture import Doc
from png_fixture import encode_png, expected_rgba8

REF_ASSETS = Path("/root/reference/assets")


def _jpeg(rgb, quality=90, subsampling=2, progressive=False, restart=0):
    import io

    from PIL import Image

    b = io.BytesIO()
    mode = "L" if rgb.ndim == 2 else "RGB"
    kw = dict(quality=quality, progressive=progressive)
    if mode == "RGB":
        kw["subsampling"] = subsampling
    if restart:
        kw["restart_marker_blocks"] = restart
    Image.fromarray(np.ascontiguousarray(rgb, np.uint8), mode).save(b, "JPEG", **kw)
    return b.getvalue()


def _pil_rgba(data: bytes) -> np.ndarray:
    import io

    from PIL import Image

    return np.asarray(Image.open(io.BytesIO(data)).convert("RGBA"))


def _images_doc(images: list[tuple[bytes, str]]):
    """One triangle per image, each with a base-colour texture of it."""
    d = Doc()
    pa = d.accessor(np.array([[0, 0, 0], [1, 0, 0], [1, 0, -1]], np.float32))
    ia = d.accessor(np.array([0, 1, 2], np.uint16))
    prims = []
    for data, mime in images:
        t = d.texture(d.image(data, mime=mime))
        prims.append({"attributes": {"POSITION": pa}, "indices": ia,
                      "material": d.material({"pbrMetallicRoughness": {"baseColorTexture": {"index": t}}})})
    d.node(mesh=d.mesh(prims))
    d.j["scenes"] = [{"nodes": [0]}]
    d.j["scene"] = 0
    return d


def _quad_doc(materials: list[dict], uv=None):
    """One quad (two triangles) per material, with NORMAL and TEXCOORD_0."""
    d = Doc()
    prims = []
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 0, -1], [0, 0, -1]], np.float32)  # glTF y-up: the z = 0 plane after import
    N = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    U = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32) if uv is None else np.asarray(uv, np.float32)
    pa, na = d.interleaved(P, N)
    ua = d.accessor(U)
    for m in materials:
        mi = d.material(m)
        prims.append({"attributes": {"POSITION": pa, "NORMAL": na, "TEXCOORD_0": ua},
                      "indices": d.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint16)), "material": mi})
    d.node(mesh=d.mesh(prims))
    d.j["scenes"] = [{"nodes": [0]}]
    d.j["scene"] = 0
    return d


def _import(d: Doc, tmp_path, form="data_uri"):
    from violet_amd import scene

    path = {"gltf": lambda: d.save_gltf(tmp_path / "t.gltf"), "data_uri": lambda: d.save_data_uri(tmp_path / "t.gltf"),
            "glb": lambda: d.save_glb(tmp_path / "t.glb")}[form]()
    return scene.load_gltf(path)


def _png_cases():
    rng = np.random.default_rng(7)
    cases = []
    for ct, depths in ((0, (1, 2, 4, 8, 16)), (2, (8, 16)), (3, (1, 2, 4, 8)), (4, (8, 16)), (6, (8, 16))):
        for depth in depths:
            for (h, w) in ((1, 1), (7, 13), (17, 33)):
                ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ct]
                hi = min(1 << depth, 256 if ct == 3 else 1 << depth)
                s = rng.integers(0, hi, (h, w, ch))
                pal = trns = None
                if ct == 3:
                    pal = rng.integers(0, 256, (hi, 3))
                    trns = rng.integers(0, 256, hi // 2 + 1) if w % 2 else None
                elif ct == 0 and w > 1:
                    trns = [int(s[0, 0, 0])]
                elif ct == 2 and w > 1:
                    trns = [int(v) for v in s[0, 1]]
                cases.append((ct, depth, s, pal, trns))
    return cases


CASES = _png_cases()
ZMODES = [(0, zlib.Z_DEFAULT_STRATEGY), (6, zlib.Z_FIXED), (9, zlib.Z_DEFAULT_STRATEGY), (6, zlib.Z_HUFFMAN_ONLY),
          (6, zlib.Z_RLE)]


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("zmode", range(len(ZMODES)))
def test_png_decoding_every_format(tmp_path, interlace, zmode):
    level, strategy = ZMODES[zmode]
    d = Doc()
    expected = []
    mats = []
    for k, (ct, depth, s, pal, trns) in enumerate(CASES):
        filt = [(k + j) % 5 for j in range(5)]
        data = encode_png(s, ct, depth, filters=filt, interlace=interlace, palette=pal, trns=trns, level=level,
                          strategy=strategy, idat_split=1 + k % 3)
        t = d.texture(d.image(data, in_buffer=bool(k % 2)))
        mats.append({"pbrMetallicRoughness": {"baseColorTexture": {"index": t}}})
        expected.append(expected_rgba8(s, ct, depth, pal, trns))
    d2 = _quad_doc([])
    d.j["meshes"], d.j["nodes"], d.j["scenes"], d.j["scene"] = [], [], [], 0
    # one quad per material, sharing the fixture's geometry
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 0, -1]], np.float32)
    pa = d.accessor(P)
    ia = d.accessor(np.array([0, 1, 2], np.uint16))
    prims = []
    for m in mats:
        prims.append({"attributes": {"POSITION": pa}, "indices": ia, "material": d.material(m)})
    d.node(mesh=d.mesh(prims))
    d.j["scenes"] = [{"nodes": [0]}]
    del d2
    mesh, st = _import(d, tmp_path, form="glb" if interlace else "data_uri")
    assert st.images == len(CASES) and st.textures == len(CASES) and st.images_unsupported == 0
    for k, want in enumerate(expected):
        got = mesh.textures[k].levels[0]
        assert got.shape == want.shape and np.array_equal(got, want), (k, CASES[k][:2])


def test_corrupted_png_is_refused(tmp_path):
    from violet_amd import _lib

    rng = np.random.default_rng(3)
    good = encode_png(rng.integers(0, 256, (9, 11, 4)), 6, 8)
    bad = [good[:40], good[:-20], good[:33] + bytes([good[33] ^ 0x40]) + good[34:]]  # truncated, CRC broken
    for b in bad:
        d = _quad_doc([{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}])
        d.texture(d.image(b))
        with pytest.raises(_lib.VgiError):
            _import(d, tmp_path)
    # random byte damage: either refused or decoded, never a crash
    for k in range(40):
        b = bytearray(good)
        for _ in range(int(rng.integers(1, 6))):
            b[int(rng.integers(8, len(b)))] = int(rng.integers(0, 256))
        d = _quad_doc([{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}])
        d.texture(d.image(bytes(b)))
        try:
            _import(d, tmp_path)
        except _lib.VgiError:
            pass


@pytest.mark.parametrize("size", [(1, 1), (1, 7), (64, 64), (37, 53), (600, 3)])
def test_mip_chain_sizes_constant_and_ramp(tmp_path, size):
    from violet_amd import scene

    h, w = size
    const = np.zeros((h, w, 4), np.uint8)
    const[:] = (17, 130, 250, 77)
    ramp = np.zeros((h, w, 4), np.uint8)
    ramp[..., 0] = np.round(np.arange(w) * (255 / max(1, w - 1))).astype(np.uint8)[None, :]
    ramp[..., 3] = 255
    d = _quad_doc([{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}},
                   {"pbrMetallicRoughness": {"baseColorTexture": {"index": 1}}}])
    d.texture(d.image(encode_png(const, 6, 8)))
    d.texture(d.image(encode_png(ramp, 6, 8)))
    mesh, st = _import(d, tmp_path)
    for t in mesh.textures:
        assert len(t.levels) == scene.mip_count(w, h)
        for l, lv in enumerate(t.levels):
            assert lv.shape == (max(1, h >> l), max(1, w >> l), 4)
    for lv in mesh.textures[0].levels:
        assert (lv == const[0, 0]).all()
    assert np.array_equal(mesh.textures[1].levels[0], ramp)
    if w >= 32:
        lv = mesh.textures[1].levels[1].astype(np.float64)
        x = np.arange(lv.shape[1])
        centre = (x + 0.5) * (w / lv.shape[1]) - 0.5  # the output texel's centre in input texels
        want = centre * (255 / (w - 1))
        inner = slice(3, lv.shape[1] - 3)
        assert np.abs(lv[:, inner, 0] - want[inner]).max() <= 1.0


def test_material_slots_swizzle_srgb_and_defaults(tmp_path):
    rng = np.random.default_rng(5)
    imgs = [rng.integers(0, 256, (4, 6, 4)).astype(np.uint8) for _ in range(5)]
    d = _quad_doc([
        {"pbrMetallicRoughness": {"baseColorTexture": {"index": 1}, "metallicRoughnessTexture": {"index": 2}},
         "normalTexture": {"index": 3}},
        {"pbrMetallicRoughness": {"baseColorTexture": {"index": 4}}, "normalTexture": {"index": 4}},  # two uses
        {"pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.5, 0.5, 1.0]}},  # untextured
        {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}},  # a progressive JPEG
    ])
    d.texture(d.image(_jpeg(imgs[0][..., :3], progressive=True), mime="image/jpeg"))  # texture 0 -> image 0
    for k in range(4):  # textures 1..4 -> images 1..4
        d.texture(d.image(encode_png(imgs[k], 6, 8)))
    d.image(encode_png(imgs[4], 6, 8))  # image 5: not used by any material
    mesh, st = _import(d, tmp_path)
    assert (st.images, st.images_unsupported, st.textures) == (5, 1, 4)
    # used images in image order: 1 -> texture 2, 2 -> 3, 3 -> 4, 4 -> 5
    assert mesh.material_textures.tolist() == [[2, 3, 4], [5, 0, 5], [0, 0, 1], [0, 0, 1]]
    t = mesh.textures
    assert [x.srgb for x in t] == [True, False, False, False]
    assert np.array_equal(t[0].levels[0], imgs[0])
    assert np.array_equal(t[1].levels[0], imgs[1][..., [2, 1, 0, 3]])  # metallic (B) -> R, roughness (G) -> G
    assert np.array_equal(t[2].levels[0], imgs[2])
    assert np.array_equal(t[3].levels[0], imgs[3])  # base colour and normal map: not swizzled, not sRGB


@pytest.mark.parametrize("flip_v", [False, True])
def test_texcoords_and_generated_tangents(tmp_path, flip_v):
    uv = np.array([[0, 0], [2, 0], [2, 3], [0, 3]], np.float32)
    if flip_v:
        uv[:, 1] = -uv[:, 1]
    d = _quad_doc([{"pbrMetallicRoughness": {}}], uv=uv)
    mesh, st = _import(d, tmp_path)
    assert mesh.texcoords is not None and mesh.tangents is not None
    assert np.array_equal(mesh.texcoords, uv)
    # u grows along glTF +x, which stays +x after the Y-up -> Z-up rotation
    np.testing.assert_allclose(mesh.tangents[:, :3], np.tile([1.0, 0.0, 0.0], (4, 1)), atol=1e-6)
    # v grows along glTF -z = world +y; the bitangent cross(N, T) = (0,0,1) x (1,0,0) = +y
    assert (mesh.tangents[:, 3] == (-1.0 if flip_v else 1.0)).all()


def test_duck_asset_texture():
    from PIL import Image
    import base64
    import io
    import json

    from violet_amd import scene

    path = REF_ASSETS / "Duck" / "Duck.gltf"
    if not path.exists():
        pytest.skip("reference assets absent")
    doc = json.loads(path.read_text())
    uri = doc["images"][0]["uri"]
    png = base64.b64decode(uri.split(",", 1)[1])
    want = np.asarray(Image.open(io.BytesIO(png)).convert("RGBA"))
    mesh, st = scene.load_gltf(path)
    assert (st.images, st.textures, st.images_unsupported) == (1, 1, 0)
    t = mesh.textures[0]
    assert t.srgb and np.array_equal(t.levels[0], want)
    assert len(t.levels) == scene.mip_count(want.shape[1], want.shape[0])
    assert mesh.material_textures.tolist() == [[2, 0, 1]]
    assert mesh.texcoords is not None and np.isfinite(mesh.texcoords).all()
    assert mesh.tangents is not None
    # generated in the model's space as unit vectors orthogonal to the unit
    # normals; the node's uniform scale then scales both alike through the
    # normal transform
    tl = np.linalg.norm(mesh.tangents[:, :3], axis=1)
    nl = np.linalg.norm(mesh.normals, axis=1)
    assert np.abs(tl / nl - 1).max() < 1e-3
    cos = np.sum(mesh.tangents[:, :3] * mesh.normals, axis=1) / (tl * nl)
    assert np.abs(cos).max() < 1e-3


def _smooth_rgb(rng, h, w):
    g = rng.random((h // 8 + 2, w // 8 + 2, 3))
    img = np.kron(g, np.ones((8, 8, 1)))[:h, :w]
    return (np.clip(img * 0.85 + 0.15 * rng.random((h, w, 3)), 0, 1) * 255).astype(np.uint8)


def test_baseline_jpeg_matches_libjpeg(tmp_path):
    rng = np.random.default_rng(11)
    cases = []
    for (h, w) in ((33, 47), (17, 9), (64, 64), (120, 200)):
        for sub in (0, 1, 2):
            for q, rst in ((50, 3), (95, 0)):
                cases.append(_jpeg(_smooth_rgb(rng, h, w), quality=q, subsampling=sub, restart=rst))
        cases.append(_jpeg(_smooth_rgb(rng, h, w)[..., 1], quality=80))
    mesh, st = _import(_images_doc([(c, "image/jpeg") for c in cases]), tmp_path, form="glb")
    assert (st.images, st.textures, st.images_unsupported) == (len(cases), len(cases), 0)
    for k, c in enumerate(cases):
        got, want = mesh.textures[k].levels[0].astype(int), _pil_rgba(c).astype(int)
        assert got.shape == want.shape
        dd = np.abs(got - want)
        assert dd.max() <= 3 and dd.mean() < 0.1, (k, dd.max(), dd.mean())


def test_damaged_jpeg_is_refused_or_decoded(tmp_path):
    from violet_amd import _lib

    rng = np.random.default_rng(4)
    good = _jpeg(_smooth_rgb(rng, 40, 56), quality=85, restart=2)
    with pytest.raises(_lib.VgiError):
        _import(_images_doc([(good[:len(good) // 3], "image/jpeg")]), tmp_path)
    for k in range(60):
        b = bytearray(good)
        for _ in range(int(rng.integers(1, 8))):
            b[int(rng.integers(2, len(b)))] = int(rng.integers(0, 256))
        try:
            _import(_images_doc([(bytes(b), "image/jpeg")]), tmp_path)
        except _lib.VgiError:
            pass


def test_sponza_jpeg_textures():
    import glob

    from violet_amd import _lib  # noqa: F401

    files = sorted(glob.glob(str(REF_ASSETS / "Sponza" / "*.jpg")))[:4]
    if not files:
        pytest.skip("reference assets absent")
    import tempfile

    tmp = Path(tempfile.mkdtemp())
    data = [Path(f).read_bytes() for f in files]
    mesh, st = _import(_images_doc([(b, "image/jpeg") for b in data]), tmp, form="glb")
    assert st.textures == len(files)
    for k, b in enumerate(data):
        got, want = mesh.textures[k].levels[0].astype(int), _pil_rgba(b).astype(int)
        dd = np.abs(got - want)
        assert dd.max() <= 3 and dd.mean() < 0.1, (files[k], dd.max(), dd.mean())
