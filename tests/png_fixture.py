"""A PNG writer for the importer tests (ISO/IEC 15948): every colour type and
bit depth, per-row filter choice, Adam7 interlacing, several IDAT chunks and
zlib's block strategies, so the decoder meets stored, fixed-Huffman and
dynamic-Huffman blocks.  The expected RGBA8 texels are computed from the
same samples by `expected_rgba8`, independently of the decoder."""
from __future__ import annotations

import struct
import zlib

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def _pack_row(samples: np.ndarray, depth: int) -> bytes:
    s = samples.reshape(-1).astype(np.uint32)
    if depth == 8:
        return s.astype(np.uint8).tobytes()
    if depth == 16:
        return s.astype(">u2").tobytes()
    per = 8 // depth
    out = bytearray((len(s) * depth + 7) // 8)
    for i, v in enumerate(s):
        out[i // per] |= int(v) << (8 - depth * (i % per + 1))
    return bytes(out)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter(rows: list[bytes], bpp: int, filters) -> bytes:
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for y, row in enumerate(rows):
        f = filters[y % len(filters)]
        out.append(f)
        for i, x in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = (0, a, b, (a + b) >> 1, _paeth(a, b, c))[f]
            out.append((x - pred) & 0xFF)
        prev = row
    return bytes(out)


def encode_png(samples: np.ndarray, color_type: int, depth: int, *, filters=(0, 1, 2, 3, 4), interlace=False,
               palette=None, trns=None, level=6, strategy=zlib.Z_DEFAULT_STRATEGY, idat_split=3) -> bytes:
    """samples: (h, w) or (h, w, channels) integers below 2**depth."""
    samples = np.asarray(samples)
    h, w = samples.shape[:2]
    ch = CHANNELS[color_type]
    samples = samples.reshape(h, w, ch)
    bpp = max(1, ch * depth // 8)
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    raw = b""
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        raw += _filter([_pack_row(r, depth) for r in sub], bpp, filters)
    co = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
    z = co.compress(raw) + co.flush()
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, int(interlace)))
    out += _chunk(b"tEXt", b"Comment\x00ancillary chunk")
    if palette is not None:
        out += _chunk(b"PLTE", np.asarray(palette, np.uint8).reshape(-1).tobytes())
    if trns is not None:
        if color_type == 3:
            out += _chunk(b"tRNS", np.asarray(trns, np.uint8).tobytes())
        else:
            out += _chunk(b"tRNS", np.asarray(trns, ">u2").tobytes())
    step = max(1, -(-len(z) // idat_split))
    for i in range(0, len(z), step):
        out += _chunk(b"IDAT", z[i:i + step])
    return out + _chunk(b"IEND", b"")


def expected_rgba8(samples: np.ndarray, color_type: int, depth: int, palette=None, trns=None) -> np.ndarray:
    samples = np.asarray(samples).astype(np.int64)
    h, w = samples.shape[:2]
    s = samples.reshape(h, w, CHANNELS[color_type])

    def to8(v):
        if depth == 8:
            return v
        if depth == 16:
            return (2 * v + 257) // 514
        return v * 255 // ((1 << depth) - 1)

    out = np.zeros((h, w, 4), np.int64)
    if color_type == 3:
        pal = np.asarray(palette, np.int64).reshape(-1, 3)
        out[..., :3] = pal[s[..., 0]]
        a = np.full(len(pal), 255, np.int64)
        if trns is not None:
            a[:len(trns)] = trns
        out[..., 3] = a[s[..., 0]]
    elif color_type in (0, 4):
        out[..., :3] = to8(s[..., :1])
        out[..., 3] = to8(s[..., 1]) if color_type == 4 else 255
        if color_type == 0 and trns is not None:
            out[..., 3] = np.where(s[..., 0] == trns[0], 0, 255)
    else:
        out[..., :3] = to8(s[..., :3])
        out[..., 3] = to8(s[..., 3]) if color_type == 6 else 255
        if color_type == 2 and trns is not None:
            out[..., 3] = np.where(np.all(s[..., :3] == np.asarray(trns), axis=-1), 0, 255)
    return out.astype(np.uint8)
