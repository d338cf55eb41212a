"""Independent numpy restatements of violet's integer streams and encodings.

Written from the HLSL text (rand.hlsl:9-46,82-112; enc.inc.hlsl:11-63;
restir/reservoir.hlsl:52-117; restir/hash_grid.inc.hlsl:32-105) separately
from the C oracle, so that an agreement between the two pins transcription
mistakes in either.  Vectorised over uint32 arrays.
"""
from __future__ import annotations

import numpy as np

U = np.uint32


def jenkins(x):
    x = np.asarray(x, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        x = x + (x << U(10))
        x = x ^ (x >> U(6))
        x = x + (x << U(3))
        x = x ^ (x >> U(11))
        x = x + (x << U(15))
    return x


def lcg_init(px, py, seed):
    return jenkins(np.asarray(px, np.uint32) ^ (np.asarray(py, np.uint32) << U(16)) ^ U(seed))


def xorshift(x):
    x = np.asarray(x, dtype=np.uint32).copy()
    x ^= x << U(13)
    x ^= x >> U(17)
    x ^= x << U(5)
    return x


def u2f(x):
    bits = (np.asarray(x, np.uint32) >> U(9)) | U(0x3F800000)
    return bits.view(np.float32) - np.float32(1.0)


def xxh(words):
    """XXHash32 variant: init(w0).add(w1)...eval()."""
    p2, p3, p4, p5 = U(2246822519), U(3266489917), U(668265263), U(374761393)

    def rotl(h, r):
        return (h << U(r)) | (h >> U(32 - r))

    with np.errstate(over="ignore"):
        h = np.asarray(words[0], np.uint32) + p5
        for w in words[1:]:
            h = h + np.asarray(w, np.uint32) * p3
            h = p4 * rotl(h, 17)
        h = p4 * rotl(h, 17)
        h = p2 * (h ^ (h >> U(15)))
        h = p3 * (h ^ (h >> U(13)))
        return h ^ (h >> U(16))


def oct_decode(fx, fy):
    fx = np.asarray(fx, np.float32) * np.float32(2) - np.float32(1)
    fy = np.asarray(fy, np.float32) * np.float32(2) - np.float32(1)
    nz = (np.float32(1) - np.abs(fx)) - np.abs(fy)
    t = np.clip(-nz, 0, 1).astype(np.float32)
    nx = fx + np.where(fx >= 0, -t, t)
    ny = fy + np.where(fy >= 0, -t, t)
    n = np.stack([nx, ny, nz], -1).astype(np.float64)
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


def reservoir_decode_32b(u):
    u = np.asarray(u, np.uint32)
    M = u & U(0xFF)
    W = ((u >> U(1)) & U(0x7FFFFFE0)).view(np.float32)
    return M, W


def reservoir_encode_32b(M, W):
    M = np.minimum(np.asarray(M, np.uint32), U(255))
    return M | ((np.asarray(W, np.float32).view(np.uint32) << U(1)) & U(0xFFFFFF00))


def hash_grid_key(pos, cam):
    """(hash & 0xFFFF, checksum | 1, lod) of VertexDescriptor for f32 positions (N,3)."""
    pos = np.asarray(pos, np.float32)
    d = pos - np.asarray(cam, np.float32)
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    with np.errstate(divide="ignore"):
        lod = np.floor(np.maximum(0.5 * np.log2(d2.astype(np.float64)), 0.0)).astype(np.uint32)
    cell = np.float32(1.0 / 16.0) * (U(1) << lod).astype(np.float32)
    q = np.floor(pos * (np.float32(1) / cell)[:, None]).astype(np.int32).view(np.uint32)
    px, py, pz = q[:, 0], q[:, 1], q[:, 2]
    packed_x = px ^ (py << U(22))
    packed_y = (py >> U(10)) ^ (pz << U(22))
    h = xxh([packed_x, packed_y, lod]) & U(0xFFFF)
    c = xxh([py, pz, px, lod]) | U(1)
    return h, c, lod
