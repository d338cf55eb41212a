"""Pins of the oracle's integer streams and encodings (CPU).

The reference ships no test vectors (violet has no tests), so each function
is pinned by (a) public known-answer values of the published algorithm it
is an instance of, (b) an independent numpy restatement (tests/np_ref.py),
(c) exhaustive or closed-form checks, and (d) the committed golden file
tests/golden/kat_vectors.json written by tests/golden/make_kat.py.
"""
from __future__ import annotations

import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

import np_ref

GOLDEN = Path(__file__).parent / "golden" / "kat_vectors.json"


def test_jenkins_matches_one_at_a_time_public_vector(oracle64):
    # jenkins_hash(x) for x < 256 is Bob Jenkins' one-at-a-time hash of the
    # single byte x; the published OAAT("a") = 0xca2e9442.
    assert oracle64.lib.oracle_jenkins(ord("a")) == 0xCA2E9442
    assert oracle64.lib.oracle_jenkins(0) == 0


def test_xorshift_marsaglia_vector(oracle64):
    # Marsaglia (2003) xor32 with shifts (13,17,5) from y=2463534242 -> 723471715
    st = (C.c_uint32 * 1)()
    rr = (C.c_double * 1)()
    oracle64.lib.oracle_lcg_stream(2463534242, 1, st, rr)
    assert st[0] == 723471715
    assert rr[0] == (723471715 >> 9) * 2.0 ** -23


def test_rng_streams_match_numpy(oracle64, oracle32):
    rng = np.random.default_rng(1)
    xs = rng.integers(0, 2**32, 2000, dtype=np.uint64).astype(np.uint32)
    ref = np_ref.jenkins(xs)
    got = np.array([oracle64.lib.oracle_jenkins(int(x)) for x in xs], np.uint32)
    assert np.array_equal(ref, got)
    px, py = rng.integers(0, 4096, 500), rng.integers(0, 4096, 500)
    seed = 0x1234ABCD
    ref = np_ref.lcg_init(px, py, seed)
    got = np.array([oracle32.lib.oracle_lcg_init(int(a), int(b), seed) for a, b in zip(px, py)], np.uint32)
    assert np.array_equal(ref, got)
    for lib, dt, ct in ((oracle64, np.float64, C.c_double), (oracle32, np.float32, C.c_float)):
        n = 64
        st = (C.c_uint32 * n)()
        rr = (ct * n)()
        lib.lib.oracle_lcg_stream(12345, n, st, rr)
        s = np.uint32(12345)
        for k in range(n):
            s = np_ref.xorshift(s)
            assert st[k] == int(s)
            assert rr[k] == float(np_ref.u2f(s))
            assert 0.0 <= rr[k] < 1.0


def test_xxhash_variant_matches_numpy(oracle64):
    rng = np.random.default_rng(2)
    for n in (1, 2, 3, 4):
        for _ in range(200):
            w = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
            arr = (C.c_uint32 * n)(*[int(v) for v in w])
            assert oracle64.lib.oracle_xxh(arr, n) == int(np_ref.xxh([np.uint32(v) for v in w]))


def test_f16_exhaustive_decode_and_rne_encode(oracle64):
    lib = oracle64.lib
    # decode: every half bit pattern vs numpy's float16
    codes = np.arange(65536, dtype=np.uint16)
    ref = codes.view(np.float16).astype(np.float64)
    got = np.array([lib.oracle_f16_to_f64(int(c)) for c in codes])
    finite = np.isfinite(ref)
    assert np.array_equal(ref[finite], got[finite])
    assert np.array_equal(np.isnan(ref), np.isnan(got))
    # encode: round-half-even, including ties, subnormals and overflow
    rng = np.random.default_rng(3)
    vals = np.concatenate([
        rng.standard_normal(4000).astype(np.float32) * 100,
        (rng.random(2000).astype(np.float32) * 2 ** -14),
        np.array([65504, 65519.99, 65520, 1e9, -1e9, 2 ** -25, 2 ** -24, 3 * 2 ** -25, 0.0, -0.0], np.float32),
        (codes[:31744].view(np.float16).astype(np.float32) + np.float32(0)),  # exact halves
    ])
    # midpoints between consecutive halves (ties)
    h = codes[1:31743].view(np.float16).astype(np.float64)
    mids = ((h[:-1] + h[1:]) / 2).astype(np.float32)
    vals = np.concatenate([vals, mids[np.float64(mids) == (h[:-1] + h[1:]) / 2]])
    with np.errstate(over="ignore"):  # 65520 and +-1e9 round to +-inf on purpose
        ref = vals.astype(np.float16).view(np.uint16)
    got = np.array([lib.oracle_f32_to_f16(float(v)) for v in vals], np.uint32)
    assert np.array_equal(ref.astype(np.uint32), got)


def test_oct_decode_matches_numpy_and_unit_length(oracle64):
    lib = oracle64.lib
    out = (C.c_double * 3)()
    codes = np.arange(0, 65536, 7, dtype=np.uint32)
    for e in codes:
        lib.oracle_oct_dec_u16(int(e), out)
        ref = np_ref.oct_decode((e & 0xFF) / np.float32(255.0), ((e >> 8) & 0xFF) / np.float32(255.0))
        assert np.allclose(np.array(out[:]), ref, atol=1e-6)
        assert abs(math.sqrt(sum(v * v for v in out)) - 1.0) < 1e-12


def test_oct_closed_forms(oracle64):
    lib = oracle64.lib
    out = (C.c_double * 3)()
    # (0,0,1) -> (0.5,0.5); (1,0,0) -> (1,0.5); (0,0,-1) wraps to a corner
    for n, enc in (((0.0, 0.0, 1.0), (0.5, 0.5)), ((1.0, 0.0, 0.0), (1.0, 0.5)), ((0.0, -1.0, 0.0), (0.5, 0.0))):
        arr = (C.c_double * 3)(*n)
        e = lib.oracle_oct_enc_u32(arr)
        assert (e & 0xFFFF) == int(enc[0] * 65535.0)
        assert (e >> 16) == int(enc[1] * 65535.0)
    lib.oracle_oct_dec_u32(0, out)  # f = (-1,-1) -> (0,0,-1)
    assert np.allclose(out[:], (0, 0, -1))
    # round trip error of the 16:16 encoding on random unit vectors
    rng = np.random.default_rng(4)
    v = rng.standard_normal((500, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for n in v:
        e = lib.oracle_oct_enc_u32((C.c_double * 3)(*n))
        lib.oracle_oct_dec_u32(e, out)
        assert np.dot(np.array(out[:]), n) > 1 - 1e-8


def test_reservoir_32b_quirk(oracle64):
    lib = oracle64.lib
    M = C.c_uint32()
    W = C.c_float()
    rng = np.random.default_rng(5)
    for m in range(0, 300, 7):
        w = np.float32(rng.random() * 10)
        u = lib.oracle_rsv_enc(m, float(w))
        assert u == int(np_ref.reservoir_encode_32b(m, w))
        lib.oracle_rsv_dec(u, C.byref(M), C.byref(W))
        rm, rw = np_ref.reservoir_decode_32b(np.uint32(u))
        assert M.value == int(rm) == min(m, 255)
        assert W.value == float(rw)
        # W keeps its top 24 bits (sign dropped); M bits 6..7 leak into W's low mantissa
        wbits = np.float32(w).view(np.uint32)
        leak = (min(m, 255) >> 6) & 3
        assert np.float32(W.value).view(np.uint32) == (wbits & 0x7FFFFF80) | (leak << 5)


def test_sincos_accuracy_vs_libm(oracle32, oracle64):
    s = C.c_float()
    c = C.c_float()
    xs = np.concatenate([np.linspace(-60, 60, 20001), [0.0, math.pi / 4, math.pi / 2, math.pi, 1e-8]]).astype(np.float32)
    worst = 0.0
    for x in xs:
        oracle32.lib.oracle_sincos(float(x), C.byref(s), C.byref(c))
        for got, ref in ((s.value, math.sin(float(x))), (c.value, math.cos(float(x)))):
            ulp = np.spacing(np.float32(abs(ref))) if ref != 0 else np.float32(1e-45)
            err = abs(got - ref) / max(float(ulp), 1e-45)
            worst = max(worst, err if abs(ref) > 1e-6 else abs(got - ref) / 1e-7)
    assert worst <= 2.0


def test_log2_accuracy_vs_libm(oracle32):
    """The single-precision log2 shared with the GPU (texture LOD) against
    libm double: within 2 ulp of the result, exact at powers of two,
    including subnormal arguments."""
    f = oracle32.lib.oracle_log2
    f.restype, f.argtypes = C.c_float, [C.c_float]
    rng = np.random.default_rng(5)
    xs = np.concatenate([np.exp2(rng.uniform(-149, 127, 20000)), rng.uniform(0.5, 2.0, 20000),
                         [1.0, 2.0, 0.5, 1024.0, 2.0 ** -140, 3.0e38, 0.70710677, 1.4142135]]).astype(np.float32)
    worst = 0.0
    for x in xs[xs > 0]:
        got, ref = f(float(x)), math.log2(float(x))
        ulp = float(np.spacing(np.float32(abs(ref)))) if ref != 0 else 1e-45
        worst = max(worst, abs(got - ref) / ulp)
    assert worst <= 2.0
    for e in range(-149, 128):
        assert f(2.0 ** e) == float(e)


def test_golden_file(oracle64):
    data = json.loads(GOLDEN.read_text())
    for x, h in data["jenkins"]:
        assert oracle64.lib.oracle_jenkins(x) == h
    for (px, py, seed), v in data["lcg_init"]:
        assert oracle64.lib.oracle_lcg_init(px, py, seed) == v
    for words, h in data["xxh"]:
        arr = (C.c_uint32 * len(words))(*words)
        assert oracle64.lib.oracle_xxh(arr, len(words)) == h
