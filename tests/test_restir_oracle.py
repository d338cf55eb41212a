"""Pins of the oracle's ReSTIR and hash-grid restatements (CPU only).

* hash grid: the oracle's query select / compact / decay against pure-Python
  models written from the HLSL text (hash_grid.inc.hlsl:124-185,
  hash_grid_query_select.hlsl:22-56, hash_grid_query_compact.hlsl:23-63,
  hash_grid_decay.hlsl:13-128), on inputs chosen to overflow buckets and to
  hit the decay pass's early-out quirk;
* resampling parameters: the explicit reference constants (M cap 16, 9
  spiral taps, radius 0.05*min(w,h)) give bit-identical results to the
  defaults; a cap bounds the temporal M exactly; a different tap count or
  radius changes the spatial result.
"""
from __future__ import annotations

import numpy as np
import pytest

from np_ref import jenkins, xorshift, reservoir_decode_32b, reservoir_encode_32b

NB, BS = 65536, 16
NC = NB * BS


@pytest.fixture(scope="module")
def R32():
    import oracle
    from restir_harness import make_mesh

    return oracle.OracleRestir(32, oracle.OracleScene(32, make_mesh(200)))


# ---------------------------------------------------------------------------
# hash grid models
# ---------------------------------------------------------------------------
def select_model(queries, count, checksums, frame_hash):
    """Sequential schedule of the select pass: find-or-insert the cell of each
    query in index order, then atomic-max the selection word."""
    ck = checksums.copy()
    sel = {}
    for q in range(count):
        h, c = int(queries[q, 4]), int(queries[q, 5])
        addr = None
        for a in range(h * BS, h * BS + BS):
            if ck[a] == c:
                addr = a
                break
            if ck[a] == 0:
                ck[a] = c
                addr = a
                break
        if addr is None:
            continue  # bucket full of other cells
        r = int(xorshift(np.uint32(jenkins(np.uint32(q ^ frame_hash)))))
        word = (r & (0xFFFFFFFF << 18) & 0xFFFFFFFF) | (1 << 31) | q
        sel[addr] = max(sel.get(addr, 0), word)
    return ck, sel


def decay_model(decay, cells):
    decay = decay.copy()
    cells = cells.copy()
    for b in np.nonzero(decay.reshape(NB, BS).any(axis=1))[0]:
        off = int(b) * BS
        mask, local, s = 0, [0, 0, 0, 0], 0
        while s < 4:
            sa = off + 4 * s
            packed = int.from_bytes(decay[sa:sa + 4].tobytes(), "little")
            if packed == 0:
                break
            upd, d = 0, 0
            for i in range(4):
                d = (packed >> (8 * i)) & 0xFF
                if d == 0:
                    break
                d -= 1
                if d == 0:
                    cells[sa + i] = 0
                else:
                    upd |= d << (8 * i)
                    mask |= 1 << (4 * s + i)
            local[s] = upd
            s += 1
            if d == 0:
                break  # also taken when the set's last cell just expired
        while mask:
            last = mask.bit_length() - 1
            inv = ~mask & 0xFFFFFFFF
            first = (inv & -inv).bit_length() - 1
            if first >= last:
                break
            v = (local[last >> 2] >> (8 * (last & 3))) & 0xFF
            local[last >> 2] &= ~(0xFF << (8 * (last & 3)))
            local[first >> 2] |= v << (8 * (first & 3))
            cells[off + first] = cells[off + last]
            cells[off + last] = 0
            mask = (mask | (1 << first)) & ~(1 << last)
        for i in range(s):
            decay[off + 4 * i: off + 4 * i + 4] = np.frombuffer(int(local[i]).to_bytes(4, "little"), np.uint8)
    return decay, cells


def test_hash_grid_select_and_compact_match_models(R32):
    rng = np.random.default_rng(5)
    n = 600
    q = np.zeros((131072, 6), np.uint32)
    q[:n, 4] = rng.integers(0, 6, n)                   # 6 buckets -> collisions
    pool = (rng.integers(1, 2**31, 30) * 2 + 1).astype(np.uint32)  # 30 checksums > 16 slots
    q[:n, 5] = pool[rng.integers(0, 30, n)]
    cells = np.zeros((NC, 4), np.uint32)
    cells[2 * BS:2 * BS + 3, 0] = pool[:3]             # bucket 2 already holds 3 cells
    frame_hash = 0x9E3779B9
    got_cells, sel = R32.hg_select(q, n, cells, frame_hash)
    ck, msel = select_model(q, n, cells[:, 0], frame_hash)
    assert np.array_equal(got_cells[:, 0], ck)
    assert np.array_equal(got_cells[:, 1:], cells[:, 1:])  # radiance untouched
    nz = np.nonzero(sel)[0]
    assert dict(zip(nz.tolist(), sel[nz].tolist())) == msel
    # some queries found their bucket full
    assert any(all(ck[h * BS + k] != 0 for k in range(BS)) for h in range(6))
    idx, addr = R32.hg_compact(sel)
    assert sorted(zip(addr.tolist(), idx.tolist())) == sorted((a, w & 0x1FFFF) for a, w in msel.items())
    # every selected query belongs to its cell
    for a, i in zip(addr.tolist(), idx.tolist()):
        assert ck[a] == q[i, 5] and a // BS == q[i, 4]


def test_hash_grid_decay_matches_model(R32):
    rng = np.random.default_rng(7)
    decay = np.zeros(NC, np.uint8)
    cells = np.zeros((NC, 4), np.uint32)
    for b in range(200):
        k = int(rng.integers(0, BS + 1))
        off = b * BS
        decay[off:off + k] = rng.integers(1, 4, k)
        cells[off:off + k, 0] = rng.integers(1, 2**32 - 1, k, dtype=np.uint64).astype(np.uint32) | 1
        cells[off:off + k, 1:] = rng.integers(0, 2**30, (k, 3))
    # a bucket whose 4th cell of the first set expires: the pass stops there
    # (quirk) and leaves the second set untouched
    decay[300 * BS:300 * BS + 8] = [3, 3, 3, 1, 2, 2, 2, 2]
    cells[300 * BS:300 * BS + 8, 0] = np.arange(1, 17, 2)
    d_model, c_model = decay_model(decay, cells)
    d_got, c_got = R32.hg_decay(decay, cells)
    assert np.array_equal(d_got, d_model)
    assert np.array_equal(c_got, c_model)
    assert list(d_got[300 * BS:300 * BS + 8]) == [2, 2, 2, 0, 2, 2, 2, 2]
    # a live cell keeps its checksum; an expired or emptied one is cleared
    assert np.all(c_got[d_got > 0, 0] != 0)
    assert np.all(c_got[(d_got == 0) & (decay > 0), 0] == 0)


# ---------------------------------------------------------------------------
# resampling parameters
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def frame_inputs(R32):
    from violet_amd import synth
    from restir_harness import orbit_view

    w, h = 40, 28
    v0, v1 = orbit_view(0, w, h), orbit_view(1, w, h)
    fp0 = synth.make_frame_params(v0, None)
    fp1 = synth.make_frame_params(v1, v0)
    d0, g0 = R32.gbuffer(fp0, w, h)
    d1, g1 = R32.gbuffer(fp1, w, h)
    cells = np.zeros((NC, 4), np.uint32)
    qb = np.zeros((131072, 6), np.uint32)
    cnt = np.zeros(1, np.uint32)
    tr = R32.trace_res(cells, qb, cnt)
    hit, rad = R32.diff_sample_gen(fp1, tr, g1, d1, 77, False)
    rng = np.random.default_rng(3)
    M = rng.integers(0, 200, (h, w))
    W = rng.random((h, w)).astype(np.float32)
    prev_rsv = reservoir_encode_32b(M, W)
    prev_hit, prev_rad = R32.diff_sample_gen(fp0, tr, g0, d0, 78, False)
    return dict(fp=fp1, d0=d0, d1=d1, g1=g1, hit=hit, rad=rad, prev=(prev_rsv, prev_hit, prev_rad), M=M)


def test_explicit_reference_constants_equal_defaults(R32, frame_inputs):
    f = frame_inputs
    a = R32.diff_temporal(f["fp"], f["d0"], f["d1"], f["hit"], f["rad"], *f["prev"], 5, True, True)
    b = R32.diff_temporal(f["fp"], f["d0"], f["d1"], f["hit"], f["rad"], *f["prev"], 5, True, True, m_cap=16)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    ao = np.ones(f["d1"].shape + (2,), np.float32)
    s0 = R32.diff_spatial(f["fp"], f["g1"], f["d1"], ao, *a, 9)
    h, w = f["d1"].shape
    s1 = R32.diff_spatial(f["fp"], f["g1"], f["d1"], ao, *a, 9, num_taps=9, radius_px=np.float32(0.05) * min(w, h))
    for x, y in zip(s0, s1):
        assert np.array_equal(x, y)
    s2 = R32.diff_spatial(f["fp"], f["g1"], f["d1"], ao, *a, 9, num_taps=5, radius_px=12.0)
    assert not np.array_equal(s0[0], s2[0])


@pytest.mark.parametrize("cap", [2, 30])
def test_temporal_m_cap_bounds_history(R32, frame_inputs, cap):
    f = frame_inputs
    rsv, _, _ = R32.diff_temporal(f["fp"], f["d0"], f["d1"], f["hit"], f["rad"], *f["prev"], 5, True, True,
                                  m_cap=cap)
    M_out, _ = reservoir_decode_32b(rsv)
    geo = f["d1"] > 0
    assert M_out[geo].max() <= cap + 1
    assert (M_out[geo] == cap + 1).any()
