"""Pins of the oracle's order-independent cache update (CPU only).

The library's `vgi_pass_hash_grid_*_ordered` passes have no reference
counterpart: they replace the parts of the reference's cache update that
depend on the GPU's schedule (slot of a new key, selected query, compaction
order, cache values the cache rays read).  What they must keep is what the
reference guarantees for every schedule.  Checked here on the oracle:

* every permutation of the query buffer gives the same cells, selection
  words, selected query contents and compact lists (the point of the passes);
* the cells equal the reference's select (hash_grid_query_select.hlsl:22-56,
  the oracle's sequential restatement) run on the queries sorted by
  checksum, one of the reference's schedules;
* a key queried again behind a slot the decay pass emptied: with a smaller
  new key in front of it the hole goes to that key and the older copy is
  selected; alone, it is inserted into the hole in front of its older copy;
* the selected query is uniform over a cell's distinct queries (chi-square
  over frame seeds), as the reference's index-seeded draw is;
* compaction lists the selected cells in cell order, once each;
* the query threshold word: a candidate is kept when its hash is at least
  the word, and the word set from one frame's kept count keeps about
  2^16 of the next frame's candidates (binomial bounds), whatever the
  previous word was; 0 (keep all) when the candidates fit.
"""
from __future__ import annotations

import numpy as np
import pytest

NB, BS = 65536, 16
NC = NB * BS


@pytest.fixture(scope="module")
def R32():
    import oracle
    from restir_harness import make_mesh

    return oracle.OracleRestir(32, oracle.OracleScene(32, make_mesh(200)))


def _queries(rng, n_keys, n_queries, buckets):
    """Queries over `n_keys` keys in a few buckets (collisions), several
    queries per key with distinct positions, plus exact duplicates."""
    hashes = rng.choice(buckets, n_keys).astype(np.uint32)
    checks = rng.integers(1, 2**32, n_keys, dtype=np.uint64).astype(np.uint32)
    k = rng.integers(0, n_keys, n_queries)
    q = np.zeros((n_queries, 6), np.uint32)
    q[:, :3] = rng.random((n_queries, 3)).astype(np.float32).view(np.uint32)
    q[:, 3] = rng.integers(0, 2**32, n_queries, dtype=np.uint64).astype(np.uint32)
    q[:, 4], q[:, 5] = hashes[k], checks[k]
    dup = rng.random(n_queries) < 0.1
    src = rng.integers(0, n_queries, n_queries)
    q[dup] = q[src[dup]]
    return q


def _initial_cells(rng, buckets):
    """Older cells: a prefix of each bucket, with one slot emptied by the
    decay pass (a hole) in some buckets."""
    cells = np.zeros((NC, 4), np.uint32)
    for b in buckets:
        n = int(rng.integers(0, 6))
        for s in range(n):
            a = b * BS + s
            cells[a, 0] = rng.integers(1, 2**32, dtype=np.uint64)
            cells[a, 1:] = rng.random(3).astype(np.float32).view(np.uint32)
        if n > 2 and rng.random() < 0.5:
            cells[b * BS + int(rng.integers(0, n - 1))] = 0
    return cells


def _bucket_sets(cells):
    b = np.nonzero(cells[:, 0])[0]
    return sorted(zip((b // BS).tolist(), cells[b, 0].tolist()))


def _run(R, q, cells, fh):
    c, sel, win = R.hg_select_ordered(q, len(q), cells, fh)
    idx, addr = R.hg_compact_ordered(sel, win)
    return c, sel, win, idx, addr


def test_every_query_order_gives_the_same_update(R32):
    rng = np.random.default_rng(5)
    buckets = rng.choice(NB, 150, replace=False)
    q = _queries(rng, 300, 3000, buckets)
    cells = _initial_cells(rng, buckets)
    # older copies of some queried keys behind a hole
    for b in buckets[:10]:
        mine = q[q[:, 4] == b]
        free = np.nonzero(cells[b * BS: b * BS + BS, 0] == 0)[0]
        if len(mine) and len(free) > 2:
            cells[b * BS + free[1], 0] = mine[0, 5]
    fh = 0x9E3779B9
    c0, s0, w0, i0, a0 = _run(R32, q, cells, fh)
    sel_cells = np.nonzero(s0)[0]
    assert len(sel_cells) > 200
    full = (c0[:, 0] != 0).reshape(NB, BS).all(1)
    assert not full.any(), "an overflowing bucket keeps schedule-dependent keys"
    for seed in range(4):
        p = np.random.default_rng(100 + seed).permutation(len(q))
        c, s, w, i, a = _run(R32, q[p], cells, fh)
        assert np.array_equal(c, c0) and np.array_equal(s, s0)
        assert np.array_equal(q[p][w[sel_cells]], q[w0[sel_cells]])
        assert np.array_equal(a, a0) and np.array_equal(q[p][i], q[i0])


def test_cells_are_the_reference_select_in_checksum_order(R32):
    rng = np.random.default_rng(7)
    buckets = rng.choice(NB, 60, replace=False)
    q = _queries(rng, 400, 2500, buckets)
    cells = _initial_cells(rng, buckets)
    for b in buckets[:30]:  # older copies of queried keys behind holes
        mine = q[q[:, 4] == b]
        free = np.nonzero(cells[b * BS: b * BS + BS, 0] == 0)[0]
        if len(mine) and len(free) > 2:
            cells[b * BS + free[1], 0] = mine[0, 5]
    fh = 12345
    c_ord, s_ord, _ = R32.hg_select_ordered(q, len(q), cells, fh)
    srt = q[np.argsort(q[:, 5], kind="stable")]
    c_ref, s_ref = R32.hg_select(srt, len(srt), cells, fh)
    assert np.array_equal(c_ord, c_ref)
    # the same cells are selected (one per key that got a cell)
    assert np.array_equal(s_ord != 0, s_ref != 0)
    # the words carry bit 31 and radiance is untouched
    assert np.all(s_ord[s_ord != 0] >> 31 == 1)
    assert np.array_equal(c_ord[:, 1:], cells[:, 1:])


def test_key_behind_a_hole_is_inserted_in_front_of_its_older_copy(R32):
    b = 777
    cells = np.zeros((NC, 4), np.uint32)
    # slots 0, 1 live; 2 emptied by the decay; 3 an older copy of key K; 4 live
    cells[b * BS + 0, 0], cells[b * BS + 1, 0] = 11, 22
    K = 0xDEADBEEF
    cells[b * BS + 3, 0], cells[b * BS + 4, 0] = K, 44
    cells[b * BS + 3, 1:] = np.float32([1.0, 2.0, 3.0]).view(np.uint32)
    new = [0x1000, 0x0500, 0x2000]  # new keys, inserted after K in query order below
    rows = []
    for i, key in enumerate([*new, K, K]):
        rows.append([np.float32(i).view(np.uint32), 0, 0, 0, b, key])
    q = np.array(rows, np.uint32)
    for p in ([0, 1, 2, 3, 4], [4, 3, 2, 1, 0], [3, 0, 4, 2, 1]):
        c, s, w = R32.hg_select_ordered(q[p], len(q), cells, 99)
        bucket = c[b * BS: b * BS + BS, 0]
        # 0x0500 goes first and takes the hole; K then finds its older copy
        # at 3 (with its radiance), the others take the next free slots
        assert bucket[:8].tolist() == [11, 22, 0x0500, K, 44, 0x1000, 0x2000, 0]
        assert s[b * BS + 3] >> 31 and q[p][w[b * BS + 3], 5] == K
    # K alone: inserted into the hole in front of its older copy, which the
    # lookup (first match) no longer reaches
    c, s, w = R32.hg_select_ordered(q[3:], 2, cells, 99)
    assert c[b * BS: b * BS + 6, 0].tolist() == [11, 22, K, K, 44, 0]
    assert s[b * BS + 2] >> 31 and s[b * BS + 3] == 0


def test_selected_query_is_uniform_over_a_cells_queries(R32):
    b, key = 4242, 0xABCDEF
    n = 5
    q = np.zeros((n, 6), np.uint32)
    q[:, 0] = np.arange(n, dtype=np.float32).view(np.uint32)
    q[:, 4], q[:, 5] = b, key
    counts = np.zeros(n)
    cells = np.zeros((NC, 4), np.uint32)
    trials = 1500
    for t in range(trials):
        _, s, w = R32.hg_select_ordered(q, n, cells, (t * 2654435761) & 0xFFFFFFFF)
        counts[w[b * BS]] += 1
    chi2 = ((counts - trials / n) ** 2 / (trials / n)).sum()
    assert chi2 < 18.5, f"selection is not uniform over the cell's queries: {counts}"  # chi2(4) p = 0.001


def test_compaction_lists_selected_cells_in_cell_order(R32):
    rng = np.random.default_rng(11)
    buckets = rng.choice(NB, 100, replace=False)
    q = _queries(rng, 500, 4000, buckets)
    c, s, w = R32.hg_select_ordered(q, len(q), np.zeros((NC, 4), np.uint32), 3)
    idx, addr = R32.hg_compact_ordered(s, w)
    assert np.array_equal(addr, np.nonzero(s)[0])
    assert np.array_equal(idx, w[addr])
    # each listed query maps to its cell
    assert np.array_equal(q[idx, 5], c[addr, 0])


@pytest.mark.parametrize("candidates", [50_000, 400_000, 2_000_000])
def test_query_threshold_keeps_about_the_target(R32, candidates):
    target = 131072 // 2
    rng = np.random.default_rng(candidates)
    thr0 = R32.hg_query_threshold_for(3 * candidates)  # a first guess 3x too high
    h = rng.integers(0, 2**32, candidates, dtype=np.uint64)
    kept0 = int((h >= thr0).sum())
    # the select pass's update from the kept count
    thr1 = R32.hg_next_query_threshold(kept0, thr0)
    h1 = rng.integers(0, 2**32, candidates, dtype=np.uint64)
    kept1 = int((h1 >= thr1).sum())
    if candidates <= target:
        assert thr1 == 0 and kept1 == candidates
    else:
        sd = np.sqrt(target)
        assert abs(kept1 - target) < 6 * sd + 0.01 * target, (kept0, thr1, kept1)
        assert kept1 < 131072
