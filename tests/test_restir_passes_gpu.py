"""Per-pass replay through the C ABI of the ReSTIR passes whose results the
frame tests only see folded into later passes: the hash-grid cache update
(decay, query select, compaction, cache rays; the reference's passes and
the order-independent ones, which must match the oracle bit for bit
everywhere), the diffuse and specular
sample validation, specular sample generation with a screen cache, and the
specular temporal resampling with history (its balance-heuristic MIS merge).

A renderer runs 9 frames of the orbit (validation frames 4 and 8).  Each pass
is then called directly (vgi_pass_*) on device copies of the inputs it saw in
frame k, and its outputs are compared with the oracle's on the same inputs:
* decay: bit-exact (one lane per bucket, deterministic).
* query select: the slot a CAS insertion takes and the query index whose
  atomicMax wins depend on scheduling only through the slot order inside a
  bucket, so each bucket is compared as the set of its cells
  {checksum, radiance bits, selection word}.
* compaction: the set of (query index, cell) pairs; the output order follows
  the order in which waves reserve their output ranges.
* cache rays: decay bytes and the appended queries exact; each updated
  cell's radiance equals the oracle's blend computed with the pass's cache
  lookups reading either the cells before the pass or the GPU's cells after
  it.  The reference reads and writes the one storage buffer in the same
  dispatch (hash_grid_raygen.hlsl:74-86 after trace_radiance's lookup,
  raytrace.inc.hlsl:172-188), so a ray whose hit lands in a cell another ray
  updates may see either value; each ray makes one lookup, so no third value
  exists.
* validation, sample generation, temporal resampling: bit-exact with the
  FP32 twin, and the GPU replay equals the renderer's own frame bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from parity import assert_close_frac
from hash_grid_forms import decay_closed_form
from restir_harness import make_mesh, multiset, run_gpu_frames, seeds_for_frame

pytestmark = pytest.mark.gpu

W, H = 96, 64
FRAMES = 9
NCELLS = 65536 * 16
NQ = 128 * 1024


@pytest.fixture(scope="module")
def run():
    """The renderer with the reference's cache passes."""
    from violet_amd import renderer

    mesh = make_mesh()
    fps, outs, snaps, r, sc = run_gpu_frames(mesh, W, H, FRAMES,
                                             config=renderer.RestirConfig(ordered_cache_update=False))
    return dict(mesh=mesh, fps=fps, snaps=snaps, r=r, sc=sc)


@pytest.fixture(scope="module")
def run_ordered(run):
    """The renderer with its default, order-independent cache passes."""
    from violet_amd import renderer

    fps, outs, snaps, r, sc = run_gpu_frames(run["mesh"], W, H, FRAMES, config=renderer.RestirConfig())
    return dict(mesh=run["mesh"], fps=fps, snaps=snaps, r=r, sc=sc)


@pytest.fixture(scope="module")
def o32(run):
    import oracle

    return oracle.OracleRestir(32, oracle.OracleScene(32, run["mesh"]))


@pytest.fixture(scope="module")
def o64(run):
    import oracle

    return oracle.OracleRestir(64, oracle.OracleScene(64, run["mesh"]))


class Dev:
    """Device copies of host arrays for direct entry-point calls."""

    def __init__(self):
        self.keep = []

    def put(self, a: np.ndarray) -> int:
        from violet_amd import gpu

        t = gpu.dev(np.ascontiguousarray(a))
        self.keep.append(t)
        return t.data_ptr()

    def zeros(self, nbytes: int) -> int:
        from violet_amd import gpu

        t = gpu.zeros(nbytes)
        self.keep.append(t)
        return t.data_ptr()

    @staticmethod
    def get(p: int, dtype, shape) -> np.ndarray:
        import torch

        from violet_amd import gpu

        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        torch.cuda.synchronize()
        return gpu.copy_from_ptr(p, n).cpu().numpy().view(dtype).reshape(shape).copy()


def call(name, *args):
    from violet_amd import _lib, gpu

    _lib.call(name, *args, gpu.stream())


def gpu_trace_res(dev, run, storage_ptr, prev):
    """Trace resources of a frame-k pass: the cache, the previous frame's
    denoised diffuse and depth (screen cache), fresh query buffer + counter."""
    from violet_amd import _abi

    tr = _abi.TraceResources()
    tr.scene_tlas = run["sc"].handle
    tr.sky_radiance[:] = [1.0, 1.0, 1.0]
    tr.storage = 0
    tr.prev_indirect_diffuse_texture = dev.put(prev["out_diffuse"])
    tr.prev_depth_texture = dev.put(prev["depth"])
    tr.prev_width, tr.prev_height = W, H
    tr.hash_grid_storage_buffer = storage_ptr
    q, cnt = dev.zeros(NQ * 24), dev.zeros(4)
    tr.rw_hash_grid_query_buffer = q
    tr.rw_hash_grid_query_counter_buffer = cnt
    return tr, q, cnt


def read_queries(q, cnt):
    n = int(Dev.get(cnt, np.uint32, (1,))[0])
    rows = Dev.get(q, np.uint32, (NQ, 6))
    return rows[: min(n, NQ)], n


def bits_equal(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    neq = a.view(np.uint32) != b.view(np.uint32) if a.dtype.itemsize == 4 else a != b
    frac = float(np.mean(neq))
    assert frac == 0.0, f"{what}: {frac:.4%} of words differ (first at {np.argwhere(neq)[:3].tolist()})"


def hg_frames(run):
    return [k for k in range(1, FRAMES) if run["snaps"][k]["flags"]["hg"]]


def bucket_sets(cells: np.ndarray, sel: np.ndarray) -> np.ndarray:
    """Occupied cells as rows (bucket, checksum, r, g, b, selection word) in a
    canonical order: the content of each bucket without its slot order."""
    occ = np.nonzero(cells[:, 0])[0]
    rows = np.column_stack([occ >> 4, cells[occ], sel[occ]]).astype(np.uint64)
    return rows[np.lexsort(rows.T[::-1])]


def test_hash_grid_decay_expires_and_compacts_buckets(vgi, o32):
    """Synthetic cache: every bucket holds a random prefix of occupied cells
    with decay 1..4, so a quarter of the cells expire and the pass moves the
    bucket's last cells into the holes.  GPU == oracle bit for bit, and both
    equal the closed form of decay_closed_form()."""
    rng = np.random.default_rng(7)
    n_occ = rng.integers(0, 17, 65536)
    slot = np.tile(np.arange(16), 65536)
    occupied = slot < np.repeat(n_occ, 16)
    decay = np.where(occupied, rng.integers(1, 5, NCELLS), 0).astype(np.uint8)
    cells = np.zeros((NCELLS, 4), np.uint32)
    cells[:, 0] = np.where(occupied, rng.integers(1, 2**32, NCELLS, dtype=np.uint64), 0).astype(np.uint32)
    cells[:, 1:] = np.where(occupied[:, None], rng.random((NCELLS, 3), np.float32), 0).view(np.uint32)
    from violet_amd import _abi

    dev = Dev()
    gd, gc = dev.put(decay), dev.put(cells)
    call("vgi_pass_hash_grid_decay", C.byref(_abi.HGDecayIO(gd, gc)))
    g_decay, g_cells = Dev.get(gd, np.uint8, (NCELLS,)), Dev.get(gc, np.uint32, (NCELLS, 4))
    o_decay, o_cells = o32.hg_decay(decay, cells)
    bits_equal(g_cells, o_cells, "decayed cells")
    bits_equal(g_decay, o_decay, "decay bytes")

    scanned, head, want = decay_closed_form(decay, cells)
    assert 0.15 < (decay == 1).sum() / occupied.sum() < 0.35
    assert (~scanned & occupied.reshape(65536, 16)).any(), "no bucket took the early out"
    sc, hd = scanned.ravel(), head.ravel()
    # the survivors, as a multiset per bucket, fill the bucket's first slots
    got = np.column_stack([np.nonzero(hd)[0] >> 4, g_cells[hd], g_decay[hd]]).astype(np.uint64)
    srt = lambda a: a[np.lexsort(a.T[::-1])]  # noqa: E731
    assert np.array_equal(srt(want), srt(got))
    # the rest of the scanned sets is cleared; unscanned sets are untouched
    assert np.all(g_cells[sc & ~hd] == 0) and np.all(g_decay[sc & ~hd] == 0)
    assert np.array_equal(g_cells[~sc], cells[~sc]) and np.array_equal(g_decay[~sc], decay[~sc])


@pytest.mark.parametrize("k", [2, 5, 8])
def test_hash_grid_cache_update_replayed(vgi, run, o32, k):
    """Frame k's cache update (decay, select, compact, raygen) on the
    previous frame's cache and query buffer, pass by pass."""
    from violet_amd import _abi

    snaps, fp = run["snaps"], run["fps"][k]
    assert snaps[k]["flags"]["hg"], f"frame {k} should update the cache"
    p = snaps[k - 1]
    seeds = seeds_for_frame(k, True, snaps[k]["flags"]["validation"])
    dev = Dev()
    cells0, decay0 = p["hash_grid_storage"], p["hash_grid_decay"]
    queries, counter = p["hash_grid_queries"], p["hash_grid_query_counter"]
    nq = int(counter[0])
    assert nq > 500, f"too few queries to exercise the cache ({nq})"

    # decay (restir_renderer.rs:322-329)
    gc, gd = dev.put(cells0), dev.put(decay0)
    call("vgi_pass_hash_grid_decay", C.byref(_abi.HGDecayIO(gd, gc)))
    cells1, decay1 = Dev.get(gc, np.uint32, (NCELLS, 4)), Dev.get(gd, np.uint8, (NCELLS,))
    o_decay1, o_cells1 = o32.hg_decay(decay0, cells0)
    bits_equal(cells1, o_cells1, f"decay cells[{k}]")
    bits_equal(decay1, o_decay1, f"decay bytes[{k}]")

    # query select (:361-373)
    gq, gn, gs = dev.put(queries), dev.put(counter), dev.zeros(NCELLS * 4)
    call("vgi_pass_hash_grid_query_select", C.byref(_abi.HGSelectIO(gq, gn, gc, gs)),
         C.byref(_abi.PCFrameHash(seeds["hg_select"])))
    cells2, sel = Dev.get(gc, np.uint32, (NCELLS, 4)), Dev.get(gs, np.uint32, (NCELLS,))
    o_cells2, o_sel = o32.hg_select(queries, nq, cells1, seeds["hg_select"])
    full = (o_cells2[:, 0] != 0).reshape(65536, 16).all(1)
    assert not full.any(), "a full bucket makes the inserted set depend on scheduling"
    assert np.array_equal(bucket_sets(cells2, sel), bucket_sets(o_cells2, o_sel)), f"select buckets differ [{k}]"
    inserted = int((cells2[:, 0] != 0).sum() - (cells1[:, 0] != 0).sum())
    selected = int((sel != 0).sum())
    assert selected > 100 and inserted >= 0

    # compaction (:376-391) of the GPU's own selection
    gi, ga, gcc = dev.zeros(NQ * 4), dev.zeros(NQ * 4), dev.zeros(4)
    call("vgi_pass_hash_grid_query_compact", C.byref(_abi.HGCompactIO(gs, gi, ga, gcc)))
    n = int(Dev.get(gcc, np.uint32, (1,))[0])
    idx, addr = Dev.get(gi, np.uint32, (NQ,))[:n], Dev.get(ga, np.uint32, (NQ,))[:n]
    o_idx, o_addr = o32.hg_compact(sel)
    assert n == len(o_addr) == selected
    order = np.argsort(addr)
    assert np.array_equal(addr[order], o_addr) and np.array_equal(idx[order], o_idx)

    # cache rays (:393-432) on the GPU's compact lists, in the GPU's order
    tr, q_new, c_new = gpu_trace_res(dev, run, gc, p)
    io = _abi.HGRaygenIO(gc, gq, gcc, gi, ga, gd)
    call("vgi_pass_hash_grid_raygen", C.byref(fp), C.byref(tr), C.byref(io), C.byref(_abi.PCFrameHash(seeds["hg_raygen"])))
    cells3, decay3 = Dev.get(gc, np.uint32, (NCELLS, 4)), Dev.get(gd, np.uint8, (NCELLS,))
    g_newq, _ = read_queries(q_new, c_new)

    def oracle_rays(lookup_cells):
        oq, oc = np.zeros((NQ, 6), np.uint32), np.zeros(1, np.uint32)
        look = np.ascontiguousarray(lookup_cells)
        otr = o32.trace_res(look, oq, oc, prev_ind_diff=p["out_diffuse"], prev_depth=p["depth"])
        res = o32.hg_raygen(fp, otr, cells2, queries, idx, addr, decay1, seeds["hg_raygen"])
        return res, oq[: int(oc[0])]

    (a_cells, a_decay), a_q = oracle_rays(cells2)   # lookups before the pass
    (b_cells, _), _ = oracle_rays(cells3)            # lookups after it
    bits_equal(decay3, a_decay, f"raygen decay[{k}]")
    assert np.all(decay3[addr] == 10)
    same_a = (cells3 == a_cells).all(1)
    same_b = (cells3 == b_cells).all(1)
    bad = np.nonzero(~(same_a | same_b))[0]
    assert bad.size == 0, f"raygen cells match neither lookup order [{k}]: {bad[:5].tolist()}"
    changed = (cells3[addr, 1:] != cells2[addr, 1:]).any(1).mean()
    assert changed > 0.5, f"cache rays left most cells unchanged ({changed:.2f})"
    assert multiset(g_newq) == multiset(a_q), f"raygen queries differ [{k}]"
    # the replay reproduces the renderer's frame up to slot order in buckets
    s = snaps[k]
    assert np.array_equal(bucket_sets(cells3, np.zeros(NCELLS, np.uint32))[:, :2],
                          bucket_sets(s["hash_grid_storage"], np.zeros(NCELLS, np.uint32))[:, :2])


@pytest.mark.parametrize("k", [2, 5, 8])
def test_hash_grid_ordered_update_replayed(vgi, run_ordered, o32, k):
    """Frame k's order-independent cache update (decay, ordered select,
    ordered compaction, ordered cache rays) pass by pass: every output bit-
    exact with the oracle's FP32 twin, and the replay's cache equals the
    renderer's frame k cache bit for bit."""
    from violet_amd import _abi

    run = run_ordered
    snaps, fp = run["snaps"], run["fps"][k]
    assert snaps[k]["flags"]["hg"], f"frame {k} should update the cache"
    p = snaps[k - 1]
    seeds = seeds_for_frame(k, True, snaps[k]["flags"]["validation"])
    dev = Dev()
    cells0, decay0 = p["hash_grid_storage"], p["hash_grid_decay"]
    queries, counter = p["hash_grid_queries"], p["hash_grid_query_counter"]
    nq = int(counter[0])
    assert nq > 500, f"too few queries to exercise the cache ({nq})"

    gc, gd = dev.put(cells0), dev.put(decay0)
    call("vgi_pass_hash_grid_decay", C.byref(_abi.HGDecayIO(gd, gc)))
    cells1, decay1 = Dev.get(gc, np.uint32, (NCELLS, 4)), Dev.get(gd, np.uint8, (NCELLS,))

    gq, gn, gs, gw = dev.put(queries), dev.put(counter), dev.zeros(NCELLS * 4), dev.zeros(NCELLS * 4)
    thr0 = p["hash_grid_query_threshold"]
    gt = dev.put(thr0)
    call("vgi_pass_hash_grid_query_select_ordered", C.byref(_abi.HGOSelectIO(gq, gn, gc, gs, gw, gt)),
         C.byref(_abi.PCFrameHash(seeds["hg_select"])))
    cells2, sel, win = (Dev.get(gc, np.uint32, (NCELLS, 4)), Dev.get(gs, np.uint32, (NCELLS,)),
                        Dev.get(gw, np.uint32, (NCELLS,)))
    o_thr = thr0.copy()
    o_cells2, o_sel, o_win = o32.hg_select_ordered(queries, nq, cells1, seeds["hg_select"], o_thr)
    thr1 = Dev.get(gt, np.uint32, (1,))
    assert thr1[0] == o_thr[0] and thr1[0] == snaps[k]["hash_grid_query_threshold"][0]
    bits_equal(cells2, o_cells2, f"ordered select cells[{k}]")
    bits_equal(sel, o_sel, f"ordered select words[{k}]")
    chosen = np.nonzero(sel)[0]
    assert len(chosen) > 100
    assert np.array_equal(queries[win[chosen]], queries[o_win[chosen]]), f"selected queries differ [{k}]"
    inserted = int((cells2[:, 0] != 0).sum() - (cells1[:, 0] != 0).sum())
    assert inserted > 0

    gi, ga, gcc = dev.zeros(NQ * 4), dev.zeros(NQ * 4), dev.zeros(4)
    gws = dev.zeros(_abi.HG_COMPACT_WORKSPACE_WORDS * 4)
    call("vgi_pass_hash_grid_query_compact_ordered", C.byref(_abi.HGOCompactIO(gs, gw, gi, ga, gcc, gws)))
    n = int(Dev.get(gcc, np.uint32, (1,))[0])
    idx, addr = Dev.get(gi, np.uint32, (NQ,))[:n], Dev.get(ga, np.uint32, (NQ,))[:n]
    o_idx, o_addr = o32.hg_compact_ordered(o_sel, o_win)
    assert np.array_equal(addr, o_addr), f"compaction order differs [{k}]"
    assert np.array_equal(queries[idx], queries[o_idx])
    # the compaction leaves both buffers zero for the next frame's select
    assert not Dev.get(gs, np.uint32, (NCELLS,)).any() and not Dev.get(gw, np.uint32, (NCELLS,)).any()

    tr, q_new, c_new = gpu_trace_res(dev, run, gc, p)
    tr.hash_grid_query_threshold = gt
    gst = dev.zeros(NQ * 16)
    io = _abi.HGORaygenIO(_abi.HGRaygenIO(gc, gq, gcc, gi, ga, gd), gst)
    call("vgi_pass_hash_grid_raygen_ordered", C.byref(fp), C.byref(tr), C.byref(io),
         C.byref(_abi.PCFrameHash(seeds["hg_raygen"])))
    cells3, decay3 = Dev.get(gc, np.uint32, (NCELLS, 4)), Dev.get(gd, np.uint8, (NCELLS,))
    g_newq, _ = read_queries(q_new, c_new)
    oq, oc = np.zeros((NQ, 6), np.uint32), np.zeros(1, np.uint32)
    before = np.ascontiguousarray(cells2)  # the cache rays read the cache from before the pass
    otr = o32.trace_res(before, oq, oc, prev_ind_diff=p["out_diffuse"], prev_depth=p["depth"], threshold=o_thr)
    o_cells3, o_decay3 = o32.hg_raygen(fp, otr, cells2, queries, idx, addr, decay1, seeds["hg_raygen"])
    bits_equal(cells3, o_cells3, f"ordered cache rays cells[{k}]")
    bits_equal(decay3, o_decay3, f"ordered cache rays decay[{k}]")
    assert multiset(g_newq) == multiset(oq[: int(oc[0])]), f"cache-ray queries differ [{k}]"
    # the renderer's own frame k
    bits_equal(cells3, snaps[k]["hash_grid_storage"], f"renderer cache[{k}]")
    bits_equal(decay3, snaps[k]["hash_grid_decay"], f"renderer decay[{k}]")


@pytest.mark.parametrize("k", [4, 8])
def test_diffuse_validation_replayed(vgi, run, o32, o64, k):
    """Validation frame k (frame_index & 3 == 0, restir_renderer.rs:450-455):
    re-trace the previous frame's spatial reservoirs in place
    (ind_diff_sample_validate.hlsl:23-78), then the diffuse temporal pass
    without a new sample reads them."""
    from violet_amd import _abi

    snaps, fp = run["snaps"], run["fps"][k]
    s, p = snaps[k], snaps[k - 1]
    assert s["flags"]["validation"], f"frame {k} should validate"
    seeds = seeds_for_frame(k, s["flags"]["hg"], True)
    dev = Dev()
    tr, q, cnt = gpu_trace_res(dev, run, dev.put(s["hash_grid_storage"]), p)
    g_rsv, g_hit, g_rad = dev.put(p["diff_reservoir"]), dev.put(p["diff_hit_pos_normal"]), dev.put(p["diff_hit_radiance"])
    io = _abi.DiffSampleValidateIO(W, H, 0, dev.put(p["depth"]), g_rsv, g_hit, g_rad)
    call("vgi_pass_ind_diff_sample_validate", C.byref(fp), C.byref(tr), C.byref(io))
    v_rsv = Dev.get(g_rsv, np.uint32, (H, W))
    v_hit = Dev.get(g_hit, np.uint32, (H, W, 2))
    v_rad = Dev.get(g_rad, np.float32, (H, W, 3))
    gq, _ = read_queries(q, cnt)

    oq, oc = np.zeros((NQ, 6), np.uint32), np.zeros(1, np.uint32)
    cells = np.ascontiguousarray(s["hash_grid_storage"])
    otr = o32.trace_res(cells, oq, oc, prev_ind_diff=p["out_diffuse"], prev_depth=p["depth"])
    o_rsv, o_hit, o_rad = o32.diff_sample_validate(fp, otr, p["depth"], p["diff_reservoir"],
                                                   p["diff_hit_pos_normal"], p["diff_hit_radiance"])
    bits_equal(v_rsv, o_rsv, f"validated reservoir[{k}]")
    bits_equal(v_hit, o_hit, f"validated hit[{k}]")
    bits_equal(v_rad, o_rad, f"validated radiance[{k}]")
    assert multiset(gq) == multiset(oq[: int(oc[0])])
    M0 = p["diff_reservoir"] & 0xFF
    M1 = v_rsv & 0xFF
    assert np.all(M1 <= M0)
    assert (M1 < M0).mean() > 0.001, "validation lowered no reservoir's M"
    assert (v_rad != p["diff_hit_radiance"]).any(-1).mean() > 0.1, "validation refreshed no radiance"

    # diffuse temporal on the validated reservoirs (has_new_sample = 0, Q13)
    t_rsv, t_hit, t_rad = dev.zeros(W * H * 4), dev.zeros(W * H * 8), dev.zeros(W * H * 12)
    tio = _abi.DiffTemporalIO(W, H, 0, dev.put(p["depth"]), dev.put(s["depth"]), None, None, g_rsv, g_hit, g_rad,
                              t_rsv, t_hit, t_rad)
    call("vgi_pass_ind_diff_temporal_resample", C.byref(fp), C.byref(tio),
         C.byref(_abi.PCDiffTemporal(seeds["diff_temporal"], 1, 0)), None)
    got = Dev.get(t_rsv, np.uint32, (H, W))
    o_t = o32.diff_temporal(fp, p["depth"], s["depth"], None, None, o_rsv, o_hit, o_rad, seeds["diff_temporal"], True,
                            False, 16)
    bits_equal(got, o_t[0], f"temporal after validation[{k}]")
    bits_equal(Dev.get(t_hit, np.uint32, (H, W, 2)), o_t[1], f"temporal hit after validation[{k}]")
    bits_equal(got, s["diff_temporal_reservoir"], f"replay vs renderer temporal reservoir[{k}]")
    bits_equal(Dev.get(t_rad, np.float32, (H, W, 3)), s["diff_temporal_hit_radiance"],
               f"replay vs renderer temporal radiance[{k}]")
    # FP64 validation: same M decisions away from the rounding of the thresholds
    oq64, oc64 = np.zeros((NQ, 6), np.uint32), np.zeros(1, np.uint32)
    otr64 = o64.trace_res(cells, oq64, oc64, prev_ind_diff=p["out_diffuse"], prev_depth=p["depth"])
    r64, h64, d64 = o64.diff_sample_validate(fp, otr64, p["depth"], p["diff_reservoir"], p["diff_hit_pos_normal"],
                                             p["diff_hit_radiance"])
    assert np.mean(r64 != o_rsv) < 0.01
    same = (h64 == o_hit).all(-1) & (r64 == o_rsv)
    assert_close_frac(v_rad[same], d64[same], f"validated radiance vs FP64 [{k}]")


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("params", [None, (30, 5, 30.0)], ids=["reference", "m_cap30"])
def test_specular_gen_validate_temporal_replayed(vgi, run, o32, o64, k, params):
    """Frame k's specular chain with history: validation of last frame's
    samples (ind_spec_sample_validate.hlsl:17-63), new samples traced with the
    screen cache (ind_spec_sample_gen.hlsl:29-128), then the temporal pass's
    MIS merge of the two (ind_spec_temporal_resample.hlsl:147-362)."""
    from violet_amd import _abi

    snaps, fp = run["snaps"], run["fps"][k]
    s, p = snaps[k], snaps[k - 1]
    fl = s["flags"]
    seeds = seeds_for_frame(k, fl["hg"], fl["validation"])
    m_cap = 16 if params is None else params[0]
    rp = None if params is None else C.byref(_abi.ResampleParams(*params))
    dev = Dev()
    cells = np.ascontiguousarray(s["hash_grid_storage"])
    g_cells = dev.put(cells)
    geo, pgeo = s["depth"] > 0, p["depth"] > 0

    # validation (runs when the cache was updated: restir_renderer.rs:722-737)
    prev_rsv, prev_hp, prev_rad = p["spec_reservoir"], p["spec_hit_pos"], p["spec_hit_radiance"]
    g_prsv, g_php, g_prad = dev.put(prev_rsv), dev.put(prev_hp), dev.put(prev_rad)
    if fl["hg"]:
        tr, q, cnt = gpu_trace_res(dev, run, g_cells, p)
        io = _abi.SpecSampleValidateIO(W, H, 0, g_prsv, g_php, g_prad)
        call("vgi_pass_ind_spec_sample_validate", C.byref(fp), C.byref(tr), C.byref(io))
        v_rsv, v_hp = Dev.get(g_prsv, np.uint32, (H, W, 2)), Dev.get(g_php, np.float32, (H, W, 4))
        v_rad = Dev.get(g_prad, np.float32, (H, W, 3))
        oq, oc = np.zeros((NQ, 6), np.uint32), np.zeros(1, np.uint32)
        otr = o32.trace_res(cells, oq, oc, prev_ind_diff=p["out_diffuse"], prev_depth=p["depth"])
        o_rsv, o_hp, o_rad = o32.spec_sample_validate(fp, otr, prev_rsv, prev_hp, prev_rad)
        bits_equal(v_rsv, o_rsv, f"spec validated reservoir[{k}]")
        bits_equal(v_hp[pgeo], o_hp[pgeo], f"spec validated hit pos[{k}]")
        bits_equal(v_rad[pgeo], o_rad[pgeo], f"spec validated radiance[{k}]")
        assert multiset(read_queries(q, cnt)[0]) == multiset(oq[: int(oc[0])])
        prev_rsv, prev_hp, prev_rad = v_rsv, v_hp, v_rad

    # new samples (the temporal pass resamples them in place)
    tr, q, cnt = gpu_trace_res(dev, run, g_cells, p)
    g_hp, g_rad = dev.zeros(W * H * 16), dev.zeros(W * H * 12)
    gio = _abi.SpecSampleGenIO(W, H, 0, dev.put(s["gbuffer"]), dev.put(s["depth"]), g_hp, g_rad)
    call("vgi_pass_ind_spec_sample_gen", C.byref(fp), C.byref(tr), C.byref(gio),
         C.byref(_abi.PCSampleGen(seeds["spec_gen"], 1)))
    n_hp, n_rad = Dev.get(g_hp, np.float32, (H, W, 4)), Dev.get(g_rad, np.float32, (H, W, 3))
    oq, oc = np.zeros((NQ, 6), np.uint32), np.zeros(1, np.uint32)
    otr = o32.trace_res(cells, oq, oc, prev_ind_diff=p["out_diffuse"], prev_depth=p["depth"])
    o_hp, o_rad = o32.spec_sample_gen(fp, otr, s["gbuffer"], s["depth"], seeds["spec_gen"], True)
    bits_equal(n_hp[geo], o_hp[geo], f"spec gen hit pos[{k}]")
    bits_equal(n_rad[geo], o_rad[geo], f"spec gen radiance[{k}]")
    assert multiset(read_queries(q, cnt)[0]) == multiset(oq[: int(oc[0])])

    # temporal resampling with history
    g_trsv = dev.zeros(W * H * 8)
    tio = _abi.SpecTemporalIO(W, H, 0, dev.put(p["gbuffer"]), dev.put(p["depth"]), dev.put(s["gbuffer"]),
                              dev.put(s["depth"]), dev.put(prev_rsv), dev.put(prev_hp), dev.put(prev_rad), g_trsv,
                              g_hp, g_rad)
    call("vgi_pass_ind_spec_temporal_resample", C.byref(fp), C.byref(tio),
         C.byref(_abi.PCSpecTemporal(seeds["spec_temporal"], 1)), rp)
    t_rsv = Dev.get(g_trsv, np.uint32, (H, W, 2))
    t_hp, t_rad = Dev.get(g_hp, np.float32, (H, W, 4)), Dev.get(g_rad, np.float32, (H, W, 3))
    args = (fp, p["gbuffer"], p["depth"], s["gbuffer"], s["depth"], prev_rsv, prev_hp, prev_rad, n_hp, n_rad,
            seeds["spec_temporal"], True, m_cap)
    u_rsv, u_hp, u_rad = o32.spec_temporal(*args)
    bits_equal(t_rsv, u_rsv, f"spec temporal reservoir[{k}]")
    bits_equal(t_hp[geo], u_hp[geo], f"spec temporal hit pos[{k}]")
    bits_equal(t_rad[geo], u_rad[geo], f"spec temporal radiance[{k}]")
    # the MIS branch ran: history was found for most pixels, and some kept it
    prev_m = t_rsv[..., 0][geo]
    assert (prev_m > 1).mean() > 0.3, f"history M > 1 at only {(prev_m > 1).mean():.2f} of pixels"
    kept_prev = (t_hp[..., :3] != n_hp[..., :3]).any(-1)[geo]
    assert kept_prev.mean() > 0.05, "no pixel kept its previous sample"
    if params is None:
        bits_equal(t_rsv, s["spec_reservoir"], f"replay vs renderer spec reservoir[{k}]")
        bits_equal(t_hp[geo], s["spec_hit_pos"][geo], f"replay vs renderer spec hit pos[{k}]")
    # FP64: where the same sample is selected, W and the target function agree
    w_rsv, w_hp, _ = o64.spec_temporal(*args)
    same = geo & (w_hp[..., :3].astype(np.float32) == u_hp[..., :3]).all(-1) & (w_rsv[..., 0] == u_rsv[..., 0])
    assert same[geo].mean() > 0.98
    W32 = t_rsv[..., 1].view(np.float32)[same].astype(np.float64)
    W64 = w_rsv[..., 1].view(np.float32)[same].astype(np.float64)
    assert_close_frac(W32[:, None], W64[:, None], f"spec temporal W vs FP64 [{k}]")
