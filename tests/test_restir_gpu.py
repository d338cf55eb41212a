"""ReSTIR GI pass parity on the GPU (through the C ABI and the renderer
sequencer) against the oracle.

* Frame 0 (no history): every per-pixel output of the GPU frame - G-buffer,
  new diffuse/specular samples (BVH hits, shading, sun shadow rays), AO,
  temporal and spatial reservoirs, raw lighting - equals the FP32 twin bit
  for bit; the appended hash-grid queries are the same multiset.
* Later frames: each pass is replayed by the FP32 twin on the GPU run's
  recorded inputs; reservoir words, selected samples and hit encodings must
  be bit-identical, float outputs bit-identical (they use IEEE operations
  only), and within the FP64 oracle's tolerance.

Every test runs with the reference's resampling constants (M cap 16, 9
spiral taps, radius 0.05*min(w,h)), with the 1280x720 frame-loop workload's
(M cap 30, 5 taps, 30 px), with the latter on a single stream (no
concurrent lanes: the renderer's default overlaps independent pass chains),
and with the reference constants on a textured scene (material textures
sampled at LOD 2 by the indirect rays and with quad derivatives and
anisotropy by the G-buffer).
"""
from __future__ import annotations

import numpy as np
import pytest

from parity import assert_close, assert_close_frac
from restir_harness import F32, make_mesh, multiset, run_gpu_frames, seeds_for_frame

pytestmark = pytest.mark.gpu

W, H = 96, 64


PARAMS = {
    "reference": dict(temporal_m_cap=0, diff_spatial_taps=0, spatial_radius_px=0.0),
    "frame_loop": dict(temporal_m_cap=30, diff_spatial_taps=5, spatial_radius_px=30.0),
    "frame_loop_one_stream": dict(temporal_m_cap=30, diff_spatial_taps=5, spatial_radius_px=30.0,
                                  concurrent_lanes=False),
    # reference constants on the textured scene (texture coordinates, tangents,
    # base-colour / metallic-roughness / normal maps with mips)
    "textured": dict(temporal_m_cap=0, diff_spatial_taps=0, spatial_radius_px=0.0),
}


@pytest.fixture(scope="module", params=sorted(PARAMS))
def run(request):
    from violet_amd import renderer

    mesh = make_mesh(textured=request.param == "textured")
    prm = PARAMS[request.param]
    fps, outs, snaps, r, sc = run_gpu_frames(mesh, W, H, 6, config=renderer.RestirConfig(**prm))
    return dict(mesh=mesh, fps=fps, outs=outs, snaps=snaps, prm=prm)


@pytest.fixture(scope="module")
def o32(run):
    import oracle

    return oracle.OracleRestir(32, oracle.OracleScene(32, run["mesh"]))


@pytest.fixture(scope="module")
def o64(run):
    import oracle

    return oracle.OracleRestir(64, oracle.OracleScene(64, run["mesh"]))


def bits_equal(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    neq = a.view(np.uint32) != b.view(np.uint32) if a.dtype.itemsize == 4 else a != b
    frac = float(np.mean(neq))
    assert frac == 0.0, f"{what}: {frac:.4%} of words differ (first at {np.argwhere(neq)[:3].tolist()})"


def test_frame0_bit_exact_pipeline(run, o32):
    s, fp = run["snaps"][0], run["fps"][0]
    assert not s["flags"]["denoised"] and not s["flags"]["hg"]
    seeds = seeds_for_frame(0, False, False)
    depth, gb = o32.gbuffer(fp, W, H)
    bits_equal(s["depth"], depth, "depth")
    bits_equal(s["gbuffer"], gb, "gbuffer")
    assert (depth > 0).mean() > 0.3
    cells = np.zeros((65536 * 16, 4), np.uint32)
    q = np.zeros((128 * 1024, 6), np.uint32)
    cnt = np.zeros(1, np.uint32)
    tr = o32.trace_res(cells, q, cnt, threshold=s["hash_grid_query_threshold"])
    hit, rad = o32.diff_sample_gen(fp, tr, gb, depth, seeds["diff_gen"], False)
    bits_equal(s["diff_new_hit_pos_normal"], hit, "diff new hit")
    bits_equal(s["diff_new_hit_radiance"], rad, "diff new radiance")
    ao = o32.ao_gen(fp, hit, depth, None, True, False)
    bits_equal(s["ao"], ao, "ao")
    aof = o32.ao_filter(depth, ao)
    bits_equal(s["ao_filtered"], aof, "ao filtered")
    P = run["prm"]
    trsv, thit, trad = o32.diff_temporal(fp, None, depth, hit, rad, None, None, None, seeds["diff_temporal"], False, True,
                                         P["temporal_m_cap"])
    bits_equal(s["diff_temporal_reservoir"], trsv, "diff temporal reservoir")
    bits_equal(s["diff_temporal_hit_pos_normal"], thit, "diff temporal hit")
    srsv, shit, srad, slight = o32.diff_spatial(fp, gb, depth, aof, trsv, thit, trad, seeds["diff_spatial"],
                                                P["diff_spatial_taps"], P["spatial_radius_px"])
    bits_equal(s["diff_reservoir"], srsv, "diff spatial reservoir")
    bits_equal(s["diff_hit_pos_normal"], shit, "diff spatial hit")
    bits_equal(s["diff_hit_radiance"], srad, "diff spatial radiance")
    bits_equal(s["indirect_diffuse_raw"], slight, "diff lighting")
    hp, sr = o32.spec_sample_gen(fp, tr, gb, depth, seeds["spec_gen"], False)
    geo = depth > 0
    bits_equal(s["spec_hit_radiance"][geo], sr[geo], "spec new radiance")
    ursv, uhp, urad = o32.spec_temporal(fp, None, None, gb, depth, None, None, None, hp, sr, seeds["spec_temporal"], False,
                                        P["temporal_m_cap"])
    bits_equal(s["spec_reservoir"], ursv, "spec temporal reservoir")
    bits_equal(s["spec_hit_pos"][geo], uhp[geo], "spec temporal hit pos")
    light, rl = o32.spec_spatial(fp, gb, depth, ursv, uhp, urad, seeds["spec_spatial"], P["spatial_radius_px"])
    bits_equal(s["indirect_specular_raw"], light, "spec lighting")
    bits_equal(s["spec_ray_len"], rl, "spec ray length")
    # hash-grid queries appended by the two trace passes: same multiset
    n = int(cnt[0])
    assert int(s["hash_grid_query_counter"][0]) == n
    assert multiset(s["hash_grid_queries"][:n]) == multiset(q[:n])


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_resampling_passes_replayed(run, o32, o64, k):
    """Per-pass replay on recorded inputs of frame k (bit-exact vs FP32 twin,
    tolerance vs FP64 oracle)."""
    s, p, fp = run["snaps"][k], run["snaps"][k - 1], run["fps"][k]
    P = run["prm"]
    fl = s["flags"]
    seeds = seeds_for_frame(k, fl["hg"], fl["validation"])
    depth, gb = s["depth"], s["gbuffer"]
    geo = depth > 0
    # AO
    has_new = not fl["validation"]
    ao = o32.ao_gen(fp, s["diff_new_hit_pos_normal"], depth, p["ao"], has_new, True)
    bits_equal(s["ao"], ao, f"ao[{k}]")
    bits_equal(s["ao_filtered"], o32.ao_filter(depth, s["ao"]), f"ao filtered[{k}]")
    # diffuse temporal (the previous spatial reservoir is the input unless the
    # validation pass rewrote it in place this frame)
    if not fl["validation"]:
        args = (fp, p["depth"], depth, s["diff_new_hit_pos_normal"], s["diff_new_hit_radiance"], p["diff_reservoir"],
                p["diff_hit_pos_normal"], p["diff_hit_radiance"], seeds["diff_temporal"], True, True, P["temporal_m_cap"])
        t32 = o32.diff_temporal(*args)
        bits_equal(s["diff_temporal_reservoir"], t32[0], f"diff temporal reservoir[{k}]")
        bits_equal(s["diff_temporal_hit_pos_normal"], t32[1], f"diff temporal hit[{k}]")
        bits_equal(s["diff_temporal_hit_radiance"], t32[2], f"diff temporal radiance[{k}]")
    # diffuse spatial
    sp = (fp, gb, depth, s["ao_filtered"], s["diff_temporal_reservoir"], s["diff_temporal_hit_pos_normal"],
          s["diff_temporal_hit_radiance"], seeds["diff_spatial"], P["diff_spatial_taps"], P["spatial_radius_px"])
    d32 = o32.diff_spatial(*sp)
    bits_equal(s["diff_reservoir"], d32[0], f"diff spatial reservoir[{k}]")
    bits_equal(s["diff_hit_pos_normal"], d32[1], f"diff spatial hit[{k}]")
    bits_equal(s["indirect_diffuse_raw"], d32[3], f"diff lighting[{k}]")
    d64 = o64.diff_spatial(*sp)
    # same selected sample in FP64: compare the hit position halves only (the
    # oct-u16 normal is re-encoded from a decoded normal, and that truncating
    # round trip may move by one code between precisions)
    same = (d64[1][..., 0] == d32[1][..., 0]) & ((d64[1][..., 1] & 0xFFFF) == (d32[1][..., 1] & 0xFFFF))
    assert same.mean() > 0.98, f"FP64 selection agreement {same.mean():.4f}"
    assert_close_frac(s["indirect_diffuse_raw"][same], d64[3][same], f"diff lighting vs FP64 [{k}]")
    # specular spatial (its inputs are the in-place temporal results)
    ss = (fp, gb, depth, s["spec_reservoir"], s["spec_hit_pos"], s["spec_hit_radiance"], seeds["spec_spatial"],
          P["spatial_radius_px"])
    l32, r32 = o32.spec_spatial(*ss)
    bits_equal(s["indirect_specular_raw"], l32, f"spec lighting[{k}]")
    bits_equal(s["spec_ray_len"], r32, f"spec ray len[{k}]")
    l64, r64 = o64.spec_spatial(*ss)
    # same selected neighbour sample in FP64 <=> same ray length
    sel = geo & (np.abs(r64 - r32) <= 1e-4 * (np.abs(r64) + 1e-3))
    assert sel[geo].mean() > 0.99, f"FP64 specular selection agreement {sel[geo].mean():.4f}"
    assert_close_frac(s["indirect_specular_raw"][sel], l64[sel], f"spec lighting vs FP64 [{k}]")


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_trace_passes_replayed(run, o32, k):
    """Sample generation on frame k's recorded G-buffer, previous depth,
    previous denoised diffuse (screen cache) and hash-grid storage."""
    s, p, fp = run["snaps"][k], run["snaps"][k - 1], run["fps"][k]
    fl = s["flags"]
    seeds = seeds_for_frame(k, fl["hg"], fl["validation"])
    q = np.zeros((128 * 1024, 6), np.uint32)
    cnt = np.zeros(1, np.uint32)
    tr = o32.trace_res(s["hash_grid_storage"], q, cnt, prev_ind_diff=p["out_diffuse"], prev_depth=p["depth"],
                       threshold=s["hash_grid_query_threshold"])
    if not fl["validation"]:
        hit, rad = o32.diff_sample_gen(fp, tr, s["gbuffer"], s["depth"], seeds["diff_gen"], True)
        bits_equal(s["diff_new_hit_pos_normal"], hit, f"diff new hit[{k}]")
        bits_equal(s["diff_new_hit_radiance"], rad, f"diff new radiance[{k}]")
    n = int(cnt[0])
    gq = s["hash_grid_queries"][: min(int(s["hash_grid_query_counter"][0]), 128 * 1024)]
    gm, om = multiset(gq), multiset(q[:n])
    assert all(gm[key] >= c for key, c in om.items()), "oracle queries not all found in the GPU query buffer"


def test_renderer_state_export_import_resumes_bit_exact(run):
    """Checkpoint/resume: N frames, export, M frames == import + M frames."""
    import torch

    from violet_amd import renderer, scene, synth
    from restir_harness import orbit_view

    mesh = run["mesh"]
    sc = scene.Scene(mesh)
    cfg = renderer.RestirConfig(hash_grid_cache_decay=True)
    a = renderer.RestirRenderer(W, H, cfg)
    views = [orbit_view(f, W, H) for f in range(5)]
    fps = [synth.make_frame_params(views[f], views[f - 1] if f else None) for f in range(5)]
    for f in range(3):
        a.add_passes(sc, fps[f], f)
    torch.cuda.synchronize()
    blob = a.export_state()
    b = renderer.RestirRenderer(W, H, cfg)
    b.import_state(blob)
    for f in (3, 4):
        oa = a.add_passes(sc, fps[f], f)
        ob = b.add_passes(sc, fps[f], f)
        torch.cuda.synchronize()
        # the cache update's slot choice is scheduling dependent; the reservoirs
        # before the first hash-grid read must still agree exactly
        da = a.read("diff_new_hit_pos_normal", np.uint32, (H, W, 2))
        db = b.read("diff_new_hit_pos_normal", np.uint32, (H, W, 2))
        assert np.mean(da != db) < 0.02
        assert oa.denoised == ob.denoised and oa.validation_frame == ob.validation_frame
