"""ReSTIR passes with FP16 storage on the GPU: every stored binary16 value
equals the FP32 twin's result on the same (binary16-decoded) inputs rounded
to binary16, bit for bit; integer outputs (reservoirs, hit encodings) and
the FP32 specular hit positions are bit-identical.

The GPU decodes binary16 exactly, computes in FP32 in the reference's order
and rounds once on store (round to nearest even), so the twin fed the same
decoded values and rounded once must agree exactly.  Frame 0 is checked as a
whole pipeline; frames 1-3 replay the resampling passes on the recorded
inputs.
"""
from __future__ import annotations

import numpy as np
import pytest

from restir_harness import make_mesh, run_gpu_frames, seeds_for_frame

pytestmark = pytest.mark.gpu

W, H = 80, 56


@pytest.fixture(scope="module")
def run():
    from violet_amd import _lib, renderer

    mesh = make_mesh()
    cfg = renderer.RestirConfig(storage=_lib.STORAGE_F16, temporal_m_cap=30, diff_spatial_taps=5,
                                spatial_radius_px=30.0)
    fps, outs, snaps, r, sc = run_gpu_frames(mesh, W, H, 4, config=cfg)
    return dict(mesh=mesh, fps=fps, snaps=snaps, prm=dict(m=30, taps=5, radius=30.0))


@pytest.fixture(scope="module")
def o32(run):
    import oracle

    return oracle.OracleRestir(32, oracle.OracleScene(32, run["mesh"]))


def f(a):
    """binary16 image -> the float32 values the GPU computes with"""
    return np.asarray(a, np.float16).astype(np.float32)


def same16(gpu, twin, what):
    g = np.asarray(gpu, np.float16).view(np.uint16)
    t = np.asarray(twin, np.float32).astype(np.float16).view(np.uint16)
    frac = float(np.mean(g != t))
    assert frac == 0.0, f"{what}: {frac:.4%} of binary16 values differ"


def same(gpu, twin, what):
    g = np.ascontiguousarray(gpu)
    t = np.ascontiguousarray(twin, dtype=g.dtype)
    assert np.array_equal(g.view(np.uint8), t.view(np.uint8)), f"{what} differs"


def test_frame0_fp16_pipeline(run, o32):
    s, fp, P = run["snaps"][0], run["fps"][0], run["prm"]
    seeds = seeds_for_frame(0, False, False)
    depth, gb = o32.gbuffer(fp, W, H)
    same(s["depth"], depth, "depth")
    same(s["gbuffer"], gb, "gbuffer")
    cells = np.zeros((65536 * 16, 4), np.uint32)
    q = np.zeros((128 * 1024, 6), np.uint32)
    cnt = np.zeros(1, np.uint32)
    tr = o32.trace_res(cells, q, cnt)
    hit, rad = o32.diff_sample_gen(fp, tr, gb, depth, seeds["diff_gen"], False)
    same(s["diff_new_hit_pos_normal"], hit, "diff new hit")
    same16(s["diff_new_hit_radiance"], rad, "diff new radiance")
    ao = o32.ao_gen(fp, hit, depth, None, True, False)
    same16(s["ao"], ao, "ao")
    same16(s["ao_filtered"], o32.ao_filter(depth, f(s["ao"])), "ao filtered")
    t = o32.diff_temporal(fp, None, depth, hit, f(s["diff_new_hit_radiance"]), None, None, None,
                          seeds["diff_temporal"], False, True, P["m"])
    same(s["diff_temporal_reservoir"], t[0], "diff temporal reservoir")
    same(s["diff_temporal_hit_pos_normal"], t[1], "diff temporal hit")
    same16(s["diff_temporal_hit_radiance"], t[2], "diff temporal radiance")
    sp = o32.diff_spatial(fp, gb, depth, f(s["ao_filtered"]), s["diff_temporal_reservoir"],
                          s["diff_temporal_hit_pos_normal"], f(s["diff_temporal_hit_radiance"]), seeds["diff_spatial"],
                          P["taps"], P["radius"])
    same(s["diff_reservoir"], sp[0], "diff spatial reservoir")
    same(s["diff_hit_pos_normal"], sp[1], "diff spatial hit")
    same16(s["diff_hit_radiance"], sp[2], "diff spatial radiance")
    same16(s["indirect_diffuse_raw"], sp[3], "diff lighting")
    hp, sr = o32.spec_sample_gen(fp, tr, gb, depth, seeds["spec_gen"], False)
    geo = depth > 0
    # the temporal pass rewrites the sample in place; without history it keeps it
    same16(s["spec_hit_radiance"][geo], sr[geo], "spec new radiance")
    u = o32.spec_temporal(fp, None, None, gb, depth, None, None, None, hp, f(s["spec_hit_radiance"]),
                          seeds["spec_temporal"], False, P["m"])
    same(s["spec_reservoir"], u[0], "spec temporal reservoir")
    same(s["spec_hit_pos"][geo], u[1][geo], "spec temporal hit pos")
    light, rl = o32.spec_spatial(fp, gb, depth, s["spec_reservoir"], s["spec_hit_pos"], f(s["spec_hit_radiance"]),
                                 seeds["spec_spatial"], P["radius"])
    same16(s["indirect_specular_raw"], light, "spec lighting")
    same16(s["spec_ray_len"], rl, "spec ray length")


@pytest.mark.parametrize("k", [1, 2, 3])
def test_fp16_resampling_replayed(run, o32, k):
    s, p, fp, P = run["snaps"][k], run["snaps"][k - 1], run["fps"][k], run["prm"]
    fl = s["flags"]
    seeds = seeds_for_frame(k, fl["hg"], fl["validation"])
    depth, gb = s["depth"], s["gbuffer"]
    has_new = not fl["validation"]
    ao = o32.ao_gen(fp, s["diff_new_hit_pos_normal"], depth, f(p["ao"]), has_new, True)
    same16(s["ao"], ao, f"ao[{k}]")
    same16(s["ao_filtered"], o32.ao_filter(depth, f(s["ao"])), f"ao filtered[{k}]")
    if has_new:
        t = o32.diff_temporal(fp, p["depth"], depth, s["diff_new_hit_pos_normal"], f(s["diff_new_hit_radiance"]),
                              p["diff_reservoir"], p["diff_hit_pos_normal"], f(p["diff_hit_radiance"]),
                              seeds["diff_temporal"], True, True, P["m"])
        same(s["diff_temporal_reservoir"], t[0], f"diff temporal reservoir[{k}]")
        same16(s["diff_temporal_hit_radiance"], t[2], f"diff temporal radiance[{k}]")
    sp = o32.diff_spatial(fp, gb, depth, f(s["ao_filtered"]), s["diff_temporal_reservoir"],
                          s["diff_temporal_hit_pos_normal"], f(s["diff_temporal_hit_radiance"]), seeds["diff_spatial"],
                          P["taps"], P["radius"])
    same(s["diff_reservoir"], sp[0], f"diff spatial reservoir[{k}]")
    same16(s["indirect_diffuse_raw"], sp[3], f"diff lighting[{k}]")
    light, rl = o32.spec_spatial(fp, gb, depth, s["spec_reservoir"], s["spec_hit_pos"], f(s["spec_hit_radiance"]),
                                 seeds["spec_spatial"], P["radius"])
    same16(s["indirect_specular_raw"], light, f"spec lighting[{k}]")
    same16(s["spec_ray_len"], rl, f"spec ray length[{k}]")
