"""Ragged and degenerate image sizes through the whole renderer (GI,
denoiser, composition, temporal AA over jittered frames): 1x1, single
rows and columns, and sizes that are not multiples of the 8x8 tile.  Every
frame's depth, G-buffer and sun shadow mask equal the FP32 twin's bit for
bit, the final lighting and the temporal AA replayed by the twin on the
frame's own inputs are bit-identical, and every output is finite."""
from __future__ import annotations

import numpy as np
import pytest

from restir_harness import F32, make_mesh, run_gpu_frames

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (5, 3), (1, 40), (37, 1), (130, 9)]
FRAMES = 3


@pytest.fixture(scope="module")
def mesh():
    return make_mesh()


@pytest.fixture(scope="module")
def o32(mesh):
    import oracle

    return oracle.OracleRestir(32, oracle.OracleScene(32, mesh))


@pytest.mark.parametrize("w,h", SIZES)
def test_edge_size_frames(mesh, o32, w, h):
    from violet_amd import renderer

    cfg = renderer.RestirConfig(compose=True, taa=True, temporal_m_cap=30, diff_spatial_taps=5,
                                spatial_radius_px=30.0)
    fps, outs, snaps, r, sc = run_gpu_frames(mesh, w, h, FRAMES, config=cfg, jitter=True)
    for k in range(FRAMES):
        s, fp = snaps[k], fps[k]
        depth, gb = o32.gbuffer(fp, w, h)
        assert np.array_equal(s["depth"].view(np.uint32), depth.view(np.uint32)), f"depth[{k}] {w}x{h}"
        assert np.array_equal(s["gbuffer"], gb), f"gbuffer[{k}] {w}x{h}"
        assert np.array_equal(s["shadow_mask"], o32.raytraced_shadow(fp, s["depth"])), f"shadow[{k}] {w}x{h}"
        lit = o32.final_lighting(fp, s["gbuffer"], s["shadow_mask"], s["out_diffuse"].astype(F32),
                                 s["out_specular"].astype(F32))
        assert np.array_equal(s["scene_color"].view(np.uint32), lit.view(np.uint32)), f"lighting[{k}] {w}x{h}"
        hist = None if k == 0 else snaps[k - 1]["out_color"].astype(F32)
        taa = o32.temporal_aa(fp, s["depth"], s["scene_color"], hist)
        assert np.array_equal(s["out_color"].view(np.uint32), taa.view(np.uint32)), f"taa[{k}] {w}x{h}"
        for name in ("out_diffuse", "out_specular", "out_color"):
            assert np.isfinite(s[name]).all(), f"{name}[{k}] {w}x{h}"
    assert outs[-1].denoised
