"""Batched views on the GPU: each view rendered in a fused multi-view frame
(one persistent launch per pass, with and without graph replay) against the
same view rendered alone.  With the order-independent cache update (the
renderer's default) nothing in a frame depends on how the GPU schedules the
lanes, so every image is bit-identical alone and in the mix, frame after
frame:

* small views (three sizes up to 96x64, FP32, composition and temporal AA):
  every per-view image of every frame over 6 frames (cache updates from
  frame 1, a diffuse validation frame at 4);
* the batched form at full size: 16 views with seeded resolutions from
  320x240 to 1920x1080 over the shared 1M-triangle scene, FP16 storage, one
  graph-replayed fused launch per pass: the denoised diffuse and specular
  and the raw diffuse of every view, every frame, over 10 frames (two
  diffuse validation frames), with every frame's sampled cache queries
  within the query buffer (beyond it, which queries are kept depends on the
  order of the appends).
"""
from __future__ import annotations

import numpy as np
import pytest

from restir_harness import F32, make_mesh, orbit_view

pytestmark = pytest.mark.gpu

SIZES = [(96, 64), (72, 40), (44, 57)]
FRAMES = 6
EXACT_ALWAYS = ["depth", "gbuffer", "shadow_mask"]
EXACT_FRAME0 = ["diff_reservoir", "diff_hit_pos_normal", "diff_hit_radiance", "indirect_diffuse_raw",
                "spec_reservoir", "indirect_specular_raw", "spec_ray_len", "ao", "ao_filtered"]
SHAPES = {
    "depth": (F32, lambda w, h: (h, w)), "gbuffer": (np.uint32, lambda w, h: (4, h, w)),
    "shadow_mask": (np.uint8, lambda w, h: (h, w)),
    "diff_reservoir": (np.uint32, lambda w, h: (h, w)), "diff_hit_pos_normal": (np.uint32, lambda w, h: (h, w, 2)),
    "diff_hit_radiance": (F32, lambda w, h: (h, w, 3)), "indirect_diffuse_raw": (F32, lambda w, h: (h, w, 3)),
    "spec_reservoir": (np.uint32, lambda w, h: (h, w, 2)), "indirect_specular_raw": (F32, lambda w, h: (h, w, 3)),
    "spec_ray_len": (F32, lambda w, h: (h, w)), "ao": (F32, lambda w, h: (h, w, 2)),
    "ao_filtered": (F32, lambda w, h: (h, w, 2)), "diff_new_hit_pos_normal": (np.uint32, lambda w, h: (h, w, 2)),
}


def _views():
    out = []
    for k, (w, h) in enumerate(SIZES):
        vs = [orbit_view(f * 2 + 17 * k, w, h) for f in range(FRAMES)]
        out.append((w, h, vs))
    return out


def _snap(r, w, h, o):
    s = {n: r.read(n, dt, shp(w, h)).copy() for n, (dt, shp) in SHAPES.items()}
    s["out_diffuse"] = r.read_ptr(o.indirect_diffuse, w * h * 12, F32, (h, w, 3)).copy()
    s["out_color"] = r.read_ptr(o.color, w * h * 12, F32, (h, w, 3)).copy()
    s["validation"] = o.validation_frame
    return s


@pytest.mark.parametrize("graph", [False, True])
def test_views_same_alone_and_in_the_mix(vgi, graph):
    import torch

    from violet_amd import batch, renderer, scene, synth

    mesh = make_mesh()
    sc = scene.Scene(mesh)
    cfg = renderer.RestirConfig(temporal_m_cap=30, diff_spatial_taps=5, spatial_radius_px=30.0, compose=True,
                                taa=True)
    views = _views()
    fps = [[synth.make_frame_params(vs[f], vs[f - 1] if f else None) for f in range(FRAMES)] for _, _, vs in views]
    alone = [renderer.RestirRenderer(w, h, cfg) for w, h, _ in views]
    mixed = [renderer.RestirRenderer(w, h, cfg) for w, h, _ in views]
    b = batch.Batch(batch.BatchConfig(max_views=8, pixel_groups=96, trace_groups=128, atrous_groups=32,
                                      use_graph=graph))
    for f in range(FRAMES):
        outs_a = [r.add_passes(sc, fps[k][f], f) for k, r in enumerate(alone)]
        torch.cuda.synchronize()
        snap_a = [_snap(r, w, h, o) for r, (w, h, _), o in zip(alone, views, outs_a)]
        outs_b = b.add_passes(sc, mixed, [fps[k][f] for k in range(len(views))], [f] * len(views))
        torch.cuda.synchronize()
        snap_b = [_snap(r, w, h, o) for r, (w, h, _), o in zip(mixed, views, outs_b)]
        for k, (sa, sb) in enumerate(zip(snap_a, snap_b)):
            where = f"view {k} {SIZES[k]} frame {f}"
            assert outs_a[k].denoised == outs_b[k].denoised and sa["validation"] == sb["validation"], where
            names = EXACT_ALWAYS + EXACT_FRAME0 + ["out_diffuse", "out_color"]
            if not sa["validation"]:
                names = names + ["diff_new_hit_pos_normal"]
            for n in names:
                assert np.array_equal(sa[n].view(np.uint8), sb[n].view(np.uint8)), f"{n} differs: {where}"
        assert any(o.hash_grid_updated for o in outs_a) or f == 0


CONFIG_VIEWS = 16
CONFIG_FRAMES = 10
CONFIG_TRIANGLES = 1_000_000


def test_batched_form_views_bit_identical_alone_and_in_the_mix(vgi):
    import torch

    from violet_amd import _lib, batch, gpu, renderer, scene, serve

    sc = scene.Scene(scene.make_scene(CONFIG_TRIANGLES, seed=13))
    specs = serve.batch_views(CONFIG_VIEWS, seed=21)
    cfg = renderer.RestirConfig(storage=_lib.STORAGE_F16)
    print(f"batched form: {CONFIG_VIEWS} views, {sum(s.width * s.height for s in specs) / 1e6:.1f} Mpix", flush=True)

    def outputs(r, spec, o):
        n = spec.width * spec.height * 6  # 3 x binary16
        return [gpu.copy_from_ptr(o.indirect_diffuse, n), gpu.copy_from_ptr(o.indirect_specular, n),
                gpu.copy_from_ptr(r.image("indirect_diffuse_raw")[0], n)]

    alone = []
    for spec in specs:
        r = renderer.RestirRenderer(spec.width, spec.height, cfg)
        frames = []
        for f in range(CONFIG_FRAMES):
            o = r.add_passes(sc, serve.view_frame_params(spec, f), f)
            frames.append(outputs(r, spec, o))
            n = int(r.read("hash_grid_query_counter", np.uint32, (1,))[0])
            assert n < 131072, f"{spec.width}x{spec.height} frame {f}: {n} sampled queries overflow the buffer"
        torch.cuda.synchronize()
        alone.append(frames)
        r.close()
    rs = [renderer.RestirRenderer(s.width, s.height, cfg) for s in specs]
    b = batch.Batch(batch.BatchConfig(max_views=CONFIG_VIEWS, use_graph=True))
    updated = 0
    for f in range(CONFIG_FRAMES):
        outs = b.add_passes(sc, rs, [serve.view_frame_params(s, f) for s in specs], [f] * len(specs))
        got = [outputs(r, s, o) for r, s, o in zip(rs, specs, outs)]
        torch.cuda.synchronize()
        updated += sum(o.hash_grid_updated for o in outs)
        for v, (a, g) in enumerate(zip(alone, got)):
            for name, x, y in zip(("denoised diffuse", "denoised specular", "raw diffuse"), a[f], g):
                assert torch.equal(x, y), f"{name} differs: view {v} {specs[v].width}x{specs[v].height} frame {f}"
    assert updated == CONFIG_VIEWS * (CONFIG_FRAMES - 1)
    b.close()
    for r in rs:
        r.close()


def test_batch_rejects_bad_calls(vgi):
    import ctypes as C

    from violet_amd import _abi, _lib, batch, renderer, scene

    c = _abi.BatchConfig()
    _lib.call("vgi_batch_default_config", C.byref(c))
    c.max_views = 0
    h = C.c_void_p()
    assert _lib.lib().vgi_batch_create(C.byref(c), C.byref(h)) == _lib.VGI_E_INVALID
    b = batch.Batch(batch.BatchConfig(max_views=2))
    sc = scene.Scene(make_mesh(200))
    r32 = renderer.RestirRenderer(32, 24)
    r16 = renderer.RestirRenderer(32, 24, renderer.RestirConfig(storage=_lib.STORAGE_F16))
    from violet_amd import synth

    fp = synth.make_frame_params(orbit_view(0, 32, 24), None)
    n = 3
    hv = (C.c_void_p * n)(r32._h.value, r16._h.value, r32._h.value)
    fa = (_lib.FrameParams * n)(fp, fp, fp)
    ia = (C.c_uint32 * n)(0, 0, 0)
    oa = (_abi.RendererOutput * n)()
    lib = _lib.lib()
    assert lib.vgi_batch_add_passes(b._h, sc.handle, 3, hv, fa, ia, oa, None) == _lib.VGI_E_INVALID  # > max_views
    assert lib.vgi_batch_add_passes(b._h, sc.handle, 2, hv, fa, ia, oa, None) == _lib.VGI_E_INVALID  # mixed storage
    hv2 = (C.c_void_p * 2)(r32._h.value, r32._h.value)
    assert lib.vgi_batch_add_passes(b._h, sc.handle, 2, hv2, fa, ia, oa, None) == _lib.VGI_E_INVALID  # duplicate
