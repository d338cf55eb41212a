"""Direct light and composition on the GPU (raytraced_shadow.hlsl:9-53,
restir/final_lighting.hlsl:16-62), through the renderer with `compose` on
and through the two C entry points.

* The sun shadow mask of every frame equals the FP32 twin's bit for bit (the
  twin traces its own BVH: only the any-hit outcome is compared).
* The final lighting, replayed by the FP32 twin on the frame's recorded
  G-buffer, mask and denoised indirect lighting, is bit-identical (FP16
  storage: the twin's f32 result rounded to half), and within the parity
  tolerance of the FP64 oracle.
* `compose` leaves the GI outputs untouched (bit-identical first frame,
  identical G-buffer and depth on every frame).
* vgi_pass_raytraced_shadow / vgi_pass_final_lighting called directly on the
  renderer's images reproduce the renderer's mask and colour.
* Temporal AA over a jittered orbit (Halton(2,3), cycle 8): each frame's
  resolve, replayed by the FP32 twin on the frame's lit image, depth and the
  previous resolve, is bit-identical (FP16: rounded to half) and within the
  FP64 oracle's tolerance (at most 1% of pixels above the per-pixel bound:
  the variance clip cancels in f32); vgi_pass_temporal_aa called directly
  reproduces it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from parity import assert_close, assert_close_frac
from restir_harness import F32, make_mesh, run_gpu_frames

pytestmark = pytest.mark.gpu

W, H = 96, 64
FRAMES = 4
SKY = (0.4, 0.6, 1.5)


@pytest.fixture(scope="module", params=["f32", "f16"])
def run(request):
    from violet_amd import _lib, renderer

    mesh = make_mesh()
    storage = _lib.STORAGE_F16 if request.param == "f16" else _lib.STORAGE_F32
    cfg = renderer.RestirConfig(storage=storage, compose=True, sky_radiance=SKY)
    fps, outs, snaps, r, sc = run_gpu_frames(mesh, W, H, FRAMES, config=cfg)
    return dict(mesh=mesh, fps=fps, outs=outs, snaps=snaps, r=r, sc=sc, storage=storage, f16=request.param == "f16")


@pytest.fixture(scope="module")
def o32(run):
    import oracle

    return oracle.OracleRestir(32, oracle.OracleScene(32, run["mesh"]), sky=SKY)


@pytest.fixture(scope="module")
def o64(run):
    import oracle

    return oracle.OracleRestir(64, oracle.OracleScene(64, run["mesh"]), sky=SKY)


def bits_equal(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    neq = a.view(np.uint8 if a.dtype.itemsize == 1 else np.uint16 if a.dtype.itemsize == 2 else np.uint32) != \
        b.view(np.uint8 if b.dtype.itemsize == 1 else np.uint16 if b.dtype.itemsize == 2 else np.uint32)
    frac = float(np.mean(neq))
    assert frac == 0.0, f"{what}: {frac:.4%} of words differ (first at {np.argwhere(neq)[:3].tolist()})"


@pytest.mark.parametrize("k", range(FRAMES))
def test_shadow_mask_matches_twin(run, o32, k):
    s, fp = run["snaps"][k], run["fps"][k]
    mask = o32.raytraced_shadow(fp, s["depth"])
    bits_equal(s["shadow_mask"], mask, f"shadow mask[{k}]")
    geo = s["depth"] > 0
    assert set(np.unique(s["shadow_mask"])) <= {0, 255}
    assert (s["shadow_mask"][~geo] == 0).all()
    assert (s["shadow_mask"][geo] == 0).any() and (s["shadow_mask"][geo] == 255).any()


@pytest.mark.parametrize("k", range(FRAMES))
def test_final_lighting_matches_twin_and_fp64(run, o32, o64, k):
    s, fp = run["snaps"][k], run["fps"][k]
    ind_d = s["out_diffuse"].astype(F32)
    ind_s = s["out_specular"].astype(F32)
    c32 = o32.final_lighting(fp, s["gbuffer"], s["shadow_mask"], ind_d, ind_s)
    if run["f16"]:
        c32 = c32.astype(np.float16)
    bits_equal(s["out_color"], c32, f"final lighting[{k}]")
    c64 = o64.final_lighting(fp, s["gbuffer"], s["shadow_mask"], ind_d, ind_s)
    tol = dict(rel_tol=2e-3, pix_tol=2e-3) if run["f16"] else {}
    assert_close(s["out_color"].astype(np.float64), c64, f"final lighting vs FP64 [{k}]", **tol)
    sky = s["gbuffer"][3] == 0
    assert sky.any() and np.allclose(s["out_color"][sky].astype(np.float64), np.array(SKY, np.float16 if run["f16"]
                                                                                      else F32).astype(np.float64))


def test_compose_leaves_gi_outputs_unchanged(run):
    from violet_amd import renderer

    cfg = renderer.RestirConfig(storage=run["storage"], compose=False, sky_radiance=SKY)
    fps, outs, snaps, r, sc = run_gpu_frames(run["mesh"], W, H, FRAMES, config=cfg)
    assert all(not o.color for o in outs)
    for name in ("out_diffuse", "out_specular", "indirect_diffuse_raw", "indirect_specular_raw"):
        bits_equal(run["snaps"][0][name], snaps[0][name], f"{name}[0]")
    for k in range(FRAMES):
        bits_equal(run["snaps"][k]["depth"], snaps[k]["depth"], f"depth[{k}]")
        bits_equal(run["snaps"][k]["gbuffer"], snaps[k]["gbuffer"], f"gbuffer[{k}]")


def test_direct_entry_points_reproduce_renderer(run):
    import torch

    from violet_amd import _abi, _lib, gpu

    k = FRAMES - 1
    o, s, fp, r = run["outs"][k], run["snaps"][k], run["fps"][k], run["r"]
    mask = gpu.zeros(W * H)
    sio = _abi.RaytracedShadowIO(W, H, o.depth, mask.data_ptr())
    _lib.call("vgi_pass_raytraced_shadow", C.byref(fp), run["sc"].handle, C.byref(sio), gpu.stream())
    rb = W * H * (6 if run["f16"] else 12)
    color = gpu.zeros(rb)
    fio = _abi.FinalLightingIO(W, H, run["storage"], o.gbuffer, mask.data_ptr(), o.indirect_diffuse,
                               o.indirect_specular, color.data_ptr())
    fio.sky_radiance[:] = list(SKY)
    _lib.call("vgi_pass_final_lighting", C.byref(fp), C.byref(fio), gpu.stream())
    torch.cuda.synchronize()
    bits_equal(gpu.host(mask, np.uint8, (H, W)), s["shadow_mask"], "direct shadow mask")
    dt = np.float16 if run["f16"] else F32
    bits_equal(gpu.host(color, dt, (H, W, 3)), s["out_color"], "direct final lighting")
    # missing bindings are rejected on the host
    bad = _abi.FinalLightingIO(W, H, run["storage"], o.gbuffer, None, o.indirect_diffuse, o.indirect_specular,
                               color.data_ptr())
    with pytest.raises(_lib.VgiError):
        _lib.call("vgi_pass_final_lighting", C.byref(fp), C.byref(bad), gpu.stream())


@pytest.fixture(scope="module", params=["f32", "f16"])
def taa_run(request):
    from violet_amd import _lib, renderer

    mesh = make_mesh()
    storage = _lib.STORAGE_F16 if request.param == "f16" else _lib.STORAGE_F32
    cfg = renderer.RestirConfig(storage=storage, compose=True, taa=True, sky_radiance=SKY)
    fps, outs, snaps, r, sc = run_gpu_frames(mesh, W, H, FRAMES + 1, config=cfg, jitter=True)
    return dict(mesh=mesh, fps=fps, outs=outs, snaps=snaps, r=r, sc=sc, storage=storage, f16=request.param == "f16")


@pytest.mark.parametrize("k", range(FRAMES + 1))
def test_temporal_aa_matches_twin_and_fp64(taa_run, k):
    import oracle

    run = taa_run
    s, fp = run["snaps"][k], run["fps"][k]
    assert any(fp.jitter[:2]), "frames are jittered"
    hist = None if k == 0 else run["snaps"][k - 1]["out_color"].astype(F32)
    src = s["scene_color"].astype(F32)
    o32 = oracle.OracleRestir(32, oracle.OracleScene(32, run["mesh"]), sky=SKY)
    t32 = o32.temporal_aa(fp, s["depth"], src, hist)
    if run["f16"]:
        t32 = t32.astype(np.float16)
    bits_equal(s["out_color"], t32, f"temporal AA[{k}]")
    o64 = oracle.OracleRestir(64, oracle.OracleScene(64, run["mesh"]), sky=SKY)
    t64 = o64.temporal_aa(fp, s["depth"], src, hist)
    # The variance clip's sigma = sqrt(|E[x^2] - E[x]^2|) cancels in f32 (as in
    # the reference) for nearly flat neighbourhoods, and history positions on
    # the previous view's edge can take the other branch: bar as for outputs
    # downstream of a selection, at most 1% of pixels above the per-pixel bound.
    tol = dict(rel_tol=2e-3, pix_tol=2e-3) if run["f16"] else {}
    assert_close_frac(s["out_color"].astype(np.float64), t64, f"temporal AA vs FP64 [{k}]", frac=0.01, **tol)
    if k > 0:  # the history takes part
        alone = o32.temporal_aa(fp, s["depth"], src).astype(s["out_color"].dtype)
        assert not np.array_equal(s["out_color"], alone)


def test_temporal_aa_entry_point_reproduces_renderer(taa_run):
    import torch

    from violet_amd import _abi, _lib, gpu

    run = taa_run
    k = FRAMES
    o, s, fp, r = run["outs"][k], run["snaps"][k], run["fps"][k], run["r"]
    rb = W * H * (6 if run["f16"] else 12)
    out = gpu.zeros(rb)
    src, _ = r.image("color")
    io = _abi.TemporalAAIO(W, H, run["storage"], o.depth, src, run["outs"][k - 1].color, out.data_ptr())
    _lib.call("vgi_pass_temporal_aa", C.byref(fp), C.byref(io), C.byref(_abi.PCTemporalAA(1)), gpu.stream())
    torch.cuda.synchronize()
    dt = np.float16 if run["f16"] else F32
    bits_equal(gpu.host(out, dt, (H, W, 3)), s["out_color"], "direct temporal AA")
    alias = _abi.TemporalAAIO(W, H, run["storage"], o.depth, src, run["outs"][k - 1].color, src)
    with pytest.raises(_lib.VgiError):
        _lib.call("vgi_pass_temporal_aa", C.byref(fp), C.byref(alias), C.byref(_abi.PCTemporalAA(1)), gpu.stream())
