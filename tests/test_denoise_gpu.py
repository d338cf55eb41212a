"""GPU parity of the denoiser passes (HIP library through the C ABI) against
the FP64 oracle and the FP32 twin, on the seeded plane+spheres workload.

Bars: radiance / moments / variance within rel_mean <= 1e-3 and
max_pixel <= 1e-2 of the FP64 oracle (tests/parity.py); the 8-bit history
length identical to the FP32 twin.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from parity import assert_close, rel_mean

pytestmark = pytest.mark.gpu


def _seq(w, h, frames, seed=1):
    from violet_amd import synth

    return synth.denoise_sequence(w, h, frames, seed=seed)


class Pass:
    """Runs single denoiser passes on the GPU from host numpy arrays."""

    def __init__(self, storage=0):
        import torch

        from violet_amd import _lib, gpu

        self.torch, self.L, self.gpu, self.storage = torch, _lib, gpu, storage
        self.keep = []

    def d(self, a):
        if a is None:
            return None
        if a.dtype in (np.float32, np.float64) and self.storage == 1:
            a = a.astype(np.float16)
        elif a.dtype == np.float64:
            a = a.astype(np.float32)
        t = self.gpu.dev(a)
        self.keep.append(t)
        return t.data_ptr()

    def out(self, shape, dtype):
        import torch

        if dtype == np.float32 and self.storage == 1:
            dtype = np.float16
        t = torch.zeros(int(np.prod(shape)) * np.dtype(dtype).itemsize, dtype=torch.uint8, device="cuda")
        self.keep.append(t)
        return t, shape, dtype

    def get(self, o):
        t, shape, dtype = o
        self.torch.cuda.synchronize()
        return t.cpu().numpy().view(dtype).reshape(shape).astype(np.float64 if dtype != np.uint8 else np.uint8)

    def temporal(self, fp, depth, gb, ds, ss, rl, pd, pgb, dh, sh, mh, lh, has_history):
        h, w = depth.shape
        od, osp, om, ol = self.out((h, w, 3), np.float32), self.out((h, w, 3), np.float32), self.out(
            (h, w, 2), np.float32), self.out((h, w), np.uint8)
        io = self.L.DenoiseTemporalIO(w, h, self.storage, self.d(depth), self.d(gb), self.d(ds), self.d(ss),
                                      self.d(rl), self.d(pd), self.d(pgb), self.d(dh), self.d(sh), self.d(mh),
                                      self.d(lh), None, None, od[0].data_ptr(), osp[0].data_ptr(), om[0].data_ptr(),
                                      ol[0].data_ptr(), None, None)
        pc = self.L.PCDenoiseTemporal((C.c_float * 2)(w, h), int(has_history), 0)
        self.L.call("vgi_pass_denoise_temporal", C.byref(fp), C.byref(io), C.byref(pc), self.gpu.stream())
        return dict(diff=self.get(od), spec=self.get(osp), moments=self.get(om), history_len=self.get(ol))

    def spatial(self, name, depth, gb, lt, diff, spec, mom):
        h, w = depth.shape
        od, osp, om = self.out((h, w, 3), np.float32), self.out((h, w, 3), np.float32), self.out((h, w, 2), np.float32)
        io = self.L.DenoiseSpatialIO(w, h, self.storage, self.d(depth), self.d(gb), self.d(lt), self.d(diff),
                                     self.d(spec), self.d(mom), od[0].data_ptr(), osp[0].data_ptr(), om[0].data_ptr())
        self.L.call(name, C.byref(io), self.gpu.stream())
        return self.get(od), self.get(osp), self.get(om)

    def atrous(self, step, depth, gb, diff, spec, var):
        h, w = depth.shape
        od, osp, ov = self.out((h, w, 3), np.float32), self.out((h, w, 3), np.float32), self.out((h, w, 2), np.float32)
        io = self.L.DenoiseAtrousIO(w, h, self.storage, self.d(depth), self.d(gb), self.d(diff), self.d(spec),
                                    self.d(var), od[0].data_ptr(), osp[0].data_ptr(), ov[0].data_ptr())
        self.L.call("vgi_pass_denoise_atrous", C.byref(io), C.byref(self.L.PCDenoiseAtrous(step)), self.gpu.stream())
        return dict(diff=self.get(od), spec=self.get(osp), variance=self.get(ov))


def _f32(a):
    return None if a is None else np.asarray(a, np.float32)


@pytest.mark.parametrize("w,h", [(160, 90), (77, 51)])
def test_denoise_passes_vs_oracle(vgi, oracle64, oracle32, w, h):
    """Each pass fed the same f32 inputs on GPU and oracle (3 frames of history)."""
    import oracle as O

    seq = _seq(w, h, 3)
    ref = O.OracleDenoiser(64, atrous_iterations=5)
    g = Pass()
    prev = None
    for fi, f in enumerate(seq):
        hist = ref.history
        has_h = hist is not None
        args = (f.fp, f.depth, f.gbuffer, f.diffuse, f.specular, f.ray_len,
                None if prev is None else prev.depth, None if prev is None else prev.gbuffer,
                *( (_f32(hist["diff"]), _f32(hist["spec"]), _f32(hist["moments"]), hist["history_len"]) if has_h else (None,) * 4),
                has_h)
        gt = g.temporal(*args)
        ot = oracle64.denoise_temporal(*args)
        tt = oracle32.denoise_temporal(*args)
        for k in ("diff", "spec", "moments"):
            assert_close(gt[k], ot[k], f"temporal.{k} frame {fi}")
        mism = np.mean(gt["history_len"] != tt["history_len"])
        assert mism == 0.0, f"history_len differs from the FP32 twin at {mism:.2%} of pixels"
        # spatial passes on the oracle's temporal output (cast to f32)
        lt = ot["history_len"]
        d32, s32, m32 = _f32(ot["diff"]), _f32(ot["spec"]), _f32(ot["moments"])
        gx = g.spatial("vgi_pass_denoise_disocclusion_fix", f.depth, f.gbuffer, lt, d32, s32, m32)
        ox = oracle64.denoise_disocclusion(f.depth, f.gbuffer, lt, d32, s32, m32)
        for a, k in zip(gx, ("diff", "spec", "moments")):
            assert_close(a, ox[k], f"disocclusion.{k} frame {fi}")
        fx = (_f32(ox["diff"]), _f32(ox["spec"]), _f32(ox["moments"]))
        gf = g.spatial("vgi_pass_denoise_spatial_fallback", f.depth, f.gbuffer, lt, *fx)
        of = oracle64.denoise_fallback(f.depth, f.gbuffer, lt, *fx)
        for a, k in zip(gf, ("diff", "spec", "variance")):
            assert_close(a, of[k], f"fallback.{k} frame {fi}")
        d, s, v = _f32(of["diff"]), _f32(of["spec"]), _f32(of["variance"])
        for it in range(5):
            ga = g.atrous(1 << it, f.depth, f.gbuffer, d, s, v)
            oa = oracle64.denoise_atrous(1 << it, f.depth, f.gbuffer, d, s, v)
            for k in ("diff", "spec", "variance"):
                assert_close(ga[k], oa[k], f"atrous[{it}].{k} frame {fi}")
            d, s, v = _f32(oa["diff"]), _f32(oa["spec"]), _f32(oa["variance"])
        ref.add_passes(f.fp, f.depth, f.gbuffer, None if prev is None else prev.depth,
                       None if prev is None else prev.gbuffer, f.diffuse, f.specular, f.ray_len)
        prev = f
        g.keep.clear()


@pytest.mark.parametrize("fast_history_clamp", [False, True], ids=["no_fast_clamp", "fast_clamp"])
def test_denoiser_sequence_vs_oracle(vgi, fast_history_clamp):
    """Denoiser (vgi_denoiser_add_passes) over an 8-frame pan vs the FP64
    oracle sequencer, with and without the fast-history clamp
    (relax_fast_history_clamp.hlsl:12-75, denoising.rs:192-207)."""
    import oracle as O

    from violet_amd import denoiser as D
    from violet_amd import gpu

    w, h = 320, 180
    seq = _seq(w, h, 8, seed=3)
    den = D.Denoiser(w, h)
    den.atrous_iterations = 5
    den.fast_history_clamp = fast_history_clamp
    ref = O.OracleDenoiser(64, atrous_iterations=5, fast_history_clamp=fast_history_clamp)
    prev = None
    prev_t = None
    for fi, f in enumerate(seq):
        t = dict(depth=gpu.dev(f.depth), gbuffer=gpu.dev(f.gbuffer), diffuse=gpu.dev(f.diffuse),
                 specular=gpu.dev(f.specular), ray_len=gpu.dev(f.ray_len))
        out = den.add_passes(f.fp, D.Input(t["depth"], t["gbuffer"], None if prev_t is None else prev_t["depth"],
                                           None if prev_t is None else prev_t["gbuffer"], t["diffuse"], t["specular"],
                                           t["ray_len"]))
        gd = gpu.copy_from_ptr(out.diffuse, w * h * 12).cpu().numpy().view(np.float32).reshape(h, w, 3)
        gs = gpu.copy_from_ptr(out.specular, w * h * 12).cpu().numpy().view(np.float32).reshape(h, w, 3)
        gl = gpu.copy_from_ptr(out.history_len, w * h).cpu().numpy().reshape(h, w)
        r = ref.add_passes(f.fp, f.depth, f.gbuffer, None if prev is None else prev.depth,
                           None if prev is None else prev.gbuffer, f.diffuse, f.specular, f.ray_len)
        assert_close(gd, r["diffuse"], f"sequence diffuse frame {fi}")
        assert_close(gs, r["specular"], f"sequence specular frame {fi}")
        assert np.mean(gl != r["history_len"]) < 1e-3
        prev, prev_t = f, t


@pytest.mark.parametrize("storage", [0, 1], ids=["f32", "f16"])
def test_fast_history_clamp_pass_vs_oracle(vgi, oracle64, storage):
    """vgi_pass_denoise_fast_history_clamp on its own: the colour is clamped
    into the fast history's 3x3 YCoCg mean +- 0.75 sigma box
    (relax_fast_history_clamp.hlsl:12-75).  Inputs are a smooth fast history
    and a noisy slow colour, so most pixels move.  FP32 and FP16 storage vs
    the FP64 oracle on the storage-rounded inputs."""
    from violet_amd import _lib

    w, h = 173, 97
    rng = np.random.default_rng(11)
    yy, xx = np.mgrid[0:h, 0:w]
    base = np.stack([np.sin(xx / 9.0) + 1.5, np.cos(yy / 7.0) + 1.5, np.sin((xx + yy) / 13.0) + 1.5], -1)
    fast_d = base * (1.0 + 0.05 * rng.standard_normal((h, w, 3)))
    fast_s = 0.5 * base[..., ::-1] * (1.0 + 0.05 * rng.standard_normal((h, w, 3)))
    diff = base * rng.exponential(1.0, (h, w, 3))
    spec = 0.5 * base[..., ::-1] * rng.exponential(1.0, (h, w, 3))
    dt = np.float16 if storage else np.float32
    fast_d, fast_s, diff, spec = (a.astype(dt) for a in (fast_d, fast_s, diff, spec))
    g = Pass(storage)
    od, osp = g.out((h, w, 3), np.float32), g.out((h, w, 3), np.float32)
    od[0].copy_(g.torch.from_numpy(np.ascontiguousarray(diff).view(np.uint8).ravel()).cuda())
    osp[0].copy_(g.torch.from_numpy(np.ascontiguousarray(spec).view(np.uint8).ravel()).cuda())
    io = _lib.DenoiseFastClampIO(w, h, storage, g.d(fast_d), g.d(fast_s), od[0].data_ptr(), osp[0].data_ptr())
    _lib.call("vgi_pass_denoise_fast_history_clamp", C.byref(io), g.gpu.stream())
    gd, gs = g.get(od), g.get(osp)
    rd, rs = oracle64.denoise_fast_clamp(fast_d, fast_s, diff, spec)
    assert np.mean(np.abs(rd - diff.astype(np.float64)) > 1e-3) > 0.3, "the clamp moved too few pixels"
    # f16 stores round each output once (half an ulp, 4.9e-4 relative)
    tol = dict(rel_tol=1e-3, pix_tol=1e-2) if storage == 0 else dict(rel_tol=1e-3, pix_tol=1.5e-3)
    assert_close(gd, rd, "fast clamp diffuse", **tol)
    assert_close(gs, rs, "fast clamp specular", **tol)


def test_denoise_invalid_arguments_do_not_launch(vgi):
    from violet_amd import _lib

    io = _lib.DenoiseAtrousIO(0, 0, 0)
    st = vgi.vgi_pass_denoise_atrous(C.byref(io), C.byref(_lib.PCDenoiseAtrous(1)), None)
    assert st == _lib.VGI_E_INVALID
