"""vgi_denoise_frame_static, the allocation-free denoiser frame, against
vgi_denoiser_add_passes: both run the same passes on the same inputs, so
every output - denoised diffuse and specular, the spatial fallback's
variance, the history length - must agree bit for bit, frame after frame,
with the history carried by the caller's buffers.

Each storage mode runs two sequences in a row on the one scratch array: the
first at the largest image the array holds, the second at a smaller size,
so the second call sees the first one's leftovers in the array.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {
    "f32": (0, [(1920, 1080, 2), (320, 180, 3)]),
    "f16": (1, [(2560, 1440, 2), (160, 90, 3)]),
}


def _sequence(storage, w, h, frames, seed):
    import torch

    from violet_amd import _lib, gpu, synth
    from violet_amd import denoiser as D

    rdt = np.float16 if storage else np.float32
    seq = synth.denoise_sequence(w, h, frames, seed=seed)
    den = D.Denoiser(w, h, storage)
    R, Pb = (6 if storage else 12) * w * h, (4 if storage else 8) * w * h
    comp = storage == 1

    def history_set():
        return dict(diff=gpu.zeros(R), spec=gpu.zeros(R), mom=gpu.zeros(Pb), comp=gpu.zeros(Pb) if comp else None,
                    len=gpu.zeros(w * h))

    hist = [history_set(), history_set()]
    outs = dict(diff=gpu.zeros(R), spec=gpu.zeros(R), var=gpu.zeros(Pb))
    prev = None
    for fi, f in enumerate(seq):
        t = dict(depth=gpu.dev(f.depth), gbuffer=gpu.dev(f.gbuffer), diffuse=gpu.dev(f.diffuse.astype(rdt)),
                 specular=gpu.dev(f.specular.astype(rdt)), ray_len=gpu.dev(f.ray_len.astype(rdt)))
        o = den.add_passes(f.fp, D.Input(t["depth"], t["gbuffer"], None if prev is None else prev["depth"],
                                         None if prev is None else prev["gbuffer"], t["diffuse"], t["specular"],
                                         t["ray_len"]))
        hin, hout = hist[fi % 2], hist[1 - fi % 2]
        p = gpu.ptr
        io = _lib.DenoiseFrameIO(
            w, h, storage, p(t["depth"]), p(t["gbuffer"]), None if prev is None else p(prev["depth"]),
            None if prev is None else p(prev["gbuffer"]), p(t["diffuse"]), p(t["specular"]), p(t["ray_len"]),
            p(hin["diff"]), p(hin["spec"]), p(hin["mom"]), p(hin["comp"]), None if prev is None else p(hin["len"]),
            p(hout["diff"]), p(hout["spec"]), p(hout["mom"]), p(hout["comp"]), p(hout["len"]),
            p(outs["diff"]), p(outs["spec"]), p(outs["var"]))
        cfg = _lib.DenoiserConfig(1, 0, 3, 1)
        print(f"no-alloc denoiser frame: storage={storage} {w}x{h} frame {fi}", flush=True)
        _lib.call("vgi_denoise_frame_static", C.byref(f.fp), C.byref(io), C.byref(cfg), gpu.stream())
        torch.cuda.synchronize()
        want = {
            "diffuse": gpu.copy_from_ptr(o.diffuse, R), "specular": gpu.copy_from_ptr(o.specular, R),
            "variance": gpu.copy_from_ptr(o.variance, Pb), "history_len": gpu.copy_from_ptr(o.history_len, w * h),
        }
        got = {"diffuse": outs["diff"], "specular": outs["spec"], "variance": outs["var"], "history_len": hout["len"]}
        for k in want:
            a, b = got[k].cpu().numpy(), want[k].cpu().numpy()
            assert np.array_equal(a, b), f"{k} differs at {w}x{h} frame {fi}: {np.mean(a != b):.3%} of bytes"
        assert np.count_nonzero(want["diffuse"].cpu().numpy()) > 0
        prev = t


@pytest.mark.parametrize("case", sorted(CASES))
def test_static_frame_matches_denoiser(vgi, case):
    storage, runs = CASES[case]
    for i, (w, h, frames) in enumerate(runs):
        _sequence(storage, w, h, frames, seed=3 + i)


def test_static_frame_rejects_what_it_cannot_hold(vgi):
    from violet_amd import _lib, gpu, synth

    f = synth.denoise_sequence(64, 32, 1)[0]
    io = _lib.DenoiseFrameIO(4096, 4096, 0)
    assert vgi.vgi_denoise_frame_static(C.byref(f.fp), C.byref(io), None, gpu.stream()) == _lib.VGI_E_INVALID
    assert b"above the maximum" in vgi.vgi_last_error()
    cfg = _lib.DenoiserConfig(1, 1, 3, 1)
    io = _lib.DenoiseFrameIO(64, 32, 0)
    assert vgi.vgi_denoise_frame_static(C.byref(f.fp), C.byref(io), C.byref(cfg), gpu.stream()) == _lib.VGI_E_INVALID
    assert b"fast_history_clamp" in vgi.vgi_last_error()
