"""Python mirror of violet's `Denoiser` (crates/violet-restir/src/denoising.rs:41-301).

Same configuration fields and defaults (disocclusion_fix=True,
fast_history_clamp=False, atrous_iterations=3), same `reset()` and
`add_passes(input) -> Output`, backed by the HIP library's
vgi_denoiser_* entry points: the history lives in HBM inside the library
object, like violet's RGTemporal textures.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib
from . import gpu


@dataclass
class Input:
    """denoising.rs:10-18 (device tensors)."""

    depth: object
    gbuffer: object
    prev_depth: object
    prev_gbuffer: object
    diffuse: object
    specular: object
    specular_ray_len: object


@dataclass
class Output:
    """denoising.rs:27-31 + DebugViews (:20-25); raw device pointers owned by
    the denoiser, valid until the next add_passes()."""

    diffuse: int
    specular: int
    history_len: int
    variance: int


class Denoiser:
    def __init__(self, width: int, height: int, storage: int = _lib.STORAGE_F32, compensated_moments: bool = True):
        gpu.require_gpu()
        self.width, self.height, self.storage = width, height, storage
        self.disocclusion_fix = True
        self.fast_history_clamp = False
        self.atrous_iterations = 3
        # FP16 storage: second moments kept as hi + lo f16 pairs (ignored for FP32)
        self.compensated_moments = compensated_moments
        h = C.c_void_p()
        cfg = self._cfg()
        _lib.call("vgi_denoiser_create", width, height, storage, C.byref(cfg), C.byref(h))
        self._h = h

    def _cfg(self) -> _lib.DenoiserConfig:
        return _lib.DenoiserConfig(int(self.disocclusion_fix), int(self.fast_history_clamp),
                                   int(self.atrous_iterations), int(self.compensated_moments))

    def reset(self) -> None:
        _lib.call("vgi_denoiser_reset", self._h)

    def add_passes(self, fp: _lib.FrameParams, inp: Input, stream: int | None = None) -> Output:
        cfg = self._cfg()
        _lib.call("vgi_denoiser_set_config", self._h, C.byref(cfg))
        i = _lib.DenoiserInput(gpu.ptr(inp.depth), gpu.ptr(inp.gbuffer), gpu.ptr(inp.prev_depth),
                               gpu.ptr(inp.prev_gbuffer), gpu.ptr(inp.diffuse), gpu.ptr(inp.specular),
                               gpu.ptr(inp.specular_ray_len))
        o = _lib.DenoiserOutput()
        _lib.call("vgi_denoiser_add_passes", self._h, C.byref(fp), C.byref(i), C.byref(o),
                  gpu.stream() if stream is None else stream)
        return Output(o.diffuse, o.specular, o.history_len, o.variance)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().vgi_denoiser_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
