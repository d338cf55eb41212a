"""Batched views: one frame of many RestirRenderer views with every pass of
all views fused into one persistent launch (vgi_batch_* in include/vgi.h).

No reference counterpart: violet renders one view per frame
(crates/violet-restir/src/restir_render_loop.rs:262-268).  A view rendered
in a batch advances exactly as RestirRenderer.add_passes would advance it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _abi, _lib, gpu
from .renderer import FrameOutput, RestirRenderer
from .scene import Scene


@dataclass
class BatchConfig:
    """Persistent group counts (0 = the library default for this device),
    whether the frame is captured once as a HIP graph and replayed, and
    whether independent pass chains run on a second stream."""

    max_views: int = 64
    pixel_groups: int = 0
    trace_groups: int = 0
    atrous_groups: int = 0
    use_graph: bool = False
    concurrent_lanes: bool = True

    def to_c(self) -> _abi.BatchConfig:
        c = _abi.BatchConfig()
        _lib.call("vgi_batch_default_config", C.byref(c))
        c.max_views = self.max_views
        if self.pixel_groups:
            c.pixel_groups = self.pixel_groups
        if self.trace_groups:
            c.trace_groups = self.trace_groups
        if self.atrous_groups:
            c.atrous_groups = self.atrous_groups
        c.use_graph = int(self.use_graph)
        c.concurrent_lanes = int(self.concurrent_lanes)
        return c


class Batch:
    def __init__(self, config: BatchConfig | None = None):
        gpu.require_gpu()
        self.config = config or BatchConfig()
        c = self.config.to_c()
        self.resolved = {f: getattr(c, f) for f, _ in c._fields_}
        h = C.c_void_p()
        _lib.call("vgi_batch_create", C.byref(c), C.byref(h))
        self._h = h

    def add_passes(self, scene: Scene, views: list[RestirRenderer], fps: list, frame_indices: list[int],
                   stream: int | None = None) -> list[FrameOutput]:
        n = len(views)
        hv = (C.c_void_p * n)(*[v._h.value for v in views])
        fa = (_lib.FrameParams * n)(*fps)
        ia = (C.c_uint32 * n)(*[int(i) & 0xFFFFFFFF for i in frame_indices])
        oa = (_abi.RendererOutput * n)()
        _lib.call("vgi_batch_add_passes", self._h, scene.handle, n, hv, fa, ia, oa,
                  gpu.stream() if stream is None else stream)
        return [FrameOutput(o.indirect_diffuse, o.indirect_specular, o.depth, o.gbuffer, bool(o.denoised),
                            bool(o.validation_frame), bool(o.hash_grid_updated), o.color) for o in oa]

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().vgi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
