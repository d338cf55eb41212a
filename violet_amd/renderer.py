"""Python mirror of violet's `RestirRenderer`
(crates/violet-restir/src/restir_renderer.rs:146-1028) over the HIP library.

`RestirRenderer(width, height, config)` keeps one view's temporal state in
HBM; `add_passes(scene, frame_params, frame_index)` records the frame's pass
chain (G-buffer from primary rays, cache update, diffuse/specular ReSTIR,
AO, denoiser) and returns the indirect diffuse/specular images.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi, _lib, gpu
from .scene import Scene


@dataclass
class RestirConfig:
    """restir_renderer.rs:51-74 `RestirConfig` (+ Denoiser fields, denoising.rs:41-60)."""

    storage: int = _lib.STORAGE_F32
    denoise: bool = True
    ind_diff_validate: bool = True
    ind_spec_validate: bool = True
    hash_grid_cache_decay: bool = True
    ao_radius: float = 2.0
    sky_radiance: tuple = (1.0, 1.0, 1.0)
    # resampling parameters, 0 = the reference's constants: temporal M clamp
    # (16; specular uint(cap * roughness)), diffuse spatial spiral taps (9),
    # spatial search radius in pixels (0.05 * min(w, h))
    temporal_m_cap: int = 0
    diff_spatial_taps: int = 0
    spatial_radius_px: float = 0.0
    # independent pass chains on a second stream (hash-grid update beside the
    # G-buffer, specular beside diffuse), joined into the caller's stream
    concurrent_lanes: bool = True
    # also run the sun shadow mask and the final lighting (the lit image)
    compose: bool = False
    # temporal AA of the lit image (needs compose; jitter the frame params with
    # synth.jitter_ndc(frame_index % jitter_cycle, w, h))
    taa: bool = False
    disocclusion_fix: bool = True
    fast_history_clamp: bool = False
    atrous_iterations: int = 3
    compensated_moments: bool = True  # FP16 storage: moments as hi + lo f16 pairs
    # order-independent cache update: the same frames run after run and alone
    # or in a batch (False: the reference's schedule-dependent cache passes)
    ordered_cache_update: bool = True

    def to_c(self) -> _abi.RendererConfig:
        c = _abi.RendererConfig()
        c.storage = self.storage
        c.denoise = int(self.denoise)
        c.ind_diff_validate = int(self.ind_diff_validate)
        c.ind_spec_validate = int(self.ind_spec_validate)
        c.hash_grid_cache_decay = int(self.hash_grid_cache_decay)
        c.ao_radius = self.ao_radius
        c.sky_radiance[:] = list(self.sky_radiance)
        c.temporal_m_cap = self.temporal_m_cap
        c.diff_spatial_taps = self.diff_spatial_taps
        c.spatial_radius_px = self.spatial_radius_px
        c.concurrent_lanes = int(self.concurrent_lanes)
        c.compose = int(self.compose)
        c.taa = int(self.taa)
        c.denoiser = _lib.DenoiserConfig(int(self.disocclusion_fix), int(self.fast_history_clamp),
                                         int(self.atrous_iterations), int(self.compensated_moments))
        c.ordered_cache_update = int(self.ordered_cache_update)
        return c


@dataclass
class FrameOutput:
    indirect_diffuse: int
    indirect_specular: int
    depth: int
    gbuffer: int
    denoised: bool
    validation_frame: bool
    hash_grid_updated: bool
    color: int | None = None  # final lighting when the config composes


class RestirRenderer:
    def __init__(self, width: int, height: int, config: RestirConfig | None = None):
        gpu.require_gpu()
        self.width, self.height = width, height
        self.config = config or RestirConfig()
        h = C.c_void_p()
        c = self.config.to_c()
        _lib.call("vgi_renderer_create", width, height, C.byref(c), C.byref(h))
        self._h = h

    def reset(self) -> None:
        _lib.call("vgi_renderer_reset", self._h)

    def add_passes(self, scene: Scene, fp: _lib.FrameParams, frame_index: int, stream: int | None = None) -> FrameOutput:
        o = _abi.RendererOutput()
        _lib.call("vgi_renderer_add_passes", self._h, scene.handle, C.byref(fp), frame_index & 0xFFFFFFFF, C.byref(o),
                  gpu.stream() if stream is None else stream)
        return FrameOutput(o.indirect_diffuse, o.indirect_specular, o.depth, o.gbuffer, bool(o.denoised),
                           bool(o.validation_frame), bool(o.hash_grid_updated), o.color)

    def image(self, name: str) -> tuple[int, int]:
        p = C.c_void_p()
        n = C.c_size_t()
        _lib.call("vgi_renderer_get_image", self._h, name.encode(), C.byref(p), C.byref(n))
        return p.value, n.value

    def read(self, name: str, dtype, shape) -> np.ndarray:
        p, n = self.image(name)
        t = gpu.copy_from_ptr(p, n)
        return t.cpu().numpy().view(dtype)[: int(np.prod(shape))].reshape(shape)

    def read_ptr(self, ptr: int, nbytes: int, dtype, shape) -> np.ndarray:
        t = gpu.copy_from_ptr(ptr, nbytes)
        return t.cpu().numpy().view(dtype).reshape(shape)

    def state_size(self) -> int:
        return _lib.lib().vgi_renderer_state_size(self._h)

    def device_bytes(self) -> int:
        """Device memory the view holds (state plus per-frame scratch images)."""
        return _lib.lib().vgi_renderer_device_bytes(self._h)

    def export_state(self) -> np.ndarray:
        n = self.state_size()
        blob = np.empty(n, np.uint8)
        _lib.call("vgi_renderer_export_state", self._h, blob.ctypes.data, n, gpu.stream())
        return blob

    def import_state(self, blob: np.ndarray) -> None:
        blob = np.ascontiguousarray(blob, np.uint8)
        _lib.call("vgi_renderer_import_state", self._h, blob.ctypes.data, blob.nbytes, gpu.stream())

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().vgi_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
