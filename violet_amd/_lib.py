"""ctypes binding of libvgi.so (the C ABI declared in include/vgi.h).

The library is loaded from the package directory only; if it is missing the
import raises, so no code path can silently fall back to something else.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libvgi.so"

VGI_OK = 0
VGI_E_INVALID = -1
VGI_E_HIP = -2
VGI_E_BUDGET = -3
VGI_E_NOMEM = -4
VGI_E_UNSUPPORTED = -5
VGI_E_TIMEOUT = -6
STORAGE_F32 = 0
STORAGE_F16 = 1

vp = C.c_void_p
u32 = C.c_uint32
f32 = C.c_float


class VgiError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"vgi error {status}: {msg}")
        self.status = status


class FrameParams(C.Structure):
    """frame_bindings.hlsl:6-25 (384 bytes, column-major matrices)."""

    _fields_ = [
        ("view_proj", f32 * 16),
        ("inv_view_proj", f32 * 16),
        ("view_pos", f32 * 4),
        ("view_ray_top_left", f32 * 4),
        ("view_ray_right_shift", f32 * 4),
        ("view_ray_down_shift", f32 * 4),
        ("prev_view_proj", f32 * 16),
        ("prev_inv_view_proj", f32 * 16),
        ("prev_view_pos", f32 * 4),
        ("jitter", f32 * 4),
        ("sun_dir", f32 * 4),
        ("sun_inten", f32 * 4),
    ]


assert C.sizeof(FrameParams) == 384


class DenoiseTemporalIO(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("storage", u32),
        ("depth_texture", vp), ("gbuffer_texture", vp),
        ("diff_source_texture", vp), ("spec_source_texture", vp), ("spec_ray_len_texture", vp),
        ("prev_depth_texture", vp), ("prev_gbuffer_texture", vp),
        ("diff_history_texture", vp), ("spec_history_texture", vp),
        ("moments_history_texture", vp), ("history_len_texture", vp),
        ("diff_fast_history_texture", vp), ("spec_fast_history_texture", vp),
        ("rw_diff_texture", vp), ("rw_spec_texture", vp), ("rw_moments_texture", vp),
        ("rw_history_len_texture", vp),
        ("rw_diff_fast_history_texture", vp), ("rw_spec_fast_history_texture", vp),
        ("moments_comp_history_texture", vp), ("rw_moments_comp_texture", vp),
    ]


class PCDenoiseTemporal(C.Structure):
    _fields_ = [("buffer_size", f32 * 2), ("has_history", u32), ("fast_history", u32)]


class DenoiseSpatialIO(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("storage", u32),
        ("depth_texture", vp), ("gbuffer_texture", vp), ("history_len_texture", vp),
        ("diff_texture", vp), ("spec_texture", vp), ("moments_texture", vp),
        ("rw_diff_texture", vp), ("rw_spec_texture", vp), ("rw_moments_or_variance_texture", vp),
        ("moments_comp_texture", vp), ("rw_moments_comp_texture", vp),
    ]


class DenoiseFastClampIO(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("storage", u32),
        ("diff_fast_history_texture", vp), ("spec_fast_history_texture", vp),
        ("rw_diff_texture", vp), ("rw_spec_texture", vp),
    ]


class DenoiseAtrousIO(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("storage", u32),
        ("depth_texture", vp), ("gbuffer_texture", vp),
        ("diff_texture", vp), ("spec_texture", vp), ("variance_texture", vp),
        ("rw_diff_texture", vp), ("rw_spec_texture", vp), ("rw_variance_texture", vp),
    ]


class PCDenoiseAtrous(C.Structure):
    _fields_ = [("step_size", C.c_int32)]


class DenoiserConfig(C.Structure):
    _fields_ = [("disocclusion_fix", u32), ("fast_history_clamp", u32), ("atrous_iterations", u32),
                ("compensated_moments", u32)]


class DenoiserInput(C.Structure):
    _fields_ = [
        ("depth", vp), ("gbuffer", vp), ("prev_depth", vp), ("prev_gbuffer", vp),
        ("diffuse", vp), ("specular", vp), ("specular_ray_len", vp),
    ]


class DenoiserOutput(C.Structure):
    _fields_ = [("diffuse", vp), ("specular", vp), ("history_len", vp), ("variance", vp)]


class DenoiseFrameIO(C.Structure):
    """vgi_denoise_frame_io: one allocation-free denoiser frame."""

    _fields_ = [
        ("width", u32), ("height", u32), ("storage", u32),
        ("depth", vp), ("gbuffer", vp), ("prev_depth", vp), ("prev_gbuffer", vp),
        ("diffuse", vp), ("specular", vp), ("specular_ray_len", vp),
        ("history_diffuse", vp), ("history_specular", vp), ("history_moments", vp), ("history_moments_comp", vp),
        ("history_len", vp),
        ("rw_history_diffuse", vp), ("rw_history_specular", vp), ("rw_history_moments", vp),
        ("rw_history_moments_comp", vp), ("rw_history_len", vp),
        ("rw_diffuse", vp), ("rw_specular", vp), ("rw_variance", vp),
    ]


DENOISE_STATIC_SCRATCH_BYTES = 128 << 20  # VGI_DENOISE_STATIC_SCRATCH_BYTES


class ProfileEntry(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("total_ms", C.c_double), ("count", C.c_uint64), ("bytes", C.c_double),
                ("rays_closest", C.c_double), ("rays_any", C.c_double), ("nodes", C.c_double), ("tris", C.c_double)]


BVH_NODE_BYTES = 64  # VGI_BVH_NODE_BYTES
BVH_TRI_BYTES = 48  # VGI_BVH_TRI_BYTES


class Tuning(C.Structure):
    """vgi_tuning: process-wide launch shapes (results do not depend on them)."""

    _fields_ = [("atrous_block", u32), ("trace_tile_width", u32)]


# name -> (restype, argtypes)
_SIGNATURES: dict[str, tuple] = {
    "vgi_last_error": (C.c_char_p, []),
    "vgi_version": (C.c_char_p, []),
    "vgi_profiler_enable": (C.c_int32, [u32]),
    "vgi_profiler_reset": (C.c_int32, []),
    "vgi_profiler_read": (C.c_int32, [C.POINTER(ProfileEntry), C.c_int32]),
    "vgi_trace_stats_enable": (C.c_int32, [u32]),
    "vgi_probe_dependent_fetch": (C.c_int32, [C.c_size_t, u32, u32, C.POINTER(C.c_double), vp]),
    "vgi_profiler_only": (C.c_int32, [C.c_char_p]),
    "vgi_tuning_set": (C.c_int32, [C.POINTER(Tuning)]),
    "vgi_tuning_get": (C.c_int32, [C.POINTER(Tuning)]),
    "vgi_fence_create": (C.c_int32, [C.POINTER(vp)]),
    "vgi_fence_destroy": (C.c_int32, [vp]),
    "vgi_fence_record": (C.c_int32, [vp, vp]),
    "vgi_fence_wait": (C.c_int32, [vp, u32, C.POINTER(C.c_double)]),
    "vgi_pass_denoise_temporal": (C.c_int32, [C.POINTER(FrameParams), C.POINTER(DenoiseTemporalIO),
                                              C.POINTER(PCDenoiseTemporal), vp]),
    "vgi_pass_denoise_disocclusion_fix": (C.c_int32, [C.POINTER(DenoiseSpatialIO), vp]),
    "vgi_pass_denoise_spatial_fallback": (C.c_int32, [C.POINTER(DenoiseSpatialIO), vp]),
    "vgi_pass_denoise_fast_history_clamp": (C.c_int32, [C.POINTER(DenoiseFastClampIO), vp]),
    "vgi_pass_denoise_atrous": (C.c_int32, [C.POINTER(DenoiseAtrousIO), C.POINTER(PCDenoiseAtrous), vp]),
    "vgi_denoiser_create": (C.c_int32, [u32, u32, u32, C.POINTER(DenoiserConfig), C.POINTER(vp)]),
    "vgi_denoiser_destroy": (C.c_int32, [vp]),
    "vgi_denoiser_reset": (C.c_int32, [vp]),
    "vgi_denoiser_set_config": (C.c_int32, [vp, C.POINTER(DenoiserConfig)]),
    "vgi_denoiser_add_passes": (C.c_int32, [vp, C.POINTER(FrameParams), C.POINTER(DenoiserInput),
                                            C.POINTER(DenoiserOutput), vp]),
    "vgi_denoise_frame_static": (C.c_int32, [C.POINTER(FrameParams), C.POINTER(DenoiseFrameIO),
                                             C.POINTER(DenoiserConfig), vp]),
}

_lib: C.CDLL | None = None


def register(name: str, restype, argtypes) -> None:
    """Add a signature (used by modules that bind more entry points)."""
    _SIGNATURES[name] = (restype, argtypes)
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype, fn.argtypes = restype, argtypes


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m violet_amd.build` (or __graft_entry__.build())")
        _lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_LOCAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(_lib, name)
            fn.restype, fn.argtypes = res, args
    return _lib


def call(name: str, *args) -> None:
    """Call an entry point and raise VgiError on a non-OK status."""
    st = getattr(lib(), name)(*args)
    if st != VGI_OK:
        raise VgiError(st, lib().vgi_last_error().decode())


def set_tuning(atrous_block: int = 16, trace_tile_width: int = 8) -> None:
    """Launch shapes for the passes issued from now on (vgi_tuning_set)."""
    call("vgi_tuning_set", C.byref(Tuning(atrous_block, trace_tile_width)))


def get_tuning() -> dict:
    t = Tuning()
    call("vgi_tuning_get", C.byref(t))
    return dict(atrous_block=t.atrous_block, trace_tile_width=t.trace_tile_width)


def exported_symbols() -> list[str]:
    return sorted(_SIGNATURES)


def profiler_read() -> dict[str, dict]:
    """Per-pass GPU time / launches / algorithmic bytes since the last reset,
    plus traversal counts (rays_closest, rays_any, nodes, tris) when trace
    statistics were on."""
    n = lib().vgi_profiler_read(None, 0)
    arr = (ProfileEntry * max(n, 1))()
    n = lib().vgi_profiler_read(arr, n)
    return {arr[i].name.decode(): dict(ms=arr[i].total_ms, count=arr[i].count, bytes=arr[i].bytes,
                                       rays_closest=arr[i].rays_closest, rays_any=arr[i].rays_any,
                                       nodes=arr[i].nodes, tris=arr[i].tris) for i in range(n)}
