"""violet_amd — MI355X-native (gfx950 HIP) ReSTIR GI indirect-lighting hot path.

Drop-in for violet's render-graph passes (crates/violet-restir): the C ABI in
include/vgi.h, bound here with ctypes, plus Python mirrors of violet's pass
sequencers (`Denoiser`, `RestirRenderer`).
"""
from ._lib import STORAGE_F16, STORAGE_F32, FrameParams, VgiError, lib

__all__ = ["STORAGE_F16", "STORAGE_F32", "FrameParams", "VgiError", "lib"]
