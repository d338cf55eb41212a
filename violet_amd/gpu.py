"""Device-memory plumbing for the Python mirror and the tests.

PyTorch provides HBM allocations and the current HIP stream; the compute is
always the HIP library (libvgi.so).  Images cross the C ABI as raw device
pointers, exactly as a Rust host would pass them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


def require_gpu() -> None:
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("violet_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.lib()  # fail loudly if libvgi.so is missing


def dev(a: np.ndarray) -> "torch.Tensor":
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def empty(nbytes: int) -> "torch.Tensor":
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")


def zeros(nbytes: int) -> "torch.Tensor":
    return torch.zeros(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")


def ptr(t) -> int | None:
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def host(t: "torch.Tensor", dtype, shape) -> np.ndarray:
    return t.detach().cpu().numpy().view(dtype).reshape(shape)


def radiance_dtype(storage: int):
    return np.float16 if storage == _lib.STORAGE_F16 else np.float32


def to_storage(a: np.ndarray, storage: int) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=radiance_dtype(storage))


_lib.register("vgi_memcpy_d2d", C.c_int32, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p])


def copy_from_ptr(src: int, nbytes: int) -> "torch.Tensor":
    """Copy a library-owned device image into a new torch tensor (same stream)."""
    out = empty(nbytes)
    _lib.call("vgi_memcpy_d2d", out.data_ptr(), src, nbytes, stream())
    return out


class Fence:
    """vgi_fence: marks the work issued on a stream and waits for it with a
    deadline (violet's 5 s frame fence, render_loop.rs:445-482)."""

    TIMEOUT_MS = 5000

    def __init__(self):
        self._h = C.c_void_p()
        _lib.call("vgi_fence_create", C.byref(self._h))

    def record(self, stream_handle: int | None = None) -> None:
        _lib.call("vgi_fence_record", self._h, stream() if stream_handle is None else stream_handle)

    def wait(self, timeout_ms: int = TIMEOUT_MS) -> tuple[bool, float]:
        """(finished, ms waited); finished is False after a time-out (the work
        keeps running)."""
        waited = C.c_double()
        st = _lib.lib().vgi_fence_wait(self._h, int(timeout_ms), C.byref(waited))
        if st == _lib.VGI_E_TIMEOUT:
            return False, waited.value
        if st != _lib.VGI_OK:
            raise _lib.VgiError(st, _lib.lib().vgi_last_error().decode())
        return True, waited.value

    def close(self) -> None:
        if self._h:
            _lib.lib().vgi_fence_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
