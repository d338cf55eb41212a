"""The C-ABI library loads on the CPU and exports every entry point that
include/vgi.h declares; argument validation returns VGI_E_INVALID without
touching the GPU (violet's graph warns and continues on a missing binding,
render_graph.rs:1881-1886; here nothing is launched)."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> list[str]:
    names = []
    for hdr in sorted((ROOT / "include").glob("*.h")):
        text = hdr.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(vgi_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from violet_amd import _lib

    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (vgi_\w+)", out))
    assert set(names) <= exported


def test_python_binding_covers_header():
    from violet_amd import _lib
    import violet_amd.gpu  # noqa: F401  (these modules register their entry points)
    import violet_amd.renderer  # noqa: F401
    import violet_amd.scene  # noqa: F401

    names = declared_functions()
    missing = [n for n in names if n not in _lib.exported_symbols()]
    assert not missing, f"no ctypes signature for {missing}"


def test_invalid_arguments_return_status_and_message():
    from violet_amd import _lib

    lib = _lib.lib()
    st = lib.vgi_pass_denoise_temporal(None, None, None, None)
    assert st == _lib.VGI_E_INVALID
    assert b"null" in lib.vgi_last_error()
    io = _lib.DenoiseSpatialIO(0, 10, 0)
    assert lib.vgi_pass_denoise_disocclusion_fix(C.byref(io), None) == _lib.VGI_E_INVALID
    assert b"bad size" in lib.vgi_last_error()
    io = _lib.DenoiseSpatialIO(16, 16, 7)
    assert lib.vgi_pass_denoise_spatial_fallback(C.byref(io), None) == _lib.VGI_E_INVALID
    assert b"storage" in lib.vgi_last_error()
    io = _lib.DenoiseSpatialIO(16, 16, 0)  # all bindings NULL
    assert lib.vgi_pass_denoise_spatial_fallback(C.byref(io), None) == _lib.VGI_E_INVALID
    assert b"missing binding" in lib.vgi_last_error()
    out = C.c_void_p()
    assert lib.vgi_denoiser_create(0, 0, 0, None, C.byref(out)) == _lib.VGI_E_INVALID
    assert out.value is None
