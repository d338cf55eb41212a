"""ORACLE — test infrastructure only.

CPU restatement of violet's ReSTIR GI hot path (C, oracle/*.c) compiled as
  liboracle64.so  (REAL = double: the FP64 oracle)
  liboracle32.so  (REAL = float:  the FP32 twin, same expression order as the
                   reference HLSL, for bit-exact integer/selection checks)
plus Python sequencers mirroring violet's host code.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use this package,
and only as the checker / reported CPU baseline; the product never calls it.

The reference itself (Rust + ash/Vulkan + HLSL compiled by DXC) cannot be
built or run here (no cargo/rustc, DXC, Vulkan loader or lavapipe, and the
crates are Win32-only), so there is no oracle/_ref build.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_BUILD = HERE / "_build"

P = C.c_void_p
I = C.c_int
U32 = C.c_uint32

_libs: dict[int, C.CDLL] = {}


def _np_ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def load(bits: int = 64) -> C.CDLL:
    if bits not in _libs:
        path = _BUILD / f"liboracle{bits}.so"
        if not path.exists():
            from . import build as _b

            _b.build(verbose=False)
        lib = C.CDLL(str(path))
        lib.oracle_jenkins.restype = U32
        lib.oracle_jenkins.argtypes = [U32]
        lib.oracle_lcg_init.restype = U32
        lib.oracle_lcg_init.argtypes = [U32, U32, U32]
        lib.oracle_lcg_stream.argtypes = [U32, I, P, P]
        lib.oracle_xxh.restype = U32
        lib.oracle_xxh.argtypes = [P, I]
        lib.oracle_f32_to_f16.restype = U32
        lib.oracle_f32_to_f16.argtypes = [C.c_float]
        lib.oracle_f16_to_f64.restype = C.c_double
        lib.oracle_f16_to_f64.argtypes = [U32]
        for n in ("oracle_oct_dec_u16", "oracle_oct_dec_u32", "oracle_gb_normal"):
            getattr(lib, n).argtypes = [U32, P]
        for n in ("oracle_oct_enc_u16", "oracle_oct_enc_u32"):
            getattr(lib, n).restype = U32
            getattr(lib, n).argtypes = [P]
        lib.oracle_rsv_enc.restype = U32
        lib.oracle_rsv_enc.argtypes = [U32, C.c_float]
        lib.oracle_rsv_dec.argtypes = [U32, P, P]
        real = C.c_double if bits == 64 else C.c_float
        lib.oracle_sincos.argtypes = [real, P, P]
        lib.oracle_real_bytes.restype = I
        lib.oracle_denoise_temporal.argtypes = [P, I, I] + [P] * 13 + [I, I] + [P] * 6
        lib.oracle_denoise_disocclusion.argtypes = [I, I] + [P] * 9
        lib.oracle_denoise_fallback.argtypes = [I, I] + [P] * 9
        lib.oracle_denoise_fast_clamp.argtypes = [I, I] + [P] * 4
        lib.oracle_denoise_atrous.argtypes = [I, I, I] + [P] * 8
        for n in ("oracle_trace_brute", "oracle_trace_bvh"):
            getattr(lib, n).argtypes = [P, P, U32, P, U32, I, P, P]
        lib.oracle_num_threads.restype = I
        _libs[bits] = lib
    return _libs[bits]


def real_dtype(bits: int):
    return np.float64 if bits == 64 else np.float32


class Oracle:
    """Typed wrappers around one precision build of the oracle."""

    def __init__(self, bits: int = 64):
        self.bits = bits
        self.lib = load(bits)
        self.rt = real_dtype(bits)

    def r(self, a):
        return None if a is None else np.ascontiguousarray(a, dtype=self.rt)

    # ---- denoiser passes -------------------------------------------------
    def denoise_temporal(self, fp, depth, gb, diff_src, spec_src, ray_len, prev_depth, prev_gb, diff_hist,
                         spec_hist, mom_hist, len_hist, has_history, fast=False, diff_fast=None, spec_fast=None):
        h, w = depth.shape
        n = h * w
        args = [self.r(depth), np.ascontiguousarray(gb, np.uint32), self.r(diff_src), self.r(spec_src),
                self.r(ray_len), self.r(prev_depth), None if prev_gb is None else np.ascontiguousarray(prev_gb, np.uint32),
                self.r(diff_hist), self.r(spec_hist), self.r(mom_hist),
                None if len_hist is None else np.ascontiguousarray(len_hist, np.uint8),
                self.r(diff_fast), self.r(spec_fast)]
        od = np.zeros((h, w, 3), self.rt)
        osp = np.zeros((h, w, 3), self.rt)
        om = np.zeros((h, w, 2), self.rt)
        ol = np.zeros((h, w), np.uint8)
        odf = np.zeros((h, w, 3), self.rt)
        osf = np.zeros((h, w, 3), self.rt)
        fpa = np.frombuffer(bytes(fp), dtype=np.float32).copy() if not isinstance(fp, np.ndarray) else fp
        self.lib.oracle_denoise_temporal(_np_ptr(fpa), w, h, *[_np_ptr(a) for a in args], int(has_history), int(fast),
                                         _np_ptr(od), _np_ptr(osp), _np_ptr(om), _np_ptr(ol), _np_ptr(odf),
                                         _np_ptr(osf))
        assert n == od.shape[0] * od.shape[1]
        return dict(diff=od, spec=osp, moments=om, history_len=ol, diff_fast=odf, spec_fast=osf)

    def _spatial(self, fn, depth, gb, len_tex, diff, spec, mom):
        h, w = depth.shape
        od = np.zeros((h, w, 3), self.rt)
        osp = np.zeros((h, w, 3), self.rt)
        om = np.zeros((h, w, 2), self.rt)
        fn(w, h, _np_ptr(self.r(depth)), _np_ptr(np.ascontiguousarray(gb, np.uint32)),
           _np_ptr(np.ascontiguousarray(len_tex, np.uint8)), _np_ptr(self.r(diff)), _np_ptr(self.r(spec)),
           _np_ptr(self.r(mom)), _np_ptr(od), _np_ptr(osp), _np_ptr(om))
        return od, osp, om

    def denoise_disocclusion(self, depth, gb, len_tex, diff, spec, mom):
        d, s, m = self._spatial(self.lib.oracle_denoise_disocclusion, depth, gb, len_tex, diff, spec, mom)
        return dict(diff=d, spec=s, moments=m)

    def denoise_fallback(self, depth, gb, len_tex, diff, spec, mom):
        d, s, v = self._spatial(self.lib.oracle_denoise_fallback, depth, gb, len_tex, diff, spec, mom)
        return dict(diff=d, spec=s, variance=v)

    def denoise_fast_clamp(self, diff_fast, spec_fast, diff, spec):
        h, w = diff.shape[:2]
        d = self.r(diff).copy()
        s = self.r(spec).copy()
        self.lib.oracle_denoise_fast_clamp(w, h, _np_ptr(self.r(diff_fast)), _np_ptr(self.r(spec_fast)), _np_ptr(d),
                                           _np_ptr(s))
        return d, s

    def denoise_atrous(self, step, depth, gb, diff, spec, var):
        h, w = depth.shape
        od = np.zeros((h, w, 3), self.rt)
        osp = np.zeros((h, w, 3), self.rt)
        ov = np.zeros((h, w, 2), self.rt)
        self.lib.oracle_denoise_atrous(w, h, int(step), _np_ptr(self.r(depth)), _np_ptr(np.ascontiguousarray(gb, np.uint32)),
                                       _np_ptr(self.r(diff)), _np_ptr(self.r(spec)), _np_ptr(self.r(var)),
                                       _np_ptr(od), _np_ptr(osp), _np_ptr(ov))
        return dict(diff=od, spec=osp, variance=ov)


def trace(bits: int, positions, indices, rays, any_hit: bool = False, brute: bool = True):
    """Closest/any hit of f32 rays (N,8) against the triangle soup.
    Returns (thv (N,3) REAL, ids (N,) u32)."""
    lib = load(bits)
    P_ = np.ascontiguousarray(positions, np.float32)
    I_ = np.ascontiguousarray(indices, np.uint32)
    R_ = np.ascontiguousarray(rays, np.float32)
    n = len(R_)
    thv = np.zeros((n, 3), real_dtype(bits))
    ids = np.zeros(n, np.uint32)
    fn = lib.oracle_trace_brute if brute else lib.oracle_trace_bvh
    fn(_np_ptr(P_), _np_ptr(I_), len(I_), _np_ptr(R_), n, int(any_hit), _np_ptr(thv), _np_ptr(ids))
    return thv, ids


class OracleDenoiser:
    """Mirror of denoising.rs `Denoiser::add_passes` (:66-301) over the oracle passes."""

    def __init__(self, bits: int = 64, disocclusion_fix=True, fast_history_clamp=False, atrous_iterations=3):
        self.o = Oracle(bits)
        self.disocclusion_fix = disocclusion_fix
        self.fast_history_clamp = fast_history_clamp
        self.atrous_iterations = atrous_iterations
        self.history = None
        self.fast = None

    def reset(self):
        self.history = None

    def add_passes(self, fp, depth, gb, prev_depth, prev_gb, diffuse, specular, ray_len):
        iters = max(0, min(5, self.atrous_iterations))
        has_history = self.history is not None and prev_depth is not None
        hist = self.history if has_history else dict(diff=None, spec=None, moments=None, history_len=None)
        fast = self.fast_history_clamp
        h, w = depth.shape
        df_in = sf_in = None
        if fast:
            if self.fast is not None and has_history:
                df_in, sf_in = self.fast
            else:
                df_in = np.zeros((h, w, 3))
                sf_in = np.zeros((h, w, 3))
        t = self.o.denoise_temporal(fp, depth, gb, diffuse, specular, ray_len, prev_depth if has_history else None,
                                    prev_gb if has_history else None, hist["diff"], hist["spec"], hist["moments"],
                                    hist["history_len"], has_history, fast, df_in, sf_in)
        cur = dict(diff=t["diff"], spec=t["spec"], moments=t["moments"])
        if self.disocclusion_fix:
            cur = self.o.denoise_disocclusion(depth, gb, t["history_len"], cur["diff"], cur["spec"], cur["moments"])
        if fast:
            d, s = self.o.denoise_fast_clamp(t["diff_fast"], t["spec_fast"], cur["diff"], cur["spec"])
            cur = dict(diff=d, spec=s, moments=cur["moments"])
            self.fast = (t["diff_fast"], t["spec_fast"])
        else:
            self.fast = None
        self.history = dict(diff=cur["diff"], spec=cur["spec"], moments=cur["moments"], history_len=t["history_len"])
        fb = self.o.denoise_fallback(depth, gb, t["history_len"], cur["diff"], cur["spec"], cur["moments"])
        d, s, v = fb["diff"], fb["spec"], fb["variance"]
        for i in range(iters):
            a = self.o.denoise_atrous(1 << i, depth, gb, d, s, v)
            d, s, v = a["diff"], a["spec"], a["variance"]
        return dict(diffuse=d, specular=s, history_len=t["history_len"], variance=fb["variance"],
                    temporal=t, fallback=fb)
