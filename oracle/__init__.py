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


# ---------------------------------------------------------------------------
# ReSTIR passes
# ---------------------------------------------------------------------------
class _TraceRes(C.Structure):
    _fields_ = [("scene", P), ("sky", C.c_float * 3), ("prev_ind_diff", P), ("prev_depth", P), ("pw", C.c_int32),
                ("ph", C.c_int32), ("hg_cells", P), ("queries", P), ("query_counter", P), ("query_threshold", P)]


def _restir_sigs(lib):
    lib.oracle_scene_create.restype = P
    lib.oracle_scene_create.argtypes = [P, P, P, U32, P, P]
    lib.oracle_scene_set_textures.argtypes = [P, P, P, P, P, U32]
    R = C.c_float if lib._bits == 32 else C.c_double
    lib.oracle_tex_sample.argtypes = [P, U32, R, R, R, P, P]
    lib.oracle_scene_destroy.argtypes = [P]
    TR = C.POINTER(_TraceRes)
    lib.oracle_trace_radiance.argtypes = [P, TR, P, U32, I, P]
    lib.oracle_diff_sample_gen.argtypes = [P, TR, I, I, P, P, U32, I, P, P]
    lib.oracle_diff_sample_validate.argtypes = [P, TR, I, I, P, P, P, P]
    lib.oracle_spec_sample_gen.argtypes = [P, TR, I, I, P, P, U32, I, P, P]
    lib.oracle_spec_sample_validate.argtypes = [P, TR, I, I, P, P, P]
    lib.oracle_ao_gen.argtypes = [P, I, I, P, P, P, I, I, C.c_double if lib._bits == 64 else C.c_float, P]
    lib.oracle_ao_filter.argtypes = [I, I, P, P, P]
    R = C.c_double if lib._bits == 64 else C.c_float
    lib.oracle_diff_temporal.argtypes = [P, I, I, P, P, P, P, P, P, P, U32, I, I, P, P, P, U32]
    lib.oracle_diff_spatial.argtypes = [P, I, I, P, P, P, P, P, P, U32, P, P, P, P, I, R]
    lib.oracle_spec_temporal.argtypes = [P, I, I, P, P, P, P, P, P, P, U32, I, P, P, P, U32]
    lib.oracle_spec_spatial.argtypes = [P, I, I, P, P, P, P, P, U32, P, P, R]
    lib.oracle_gbuffer.argtypes = [P, P, I, I, P, P]
    lib.oracle_hg_decay.argtypes = [P, P]
    lib.oracle_hg_select.argtypes = [P, U32, P, P, U32]
    lib.oracle_hg_compact.restype = U32
    lib.oracle_hg_compact.argtypes = [P, P, P]
    lib.oracle_hg_select_ordered.argtypes = [P, U32, P, P, P, P, U32]
    lib.oracle_hg_query_threshold_for.restype = U32
    lib.oracle_hg_query_threshold_for.argtypes = [C.c_uint64]
    lib.oracle_hg_next_query_threshold.restype = U32
    lib.oracle_hg_next_query_threshold.argtypes = [U32, U32]
    lib.oracle_hg_compact_ordered.restype = U32
    lib.oracle_hg_compact_ordered.argtypes = [P, P, P, P]
    lib.oracle_hg_raygen.argtypes = [P, TR, P, P, U32, P, P, P, U32]
    lib.oracle_raytraced_shadow.argtypes = [P, P, I, I, P, P]
    lib.oracle_final_lighting.argtypes = [P, I, I, P, P, P, P, P, P]
    lib.oracle_temporal_aa.argtypes = [P, I, I, P, P, P, I, P]


def _fp_arr(fp):
    return fp if isinstance(fp, np.ndarray) else np.frombuffer(bytes(fp), dtype=np.float32).copy()


class OracleScene:
    """The oracle's view of a scene.Mesh (own median-split BVH)."""

    def __init__(self, bits: int, mesh):
        self.lib = load(bits)
        if not hasattr(self.lib, "_restir"):
            self.lib._bits = bits
            _restir_sigs(self.lib)
            self.lib._restir = True
        self.P = np.ascontiguousarray(mesh.positions, np.float32)
        self.N = np.ascontiguousarray(mesh.normals, np.float32)
        self.I = np.ascontiguousarray(mesh.indices, np.uint32)
        self.TM = np.ascontiguousarray(mesh.tri_material, np.uint32)
        mats = np.zeros((len(mesh.albedo), 8), np.float32)
        mats[:, :3] = mesh.albedo
        mats[:, 3] = 1.0
        mats[:, 4] = mesh.metallic
        mats[:, 5] = mesh.roughness
        # texture indices (base colour, metallic-roughness, normal map) packed
        # as MaterialParams packs them (scene_bindings.hlsl:24-51); absent:
        # white, white, flat normal
        tex = getattr(mesh, "material_textures", None)
        tex = np.tile(np.array([0, 0, 1], np.uint32), (len(mats), 1)) if tex is None else np.asarray(tex, np.uint32)
        mats.view(np.uint32)[:, 6] = tex[:, 0] | (tex[:, 1] << 16)
        mats.view(np.uint32)[:, 7] = tex[:, 2]
        self.mats = mats
        self.h = self.lib.oracle_scene_create(_np_ptr(self.P), _np_ptr(self.N), _np_ptr(self.I), len(self.I),
                                              _np_ptr(self.TM), _np_ptr(self.mats))
        # texturing: texel words of every level back to back, 20 words of
        # (width, height, levels, srgb, level offsets) per texture
        self.UV = None if getattr(mesh, 'texcoords', None) is None else np.ascontiguousarray(mesh.texcoords, np.float32)
        self.TG = None if getattr(mesh, 'tangents', None) is None else np.ascontiguousarray(mesh.tangents, np.float32)
        meta = np.zeros((max(1, len(mesh.textures)), 20), np.uint32)
        words, off = [], 0
        textures = getattr(mesh, "textures", [])
        for k, t in enumerate(textures):
            meta[k, :4] = (t.width, t.height, len(t.levels), int(bool(t.srgb)))
            for l, lv in enumerate(t.levels):
                meta[k, 4 + l] = off
                w = np.ascontiguousarray(lv, np.uint8).reshape(-1).view(np.uint32)
                words.append(w)
                off += len(w)
        self.texels = np.concatenate(words) if words else np.zeros(1, np.uint32)
        self.meta = meta
        if self.UV is not None or self.TG is not None or textures:
            self.lib.oracle_scene_set_textures(self.h, _np_ptr(self.UV) if self.UV is not None else None,
                                               _np_ptr(self.TG) if self.TG is not None else None,
                                               _np_ptr(self.texels), _np_ptr(self.meta), len(textures))

    def tex_sample(self, t: int, u: float, v: float, lod: float = 0.0, grad=None) -> np.ndarray:
        """Texture t (0-based) sampled at (u, v): explicit LOD, or implicit
        from grad = (ds/dx, dt/dx, ds/dy, dt/dy)."""
        dt = np.float32 if self.lib._bits == 32 else np.float64
        out = np.zeros(4, dt)
        g = None if grad is None else np.ascontiguousarray(grad, dt)
        self.lib.oracle_tex_sample(self.h, t, u, v, lod, _np_ptr(g), _np_ptr(out))
        return out

    def __del__(self):
        try:
            self.lib.oracle_scene_destroy(self.h)
        except Exception:
            pass


class OracleRestir:
    """Typed wrappers of the ReSTIR passes for one precision build."""

    NQ = 128 * 1024

    def __init__(self, bits: int, scene: OracleScene, sky=(1.0, 1.0, 1.0)):
        self.bits = bits
        self.lib = load(bits)
        if not hasattr(self.lib, "_restir"):
            self.lib._bits = bits
            _restir_sigs(self.lib)
            self.lib._restir = True
        self.rt_dtype = real_dtype(bits)
        self.scene = scene
        self.sky = sky

    def r(self, a):
        # Texels a pass leaves unwritten (pixels without geometry in the
        # specular hit-position texture, as in ind_spec_sample_gen.hlsl:37-42)
        # may hold any bits, signalling NaNs included, whose widening to f64
        # raises the invalid flag; no pass reads those texels.
        if a is None:
            return None
        with np.errstate(invalid="ignore"):
            return np.ascontiguousarray(a, dtype=self.rt_dtype)

    def trace_res(self, hg_cells, queries, counter, prev_ind_diff=None, prev_depth=None, threshold=None):
        tr = _TraceRes()
        tr.scene = self.scene.h
        tr.sky[:] = list(self.sky)
        self._keep = [self.r(prev_ind_diff), self.r(prev_depth)]
        tr.prev_ind_diff = _np_ptr(self._keep[0])
        tr.prev_depth = _np_ptr(self._keep[1])
        if prev_depth is not None:
            tr.ph, tr.pw = prev_depth.shape
        tr.hg_cells = _np_ptr(hg_cells)
        tr.queries = _np_ptr(queries)
        tr.query_counter = _np_ptr(counter)
        tr.query_threshold = None if threshold is None else _np_ptr(threshold)
        return tr

    def trace_radiance(self, fp, tr, rays, has_prev):
        """raytrace.inc.hlsl:65-208 for a batch of rays (n, 8) = {o, tmin, d,
        tmax}: returns (n, 9) = hit position, shading normal, radiance."""
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((len(rays), 9), self.rt_dtype)
        self.lib.oracle_trace_radiance(_np_ptr(_fp_arr(fp)), C.byref(tr), _np_ptr(rays), len(rays), int(has_prev),
                                       _np_ptr(out))
        return out

    def gbuffer(self, fp, w, h):
        depth = np.zeros((h, w), self.rt_dtype)
        gb = np.zeros((4, h, w), np.uint32)
        self.lib.oracle_gbuffer(_np_ptr(_fp_arr(fp)), self.scene.h, w, h, _np_ptr(depth), _np_ptr(gb))
        return depth, gb

    def raytraced_shadow(self, fp, depth):
        h, w = depth.shape
        mask = np.zeros((h, w), np.uint8)
        self.lib.oracle_raytraced_shadow(_np_ptr(_fp_arr(fp)), self.scene.h, w, h, _np_ptr(self.r(depth)), _np_ptr(mask))
        return mask

    def final_lighting(self, fp, gb, mask, ind_diff, ind_spec):
        h, w = mask.shape
        color = np.zeros((h, w, 3), self.rt_dtype)
        sky = np.asarray(self.sky, np.float32)
        self.lib.oracle_final_lighting(_np_ptr(_fp_arr(fp)), w, h, _np_ptr(np.ascontiguousarray(gb, np.uint32)),
                                       _np_ptr(np.ascontiguousarray(mask, np.uint8)), _np_ptr(self.r(ind_diff)),
                                       _np_ptr(self.r(ind_spec)), _np_ptr(sky), _np_ptr(color))
        return color

    def temporal_aa(self, fp, depth, source, history=None):
        h, w = depth.shape
        out = np.zeros((h, w, 3), self.rt_dtype)
        hist = self.r(history)
        self.lib.oracle_temporal_aa(_np_ptr(_fp_arr(fp)), w, h, _np_ptr(self.r(depth)), _np_ptr(self.r(source)),
                                    _np_ptr(hist), int(history is not None), _np_ptr(out))
        return out

    def diff_sample_gen(self, fp, tr, gb, depth, frame_rand, has_prev):
        h, w = depth.shape
        hit = np.zeros((h, w, 2), np.uint32)
        rad = np.zeros((h, w, 3), self.rt_dtype)
        self.lib.oracle_diff_sample_gen(_np_ptr(_fp_arr(fp)), C.byref(tr), w, h, _np_ptr(np.ascontiguousarray(gb, np.uint32)),
                                        _np_ptr(self.r(depth)), frame_rand, int(has_prev), _np_ptr(hit), _np_ptr(rad))
        return hit, rad

    def diff_sample_validate(self, fp, tr, prev_depth, rsv, hit, rad):
        h, w = prev_depth.shape
        rsv, hit, rad = rsv.copy(), hit.copy(), self.r(rad).copy()
        self.lib.oracle_diff_sample_validate(_np_ptr(_fp_arr(fp)), C.byref(tr), w, h, _np_ptr(self.r(prev_depth)),
                                             _np_ptr(rsv), _np_ptr(hit), _np_ptr(rad))
        return rsv, hit, rad

    def spec_sample_gen(self, fp, tr, gb, depth, frame_rand, has_prev, hitpos_in=None):
        h, w = depth.shape
        hp = np.zeros((h, w, 4), self.rt_dtype) if hitpos_in is None else self.r(hitpos_in).copy()
        rad = np.zeros((h, w, 3), self.rt_dtype)
        self.lib.oracle_spec_sample_gen(_np_ptr(_fp_arr(fp)), C.byref(tr), w, h, _np_ptr(np.ascontiguousarray(gb, np.uint32)),
                                        _np_ptr(self.r(depth)), frame_rand, int(has_prev), _np_ptr(hp), _np_ptr(rad))
        return hp, rad

    def spec_sample_validate(self, fp, tr, rsv, hitpos, rad):
        h, w = rad.shape[:2]
        rsv, hp, rd = rsv.copy(), self.r(hitpos).copy(), self.r(rad).copy()
        self.lib.oracle_spec_sample_validate(_np_ptr(_fp_arr(fp)), C.byref(tr), w, h, _np_ptr(rsv), _np_ptr(hp), _np_ptr(rd))
        return rsv, hp, rd

    def ao_gen(self, fp, hit, depth, prev_ao, has_new, has_prev, radius=2.0):
        h, w = depth.shape
        out = np.zeros((h, w, 2), self.rt_dtype)
        self.lib.oracle_ao_gen(_np_ptr(_fp_arr(fp)), w, h, _np_ptr(hit), _np_ptr(self.r(depth)), _np_ptr(self.r(prev_ao)),
                               int(has_new), int(has_prev), radius, _np_ptr(out))
        return out

    def ao_filter(self, depth, ao):
        h, w = depth.shape
        out = np.zeros((h, w, 2), self.rt_dtype)
        self.lib.oracle_ao_filter(w, h, _np_ptr(self.r(depth)), _np_ptr(self.r(ao)), _np_ptr(out))
        return out

    def diff_temporal(self, fp, prev_depth, depth, new_hit, new_rad, prev_rsv, prev_hit, prev_rad, frame_rand, has_prev,
                      has_new, m_cap=0):
        h, w = depth.shape
        rsv = np.zeros((h, w), np.uint32)
        hit = np.zeros((h, w, 2), np.uint32)
        rad = np.zeros((h, w, 3), self.rt_dtype)
        a = lambda x: None if x is None else np.ascontiguousarray(x, np.uint32)  # noqa: E731
        self.lib.oracle_diff_temporal(_np_ptr(_fp_arr(fp)), w, h, _np_ptr(self.r(prev_depth)), _np_ptr(self.r(depth)),
                                      _np_ptr(a(new_hit)), _np_ptr(self.r(new_rad)), _np_ptr(a(prev_rsv)),
                                      _np_ptr(a(prev_hit)), _np_ptr(self.r(prev_rad)), frame_rand, int(has_prev),
                                      int(has_new), _np_ptr(rsv), _np_ptr(hit), _np_ptr(rad), int(m_cap))
        return rsv, hit, rad

    def diff_spatial(self, fp, gb, depth, ao, rsv, hit, rad, frame_rand, num_taps=0, radius_px=0.0):
        h, w = depth.shape
        o_rsv = np.zeros((h, w), np.uint32)
        o_hit = np.zeros((h, w, 2), np.uint32)
        o_rad = np.zeros((h, w, 3), self.rt_dtype)
        o_l = np.zeros((h, w, 3), self.rt_dtype)
        self.lib.oracle_diff_spatial(_np_ptr(_fp_arr(fp)), w, h, _np_ptr(np.ascontiguousarray(gb, np.uint32)),
                                     _np_ptr(self.r(depth)), _np_ptr(self.r(ao)), _np_ptr(np.ascontiguousarray(rsv, np.uint32)),
                                     _np_ptr(np.ascontiguousarray(hit, np.uint32)), _np_ptr(self.r(rad)), frame_rand,
                                     _np_ptr(o_rsv), _np_ptr(o_hit), _np_ptr(o_rad), _np_ptr(o_l), int(num_taps),
                                     float(radius_px))
        return o_rsv, o_hit, o_rad, o_l

    def spec_temporal(self, fp, prev_gb, prev_depth, gb, depth, prev_rsv, prev_hp, prev_rad, new_hp, new_rad, frame_rand,
                      has_prev, m_cap=0):
        h, w = depth.shape
        rsv = np.zeros((h, w, 2), np.uint32)
        hp = self.r(new_hp).copy()
        rad = self.r(new_rad).copy()
        a = lambda x: None if x is None else np.ascontiguousarray(x, np.uint32)  # noqa: E731
        self.lib.oracle_spec_temporal(_np_ptr(_fp_arr(fp)), w, h, _np_ptr(a(prev_gb)), _np_ptr(self.r(prev_depth)),
                                      _np_ptr(a(gb)), _np_ptr(self.r(depth)), _np_ptr(a(prev_rsv)), _np_ptr(self.r(prev_hp)),
                                      _np_ptr(self.r(prev_rad)), frame_rand, int(has_prev), _np_ptr(rsv), _np_ptr(hp),
                                      _np_ptr(rad), int(m_cap))
        return rsv, hp, rad

    def spec_spatial(self, fp, gb, depth, rsv, hp, rad, frame_rand, radius_px=0.0):
        h, w = depth.shape
        light = np.zeros((h, w, 3), self.rt_dtype)
        rl = np.zeros((h, w), self.rt_dtype)
        self.lib.oracle_spec_spatial(_np_ptr(_fp_arr(fp)), w, h, _np_ptr(np.ascontiguousarray(gb, np.uint32)),
                                     _np_ptr(self.r(depth)), _np_ptr(np.ascontiguousarray(rsv, np.uint32)),
                                     _np_ptr(self.r(hp)), _np_ptr(self.r(rad)), frame_rand, _np_ptr(light), _np_ptr(rl),
                                     float(radius_px))
        return light, rl

    # hash grid
    def hg_decay(self, decay, cells):
        decay, cells = decay.copy(), cells.copy()
        self.lib.oracle_hg_decay(_np_ptr(decay), _np_ptr(cells))
        return decay, cells

    def hg_select(self, queries, counter, cells, frame_hash):
        cells = cells.copy()
        sel = np.zeros(65536 * 16, np.uint32)
        self.lib.oracle_hg_select(_np_ptr(np.ascontiguousarray(queries, np.uint32)), int(counter), _np_ptr(cells),
                                  _np_ptr(sel), frame_hash)
        return cells, sel

    def hg_compact(self, sel):
        idx = np.zeros(self.NQ, np.uint32)
        addr = np.zeros(self.NQ, np.uint32)
        n = self.lib.oracle_hg_compact(_np_ptr(np.ascontiguousarray(sel, np.uint32)), _np_ptr(idx), _np_ptr(addr))
        return idx[:n], addr[:n]

    def hg_select_ordered(self, queries, counter, cells, frame_hash, threshold=None):
        """Order-independent select: (cells, selection words, winner indices);
        `threshold` (a 1-word array) is updated in place when given."""
        cells = cells.copy()
        sel = np.zeros(65536 * 16, np.uint32)
        win = np.zeros(65536 * 16, np.uint32)
        self.lib.oracle_hg_select_ordered(_np_ptr(np.ascontiguousarray(queries, np.uint32)), int(counter),
                                          _np_ptr(cells), _np_ptr(sel), _np_ptr(win),
                                          None if threshold is None else _np_ptr(threshold), frame_hash)
        return cells, sel, win

    def hg_query_threshold_for(self, expected: int) -> int:
        return int(self.lib.oracle_hg_query_threshold_for(int(expected)))

    def hg_next_query_threshold(self, kept: int, thr: int) -> int:
        return int(self.lib.oracle_hg_next_query_threshold(int(kept), int(thr)))

    def hg_compact_ordered(self, sel, win):
        idx = np.zeros(self.NQ, np.uint32)
        addr = np.zeros(self.NQ, np.uint32)
        n = self.lib.oracle_hg_compact_ordered(_np_ptr(np.ascontiguousarray(sel, np.uint32)),
                                               _np_ptr(np.ascontiguousarray(win, np.uint32)), _np_ptr(idx),
                                               _np_ptr(addr))
        return idx[:n], addr[:n]

    def hg_raygen(self, fp, tr, cells, queries, cidx, caddr, decay, frame_hash):
        cells, decay = cells.copy(), decay.copy()
        cidx = np.ascontiguousarray(cidx, np.uint32)
        caddr = np.ascontiguousarray(caddr, np.uint32)
        self.lib.oracle_hg_raygen(_np_ptr(_fp_arr(fp)), C.byref(tr), _np_ptr(cells), _np_ptr(np.ascontiguousarray(queries, np.uint32)),
                                  len(cidx), _np_ptr(cidx), _np_ptr(caddr), _np_ptr(decay), frame_hash)
        return cells, decay


def frame_rand_seeds(frame_index: int, n: int) -> list[int]:
    """FrameRand (restir_renderer.rs:130-143): the first n values of the jenkins chain."""
    out, s = [], frame_index & 0xFFFFFFFF
    for _ in range(n):
        s = int(load(64).oracle_jenkins(s))
        out.append(s)
    return out


class OracleRenderer:
    """Mirror of restir_renderer.rs `RestirRenderer::add_passes` (:258-1028)
    over the oracle passes (same pass order, FrameRand seed chain, history
    flags and feedback as the GPU renderer)."""

    NC = 65536 * 16

    def __init__(self, bits: int, mesh, width: int, height: int, sky=(1.0, 1.0, 1.0), denoise=True,
                 ind_diff_validate=True, ind_spec_validate=True, hash_grid_cache_decay=True, ao_radius=2.0,
                 atrous_iterations=3, temporal_m_cap=0, diff_spatial_taps=0, spatial_radius_px=0.0,
                 ordered_cache_update=True):
        self.bits = bits
        self.scene = OracleScene(bits, mesh)
        self.R = OracleRestir(bits, self.scene, sky)
        self.den = OracleDenoiser(bits, atrous_iterations=atrous_iterations)
        self.w, self.h = width, height
        self.denoise, self.ind_diff_validate, self.ind_spec_validate = denoise, ind_diff_validate, ind_spec_validate
        self.decay_on, self.ao_radius = hash_grid_cache_decay, ao_radius
        self.m_cap, self.taps, self.radius_px = temporal_m_cap, diff_spatial_taps, spatial_radius_px
        self.ordered = ordered_cache_update
        self.cells = None
        self.prev = None  # dict of last frame's state

    def add_passes(self, fp, frame_index: int):
        R, w, h = self.R, self.w, self.h
        seeds = iter(frame_rand_seeds(frame_index, 8))
        depth, gb = R.gbuffer(fp, w, h)
        if self.cells is None:
            self.cells = np.zeros((self.NC, 4), np.uint32)
            self.decay = np.zeros(self.NC, np.uint8)
            self.q = np.zeros((131072, 6), np.uint32)
            self.cnt = np.zeros(1, np.uint32)
            # ordered update: the query threshold word before the first count
            self.thr = (np.array([R.hg_query_threshold_for(2 * w * h)], np.uint32) if self.ordered
                        else None)
        if self.decay_on:
            self.decay, self.cells = R.hg_decay(self.decay, self.cells)
        p = self.prev
        has_prev = p is not None
        hg = False
        if has_prev:
            if self.ordered:
                self.cells, sel, win = R.hg_select_ordered(self.q, int(self.cnt[0]), self.cells, next(seeds),
                                                           self.thr)
                cidx, caddr = R.hg_compact_ordered(sel, win)
                lookup = self.cells.copy()  # the cache rays read the cache from before the pass
            else:
                self.cells, sel = R.hg_select(self.q, int(self.cnt[0]), self.cells, next(seeds))
                cidx, caddr = R.hg_compact(sel)
                lookup = self.cells
            nq = np.zeros((131072, 6), np.uint32)
            ncnt = np.zeros(1, np.uint32)
            tr = R.trace_res(lookup, nq, ncnt, prev_ind_diff=p["out_diffuse"], prev_depth=p["depth"],
                             threshold=self.thr)
            self.cells, self.decay = R.hg_raygen(fp, tr, self.cells, self.q, cidx, caddr, self.decay, next(seeds))
            self.q, self.cnt = nq, ncnt
            hg = True
        validation = has_prev and (frame_index & 3) == 0 and self.ind_diff_validate
        tr = R.trace_res(self.cells, self.q, self.cnt, prev_ind_diff=p["out_diffuse"] if has_prev else None,
                         prev_depth=p["depth"] if has_prev else None, threshold=self.thr)
        if not validation:
            hit, rad = R.diff_sample_gen(fp, tr, gb, depth, next(seeds), has_prev)
            prev_rsv = (p["d_rsv"], p["d_hit"], p["d_rad"]) if has_prev else (None, None, None)
        else:
            hit, rad = None, None
            prev_rsv = R.diff_sample_validate(fp, tr, p["depth"], p["d_rsv"], p["d_hit"], p["d_rad"])
        ao = R.ao_gen(fp, hit, depth, p["ao"] if has_prev else None, not validation, has_prev, self.ao_radius)
        aof = R.ao_filter(depth, ao)
        t = R.diff_temporal(fp, p["depth"] if has_prev else None, depth, hit, rad, *prev_rsv, next(seeds), has_prev,
                            not validation, self.m_cap)
        d_rsv, d_hit, d_rad, d_light = R.diff_spatial(fp, gb, depth, aof, *t, next(seeds), self.taps, self.radius_px)
        hp, srad = R.spec_sample_gen(fp, tr, gb, depth, next(seeds), has_prev)
        ps = (p["s_rsv"], p["s_hp"], p["s_rad"]) if has_prev else (None, None, None)
        if has_prev and self.ind_spec_validate:
            ps = R.spec_sample_validate(fp, tr, *ps)
        s_rsv, s_hp, s_rad = R.spec_temporal(fp, p["gb"] if has_prev else None, p["depth"] if has_prev else None, gb,
                                             depth, *ps, hp, srad, next(seeds), has_prev, self.m_cap)
        s_light, ray_len = R.spec_spatial(fp, gb, depth, s_rsv, s_hp, s_rad, next(seeds), self.radius_px)
        if self.denoise and has_prev:
            o = self.den.add_passes(fp, depth, gb, p["depth"], p["gb"], d_light, s_light, ray_len)
            out_d, out_s = o["diffuse"], o["specular"]
        else:
            self.den.reset()
            out_d, out_s = d_light, s_light
        self.prev = dict(depth=depth, gb=gb, d_rsv=d_rsv, d_hit=d_hit, d_rad=d_rad, s_rsv=s_rsv, s_hp=s_hp,
                         s_rad=s_rad, ao=ao, out_diffuse=out_d)
        return dict(diffuse=out_d, specular=out_s, depth=depth, gbuffer=gb, hash_grid_updated=hg,
                    validation=validation, diff_light=d_light, spec_light=s_light, diff_rad=d_rad, spec_rad=s_rad)
