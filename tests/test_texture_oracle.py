"""Pins of the oracle's material-texture sampling (CPU only).

violet samples material textures through sampler_linear_wrap
(render_loop.rs:199-212): bilinear filtering, linear between mips, REPEAT
addressing, anisotropy 16; indirect hits at LOD 2
(geometry_ind.rchit.hlsl:3, geometry_rchit.inc.hlsl:79-81), the G-buffer
with implicit derivatives (mesh_gbuffer.hlsl:79-81).  The checks below come
from closed forms, not from the oracle's own formulas:

* decoding: UNORM bytes are k/255; sRGB bytes follow the sRGB EOTF at known
  points (0, 10, 128, 255);
* bilinear: a texture linear in x returns the texel-space coordinate
  u*w - 1/2 between texel centres, and wraps across the edge (REPEAT) as
  the periodic extension of the ramp;
* mips: a texture whose levels are constants c_l returns
  c_floor(lod) + frac(lod) * (c_ceil - c_floor), the base level for
  negative LODs and the last level beyond it;
* implicit LOD: isotropic derivatives of k texels give log2(k); a
  footprint elongated 8:1 samples the base level (N = 8 taps), and the tap
  average of a linear ramp is the ramp at the centre; beyond 16:1 the LOD
  rises by log2(ratio / 16);
* G-buffer: a camera looking straight down on a textured plane (an affine
  screen-to-plane map) has a texel footprint of 2 D tan(fov/2) / width *
  scale * size, so the encoded base colour equals the mip blend at that
  LOD; tilted by 45 degrees the footprint grows by sqrt(2) along the view's
  y only, and anisotropic filtering keeps the LOD of the short axis;
* normal mapping: on the flat plane with tangent +x and sign w, the shading
  normal is normalize(2 * texel - 1) in the (x, w*y, z) frame; metallic and
  roughness come from the R and G channels times the factors.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

BITS = (32, 64)


def _plane_mesh(texture_levels, mat_tex, uv_scale=1.0, tangent_w=1.0, srgb=False, extent=50.0, metallic=0.0,
                roughness=1.0, albedo=(1.0, 1.0, 1.0)):
    from violet_amd import scene

    L = extent
    P = np.array([(-L, -L, 0), (L, -L, 0), (L, L, 0), (-L, L, 0)], np.float32)
    N = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    I = np.array([(0, 1, 2), (0, 2, 3)], np.uint32)
    uv = (P[:, :2] * uv_scale).astype(np.float32)
    tg = np.tile(np.array([1, 0, 0, tangent_w], np.float32), (4, 1))
    tex = [scene.Texture([np.ascontiguousarray(l, np.uint8) for l in texture_levels], srgb=srgb)]
    return scene.Mesh(P, N, I, np.zeros(2, np.uint32), np.array([albedo], np.float64), np.array([metallic]),
                      np.array([roughness]), texcoords=uv, tangents=tg, textures=tex,
                      material_textures=np.array([mat_tex], np.uint32))


def _levels_const(w, h, values):
    """Mip chain of a w x h texture whose level l is the grey byte values[l]."""
    out = []
    for l, v in enumerate(values):
        out.append(np.full((max(1, h >> l), max(1, w >> l), 4), v, np.uint8))
    return out


def _oscene(bits, mesh):
    import oracle

    return oracle.OracleScene(bits, mesh)


@pytest.mark.parametrize("bits", BITS)
def test_decoding_unorm_and_srgb(bits):
    for srgb, expect in ((False, {0: 0.0, 10: 10 / 255, 128: 128 / 255, 255: 1.0}),
                         (True, {0: 0.0, 10: 10 / 255 / 12.92, 128: 0.2158605, 255: 1.0})):
        for byte, val in expect.items():
            lv = [np.full((4, 4, 4), byte, np.uint8)]
            lv[0][..., 3] = 51  # alpha is linear even in an sRGB texture
            s = _oscene(bits, _plane_mesh(lv, (2, 0, 1), srgb=srgb))
            out = s.tex_sample(0, 0.3, 0.7, lod=0.0)
            assert np.allclose(out[:3], val, rtol=2e-6, atol=1e-7), (srgb, byte, out)
            assert abs(out[3] - 0.2) < 1e-7


@pytest.mark.parametrize("bits", BITS)
def test_bilinear_ramp_and_repeat(bits):
    w = 256
    lv = np.zeros((2, w, 4), np.uint8)
    lv[..., 0] = np.arange(w, dtype=np.uint8)[None, :]
    s = _oscene(bits, _plane_mesh([lv], (2, 0, 1)))
    # float: u * w carries |u * w| * 2^-24 of rounding (coordinates up to 1000)
    tol = 3e-4 if bits == 32 else 1e-12
    for fu in (0.0, 0.25, 17.5, 100.125, 254.75):  # texel-space coordinate between centres
        u = (fu + 0.5) / w
        assert abs(s.tex_sample(0, u, 0.5)[0] - fu / 255) < tol
    # across the edge: between texel 255 and texel 0 of the next period
    for a in (0.25, 0.5, 0.9):
        u = (w - 0.5 + a) / w
        assert abs(s.tex_sample(0, u, 0.5)[0] - (1 - a) * 255 / 255) < tol
        assert abs(s.tex_sample(0, u - 3.0, 0.5)[0] - (1 - a)) < tol  # three periods to the left
        assert abs(s.tex_sample(0, u + 2.0, 0.5)[0] - (1 - a)) < tol


@pytest.mark.parametrize("bits", BITS)
def test_mip_levels_blend_and_clamp(bits):
    vals = [0, 40, 100, 160, 220, 250, 20]
    s = _oscene(bits, _plane_mesh(_levels_const(64, 64, vals), (2, 0, 1)))
    tol = 2e-6 if bits == 32 else 1e-12
    for lod in (0.0, 0.5, 1.0, 2.0, 2.25, 3.7, 5.999):
        lo, fr = int(math.floor(lod)), lod - math.floor(lod)
        want = (vals[lo] + fr * (vals[lo + 1] - vals[lo])) / 255
        assert abs(s.tex_sample(0, 0.37, 0.81, lod=lod)[0] - want) < tol, lod
    assert abs(s.tex_sample(0, 0.1, 0.1, lod=-3.0)[0] - vals[0] / 255) < tol
    assert abs(s.tex_sample(0, 0.1, 0.1, lod=6.0)[0] - vals[6] / 255) < tol
    assert abs(s.tex_sample(0, 0.1, 0.1, lod=40.0)[0] - vals[6] / 255) < tol


@pytest.mark.parametrize("bits", BITS)
def test_implicit_lod_isotropic_and_anisotropic(bits):
    vals = [0, 40, 100, 160, 220, 250, 20]
    W = 64
    s = _oscene(bits, _plane_mesh(_levels_const(W, W, vals), (2, 0, 1)))
    tol = 2e-5 if bits == 32 else 1e-9

    def blend(lod):
        lo, fr = int(math.floor(lod)), lod - math.floor(lod)
        return (vals[lo] + fr * (vals[lo + 1] - vals[lo])) / 255

    for k in (1.0, 2.0, 3.0, 5.5, 12.0):
        g = (k / W, 0.0, 0.0, -k / W)  # rho_x = rho_y = k texels
        assert abs(s.tex_sample(0, 0.4, 0.6, grad=g)[0] - blend(math.log2(k))) < tol, k
        g = (0.0, k / W, k / W, 0.0)  # rotated: the same footprint
        assert abs(s.tex_sample(0, 0.4, 0.6, grad=g)[0] - blend(math.log2(k))) < tol, k
    # 8:1 footprint: LOD of the short axis (0 here)
    assert abs(s.tex_sample(0, 0.4, 0.6, grad=(8.0 / W, 0.0, 0.0, 1.0 / W))[0] - vals[0] / 255) < tol
    # 64:1 is clamped to 16 taps: lod = log2(64 / 16) = 2
    assert abs(s.tex_sample(0, 0.4, 0.6, grad=(64.0 / W, 0.0, 0.0, 1.0 / W))[0] - blend(2.0)) < tol
    # 24:2 -> 12 taps at lod log2(2) = 1
    assert abs(s.tex_sample(0, 0.4, 0.6, grad=(0.0, 24.0 / W, 2.0 / W, 0.0))[0] - blend(1.0)) < tol
    # zero derivatives: base level
    assert abs(s.tex_sample(0, 0.4, 0.6, grad=(0.0, 0.0, 0.0, 0.0))[0] - vals[0] / 255) < tol


@pytest.mark.parametrize("bits", BITS)
def test_anisotropic_taps_average_a_ramp_to_its_centre(bits):
    w = 256
    lv = np.zeros((4, w, 4), np.uint8)
    lv[..., 0] = np.arange(w, dtype=np.uint8)[None, :]
    s = _oscene(bits, _plane_mesh([lv], (2, 0, 1)))
    tol = 2e-6 if bits == 32 else 1e-9
    for n_taps, ratio in ((2, 1.5), (5, 4.2), (8, 8.0), (16, 16.0)):
        fu = 128.0
        g = (ratio / w, 0.0, 0.0, 1.0 / 4)  # major axis x, ratio texels; minor = 1 texel of the 4-row image
        out = s.tex_sample(0, (fu + 0.5) / w, 0.5, grad=g)[0]
        assert abs(out - fu / 255) < tol, (n_taps, out)


def _top_down_view(w, h, height, pitch_deg=-90.0, yaw_deg=0.0):
    from violet_amd import synth

    cam = synth.Camera(np.array([0.0, 0.0, height]), yaw_deg, pitch_deg)
    return synth.make_frame_params(synth.View(cam, w, h))


def _decode_color(gb):
    return np.stack([(gb[0] >> (8 * k)) & 0xFF for k in range(4)], -1)


@pytest.mark.parametrize("bits", BITS)
def test_gbuffer_lod_top_down(bits):
    import oracle

    w = h = 64
    vals = [0, 40, 100, 160, 220, 250, 20]
    for lod in (0.5, 1.5, 2.25, 3.7):
        # footprint rho = 2 D tan(45 deg) / w * scale * 64 texels
        D = 2.0 ** lod * w / (2.0 * 64.0)
        m = _plane_mesh(_levels_const(64, 64, vals), (2, 0, 1))
        o = oracle.OracleRestir(bits, oracle.OracleScene(bits, m))
        depth, gb = o.gbuffer(_top_down_view(w, h, D), w, h)
        assert (depth > 0).all()
        col = _decode_color(gb)[..., 0].astype(np.float64)
        lo, fr = int(math.floor(lod)), lod - math.floor(lod)
        want = (vals[lo] + fr * (vals[lo + 1] - vals[lo]))  # G-buffer colour: trunc(c * 255)
        assert np.abs(col - want).max() <= 1.0, (lod, np.unique(col), want)


@pytest.mark.parametrize("bits", BITS)
def test_gbuffer_anisotropy_keeps_short_axis_lod(bits):
    import oracle

    w = h = 64
    vals = [0, 40, 100, 160, 220, 250, 20]
    D = 4.0
    m = _plane_mesh(_levels_const(64, 64, vals), (2, 0, 1))
    o = oracle.OracleRestir(bits, oracle.OracleScene(bits, m))
    # pitch -45: at the image centre the ray meets the ground at 45 degrees;
    # the footprint is R * 2/w across (R = D sqrt 2) and twice D * 2/w along
    depth, gb = o.gbuffer(_top_down_view(w, h, D, pitch_deg=-45.0), w, h)
    col = _decode_color(gb)[..., 0].astype(np.float64)
    rho_x = D * math.sqrt(2.0) * 2.0 / w * 64.0
    lod = math.log2(rho_x)  # eta = sqrt(2) -> lambda = log2(rho_max / eta) = log2(rho_x)
    lo, fr = int(math.floor(lod)), lod - math.floor(lod)
    want = vals[lo] + fr * (vals[lo + 1] - vals[lo])
    c = col[h // 2 - 1:h // 2 + 1, w // 2 - 1:w // 2 + 1]  # the four pixels around the centre
    assert np.abs(c - want).max() <= 3.0, (c, want)
    iso = math.log2(rho_x * math.sqrt(2.0))
    lo, fr = int(math.floor(iso)), iso - math.floor(iso)
    assert abs(want - (vals[lo] + fr * (vals[lo + 1] - vals[lo]))) > 15  # isotropic filtering would differ


@pytest.mark.parametrize("bits", BITS)
@pytest.mark.parametrize("tw", (1.0, -1.0))
def test_gbuffer_normal_map_and_metal_rough(bits, tw):
    import oracle

    w = h = 16
    texel = np.array([200, 100, 230, 255], np.uint8)
    mr = np.array([255, 90, 0, 255], np.uint8)
    lv_n = [np.tile(texel, (8, 8, 1))]
    lv_mr = [np.tile(mr, (8, 8, 1))]
    from violet_amd import scene

    m = _plane_mesh(lv_n, (0, 3, 2), tangent_w=tw, metallic=0.5, roughness=0.8)
    m.textures.append(scene.Texture(lv_mr))
    o = oracle.OracleRestir(bits, oracle.OracleScene(bits, m))
    depth, gb = o.gbuffer(_top_down_view(w, h, 3.0), w, h)
    l1 = gb[1].reshape(-1)
    f = np.stack([(l1 & 0xFFF) / 4095.0, ((l1 >> 12) & 0xFFF) / 4095.0], 1) * 2 - 1
    nz = 1 - np.abs(f).sum(1)
    t = np.clip(-nz, 0, 1)
    n = np.stack([f[:, 0] + np.where(f[:, 0] >= 0, -t, t), f[:, 1] + np.where(f[:, 1] >= 0, -t, t), nz], 1)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    ts = texel[:3] / 255.0 * 2 - 1
    want = np.array([ts[0], tw * ts[1], ts[2]])
    want /= np.linalg.norm(want)
    assert np.abs(n - want).max() < 2e-3
    col = _decode_color(gb).reshape(-1, 4)
    assert (col[:, 3] == int(1.0 * 0.5 * 255)).all()  # metallic = R * factor
    assert (np.abs((gb[1].reshape(-1) >> 24) - 90 / 255 * 0.8 * 255) <= 1).all()  # roughness = G * factor
