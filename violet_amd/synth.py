"""Seeded synthetic inputs and view set-up (numpy, host side).

violet draws its inputs from a rasterised glTF scene and an interactive
camera; this module produces the same kinds of inputs from seeds so that the
GPU library and the oracle are fed identical bytes:

* the camera and FrameParams exactly as violet builds them
  (crates/violet-app/src/app.rs:26-55 perspective_projection (reverse Z),
  app.rs:96-133 Camera, app.rs:390-410 view matrix,
  crates/violet-app/src/render_loop.rs:89-158 FrameParams::make, halton);
* the analytic denoiser-only workload (a ground plane plus 20 spheres,
  G-buffer encoded as gbuffer.hlsl:54-77 with fwidth from 2x2 quad
  differences like mesh_gbuffer.hlsl:110-111), with a smooth clean radiance
  field multiplied by Exp(1) noise.

All randomness is counter-based (hash of seed and element index) so that any
element can be regenerated independently.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._lib import FrameParams

F32 = np.float32


# ---------------------------------------------------------------------------
# counter-based RNG
# ---------------------------------------------------------------------------
def hash_u32(seed: int, idx: np.ndarray | int) -> np.ndarray:
    """lowbias32 finaliser of (idx * golden + seed); vectorised uint32."""
    with np.errstate(over="ignore"):
        x = (np.asarray(idx, dtype=np.uint64) * np.uint64(0x9E3779B1) + np.uint64(seed & 0xFFFFFFFF)) & np.uint64(
            0xFFFFFFFF)
        x = x.astype(np.uint32)
        x ^= x >> np.uint32(16)
        x = (x.astype(np.uint64) * np.uint64(0x7FEB352D) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        x ^= x >> np.uint32(15)
        x = (x.astype(np.uint64) * np.uint64(0x846CA68B) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        x ^= x >> np.uint32(16)
    return x


def uniform(seed: int, idx, lo=0.0, hi=1.0) -> np.ndarray:
    u = (hash_u32(seed, idx) >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)
    return lo + (hi - lo) * u


# ---------------------------------------------------------------------------
# camera / FrameParams
# ---------------------------------------------------------------------------
def perspective_projection(near: float, far: float, fov_h: float, aspect: float) -> np.ndarray:
    """app.rs:26-55 (reverse Z: near -> 1, far -> 0).  Returns the 4x4 matrix."""
    ran = math.tan(fov_h / 2.0)
    width = near * ran
    m00 = near / width
    m11 = near * aspect / width
    m22 = -near / (far - near)
    m23 = near / (1.0 - near / far)
    return np.array([[m00, 0, 0, 0], [0, m11, 0, 0], [0, 0, m22, m23], [0, 0, 1, 0]], dtype=np.float64)


@dataclass
class Camera:
    """app.rs:96-133: Z-up world, camera x right, y down, z forward."""

    pos: np.ndarray
    yaw_deg: float
    pitch_deg: float

    def axes(self):
        yaw = math.radians(self.yaw_deg)
        pitch = math.radians(self.pitch_deg)
        right = np.array([math.cos(yaw), -math.sin(yaw), 0.0])
        forward = np.array([math.sin(yaw) * math.cos(pitch), math.cos(yaw) * math.cos(pitch), math.sin(pitch)])
        down = np.cross(forward, right)
        down /= np.linalg.norm(down)
        return right, down, forward

    def view(self) -> np.ndarray:
        """app.rs:390-404 (rows: right, down, forward with -axis.pos)."""
        r, d, f = self.axes()
        p = np.asarray(self.pos, dtype=np.float64)
        return np.array([[*r, -r @ p], [*d, -d @ p], [*f, -f @ p], [0, 0, 0, 1]], dtype=np.float64)


def halton(index: int, base: int) -> float:
    """render_loop.rs:617-627"""
    f, r, i = 1.0, 0.0, index
    while i > 0:
        f /= base
        r += f * (i % base)
        i //= base
    return r


def jitter_ndc(offset_index: int, width: int, height: int) -> tuple[float, float]:
    """render_loop.rs:89-106: Halton(2,3) of index+1 in [-1,1], divided by the raster size."""
    i = offset_index + 1
    return ((2.0 * halton(i, 2) - 1.0) / width, (2.0 * halton(i, 3) - 1.0) / height)


@dataclass
class View:
    camera: Camera
    width: int
    height: int
    fov_h: float = math.pi / 2
    near: float = 0.05
    far: float = 102400.0
    jitter: tuple[float, float] = (0.0, 0.0)

    def view_proj(self) -> np.ndarray:
        proj = perspective_projection(self.near, self.far, self.fov_h, self.width / self.height)
        vp = proj @ self.camera.view()
        if self.jitter != (0.0, 0.0):
            t = np.eye(4)
            t[0, 3], t[1, 3] = self.jitter
            vp = t @ vp
        return vp


def _colmajor(m: np.ndarray) -> list[float]:
    return [float(v) for v in np.asarray(m, dtype=np.float32).T.reshape(-1)]


def make_frame_params(view: View, prev: View | None = None, sun_dir=(0.3, 0.2, 0.9),
                      sun_inten=(20.0, 20.0, 17.0)) -> FrameParams:
    """render_loop.rs:108-158 FrameParams::make (matrices inverted in f64, stored f32)."""
    vp = view.view_proj().astype(np.float32).astype(np.float64)
    inv = np.linalg.inv(vp)
    pos = np.asarray(view.camera.pos, dtype=np.float64)

    def ndc_to_ray(ndc):
        h = inv @ np.asarray(ndc, dtype=np.float64)
        return h[:3] / h[3] - pos

    tl = ndc_to_ray([-1, -1, 1, 1])
    right = ndc_to_ray([1, 0, 1, 1])
    left = ndc_to_ray([-1, 0, 1, 1])
    up = ndc_to_ray([0, -1, 1, 1])
    down = ndc_to_ray([0, 1, 1, 1])
    if prev is not None:
        pvp = prev.view_proj().astype(np.float32).astype(np.float64)
        ppos = np.asarray(prev.camera.pos, dtype=np.float64)
        pj = prev.jitter
    else:
        pvp, ppos, pj = vp, pos, (0.0, 0.0)
    sd = np.asarray(sun_dir, dtype=np.float64)
    sd = sd / np.linalg.norm(sd)
    fp = FrameParams()
    fp.view_proj[:] = _colmajor(vp)
    fp.inv_view_proj[:] = _colmajor(inv)
    fp.view_pos[:] = [*pos, 0.0]
    fp.view_ray_top_left[:] = [*tl, 0.0]
    fp.view_ray_right_shift[:] = [*(right - left), 0.0]
    fp.view_ray_down_shift[:] = [*(down - up), 0.0]
    fp.prev_view_proj[:] = _colmajor(pvp)
    fp.prev_inv_view_proj[:] = _colmajor(np.linalg.inv(pvp))
    fp.prev_view_pos[:] = [*ppos, 0.0]
    fp.jitter[:] = [view.jitter[0], view.jitter[1], pj[0], pj[1]]
    fp.sun_dir[:] = [*sd, 0.0]
    fp.sun_inten[:] = [*sun_inten, 0.0]
    return fp


def frame_params_array(fp: FrameParams) -> np.ndarray:
    return np.frombuffer(bytes(fp), dtype=np.float32).copy()


# ---------------------------------------------------------------------------
# G-buffer encodings (gbuffer.hlsl:54-77, enc.inc.hlsl:11-36), vectorised f32
# ---------------------------------------------------------------------------
def oct_encode(n: np.ndarray) -> np.ndarray:
    n = n.astype(F32)
    s = (np.abs(n[..., 0]) + np.abs(n[..., 1])) + np.abs(n[..., 2])
    with np.errstate(divide="ignore", invalid="ignore"):
        n = n / s[..., None]
    x, y, z = n[..., 0], n[..., 1], n[..., 2]
    wx = (F32(1) - np.abs(y)) * np.where(x >= 0, F32(1), F32(-1))
    wy = (F32(1) - np.abs(x)) * np.where(y >= 0, F32(1), F32(-1))
    neg = ~(z >= 0)
    x = np.where(neg, wx, x)
    y = np.where(neg, wy, y)
    return np.stack([x * F32(0.5) + F32(0.5), y * F32(0.5) + F32(0.5)], axis=-1).astype(F32)


def _trunc_u32(x: np.ndarray) -> np.ndarray:
    return np.clip(np.nan_to_num(x, nan=0.0), 0, 4294967040.0).astype(np.uint32)


def encode_gbuffer(color, metallic, normal, roughness, fwidth_z, fwidth_n, shading_path) -> np.ndarray:
    """Returns a (4, H, W) uint32 array (Texture2DArray layers 0..3)."""
    color = color.astype(F32)
    c8 = _trunc_u32(color * F32(255.0))
    m8 = _trunc_u32(metallic.astype(F32) * F32(255.0))
    l0 = c8[..., 0] | (c8[..., 1] << 8) | (c8[..., 2] << 16) | (m8 << 24)
    e = oct_encode(normal)
    nu = _trunc_u32(e * F32(4095.0))
    r8 = _trunc_u32(roughness.astype(F32) * F32(255.0))
    l1 = nu[..., 0] | (nu[..., 1] << 12) | (r8 << 24)
    fz = fwidth_z.astype(np.float16).view(np.uint16).astype(np.uint32)
    fn = fwidth_n.astype(np.float16).view(np.uint16).astype(np.uint32)
    l2 = fz | (fn << 16)
    l3 = shading_path.astype(np.uint32)
    return np.stack([l0, l1, l2, l3]).astype(np.uint32)


def quad_fwidth(v: np.ndarray, valid: np.ndarray) -> np.ndarray:
    """|ddx| + |ddy| with fine differences inside each aligned 2x2 quad.

    A difference whose partner pixel has no geometry (or lies outside the
    image) counts as 0."""
    h, w = valid.shape
    v = v.astype(np.float64)
    ys, xs = np.mgrid[0:h, 0:w]
    xp = xs ^ 1
    yp = ys ^ 1
    okx = (xp < w)
    oky = (yp < h)
    xpc = np.minimum(xp, w - 1)
    ypc = np.minimum(yp, h - 1)
    vx = v[ys, xpc]
    vy = v[ypc, xs]
    okx = okx & valid[ys, xpc] & valid
    oky = oky & valid[ypc, xs] & valid
    dx = np.where(okx[..., None] if v.ndim == 3 else okx, np.abs(vx - v), 0.0)
    dy = np.where(oky[..., None] if v.ndim == 3 else oky, np.abs(vy - v), 0.0)
    return dx + dy


# ---------------------------------------------------------------------------
# Denoiser-only workload: plane + 20 spheres, analytic primary visibility
# ---------------------------------------------------------------------------
@dataclass
class AnalyticScene:
    centers: np.ndarray  # (S,3)
    radii: np.ndarray  # (S,)
    albedo: np.ndarray  # (S+1,3), index 0 = ground plane
    roughness: np.ndarray  # (S+1,)


def make_analytic_scene(seed: int = 1, n_spheres: int = 20) -> AnalyticScene:
    idx = np.arange(n_spheres)
    radii = uniform(seed, idx * 8 + 0, 0.2, 1.5)
    cx = uniform(seed, idx * 8 + 1, -6.0, 6.0)
    cy = uniform(seed, idx * 8 + 2, 1.0, 14.0)
    cz = radii + uniform(seed, idx * 8 + 3, 0.0, 1.5)
    oid = np.arange(n_spheres + 1)
    albedo = np.stack([uniform(seed + 101, oid * 4 + k, 0.1, 0.9) for k in range(3)], axis=-1)
    rough = uniform(seed + 202, oid, 0.3, 1.0)
    return AnalyticScene(np.stack([cx, cy, cz], -1), radii, albedo, rough)


def default_denoise_camera() -> Camera:
    return Camera(pos=np.array([0.0, -7.0, 2.2]), yaw_deg=0.0, pitch_deg=-12.0)


def pixel_rays(view: View) -> tuple[np.ndarray, np.ndarray]:
    """Primary ray per pixel centre (frame_bindings.hlsl:71-79 without jitter)."""
    w, h = view.width, view.height
    inv = np.linalg.inv(view.view_proj().astype(np.float32).astype(np.float64))
    ys, xs = np.mgrid[0:h, 0:w]
    ndc = np.stack([(xs + 0.5) / w * 2 - 1, (ys + 0.5) / h * 2 - 1, np.ones_like(xs, dtype=np.float64),
                    np.ones_like(xs, dtype=np.float64)], -1)
    e = ndc @ inv.T
    end = e[..., :3] / e[..., 3:4]
    o = np.asarray(view.camera.pos, dtype=np.float64)
    d = end - o
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return np.broadcast_to(o, d.shape), d


def analytic_gbuffer(scene: AnalyticScene, view: View):
    """Returns dict(depth (H,W) f32, gbuffer (4,H,W) u32, position (H,W,3), object id)."""
    o, d = pixel_rays(view)
    h, w = view.height, view.width
    t_best = np.full((h, w), np.inf)
    oid = np.full((h, w), -1, dtype=np.int64)
    # ground plane z = 0 (object 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        tp = -o[..., 2] / d[..., 2]
    hitp = (tp > 1e-4) & np.isfinite(tp)
    t_best = np.where(hitp, tp, t_best)
    oid = np.where(hitp, 0, oid)
    for s in range(len(scene.radii)):
        oc = o - scene.centers[s]
        b = np.einsum("hwk,hwk->hw", oc, d)
        c = np.einsum("hwk,hwk->hw", oc, oc) - scene.radii[s] ** 2
        disc = b * b - c
        with np.errstate(invalid="ignore"):
            t = -b - np.sqrt(np.maximum(disc, 0.0))
        hit = (disc >= 0) & (t > 1e-4) & (t < t_best)
        t_best = np.where(hit, t, t_best)
        oid = np.where(hit, s + 1, oid)
    geo = oid >= 0
    tb = np.where(geo, t_best, 0.0)
    pos = o + d * tb[..., None]
    normal = np.zeros((h, w, 3))
    normal[..., 2] = 1.0
    for s in range(len(scene.radii)):
        m = oid == s + 1
        normal[m] = (pos[m] - scene.centers[s]) / scene.radii[s]
    vp = view.view_proj().astype(np.float32).astype(np.float64)
    hp = np.concatenate([pos, np.ones((h, w, 1))], -1) @ vp.T
    depth = np.where(geo, hp[..., 2] / hp[..., 3], 0.0).astype(F32)
    depth = np.where(geo & (depth <= 0), np.float32(1e-7), depth).astype(F32)
    fw_z = quad_fwidth(depth, geo)
    fw_n = np.linalg.norm(quad_fwidth(normal, geo), axis=-1)
    safe = np.where(oid >= 0, oid, 0)
    color = np.where(geo[..., None], scene.albedo[safe], 0.0)
    rough = np.where(geo, scene.roughness[safe], 0.0)
    gb = encode_gbuffer(color, np.zeros((h, w)), np.where(geo[..., None], normal, 0.0), rough, fw_z, fw_n,
                        geo.astype(np.uint32))
    return dict(depth=depth, gbuffer=gb, position=pos, object_id=oid, geometry=geo)


def clean_field(pos: np.ndarray, lobe: int) -> np.ndarray:
    """Smooth radiance field (low-frequency sinusoids of world position)."""
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    if lobe == 0:
        base = 0.6 + 0.35 * np.sin(0.9 * x + 0.4 * y) * np.cos(0.7 * z + 0.3 * y)
        tint = np.array([1.0, 0.85, 0.7])
    else:
        base = 0.4 + 0.3 * np.sin(0.5 * x - 0.8 * y + 0.6 * z)
        tint = np.array([0.7, 0.8, 1.0])
    return base[..., None] * tint


def noisy_radiance(pos: np.ndarray, geo: np.ndarray, frame: int, lobe: int, seed: int = 7) -> np.ndarray:
    """clean field x Exp(1), one draw per pixel per lobe for all 3 channels (f32, (H,W,3))."""
    h, w = geo.shape
    idx = (np.arange(h * w, dtype=np.uint64) * np.uint64(4) + np.uint64(lobe)).reshape(h, w)
    u = uniform(seed + 7919 * frame, idx)
    e = -np.log1p(-u)  # Exp(1)
    r = clean_field(pos, lobe) * e[..., None]
    return np.where(geo[..., None], r, 0.0).astype(F32)


@dataclass
class DenoiseFrame:
    fp: FrameParams
    depth: np.ndarray
    gbuffer: np.ndarray
    diffuse: np.ndarray
    specular: np.ndarray
    ray_len: np.ndarray


def denoise_sequence(width: int = 640, height: int = 360, frames: int = 8, seed: int = 1,
                     pan_px_per_frame: float = 2.0) -> list[DenoiseFrame]:
    """Denoiser-only workload: plane + 20 spheres, camera yaw pan of
    `pan_px_per_frame` pixels per frame at the image centre (2 px = 0.28125
    degrees at 640 px and 90 degrees horizontal FOV)."""
    scene = make_analytic_scene(seed)
    cam0 = default_denoise_camera()
    yaw_step = pan_px_per_frame * 90.0 / width
    out: list[DenoiseFrame] = []
    prev_view = None
    for f in range(frames):
        cam = Camera(cam0.pos, cam0.yaw_deg + yaw_step * f, cam0.pitch_deg)
        view = View(cam, width, height)
        g = analytic_gbuffer(scene, view)
        fp = make_frame_params(view, prev_view)
        diff = noisy_radiance(g["position"], g["geometry"], f, 0, seed)
        spec = noisy_radiance(g["position"], g["geometry"], f, 1, seed)
        ray_len = np.where(g["geometry"], 1.0, 0.0).astype(F32)
        out.append(DenoiseFrame(fp, g["depth"], g["gbuffer"], diff, spec, ray_len))
        prev_view = view
    return out
