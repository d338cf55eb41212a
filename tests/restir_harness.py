"""Shared set-up for the ReSTIR parity tests: a small seeded scene, an orbit
camera, a GPU renderer run frame by frame with snapshots of every per-view
image, and the matching oracle objects."""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def orbit_view(frame: int, w: int, h: int, radius=11.0, height=3.0, deg_per_frame=1.5, jitter=False):
    from violet_amd import synth

    a = math.radians(-90.0 + deg_per_frame * frame)
    pos = np.array([radius * math.cos(a), radius * math.sin(a), height])
    # look at the origin: yaw measured from +y towards +x (app.rs:104-113)
    d = -pos / np.linalg.norm(pos)
    yaw = math.degrees(math.atan2(d[0], d[1]))
    pitch = math.degrees(math.asin(d[2]))
    j = synth.jitter_ndc(frame % 8, w, h) if jitter else (0.0, 0.0)
    return synth.View(synth.Camera(pos, yaw, pitch), w, h, jitter=j)


def make_mesh(tris=4000, seed=3, extent=8.0, textured=False):
    from violet_amd import scene

    m = scene.make_scene(tris, seed=seed, extent=extent)
    return scene.texture_scene(m) if textured else m


SNAP = {
    "depth": (F32, lambda w, h: (h, w)),
    "gbuffer": (np.uint32, lambda w, h: (4, h, w)),
    "prev_depth": (F32, lambda w, h: (h, w)),
    "diff_reservoir": (np.uint32, lambda w, h: (h, w)),
    "diff_hit_pos_normal": (np.uint32, lambda w, h: (h, w, 2)),
    "diff_hit_radiance": (F32, lambda w, h: (h, w, 3)),
    "diff_temporal_reservoir": (np.uint32, lambda w, h: (h, w)),
    "diff_temporal_hit_pos_normal": (np.uint32, lambda w, h: (h, w, 2)),
    "diff_temporal_hit_radiance": (F32, lambda w, h: (h, w, 3)),
    "diff_new_hit_pos_normal": (np.uint32, lambda w, h: (h, w, 2)),
    "diff_new_hit_radiance": (F32, lambda w, h: (h, w, 3)),
    "ao": (F32, lambda w, h: (h, w, 2)),
    "ao_filtered": (F32, lambda w, h: (h, w, 2)),
    "indirect_diffuse_raw": (F32, lambda w, h: (h, w, 3)),
    "spec_reservoir": (np.uint32, lambda w, h: (h, w, 2)),
    "spec_hit_pos": (F32, lambda w, h: (h, w, 4)),
    "spec_hit_radiance": (F32, lambda w, h: (h, w, 3)),
    "indirect_specular_raw": (F32, lambda w, h: (h, w, 3)),
    "spec_ray_len": (F32, lambda w, h: (h, w)),
    "hash_grid_storage": (np.uint32, lambda w, h: (65536 * 16, 4)),
    "hash_grid_decay": (np.uint8, lambda w, h: (65536 * 16,)),
    "hash_grid_queries": (np.uint32, lambda w, h: (128 * 1024, 6)),
    "hash_grid_query_counter": (np.uint32, lambda w, h: (1,)),
    "hash_grid_query_threshold": (np.uint32, lambda w, h: (1,)),
}


# Images whose element type follows the storage mode (binary16 in FP16 mode).
STORAGE_IMAGES = {"diff_hit_radiance", "diff_temporal_hit_radiance", "diff_new_hit_radiance", "ao", "ao_filtered",
                  "indirect_diffuse_raw", "spec_hit_radiance", "indirect_specular_raw", "spec_ray_len"}


def run_gpu_frames(mesh, w, h, frames, config=None, snap=True, jitter=False):
    """Run the GPU renderer; returns (frame params list, outputs list, snapshots list)."""
    import torch

    from violet_amd import _lib, renderer, scene, synth

    sc = scene.Scene(mesh)
    config = config or renderer.RestirConfig()
    f16 = config.storage == _lib.STORAGE_F16
    r = renderer.RestirRenderer(w, h, config)
    fps, outs, snaps = [], [], []
    prev_view = None
    for f in range(frames):
        view = orbit_view(f, w, h, jitter=jitter)
        fp = synth.make_frame_params(view, prev_view)
        o = r.add_passes(sc, fp, f)
        torch.cuda.synchronize()
        s = {}
        if snap:
            for name, (dt, shp) in SNAP.items():
                dt = np.float16 if f16 and name in STORAGE_IMAGES else dt
                s[name] = r.read(name, dt, shp(w, h)).copy()
            rb = w * h * (6 if f16 else 12)
            dt = np.float16 if f16 else F32
            s["out_diffuse"] = r.read_ptr(o.indirect_diffuse, rb, dt, (h, w, 3)).copy()
            s["out_specular"] = r.read_ptr(o.indirect_specular, rb, dt, (h, w, 3)).copy()
            s["flags"] = dict(denoised=o.denoised, validation=o.validation_frame, hg=o.hash_grid_updated)
            if o.color:
                s["shadow_mask"] = r.read("shadow_mask", np.uint8, (h, w)).copy()
                s["out_color"] = r.read_ptr(o.color, rb, dt, (h, w, 3)).copy()
                s["scene_color"] = r.read("color", dt, (h, w, 3)).copy()
        fps.append(fp)
        outs.append(o)
        snaps.append(s)
        prev_view = view
    return fps, outs, snaps, r, sc


def seeds_for_frame(frame_index: int, hg_updated: bool, validation: bool) -> dict:
    """Seeds drawn by the passes of one frame, in violet's order (Appendix C of
    the pass list: select, raygen, diff gen, diff temporal, diff spatial,
    spec gen, spec temporal, spec spatial)."""
    import oracle

    names = []
    if hg_updated:
        names += ["hg_select", "hg_raygen"]
    if not validation:
        names += ["diff_gen"]
    names += ["diff_temporal", "diff_spatial", "spec_gen", "spec_temporal", "spec_spatial"]
    vals = oracle.frame_rand_seeds(frame_index, len(names))
    return dict(zip(names, vals))


def multiset(rows: np.ndarray) -> dict:
    from collections import Counter

    return Counter(map(bytes, np.ascontiguousarray(rows)))
