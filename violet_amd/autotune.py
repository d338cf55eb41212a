"""Autotuner for batched frames: picks the launch shapes of the fused passes
on the actual views and scene, by timing whole graph-replayed batched frames.

Dimensions, in order (the trace passes dominate a frame):
  trace_tile_width  pixel tile of a 64-lane wave in the four sample passes:
                    8x8 (2x2 quads), 16x4 or 32x2 (vgi_tuning)
  atrous_block      a-trous lattice block staged in LDS per workgroup:
                    16x16 (256 threads), 8x8 (64) or 32x32 (1024) (vgi_tuning)
  trace_groups,     fused launches: one workgroup per tile (0) or a
  pixel_groups,     persistent grid of `multiplier x compute units` groups
  atrous_groups     striding over the tiles (vgi_batch_config)

Coordinate descent: for each dimension every option is timed `repeats`
times (each a median over `frames` frames after `warmup` untimed ones,
options interleaved round-robin so slow drifts hit them alike); an option
replaces the incumbent only if its median beats the incumbent's by more
than the incumbent's own spread across repeats (and at least 1%).  Shapes
do not change results, so tuning only advances the views by the frames it
renders.  The chosen launch shapes stay set (vgi_tuning_set) on return.

Results persist in a JSON cache file (TuningCache; tune_batch(cache=...)):
keyed by the device, the library version and the workload (the views'
sizes, storage and resampling parameters, the scene's triangle count), a
hit applies the stored shapes without timing anything.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path

import numpy as np

from . import _lib, batch


def _time_frames(b, scene, views, fps_of, frame0: int, warmup: int, frames: int) -> float:
    import torch

    n = len(views)
    for k in range(warmup):
        f = frame0 + k
        b.add_passes(scene, views, [fps_of(i, f) for i in range(n)], [f] * n)
    torch.cuda.synchronize()
    ts = []
    for k in range(frames):
        f = frame0 + warmup + k
        t0 = time.perf_counter()
        b.add_passes(scene, views, [fps_of(i, f) for i in range(n)], [f] * n)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


DIMENSIONS = (("trace_tile_width", (8, 16, 32)), ("atrous_block", (16, 8, 32)), ("trace_groups", (0, 16, 32)),
              ("pixel_groups", (0, 8)), ("atrous_groups", (0, 4)))
SHAPES = ("trace_tile_width", "atrous_block")
GROUPS = ("trace_groups", "pixel_groups", "atrous_groups")


class TuningCache:
    """Tuned launch shapes on disk: a JSON object {key: {"shapes": {...},
    "table": [...], "time": unix seconds}}.  Writes go through a temporary
    file and a rename, so a reader never sees a partial file."""

    def __init__(self, path):
        self.path = Path(path)

    def load(self) -> dict:
        try:
            d = json.loads(self.path.read_text())
            return d if isinstance(d, dict) else {}
        except (OSError, ValueError):
            return {}

    def get(self, key: str) -> dict | None:
        e = self.load().get(key)
        if not isinstance(e, dict) or not isinstance(e.get("shapes"), dict):
            return None
        sh = e["shapes"]
        ok = all(sh.get(k) in opts for k, opts in DIMENSIONS)
        return dict(sh) if ok else None

    def put(self, key: str, shapes: dict, table=None) -> None:
        d = self.load()
        d[key] = {"shapes": {k: int(shapes[k]) for k, _ in DIMENSIONS}, "time": time.time(),
                  "table": [[{k: int(v) for k, v in c.items()}, round(float(m), 4), round(float(sp), 4)]
                            for c, m, sp in (table or [])]}
        self.path.parent.mkdir(parents=True, exist_ok=True)
        tmp = self.path.with_name(self.path.name + f".{os.getpid()}.tmp")
        tmp.write_text(json.dumps(d, indent=1, sort_keys=True))
        os.replace(tmp, self.path)


def workload_key(views, ntris: int, device: str, extra: str = "") -> str:
    """Cache key: device, library version and the workload's shape (each
    view's size, storage and resampling configuration; the triangle
    count)."""
    def cfg(v):
        c = getattr(v, "config", None)
        return json.dumps(vars(c) if hasattr(c, "__dict__") else c, sort_keys=True, default=str)

    desc = {"device": device, "library": _lib.lib().vgi_version().decode(), "triangles": int(ntris),
            "views": [[int(v.width), int(v.height), cfg(v)] for v in views], "extra": extra}
    return hashlib.sha256(json.dumps(desc, sort_keys=True).encode()).hexdigest()[:32]


def tune_batch(scene, views, fps_of, frame0: int = 0, warmup: int = 2, frames: int = 6, repeats: int = 3,
               dimensions=DIMENSIONS, use_graph: bool = True, cache: TuningCache | None = None,
               cache_key: str | None = None):
    """fps_of(view_index, frame) -> FrameParams.  Returns (best BatchConfig,
    table of (candidate dict, median ms, spread ms), the next free frame
    index, the best shapes); the best launch shapes are left set.  With a
    cache, a stored entry for the workload is applied without timing (the
    table is then empty), and a tuned result is stored."""
    import torch

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    best = dict(trace_tile_width=8, atrous_block=16, trace_groups=0, pixel_groups=0, atrous_groups=0)
    table = []
    f = frame0

    def config(c):
        return batch.BatchConfig(max_views=max(len(views), 1), use_graph=use_graph,
                                 **{k: c[k] * (cus if k in GROUPS else 1) for k in GROUPS})

    if cache is not None:
        if cache_key is None:
            ntris = scene.stats().num_triangles if hasattr(scene, "stats") else 0
            cache_key = workload_key(views, ntris, torch.cuda.get_device_name(0))
        hit = cache.get(cache_key)
        if hit is not None:
            _lib.set_tuning(hit["atrous_block"], hit["trace_tile_width"])
            return config(hit), [], f, hit

    for key, opts in dimensions:
        cands = [dict(best, **{key: m}) for m in opts]
        runs = {i: [] for i in range(len(cands))}
        bs = {}
        for _ in range(repeats):
            for i, c in enumerate(cands):
                _lib.set_tuning(c["atrous_block"], c["trace_tile_width"])
                if i not in bs:
                    bs[i] = batch.Batch(config(c))
                runs[i].append(_time_frames(bs[i], scene, views, fps_of, f, warmup, frames) * 1e3)
                f += warmup + frames
        for b in bs.values():
            b.close()
        med = {i: float(np.median(v)) for i, v in runs.items()}
        spread = {i: float(np.max(v) - np.min(v)) for i, v in runs.items()}
        inc = next(i for i, c in enumerate(cands) if c[key] == best[key])
        win = inc
        for i in med:
            if med[i] < med[win] - max(spread[inc], 0.01 * med[inc]):
                win = i
        best = dict(cands[win])
        table += [(cands[i], med[i], spread[i]) for i in med]
    _lib.set_tuning(best["atrous_block"], best["trace_tile_width"])
    if cache is not None:
        cache.put(cache_key, best, table)
    return config(best), table, f, best
