"""Autotuner for batched frames: picks the launch shape of the fused passes
(ray-tracing passes, pixel and cache passes, a-trous passes) on the actual
views and scene, by timing whole batched frames.

Each class either launches one workgroup per tile (multiplier 0: the
hardware dispatcher balances uneven tiles) or a persistent grid of
`multiplier x compute units` groups striding over the tiles.  Coordinate
descent over the three classes, trace passes first since they dominate a
frame; every candidate runs graph-replayed frames and the fastest mean frame
time wins.  Tuning advances the views by the frames it renders.
"""
from __future__ import annotations

import time

from . import batch


def _time_frames(b, scene, views, fps_of, frame0: int, warmup: int, frames: int) -> float:
    import torch

    n = len(views)
    for k in range(warmup):
        f = frame0 + k
        b.add_passes(scene, views, [fps_of(i, f) for i in range(n)], [f] * n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(frames):
        f = frame0 + warmup + k
        b.add_passes(scene, views, [fps_of(i, f) for i in range(n)], [f] * n)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / frames


def tune_batch(scene, views, fps_of, frame0: int = 0, warmup: int = 2, frames: int = 4,
               trace=(0, 16, 32), pixel=(0, 8), atrous=(0, 4), use_graph: bool = True):
    """fps_of(view_index, frame) -> FrameParams.  Returns (best BatchConfig,
    list of (config dict, ms per frame)) and the next free frame index."""
    import torch

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    best = dict(trace_groups=0, pixel_groups=0, atrous_groups=0)
    table = []
    f = frame0
    for key, opts in (("trace_groups", trace), ("pixel_groups", pixel), ("atrous_groups", atrous)):
        scores = {}
        for m in opts:
            cand = dict(best)
            cand[key] = m * cus
            cfg = batch.BatchConfig(max_views=max(len(views), 1), use_graph=use_graph, **cand)
            b = batch.Batch(cfg)
            ms = _time_frames(b, scene, views, fps_of, f, warmup, frames) * 1e3
            f += warmup + frames
            b.close()
            scores[m] = ms
            table.append((cand, ms))
        best[key] = min(scores, key=scores.get) * cus
    return batch.BatchConfig(max_views=max(len(views), 1), use_graph=use_graph, **best), table, f
