"""Launch shapes (vgi_tuning: the sample passes' wave tile, the a-trous LDS
block) change no result: every per-view image is bit-identical under every
shape, alone and batched, and the autotuner runs its coordinate descent and
leaves the chosen shapes set."""
from __future__ import annotations

import numpy as np
import pytest

from restir_harness import F32, make_mesh, orbit_view

pytestmark = pytest.mark.gpu

SIZES = [(96, 64), (45, 37)]
FRAMES = 5
IMAGES = {"diff_reservoir": (np.uint32, 1), "diff_hit_radiance": (F32, 3), "indirect_diffuse_raw": (F32, 3),
          "spec_reservoir": (np.uint32, 2), "spec_hit_pos": (F32, 4), "indirect_specular_raw": (F32, 3)}
SHAPES = [(16, 8), (8, 8), (32, 8), (16, 16), (16, 32), (8, 32)]


def _render(sc, batched):
    import torch

    from violet_amd import batch, renderer, synth

    cfg = renderer.RestirConfig()
    views = [renderer.RestirRenderer(w, h, cfg) for w, h in SIZES]
    b = batch.Batch(batch.BatchConfig(max_views=4, use_graph=True)) if batched else None
    out = []
    for f in range(FRAMES):
        fps = []
        for k, (w, h) in enumerate(SIZES):
            fps.append(synth.make_frame_params(orbit_view(f + 9 * k, w, h), orbit_view(f - 1 + 9 * k, w, h)
                                               if f else None))
        if b:
            outs = b.add_passes(sc, views, fps, [f] * len(views))
        else:
            outs = [r.add_passes(sc, fp, f) for r, fp in zip(views, fps)]
        torch.cuda.synchronize()
        frame = []
        for r, o, (w, h) in zip(views, outs, SIZES):
            imgs = {n: r.read(n, dt, (h, w, c)).copy() for n, (dt, c) in IMAGES.items()}
            imgs["out"] = r.read_ptr(o.indirect_diffuse, w * h * 12, F32, (h, w, 3)).copy()
            frame.append(imgs)
        out.append(frame)
    if b:
        b.close()
    for r in views:
        r.close()
    return out


@pytest.mark.parametrize("batched", [False, True], ids=["alone", "batched"])
def test_results_do_not_depend_on_launch_shapes(vgi, batched):
    from violet_amd import _lib, scene

    sc = scene.Scene(make_mesh())
    try:
        _lib.set_tuning()
        ref = _render(sc, batched)
        for ab, tw in SHAPES[1:]:
            _lib.set_tuning(ab, tw)
            assert _lib.get_tuning() == dict(atrous_block=ab, trace_tile_width=tw)
            got = _render(sc, batched)
            for f in range(FRAMES):
                for k in range(len(SIZES)):
                    for n in ref[f][k]:
                        assert np.array_equal(ref[f][k][n].view(np.uint8), got[f][k][n].view(np.uint8)), \
                            f"{n} differs: a-trous block {ab}, tile width {tw}, view {k}, frame {f}"
    finally:
        _lib.set_tuning()


def test_autotuner_descends_and_sets_the_shapes(vgi, tmp_path):
    from violet_amd import _lib, autotune, renderer, scene, synth

    sc = scene.Scene(make_mesh())
    views = [renderer.RestirRenderer(w, h) for w, h in SIZES]

    def fps_of(i, f):
        w, h = SIZES[i]
        return synth.make_frame_params(orbit_view(f, w, h), orbit_view(f - 1, w, h) if f else None)

    try:
        dims = (("trace_tile_width", (8, 16)), ("atrous_block", (16, 8)), ("trace_groups", (0, 16)))
        cfg, table, f, best = autotune.tune_batch(sc, views, fps_of, warmup=1, frames=2, repeats=2, dimensions=dims)
        assert len(table) == 6 and all(ms > 0 for _, ms, _ in table)
        assert f == 6 * 2 * (1 + 2)  # every candidate's frames advanced the views
        assert _lib.get_tuning() == dict(atrous_block=best["atrous_block"], trace_tile_width=best["trace_tile_width"])
        assert cfg.trace_groups in (0, 16 * __import__("torch").cuda.get_device_properties(0).multi_processor_count)
        # the cache file: a first tuning stores the result, a second applies it without timing
        cache = autotune.TuningCache(tmp_path / "tune.json")
        _lib.set_tuning()
        cfg1, table1, f1, best1 = autotune.tune_batch(sc, views, fps_of, frame0=f, warmup=1, frames=2, repeats=1,
                                                      dimensions=dims, cache=cache)
        assert len(table1) == 6 and cache.get(autotune.workload_key(
            views, sc.stats().num_triangles, __import__("torch").cuda.get_device_name(0))) == best1
        _lib.set_tuning()
        cfg2, table2, f2, best2 = autotune.tune_batch(sc, views, fps_of, frame0=f1, dimensions=dims, cache=cache)
        assert table2 == [] and f2 == f1 and best2 == best1 and cfg2 == cfg1
        assert _lib.get_tuning() == dict(atrous_block=best1["atrous_block"], trace_tile_width=best1["trace_tile_width"])
    finally:
        _lib.set_tuning()
        for r in views:
            r.close()
