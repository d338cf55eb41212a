"""Serving loop on the GPU: requests continuously batched over persistent
views give each request the same result as rendering that view's frames
alone, bit for bit (the renderer's order-independent cache update leaves
nothing to the GPU's schedule), and admission under a small HBM budget
evicts idle views instead of failing.

* small: 24 requests over four views of at most 80x64, FP32;
* the serving form at full size: a few hundred Poisson requests at 600/s
  over 64 persistent views drawn from {256x256 .. 2560x1440}, FP16, on the
  1M-triangle scene, batches of up to 16 on two streams; each request's
  denoised diffuse and specular against the same view's frames rendered
  alone.
"""
from __future__ import annotations

import numpy as np
import pytest

from restir_harness import F32, make_mesh

pytestmark = pytest.mark.gpu

SPECS = [(64, 48, 10.0), (40, 40, 100.0), (80, 36, 200.0), (48, 64, 300.0)]


def _setup():
    from violet_amd import renderer, scene, serve

    mesh = make_mesh()
    sc = scene.Scene(mesh)
    views = [serve.ViewSpec(w, h, ph, orbit_radius=11.0, orbit_height=3.0) for w, h, ph in SPECS]
    cfg = renderer.RestirConfig(temporal_m_cap=30, diff_spatial_taps=5, spatial_radius_px=30.0)
    return sc, views, cfg


def test_requests_same_alone_and_in_the_mix(vgi):
    import torch

    from violet_amd import batch, renderer, serve

    sc, views, cfg = _setup()
    got = {}

    def keep(req, r, out):
        spec = views[req.view]
        w, h = spec.width, spec.height
        frame = sum(1 for k in got if k[0] == req.view)
        got[(req.view, frame)] = dict(depth=r.read("depth", F32, (h, w)).copy(),
                                      gbuffer=r.read("gbuffer", np.uint32, (4, h, w)).copy(),
                                      out=r.read_ptr(out.indirect_diffuse, w * h * 12, F32, (h, w, 3)).copy())

    reqs = serve.poisson_requests(24, 2000.0, len(views), seed=2)
    srv = serve.Server(sc, views, hbm_budget_bytes=50e9, config=cfg, max_batch=3, streams=2,
                       batch_config=batch.BatchConfig(pixel_groups=64, trace_groups=96, atrous_groups=16),
                       on_complete=keep)
    stats = srv.run(reqs, timeout_s=60.0)
    srv.close()
    s = stats.summary()
    assert s["requests"] == len(reqs) and s["evictions"] == 0
    assert s["mean_batch"] > 1.0, "no two requests were ever batched together"
    assert min(stats.latencies_ms) > 0.0
    counts = {v: sum(1 for r in reqs if r.view == v) for v in range(len(views))}
    for v, n in counts.items():
        spec = views[v]
        w, h = spec.width, spec.height
        r = renderer.RestirRenderer(w, h, cfg)
        for f in range(n):
            o = r.add_passes(sc, serve.view_frame_params(spec, f), f)
            torch.cuda.synchronize()
            a = dict(depth=r.read("depth", F32, (h, w)), gbuffer=r.read("gbuffer", np.uint32, (4, h, w)),
                     out=r.read_ptr(o.indirect_diffuse, w * h * 12, F32, (h, w, 3)))
            b = got[(v, f)]
            assert np.array_equal(a["depth"], b["depth"]) and np.array_equal(a["gbuffer"], b["gbuffer"]), (v, f)
            assert np.array_equal(a["out"].view(np.uint32), b["out"].view(np.uint32)), (v, f)
        r.close()


def test_serving_form_requests_bit_identical_alone_and_in_the_mix(vgi, tmp_path):
    import json

    import torch

    from violet_amd import _lib, batch, gpu, renderer, scene, serve

    sc = scene.Scene(scene.make_scene(1_000_000, seed=13))
    views = serve.serve_views(64, seed=31)
    cfg = renderer.RestirConfig(storage=_lib.STORAGE_F16)
    reqs = serve.poisson_requests(300, 600.0, len(views), seed=32)
    got = {}

    def keep(req, r, out):
        spec = views[req.view]
        n = spec.width * spec.height * 6
        frame = sum(1 for k in got if k[0] == req.view)
        got[(req.view, frame)] = (gpu.copy_from_ptr(out.indirect_diffuse, n),
                                  gpu.copy_from_ptr(out.indirect_specular, n))

    mpath = tmp_path / "metrics.jsonl"
    srv = serve.Server(sc, views, hbm_budget_bytes=200e9, config=cfg, max_batch=16, streams=2,
                       batch_config=batch.BatchConfig(), on_complete=keep, metrics=mpath)
    srv.admit_all()
    stats = srv.run(reqs, timeout_s=120.0)
    srv.close()
    s = stats.summary()
    # the metrics stream: one line per batch, every request once, then the summary
    lines = [json.loads(x) for x in mpath.read_text().splitlines()]
    batches = [x for x in lines if x["event"] == "batch"]
    assert len(batches) == s["batches"] and lines[-1]["event"] == "summary"
    assert sorted(l for x in batches for l in x["latency_ms"]) == pytest.approx(sorted(stats.latencies_ms), abs=1e-3)
    assert sum(len(x["views"]) for x in batches) == len(reqs) == lines[-1]["requests"]
    assert lines[-1]["p99_ms"] == pytest.approx(s["p99_ms"])
    print(f"serving form: {s['requests']} requests, mean batch {s['mean_batch']:.2f}, p99 {s['p99_ms']:.1f} ms",
          flush=True)
    assert s["requests"] == len(reqs) and s["evictions"] == 0 and s["mean_batch"] > 1.0
    torch.cuda.synchronize()
    for v in sorted({r.view for r in reqs}):
        spec = views[v]
        n = spec.width * spec.height * 6
        r = renderer.RestirRenderer(spec.width, spec.height, cfg)
        f = 0
        while (v, f) in got:
            o = r.add_passes(sc, serve.view_frame_params(spec, f), f)
            a = (gpu.copy_from_ptr(o.indirect_diffuse, n), gpu.copy_from_ptr(o.indirect_specular, n))
            for name, x, y in zip(("diffuse", "specular"), a, got[(v, f)]):
                assert torch.equal(x, y), f"{name} differs: view {v} {spec.width}x{spec.height} frame {f}"
            f += 1
        r.close()


def test_admission_evicts_idle_views_under_a_small_budget(vgi):
    from violet_amd import serve

    sc, views, cfg = _setup()
    srv = serve.Server(sc, views, hbm_budget_bytes=0, config=cfg, max_batch=4, streams=1)
    one = srv._estimate_bytes(views[0])
    srv.budget = 2.2 * max(srv._estimate_bytes(v) for v in views)  # room for two views
    reqs = serve.poisson_requests(16, 500.0, len(views), seed=5)
    stats = srv.run(reqs, timeout_s=60.0)
    srv.close()
    s = stats.summary()
    assert one > 0 and s["requests"] == len(reqs)
    assert s["evictions"] > 0
    assert s["peak_hbm_gb"] * 1e9 <= srv.budget


def test_up_front_admission_stops_at_the_budget(vgi):
    """admit_all makes views resident in view order while they fit; the run
    then admits and evicts the rest on demand, and every request completes."""
    from violet_amd import serve

    sc, views, cfg = _setup()
    srv = serve.Server(sc, views, hbm_budget_bytes=0, config=cfg, max_batch=4, streams=2)
    sizes = [srv._estimate_bytes(v) for v in views]
    srv.budget = sizes[0] + sizes[1] + 0.5 * min(sizes[2:])  # the first two views fit
    assert srv.admit_all() == 2 and sorted(srv.resident) == [0, 1]
    assert srv.bytes_resident <= srv.budget
    reqs = serve.poisson_requests(12, 500.0, len(views), seed=7)
    stats = srv.run(reqs, timeout_s=60.0)
    srv.close()
    s = stats.summary()
    assert s["requests"] == len(reqs) and s["peak_hbm_gb"] * 1e9 <= srv.budget


def test_batches_past_the_frame_deadline_are_counted_and_finish(vgi):
    """A zero frame deadline marks every batch still running at its first
    poll as timed out; as in violet, the server keeps waiting and every
    request completes."""
    from violet_amd import serve

    sc, views, cfg = _setup()
    srv = serve.Server(sc, views, hbm_budget_bytes=50e9, config=cfg, max_batch=4, streams=2, frame_timeout_ms=0)
    reqs = serve.poisson_requests(16, 2000.0, len(views), seed=9)
    stats = srv.run(reqs, timeout_s=60.0)
    srv.close()
    s = stats.summary()
    assert s["requests"] == len(reqs) and s["timeouts"] > 0
