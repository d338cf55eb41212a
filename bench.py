#!/usr/bin/env python3
"""Benchmark of the ReSTIR GI + denoise hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the hot path over one frame of synthetic input that is
already resident in HBM.  Prints ONE JSON line (metric, value in Mpix/s,
per-step latency percentiles, the dominant kernel's roofline from live
per-pass HIP-event timers, and the CPU baseline = the oracle's FP32 build
timed on a bounded sample of the same workload on this host's cores).

The `value` is the plain single-view frame loop at 1920x1080 (the metric's
resolution) on the seeded 1M-triangle scene, FP32 storage, violet's own
ReSTIR constants; `--width 1280 --height 720 --triangles 200000 --resample
frame_loop` gives the 1280x720 frame-loop configuration.  Beside it the line
carries the other forms on the same scene: "batched" (16 views of seeded
resolutions, plain vs one fused persistent launch per pass, graph-replayed,
autotuned), "serving" (2000 Poisson requests at 600/s over 64 persistent
views, continuously batched, p50/p99 latency) and "precision" (1920x1080
with FP16 vs FP32 storage).  --forms '' skips them.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SCENE_SEED = 13  # the seeded boxes + icospheres scene shared by every form

# Resampling parameters: the 1280x720 frame-loop workload's (temporal M cap
# 30, 5 spatial taps in a 30 px radius) or the reference's constants.
RESAMPLE = {
    "frame_loop": dict(temporal_m_cap=30, diff_spatial_taps=5, spatial_radius_px=30.0),
    "reference": dict(temporal_m_cap=0, diff_spatial_taps=0, spatial_radius_px=0.0),
}
RESAMPLE_DESC = {
    "frame_loop": "M cap 30 (specular 30*roughness), 5 spatial taps, 30 px radius",
    "reference": "reference constants (M cap 16, 9 spatial taps, radius 0.05*min(w,h))",
}


# Pass name -> kernel symbol (for counter summaries of that kernel).
PASS_KERNEL = {
    "denoise_atrous": "k_atrous", "denoise_temporal": "k_temporal", "denoise_disocclusion_fix": "k_disocclusion",
    "denoise_spatial_fallback": "k_fallback", "ind_diff_spatial_resample": "k_diff_spatial_taps",
    "ind_diff_temporal_resample": "k_diff_temporal", "ind_spec_temporal_resample": "k_spec_temporal",
    "ind_spec_spatial_resample": "k_spec_spatial", "raytraced_ao_gen": "k_ao_gen", "raytraced_ao_filter": "k_ao_filter",
    "hash_grid_query_compact": "k_hg_compact", "hash_grid_decay": "k_hg_decay", "hash_grid_query_select": "k_hg_select",
}


def calibration_copy(n: int = 4):
    """Plain device copies of a known size (256 MiB read + 256 MiB written
    each), for scaling HBM counters in a rocprofv3 --pmc run of this script."""
    import torch

    src = torch.ones((256 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(n):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst


def measured_traffic(workload: str, pass_name: str):
    """HBM bytes per launch of a pass from the committed counter summary of
    this workload (profiles/traffic_<workload>.json), or None."""
    path = ROOT / "profiles" / f"traffic_{workload}.json"
    sym = PASS_KERNEL.get(pass_name)
    if not sym or not path.exists():
        return None
    data = json.loads(path.read_text())
    for k, v in data["kernels"].items():
        if k.split("<")[0].split("::")[-1] == sym:
            return round(v["fetch_bytes"] + v["write_bytes"])
    return None


def _lib_const(name: str):
    from violet_amd import _lib

    return getattr(_lib, name)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def denoise_workload(width: int, height: int, frames: int, iterations: int):
    from violet_amd import synth

    log(f"[bench] generating {frames} frames of the plane+spheres denoise workload at {width}x{height}")
    seq = synth.denoise_sequence(width, height, frames, seed=1)
    return seq


def run_denoise(args):
    import torch

    from violet_amd import _lib, gpu
    from violet_amd import denoiser as D

    w, h = args.width, args.height
    seq = denoise_workload(w, h, args.frames, args.iterations)
    dev = [dict(depth=gpu.dev(f.depth), gbuffer=gpu.dev(f.gbuffer), diffuse=gpu.dev(f.diffuse),
                specular=gpu.dev(f.specular), ray_len=gpu.dev(f.ray_len)) for f in seq]
    den = D.Denoiser(w, h, _lib.STORAGE_F16 if args.storage == "f16" else _lib.STORAGE_F32)
    den.atrous_iterations = args.iterations

    def step(k: int):
        fi = k % len(seq)
        if fi == 0:
            den.reset()
        f, t = seq[fi], dev[fi]
        p = dev[fi - 1] if fi > 0 else None
        den.add_passes(f.fp, D.Input(t["depth"], t["gbuffer"], None if p is None else p["depth"],
                                     None if p is None else p["gbuffer"], t["diffuse"], t["specular"], t["ray_len"]))

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    # per-step latency (host-synchronised), separate from the throughput run
    lat = []
    for k in range(min(args.steps, 200)):
        t0 = time.perf_counter()
        step(k)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
    prof = profile_passes(step, 0, min(args.steps, 60))
    dname = dominant_pass(prof, set())
    elapsed, timed = timed_region(step, 0, args.steps, dname)
    return dict(elapsed=elapsed, lat=lat, prof=prof, timed=timed, dname=dname, pixels=w * h)


def cpu_baseline_denoise(args, budget_s: float = 20.0):
    """FP32 twin (oracle, OpenMP) on a bounded sample of the same frames."""
    import oracle as O

    lib = O.load(32)
    threads = lib.oracle_num_threads()
    w, h = args.cpu_width, args.cpu_height
    seq = denoise_workload(w, h, args.frames, args.iterations)
    den = O.OracleDenoiser(32, atrous_iterations=args.iterations)
    t0 = time.perf_counter()
    n = 0
    prev = None
    while True:
        f = seq[n % len(seq)]
        if n % len(seq) == 0:
            den.reset()
            prev = None
        den.add_passes(f.fp, f.depth, f.gbuffer, None if prev is None else prev.depth,
                       None if prev is None else prev.gbuffer, f.diffuse, f.specular, f.ray_len)
        prev = f
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 4 * len(seq):
            break
    el = time.perf_counter() - t0
    return dict(value=n * w * h / el / 1e6, unit="Mpix/s", cores=threads, kind="port",
                sample=f"{n} frames of the denoise workload at {w}x{h} ({args.iterations} a-trous iterations), "
                       f"FP32 twin of the oracle, {el:.1f} s")


def orbit_view(frame: int, w: int, h: int, radius=30.0, height=9.0, deg_per_frame=6.0, jitter_cycle=0):
    """Camera orbit around the scene centre (60 frames = one turn); with a
    jitter cycle, Halton(2,3) sub-pixel jitter of frame mod cycle."""
    import math

    from violet_amd import synth

    a = math.radians(-90.0 + deg_per_frame * frame)
    pos = np.array([radius * math.cos(a), radius * math.sin(a), height])
    d = -pos / np.linalg.norm(pos)
    yaw = math.degrees(math.atan2(d[0], d[1]))
    pitch = math.degrees(math.asin(d[2])) + 8.0
    j = synth.jitter_ndc(frame % jitter_cycle, w, h) if jitter_cycle else (0.0, 0.0)
    return synth.View(synth.Camera(pos, yaw, pitch), w, h, jitter=j)


def frame_params_seq(w, h, n, jitter_cycle=0):
    """Frame parameters of an n-frame cycle of the orbit; frame 0's previous
    view is frame n-1 (the frame rendered before it when the loop wraps)."""
    from violet_amd import synth

    views = [orbit_view(f, w, h, jitter_cycle=jitter_cycle) for f in range(n)]
    return [synth.make_frame_params(views[f], views[f - 1]) for f in range(n)]


def profile_passes(step, first: int, n: int) -> dict:
    """Per-pass HIP-event times of `n` steps (every pass timed)."""
    import torch

    from violet_amd import _lib

    _lib.call("vgi_profiler_reset")
    _lib.call("vgi_profiler_enable", 1)
    for k in range(n):
        step(first + k)
    torch.cuda.synchronize()
    _lib.call("vgi_profiler_enable", 0)
    return _lib.profiler_read()


def dominant_pass(prof: dict, exclude: set) -> str:
    """The roofline kernel: the pixel pass with the largest summed time among
    those not in `exclude` (the trace passes, which are latency/divergence
    bound and reported apart, and, when the frame runs two concurrent
    streams, the passes of those streams: a kernel sharing the device with
    another stream's kernels has no launch duration of its own to divide
    its bytes by)."""
    return max(((k, v) for k, v in prof.items() if k not in exclude), key=lambda kv: kv[1]["ms"])[0]


def timed_region(step, first: int, steps: int, only: str) -> tuple[float, dict]:
    """The timed throughput region: `steps` steps between two device
    synchronisations, with live HIP-event timers on the roofline pass only
    (timers on every pass would add two events per pass to the frame)."""
    import torch

    from violet_amd import _lib

    _lib.call("vgi_profiler_reset")
    _lib.call("vgi_profiler_only", only.encode())
    _lib.call("vgi_profiler_enable", 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(first + k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.call("vgi_profiler_enable", 0)
    _lib.call("vgi_profiler_only", None)
    return elapsed, _lib.profiler_read()


def run_frame(args):
    import torch

    from violet_amd import _lib, renderer, scene

    w, h = args.width, args.height
    log(f"[bench] building the {args.triangles}-triangle scene and its BVH")
    mesh = scene.make_scene(args.triangles, seed=SCENE_SEED)
    sc = scene.Scene(mesh)
    st = sc.stats()
    log(f"[bench] BVH: {st.num_nodes} nodes, depth {st.depth}, {st.device_bytes / 1e6:.1f} MB")
    cfg = renderer.RestirConfig(storage=_lib.STORAGE_F16 if args.storage == "f16" else _lib.STORAGE_F32,
                                atrous_iterations=args.iterations, concurrent_lanes=not args.no_lanes,
                                compose=args.compose or args.taa, taa=args.taa, **RESAMPLE[args.resample])
    r = renderer.RestirRenderer(w, h, cfg)
    # TAA: Halton(2,3) jitter with violet's cycle of 8 (restir_renderer.rs:65)
    fps = frame_params_seq(w, h, args.frames, jitter_cycle=8 if args.taa else 0)

    def step(k: int):
        r.add_passes(sc, fps[k % len(fps)], k)

    # the same frames on a one-stream renderer: per-pass times without the
    # overlap of the two pass chains, which pick the roofline pass
    iso_cfg = renderer.RestirConfig(**{**cfg.__dict__, "concurrent_lanes": False})
    r_iso = renderer.RestirRenderer(w, h, iso_cfg)

    def step_iso(k: int):
        r_iso.add_passes(sc, fps[k % len(fps)], k)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    if args.calibration_copy:
        calibration_copy()
    lat = []
    base = args.warmup
    for k in range(min(args.steps, 120)):
        t0 = time.perf_counter()
        step(base + k)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
    base += min(args.steps, 120)
    # per-pass times (every pass timed with HIP events; outside the timed region)
    prof = profile_passes(step, base, min(args.steps, 60))
    for k in range(args.warmup):
        step_iso(base + k)
    prof_iso = profile_passes(step_iso, base + args.warmup, min(args.steps, 60))
    r_iso.close()
    base += min(args.steps, 60)
    # traversal statistics: the instrumented trace kernels over a few more
    # frames of the same orbit (outside the timed region)
    _lib.call("vgi_profiler_reset")
    _lib.call("vgi_trace_stats_enable", 1)
    _lib.call("vgi_profiler_enable", 1)
    for k in range(args.count_frames):
        step(base + k)
    torch.cuda.synchronize()
    _lib.call("vgi_profiler_enable", 0)
    _lib.call("vgi_trace_stats_enable", 0)
    counts = _lib.profiler_read()
    base += args.count_frames
    trace_passes = {k for k, v in counts.items() if v["rays_closest"] + v["rays_any"] > 0}
    # with concurrent lanes every pass before the denoiser shares the device
    # with the other lane; the denoiser passes run alone unless composition
    # traces the sun shadow mask beside them (then only the final lighting
    # and TAA passes do)
    alone = ("final_lighting", "temporal_aa") + (() if cfg.compose else ("denoise_",))
    shared = {k for k in prof_iso if not k.startswith(alone)} if cfg.concurrent_lanes else set()
    dname = dominant_pass(prof_iso, trace_passes | shared)
    # the dependent-fetch rate over a buffer of the BVH's size (the trace passes' ceiling)
    import ctypes as C

    rate, rate_incoherent = C.c_double(), C.c_double()
    _lib.call("vgi_probe_dependent_fetch", int(st.device_bytes), 64, 64, C.byref(rate), None)
    _lib.call("vgi_probe_dependent_fetch", int(st.device_bytes), 64, 1, C.byref(rate_incoherent), None)
    elapsed, timed = timed_region(step, base, args.steps, dname)
    return dict(elapsed=elapsed, lat=lat, prof=prof, prof_iso=prof_iso, timed=timed, dname=dname, counts=counts,
                trace_passes=trace_passes, pixels=w * h, mesh=mesh, sc=sc, fetch_steps_per_s=rate.value,
                fetch_steps_per_s_incoherent=rate_incoherent.value,
                bvh_bytes=int(st.device_bytes))


def cpu_baseline_frame(args, mesh, budget_s: float = 12.0):
    """FP32 twin of the oracle (OpenMP) rendering the same orbit frames at the
    same size: frame 0 (no history) untimed, then frames with history for
    about `budget_s` seconds (at least two)."""
    import oracle as O

    threads = O.load(32).oracle_num_threads()
    w, h = args.cpu_width, args.cpu_height
    fps = frame_params_seq(w, h, args.frames)
    R = O.OracleRenderer(32, mesh, w, h, atrous_iterations=args.iterations, **RESAMPLE[args.resample])
    R.add_passes(fps[0], 0)
    t0 = time.perf_counter()
    n = 0
    while True:
        n += 1
        R.add_passes(fps[n % len(fps)], n)
        if time.perf_counter() - t0 > budget_s and n >= 2:
            break
    el = time.perf_counter() - t0
    return dict(value=n * w * h / el / 1e6, unit="Mpix/s", cores=threads, kind="port",
                sample=f"frames 1-{n} (GI trace, ReSTIR, hash grid, denoise; frame 0 untimed) of the "
                       f"{len(mesh.indices)}-triangle orbit workload at {w}x{h}, FP32 twin of the oracle "
                       f"(gcc -O3 -mavx2 -mfma, OpenMP), {el:.1f} s")


# ---------------------------------------------------------------------------
# Batched, serving and precision forms (reported beside the plain path)
# ---------------------------------------------------------------------------
def big_scene(args):
    from violet_amd import scene

    log(f"[bench] building the {args.big_triangles}-triangle scene of the batched/serving/precision forms")
    mesh = scene.make_scene(args.big_triangles, seed=SCENE_SEED)
    sc = scene.Scene(mesh)
    st = sc.stats()
    log(f"[bench] BVH: {st.num_nodes} nodes, depth {st.depth}, {st.device_bytes / 1e6:.1f} MB")
    return mesh, sc


def run_batched_form(args, sc, ntris):
    """16 views over the shared scene, FP16 storage: the plain path (each
    view's frame issued alone, one after the other) against one fused
    persistent launch per pass over all views, graph-replayed, with the
    launch shapes (sample-pass tile, a-trous block) and persistent group
    counts autotuned on these views; the default shapes are restored after."""
    import torch

    from violet_amd import _lib, autotune, batch, renderer, serve

    specs = serve.batch_views(args.batch_views, seed=21)
    px = sum(s.width * s.height for s in specs)
    cfg = renderer.RestirConfig(storage=_lib.STORAGE_F16, atrous_iterations=args.iterations, **RESAMPLE[args.resample])
    fps_cache = {}

    def fps_of(i, f):
        k = (i, f)
        if k not in fps_cache:
            fps_cache[k] = serve.view_frame_params(specs[i], f)
        return fps_cache[k]

    K = args.batch_frames
    # plain: the same views, one single-view frame after another
    views = [renderer.RestirRenderer(s.width, s.height, cfg) for s in specs]
    for f in range(3):
        for i, r in enumerate(views):
            r.add_passes(sc, fps_of(i, f), f)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(3, 3 + K):
        for i, r in enumerate(views):
            r.add_passes(sc, fps_of(i, f), f)
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t0) / K
    # per-pass GPU time per frame (all views), profiled outside the timed region
    _lib.call("vgi_profiler_reset")
    _lib.call("vgi_profiler_enable", 1)
    for f in range(3 + K, 6 + K):
        for i, r in enumerate(views):
            r.add_passes(sc, fps_of(i, f), f)
    torch.cuda.synchronize()
    _lib.call("vgi_profiler_enable", 0)
    plain_passes = {k: round(v["ms"] / 3, 3) for k, v in _lib.profiler_read().items()}
    for r in views:
        r.close()
    # batched: fresh views, autotune, then the timed frames
    views = [renderer.RestirRenderer(s.width, s.height, cfg) for s in specs]
    n = len(views)
    log(f"[bench] batched form: {n} views, {px / 1e6:.2f} Mpx per frame; autotuning")
    best, table, f0, chosen = autotune.tune_batch(sc, views, fps_of, frame0=0)
    b = batch.Batch(best)
    for f in range(f0, f0 + 3):
        b.add_passes(sc, views, [fps_of(i, f) for i in range(n)], [f] * n)
    torch.cuda.synchronize()
    f0 += 3
    t0 = time.perf_counter()
    for f in range(f0, f0 + K):
        b.add_passes(sc, views, [fps_of(i, f) for i in range(n)], [f] * n)
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t0) / K
    f0 += K
    lat = []
    for f in range(f0, f0 + min(K, 30)):
        t1 = time.perf_counter()
        b.add_passes(sc, views, [fps_of(i, f) for i in range(n)], [f] * n)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
    b.close()
    # per-pass GPU time per frame of the fused launches (same shapes, no graph)
    pb = batch.Batch(batch.BatchConfig(max_views=n, pixel_groups=best.pixel_groups, trace_groups=best.trace_groups,
                                       atrous_groups=best.atrous_groups, use_graph=False))
    f0 += min(K, 30)
    _lib.call("vgi_profiler_reset")
    _lib.call("vgi_profiler_enable", 1)
    for f in range(f0, f0 + 3):
        pb.add_passes(sc, views, [fps_of(i, f) for i in range(n)], [f] * n)
    torch.cuda.synchronize()
    _lib.call("vgi_profiler_enable", 0)
    batched_passes = {k: round(v["ms"] / 3, 3) for k, v in _lib.profiler_read().items()}
    pb.close()
    for r in views:
        r.close()
    _lib.set_tuning()  # default shapes for the other forms
    return {"workload": f"batched_{n}_views_f16", "views": n, "pixels_per_frame": px, "triangles": ntris,
            "resolutions": [[s.width, s.height] for s in specs], "frames": K,
            "plain_mpix_per_s": round(px / plain_s / 1e6, 2), "batched_mpix_per_s": round(px / batched_s / 1e6, 2),
            "speedup_vs_plain": round(plain_s / batched_s, 3), "plain_ms_per_frame": round(plain_s * 1e3, 3),
            "batched_ms_per_frame": round(batched_s * 1e3, 3),
            "batched_p50_ms": round(float(np.percentile(lat, 50)), 3),
            "batched_p99_ms": round(float(np.percentile(lat, 99)), 3),
            "graph": True, "autotuned": chosen,
            "autotune_columns": list(autotune.SHAPES + autotune.GROUPS) + ["median_ms", "spread_ms"],
            "autotune_table": [[c[k] for k in autotune.SHAPES + autotune.GROUPS] + [round(ms, 3), round(sp, 3)]
                               for c, ms, sp in table],
            "plain_pass_ms_per_frame": plain_passes, "batched_pass_ms_per_frame": batched_passes}


def run_serving_form(args, sc, ntris):
    """Poisson stream of view-render requests over persistent views of mixed
    resolution, continuously batched (violet_amd.serve)."""
    from violet_amd import _lib, batch, renderer, serve

    specs = serve.serve_views(args.serve_views, seed=31)
    cfg = renderer.RestirConfig(storage=_lib.STORAGE_F16, atrous_iterations=args.iterations, **RESAMPLE[args.resample])
    reqs = serve.poisson_requests(args.serve_requests, args.serve_rate, len(specs), seed=32)
    log(f"[bench] serving form: {len(reqs)} requests at {args.serve_rate:g}/s over {len(specs)} views")
    bcfg = batch.BatchConfig(trace_groups=args.serve_trace_groups, use_graph=args.serve_graph)
    srv = serve.Server(sc, specs, hbm_budget_bytes=args.serve_budget_gb * 1e9, config=cfg,
                       max_batch=args.serve_max_batch, streams=args.serve_streams, batch_config=bcfg,
                       max_batch_pixels=int(args.serve_max_pixels))
    admitted = srv.admit_all()
    stats = srv.run(reqs, timeout_s=600.0)
    srv.close()
    s = stats.summary()
    return {"workload": f"serving_{len(reqs)}_requests_{args.serve_rate:g}_per_s", "views": len(specs),
            "triangles": ntris, "hbm_budget_gb": args.serve_budget_gb, "max_batch": args.serve_max_batch,
            "p50_ms": round(s["p50_ms"], 3), "p99_ms": round(s["p99_ms"], 3), "mean_ms": round(s["mean_ms"], 3),
            "mpix_per_s": round(s["mpix_per_s"], 2), "wall_s": round(s["wall_s"], 3), "batches": s["batches"],
            "mean_batch": round(s["mean_batch"], 2), "evictions": s["evictions"],
            "graph": args.serve_graph, "trace_groups": args.serve_trace_groups, "streams": args.serve_streams,
            "max_batch_pixels": int(args.serve_max_pixels),
            "peak_hbm_gb": round(s["peak_hbm_gb"], 2), "admitted_up_front": admitted}


def run_precision_form(args, sc, ntris):
    """1920x1080 frame loop with FP16 storage (compensated moments) on the
    shared scene: throughput beside the same loop with FP32 storage."""
    import torch

    from violet_amd import _lib, renderer, serve

    spec = serve.ViewSpec(1920, 1080, 0.0)
    out = {"workload": "precision_1920x1080_f16", "triangles": ntris, "frames": args.precision_frames}
    for name, st in (("f16", _lib.STORAGE_F16), ("f32", _lib.STORAGE_F32)):
        r = renderer.RestirRenderer(1920, 1080, renderer.RestirConfig(storage=st, atrous_iterations=args.iterations,
                                                                      **RESAMPLE[args.resample]))
        fps = [serve.view_frame_params(spec, f) for f in range(args.precision_frames + 3)]
        for f in range(3):
            r.add_passes(sc, fps[f], f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(3, 3 + args.precision_frames):
            r.add_passes(sc, fps[f], f)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / args.precision_frames
        out[f"{name}_mpix_per_s"] = round(1920 * 1080 / el / 1e6, 2)
        out[f"{name}_ms_per_frame"] = round(el * 1e3, 3)
        r.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--workload", default="frame", choices=["frame", "denoise"])
    ap.add_argument("--width", type=int, default=None)
    ap.add_argument("--height", type=int, default=None)
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--iterations", type=int, default=None)
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16"])
    ap.add_argument("--cpu-width", type=int, default=None)
    ap.add_argument("--cpu-height", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--resample", default="reference", choices=sorted(RESAMPLE))
    ap.add_argument("--count-frames", type=int, default=8, help="frames run with traversal statistics")
    ap.add_argument("--no-lanes", action="store_true",
                    help="every pass on one stream (per-pass times without overlap between pass chains)")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="run plain 256 MiB device copies after warmup (scale for rocprofv3 --pmc counters)")
    ap.add_argument("--forms", default="batched,serving,precision",
                    help="comma list of forms reported beside the plain frame path ('' for none)")
    ap.add_argument("--compose", action="store_true",
                    help="also run the sun shadow mask and the final lighting in every frame")
    ap.add_argument("--taa", action="store_true",
                    help="composition plus temporal AA over a Halton-jittered orbit")
    ap.add_argument("--big-triangles", type=int, default=1_000_000)
    ap.add_argument("--batch-views", type=int, default=16)
    ap.add_argument("--batch-frames", type=int, default=30)
    ap.add_argument("--precision-frames", type=int, default=120)
    ap.add_argument("--serve-views", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=600.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-max-batch", type=int, default=16)
    ap.add_argument("--serve-graph", type=int, default=0, help="1: replay each batch slot's frame as a HIP graph")
    ap.add_argument("--serve-streams", type=int, default=2, help="batch slots (streams) the server keeps busy")
    ap.add_argument("--serve-max-pixels", type=float, default=4e6,
                    help="pixel cap of one batch (0 = none; the oldest queued request always goes)")
    ap.add_argument("--serve-trace-groups", type=int, default=0,
                    help="persistent groups of the fused trace passes (0 = one per tile)")
    args = ap.parse_args()
    frame = args.workload == "frame"
    # full frame: 1920x1080, 60-frame orbit, 3 a-trous iterations (the reference
    # default); denoise only: 1920x1080, 8-frame pan, 5 iterations.  The CPU
    # baseline renders the same frames at the same size (frame) or at 640x360
    # (denoise only)
    args.width = args.width or 1920
    args.height = args.height or 1080
    args.frames = args.frames or (60 if frame else 8)
    args.iterations = args.iterations if args.iterations is not None else (3 if frame else 5)
    args.cpu_width = args.cpu_width or (args.width if frame else 640)
    args.cpu_height = args.cpu_height or (args.height if frame else 360)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of this "
              "path is designed but not built", file=sys.stderr)
        sys.exit(2)

    r = run_frame(args) if frame else run_denoise(args)
    workload_name = (f"restir_gi_frame_{args.width}x{args.height}" if frame else
                     f"denoise_only_{args.width}x{args.height}")
    mpix = args.steps * r["pixels"] / r["elapsed"] / 1e6
    prof = r["prof"]
    counts = r.get("counts", {})
    trace_passes = r.get("trace_passes", set())
    # the roofline is the dominant pixel pass among those that run alone on
    # the device (chosen from the per-pass run, see dominant_pass), timed
    # live inside the timed region; trace passes are latency/divergence
    # bound and are reported in Mrays/s with their counted traversal bytes
    # and the dependent-fetch ceiling instead
    dname = r["dname"]
    d = r["timed"][dname]
    avg_ms = d["ms"] / max(d["count"], 1)
    bytes_per_launch = d["bytes"] / max(d["count"], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    lat = np.array(r["lat"])
    prof_iso = r.get("prof_iso", prof)

    def pass_table(p):
        return {k: dict(avg_us=round(v["ms"] / max(v["count"], 1) * 1e3, 2), launches=v["count"],
                        gbs=round(v["bytes"] / max(v["ms"] * 1e-3, 1e-12) / 1e9, 1),
                        share=round(v["ms"] / max(sum(x["ms"] for x in p.values()), 1e-12), 4))
                for k, v in sorted(p.items(), key=lambda kv: -kv[1]["ms"])}

    passes = pass_table(prof_iso)
    trace = {}
    for k in sorted(trace_passes, key=lambda n: -prof_iso.get(n, {"ms": 0})["ms"]):
        c, t = counts[k], prof_iso.get(k)
        if not t or not c["count"]:
            continue
        rays = (c["rays_closest"] + c["rays_any"]) / c["count"]
        tb = (c["nodes"] * _lib_const("BVH_NODE_BYTES") + c["tris"] * _lib_const("BVH_TRI_BYTES")) / c["count"]
        us = t["ms"] / max(t["count"], 1) * 1e3
        trace[k] = dict(avg_us=round(us, 2), rays_per_launch=round(rays), mrays_per_s=round(rays / us, 1),
                        closest_share=round(c["rays_closest"] / max(c["rays_closest"] + c["rays_any"], 1), 3),
                        nodes_per_ray=round(c["nodes"] / max(c["rays_closest"] + c["rays_any"], 1), 2),
                        tris_per_ray=round(c["tris"] / max(c["rays_closest"] + c["rays_any"], 1), 2),
                        traversal_bytes_per_ray=round(tb / max(rays, 1), 1),
                        traversal_gbs=round(tb / (us * 1e-6) / 1e9, 1))
    trace_roofline = None
    if trace:
        tk = next(iter(trace))
        tr_ = trace[tk]
        c = counts[tk]
        rays = c["rays_closest"] + c["rays_any"]
        steps_per_ray = (c["nodes"] + c["tris"]) / max(rays, 1)
        achieved_steps = tr_["mrays_per_s"] * 1e6 * steps_per_ray
        ceiling = r["fetch_steps_per_s"]
        trace_roofline = {
            "kernel": tk, "bound": "dependent-fetch latency",
            "achieved": round(achieved_steps / 1e9, 3), "peak": round(ceiling / 1e9, 3), "unit": "Gfetch/s",
            "frac": round(achieved_steps / ceiling, 4),
            "fetches_per_ray": round(steps_per_ray, 2), "mrays_per_s": tr_["mrays_per_s"],
            "ceiling_mrays_per_s": round(ceiling / steps_per_ray / 1e6, 1),
            "incoherent_gfetch_per_s": round(r["fetch_steps_per_s_incoherent"] / 1e9, 3),
            "model": (f"each ray makes one dependent round trip per interior node ({_lib_const('BVH_NODE_BYTES')} B) "
                      f"and per triangle tested ({_lib_const('BVH_TRI_BYTES')} B); the peak is vgi_probe_dependent_fetch's rate of dependent 128-byte "
                      "fetches with every lane of the chip busy and the 64 lanes of each wave on one chain "
                      f"(perfectly coherent rays) through {r['bvh_bytes'] / 1e6:.0f} MB (the scene's BVH size), "
                      "launched like the trace kernels; incoherent_gfetch_per_s is the same with every lane "
                      "on its own chain"),
        }
    if args.no_cpu_baseline:
        cpu = None
    else:
        cpu = cpu_baseline_frame(args, r["mesh"]) if frame else cpu_baseline_denoise(args)
    out = {
        "metric": "ReSTIR GI+denoise Mpix/s @1080p; p50/p99 frame latency, batched vs plain",
        "value": round(mpix, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["elapsed"] / args.steps * 1e3, 4),
        "p50_ms": round(float(np.percentile(lat, 50)), 4),
        "p99_ms": round(float(np.percentile(lat, 99)), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else "f16-storage/f32-math",
        "data": ("synthetic (seeded boxes + icospheres on a plane, sun + constant sky, 60-frame camera orbit)"
                 if frame else
                 "synthetic (seeded plane + 20 spheres G-buffer, clean field x Exp(1) noise, 2 px/frame pan)"),
        "config": ({"workload": f"restir_gi_frame_{args.width}x{args.height}", "width": args.width,
                    "height": args.height, "triangles": int(len(r["mesh"].indices)), "frames_cycled": args.frames,
                    "gi_rays_per_px": "1 diffuse + 1 specular (+1 shadow each)", "atrous_iterations": args.iterations,
                    "restir_params": RESAMPLE_DESC[args.resample],
                    "storage": args.storage, "parallelism": "single-gpu",
                    "concurrent_lanes": not args.no_lanes, "compose": args.compose or args.taa,
                    "taa": args.taa} if frame else
                   {"workload": f"denoise_only_{args.width}x{args.height}", "width": args.width,
                    "height": args.height, "frames_cycled": args.frames, "atrous_iterations": args.iterations,
                    "storage": args.storage, "parallelism": "single-gpu"}),
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(workload_name, dname),
                     "avg_launch_us": round(avg_ms * 1e3, 3), "algorithmic_bytes_per_launch": bytes_per_launch},
        "passes": passes,
        "passes_note": ("per-pass HIP-event times from a separate run of the same frames on one stream (the "
                        "frame itself runs the hash-grid update beside the G-buffer and specular beside diffuse "
                        "on two streams); the roofline pass is the pixel pass with the largest share here among "
                        "those that run alone on the device in the frame (the denoiser passes: the others share it "
                        "with the other stream), timed live in the timed region" if frame else
                        "per-pass HIP-event times from a separate run (the timed region times the roofline "
                        "pass only), one stream"),
        "trace": trace,
        "trace_roofline": trace_roofline,
        "cpu_baseline": cpu,
    }
    forms = [f for f in args.forms.split(",") if f] if frame else []
    if forms:
        if args.big_triangles == args.triangles:
            mesh, sc = r["mesh"], r["sc"]
        else:
            mesh, sc = big_scene(args)
        ntris = int(len(mesh.indices))
        runners = {"batched": run_batched_form, "serving": run_serving_form, "precision": run_precision_form}
        for f in forms:
            log(f"[bench] form: {f}")
            out[f] = runners[f](args, sc, ntris)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
