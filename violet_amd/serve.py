"""Serving loop: view-render requests of mixed resolutions, continuously
batched onto the GPU.

No reference counterpart: violet renders one interactive view
(crates/violet-restir/src/restir_render_loop.rs:262-268).  Here each request
asks for the next frame of one of many persistent views.  A view's state
(reservoirs, denoiser history, radiance cache) stays resident in HBM between
its requests; a view is allocated on its first request if the HBM budget
allows (least-recently-used idle views are evicted otherwise) and reset on
re-admission after an eviction; admit_all() makes the views resident up
front instead, as far as the budget allows.

Scheduling (continuous batching): requests join a FIFO queue when they
arrive.  Whenever one of `streams` batch slots is free, the scheduler takes
up to `max_batch` queued requests in arrival order - at most one per view
and none for a view that still has a frame in flight, since a view's frames
are sequential, and, with `max_batch_pixels`, no more pixels than that
(the oldest request always goes) - and issues them as one fused multi-view frame
(violet_amd.batch) on that slot's stream.  Requests that arrive while other
batches run are admitted as soon as a slot frees, so the work of different
batches overlaps on the device.  A request's latency runs from its arrival
to the host observing its batch's completion fence (vgi_fence).  As violet's
frame fence (crates/violet-app/src/render_loop.rs:445-482), a batch still
running `frame_timeout_ms` after its issue is counted as timed out and
waited for further, and one finishing after 33.3 ms as slow.

Metrics: with `metrics` (a path or an open text file) the server appends
one JSON line per completed batch (slot, issue time, batch time, views,
pixels, each request's latency, whether it timed out) and a summary line
at the end of run() (the fields of ServeStats.summary()).
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import batch, gpu, renderer, synth
from .scene import Scene


@dataclass
class Request:
    rid: int
    view: int
    arrival_s: float


@dataclass
class ViewSpec:
    width: int
    height: int
    orbit_phase_deg: float
    orbit_radius: float = 30.0
    orbit_height: float = 9.0


@dataclass
class ServeStats:
    latencies_ms: list = field(default_factory=list)
    pixels: int = 0
    wall_s: float = 0.0
    batches: int = 0
    evictions: int = 0
    peak_bytes: int = 0
    timeouts: int = 0       # batches still running frame_timeout_ms after their issue
    slow_batches: int = 0   # batches that took longer than 33.3 ms (violet's warning threshold)

    def summary(self) -> dict:
        lat = np.array(self.latencies_ms) if self.latencies_ms else np.zeros(1)
        return dict(requests=len(self.latencies_ms), p50_ms=float(np.percentile(lat, 50)),
                    p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
                    mpix_per_s=self.pixels / max(self.wall_s, 1e-9) / 1e6, wall_s=self.wall_s,
                    batches=self.batches, mean_batch=len(self.latencies_ms) / max(self.batches, 1),
                    evictions=self.evictions, peak_hbm_gb=self.peak_bytes / 1e9, timeouts=self.timeouts,
                    slow_batches=self.slow_batches)


def view_frame_params(spec: ViewSpec, frame: int, deg_per_frame: float = 6.0):
    """Camera of frame `frame` of a view: an orbit around the scene centre."""

    def view_at(k):
        a = math.radians(spec.orbit_phase_deg + deg_per_frame * k)
        pos = np.array([spec.orbit_radius * math.cos(a), spec.orbit_radius * math.sin(a), spec.orbit_height])
        d = -pos / np.linalg.norm(pos)
        yaw = math.degrees(math.atan2(d[0], d[1]))
        pitch = math.degrees(math.asin(d[2]))
        return synth.View(synth.Camera(pos, yaw, pitch), spec.width, spec.height)

    return synth.make_frame_params(view_at(frame), view_at(frame - 1) if frame > 0 else None)


class Server:
    def __init__(self, scene: Scene, views: list[ViewSpec], hbm_budget_bytes: float,
                 config: renderer.RestirConfig | None = None, max_batch: int = 16, streams: int = 2,
                 batch_config: batch.BatchConfig | None = None, on_complete=None, max_batch_pixels: int = 0,
                 frame_timeout_ms: int = gpu.Fence.TIMEOUT_MS, metrics=None):
        import torch

        gpu.require_gpu()
        self.torch = torch
        self.scene = scene
        self.views = views
        self.budget = hbm_budget_bytes
        self.config = config or renderer.RestirConfig()
        self.max_batch = max_batch
        self.max_batch_pixels = max_batch_pixels  # 0: no limit
        self.frame_timeout_ms = frame_timeout_ms
        self.on_complete = on_complete
        # JSON-lines metrics stream: a path or a writable text file, or None
        self.metrics = open(metrics, "a") if isinstance(metrics, (str, bytes)) or hasattr(metrics, "__fspath__") \
            else metrics
        self._own_metrics = self.metrics is not metrics
        bc = batch_config or batch.BatchConfig()
        bc.max_views = max(bc.max_views, max_batch)
        self.slots = [(batch.Batch(bc), torch.cuda.Stream()) for _ in range(streams)]
        self.fences = [gpu.Fence() for _ in range(streams)]
        self.busy = [None] * streams           # per slot: [issue time, [(req, outputs)], timed out]
        self.resident: dict[int, renderer.RestirRenderer] = {}
        self.frame_of: dict[int, int] = collections.defaultdict(int)
        self.last_use: dict[int, float] = {}
        self.inflight_views: set[int] = set()
        self.bytes_resident = 0
        self.view_bytes: dict[int, int] = {}
        # device bytes of a view are affine in its pixel count: measure the
        # two coefficients once on two small views
        a = renderer.RestirRenderer(8, 8, self.config)
        b = renderer.RestirRenderer(16, 16, self.config)
        sa, sb = a.device_bytes(), b.device_bytes()
        a.close()
        b.close()
        self.bytes_per_px = (sb - sa) / (256 - 64)
        self.bytes_fixed = sa - 64 * self.bytes_per_px

    # -- memory ---------------------------------------------------------------
    def _ensure_resident(self, v: int, stats: ServeStats, pinned: set) -> bool:
        if v in self.resident:
            return True
        spec = self.views[v]
        need = self._estimate_bytes(spec)
        while self.bytes_resident + need > self.budget:
            idle = [u for u in self.resident if u not in self.inflight_views and u not in pinned]
            if not idle:
                return False
            u = min(idle, key=lambda k: self.last_use.get(k, 0.0))
            self.resident.pop(u).close()
            self.bytes_resident -= self.view_bytes.pop(u)
            self.frame_of[u] = 0  # history is gone: the view restarts
            stats.evictions += 1
        r = renderer.RestirRenderer(spec.width, spec.height, self.config)
        self.resident[v] = r
        self.view_bytes[v] = r.device_bytes()
        self.bytes_resident += self.view_bytes[v]
        stats.peak_bytes = max(stats.peak_bytes, self.bytes_resident)
        return True

    def _estimate_bytes(self, spec: ViewSpec) -> int:
        # per-view images scale with the pixel count; the radiance cache
        # (2^20 cells and its query buffers) is fixed
        return int(math.ceil(self.bytes_fixed + self.bytes_per_px * spec.width * spec.height))

    # -- loop -----------------------------------------------------------------
    def _poll(self, now: float, stats: ServeStats) -> None:
        for k, b in enumerate(self.busy):
            if b is None:
                continue
            done, _ = self.fences[k].wait(0)
            if not done:
                if not b[2] and (now - b[0]) * 1e3 > self.frame_timeout_ms:
                    b[2] = True  # violet reports the time-out and waits on
                    stats.timeouts += 1
                continue
            t_done = time.perf_counter()
            batch_ms = (t_done - self.t0 - b[0]) * 1e3
            if batch_ms > 33.3:
                stats.slow_batches += 1
            if self.metrics is not None:
                self._emit({"event": "batch", "slot": k, "t_issue_s": round(b[0], 6), "batch_ms": round(batch_ms, 4),
                            "views": [req.view for req, _ in b[1]],
                            "pixels": sum(self.views[req.view].width * self.views[req.view].height for req, _ in b[1]),
                            "latency_ms": [round((t_done - self.t0 - req.arrival_s) * 1e3, 4) for req, _ in b[1]],
                            "timed_out": bool(b[2])})
            for req, outs in b[1]:
                stats.latencies_ms.append((t_done - self.t0 - req.arrival_s) * 1e3)
                spec = self.views[req.view]
                stats.pixels += spec.width * spec.height
                self.inflight_views.discard(req.view)
                self.last_use[req.view] = t_done
                if self.on_complete:
                    self.on_complete(req, self.resident[req.view], outs)
            self.busy[k] = None

    def _issue(self, k: int, queue: collections.deque, now: float, stats: ServeStats) -> None:
        picked, seen, keep = [], set(), collections.deque()
        pixels = 0
        while queue and len(picked) < self.max_batch:
            req = queue.popleft()
            spec = self.views[req.view]
            px = spec.width * spec.height
            over = self.max_batch_pixels and picked and pixels + px > self.max_batch_pixels
            if over or req.view in self.inflight_views or req.view in seen or \
                    not self._ensure_resident(req.view, stats, seen):
                keep.append(req)
                continue
            picked.append(req)
            seen.add(req.view)
            pixels += px
        queue.extendleft(reversed(keep))
        if not picked:
            return
        bt, stream = self.slots[k]
        rs = [self.resident[r.view] for r in picked]
        fps = [view_frame_params(self.views[r.view], self.frame_of[r.view]) for r in picked]
        idx = [self.frame_of[r.view] for r in picked]
        with self.torch.cuda.stream(stream):
            outs = bt.add_passes(self.scene, rs, fps, idx, stream=stream.cuda_stream)
            self.fences[k].record(stream.cuda_stream)
        for r in picked:
            self.frame_of[r.view] += 1
            self.inflight_views.add(r.view)
        self.busy[k] = [now, list(zip(picked, outs)), False]
        stats.batches += 1

    def admit_all(self, stats: ServeStats | None = None) -> int:
        """Make every view resident up front, in view order, while the HBM
        budget allows (the rest are admitted on their first request).  A
        view's device state is allocated once and kept, so this moves the
        allocations out of the request path.  Returns the resident count."""
        stats = stats or ServeStats()
        for v in range(len(self.views)):
            if v in self.resident:
                continue
            if self.bytes_resident + self._estimate_bytes(self.views[v]) > self.budget:
                break
            self._ensure_resident(v, stats, set())
        self.torch.cuda.synchronize()
        return len(self.resident)

    def run(self, requests: list[Request], timeout_s: float = 120.0) -> ServeStats:
        stats = ServeStats()
        stats.peak_bytes = self.bytes_resident
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival_s))
        queue: collections.deque = collections.deque()
        self.t0 = time.perf_counter()
        while pending or queue or any(self.busy):
            now = time.perf_counter() - self.t0
            if now > timeout_s:
                raise TimeoutError(f"serving did not finish within {timeout_s} s")
            while pending and pending[0].arrival_s <= now:
                queue.append(pending.popleft())
            self._poll(now, stats)
            issued = False
            for k in range(len(self.slots)):
                if self.busy[k] is None and queue:
                    before = stats.batches
                    self._issue(k, queue, now, stats)
                    issued = issued or stats.batches != before
            if not issued:
                time.sleep(50e-6)
        self.torch.cuda.synchronize()
        stats.wall_s = time.perf_counter() - self.t0
        if self.metrics is not None:
            self._emit(dict({"event": "summary"}, **stats.summary()))
            self.metrics.flush()
        return stats

    def _emit(self, rec: dict) -> None:
        import json

        self.metrics.write(json.dumps(rec) + "\n")

    def close(self) -> None:
        if self.metrics is not None and self._own_metrics:
            self.metrics.close()
        self.metrics = None
        for r in self.resident.values():
            r.close()
        self.resident.clear()
        for b, _ in self.slots:
            b.close()
        for f in self.fences:
            f.close()


def poisson_requests(n: int, rate_hz: float, num_views: int, seed: int) -> list[Request]:
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    v = rng.integers(0, num_views, n)
    return [Request(i, int(v[i]), float(t[i])) for i in range(n)]


SERVE_RESOLUTIONS = [(256, 256), (640, 360), (1280, 720), (1920, 1080), (2560, 1440)]
SERVE_WEIGHTS = [0.3, 0.25, 0.2, 0.15, 0.1]


def serve_views(num_views: int, seed: int) -> list[ViewSpec]:
    """Each persistent view draws its resolution once from the serving mix."""
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(SERVE_RESOLUTIONS), num_views, p=SERVE_WEIGHTS)
    phase = rng.uniform(0.0, 360.0, num_views)
    return [ViewSpec(*SERVE_RESOLUTIONS[p], float(a)) for p, a in zip(pick, phase)]


def batch_views(num_views: int, seed: int) -> list[ViewSpec]:
    """Views of the batched form: seeded resolutions w ~ U{320..1920},
    h ~ U{240..1080}, one orbit camera each."""
    rng = np.random.default_rng(seed)
    ws, hs = rng.integers(320, 1921, num_views), rng.integers(240, 1081, num_views)
    ph = rng.uniform(0.0, 360.0, num_views)
    return [ViewSpec(int(w), int(h), float(p)) for w, h, p in zip(ws, hs, ph)]


__all__ = ["Request", "ViewSpec", "Server", "ServeStats", "poisson_requests", "serve_views", "batch_views",
           "view_frame_params", "SERVE_RESOLUTIONS", "SERVE_WEIGHTS"]
