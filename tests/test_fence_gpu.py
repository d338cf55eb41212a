"""Frame fence with a deadline (vgi_fence_*): the counterpart of violet's
5 s wait on the oldest frame's fence (crates/violet-app/src/render_loop.rs:
445-482), which reports a time-out and carries on.  The time-outs here are
forced from the host (a zero deadline while a frame is still running); every
kernel finishes normally."""
from __future__ import annotations

import pytest

from restir_harness import make_mesh, orbit_view

pytestmark = pytest.mark.gpu


def test_fence_times_out_then_completes(vgi):
    import torch

    from violet_amd import _lib, gpu, renderer, scene, synth

    f = gpu.Fence()
    assert f.wait(0)[0]  # never recorded: nothing to wait for
    sc = scene.Scene(make_mesh())
    w, h = 1920, 1080
    r = renderer.RestirRenderer(w, h)
    torch.cuda.synchronize()
    prev = None
    for k in range(6):  # several milliseconds of work on the stream
        v = orbit_view(k, w, h)
        r.add_passes(sc, synth.make_frame_params(v, prev), k)
        prev = v
    f.record()
    done, waited = f.wait(0)
    assert not done, "six 1080p frames finished before the first poll"
    assert "did not finish within 0 ms" in _lib.lib().vgi_last_error().decode()
    done, waited = f.wait(gpu.Fence.TIMEOUT_MS)
    assert done and 0.0 <= waited < gpu.Fence.TIMEOUT_MS
    assert f.wait(0)[0]  # waiting again on finished work returns at once
    torch.cuda.synchronize()
    r.close()
    f.close()
