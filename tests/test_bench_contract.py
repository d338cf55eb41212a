"""CPU checks of bench.py's reporting rules: the roofline pass choice, the
traffic lookup for every pixel pass, and the one-GPU contract."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parent.parent


def _prof(**ms):
    return {k: {"ms": v, "count": 1, "bytes": 1.0} for k, v in ms.items()}


def test_roofline_pass_skips_trace_and_shared_passes():
    prof = _prof(ind_spec_sample_gen=9.0, ind_diff_spatial_resample=3.0, denoise_atrous=2.5, denoise_temporal=1.0)
    assert bench.dominant_pass(prof, set()) == "ind_spec_sample_gen"
    assert bench.dominant_pass(prof, {"ind_spec_sample_gen"}) == "ind_diff_spatial_resample"
    # with concurrent lanes every pass before the denoiser shares the device
    shared = {k for k in prof if not k.startswith("denoise_")}
    assert bench.dominant_pass(prof, {"ind_spec_sample_gen"} | shared) == "denoise_atrous"


def test_every_pixel_pass_has_its_measured_traffic():
    workload = "restir_gi_frame_1920x1080"
    data = json.loads((ROOT / "profiles" / f"traffic_{workload}.json").read_text())
    kernels = {k.split("<")[0].split("::")[-1] for k in data["kernels"]}
    for pass_name, kernel in bench.PASS_KERNEL.items():
        if pass_name.startswith("hash_grid_"):
            continue  # the frame runs the order-independent cache update (k_hgo_*)
        assert kernel in kernels, (pass_name, kernel)
        t = bench.measured_traffic(workload, pass_name)
        assert t is not None and t > 0, pass_name


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
