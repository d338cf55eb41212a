"""BVH traversal parity: GPU closest-hit ids, t and barycentrics must equal
the FP32 twin's brute-force loop over all triangles bit for bit (the BVH is
only an accelerator), and agree with the FP64 oracle's ids except at
numerically tied edges; any-hit answers equal."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(n, seed, extent=40.0, axis_aligned=0.1):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-extent, extent, (n, 3))
    o[:, 2] = rng.uniform(0.05, 6.0, n)
    d = rng.standard_normal((n, 3))
    k = int(n * axis_aligned)
    d[:k] = 0.0
    d[np.arange(k), rng.integers(0, 3, k)] = rng.choice([-1.0, 1.0], k)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmin = np.full((n, 1), 5e-4)
    tmax = np.where(rng.random((n, 1)) < 0.2, rng.uniform(0.1, 20.0, (n, 1)), 1000.0)
    return np.concatenate([o, tmin, d, tmax], 1).astype(np.float32)


def _gpu_trace(sc, rays, any_hit):
    import torch

    from violet_amd import gpu

    r = gpu.dev(rays)
    hits = torch.zeros(len(rays) * 4, dtype=torch.int32, device="cuda")
    sc.trace(r, len(rays), any_hit, hits)
    torch.cuda.synchronize()
    h = hits.cpu().numpy().view(np.uint32).reshape(-1, 4)
    return h[:, :3].view(np.float32), h[:, 3]


@pytest.mark.parametrize("tris,seed", [(600, 1), (4000, 2)])
def test_closest_hit_bit_exact_vs_brute_force(vgi, tris, seed):
    import oracle
    from violet_amd import scene

    mesh = scene.make_scene(tris, seed=seed, extent=8.0)
    sc = scene.Scene(mesh)
    rays = _rays(20000, seed, extent=8.0)
    g_thv, g_id = _gpu_trace(sc, rays, False)
    o_thv, o_id = oracle.trace(32, mesh.positions, mesh.indices, rays, brute=True)
    assert np.array_equal(g_id, o_id), f"hit ids differ on {np.mean(g_id != o_id):.3%} of rays"
    hit = g_id != 0xFFFFFFFF
    assert hit.mean() > 0.2
    assert np.array_equal(g_thv[hit].view(np.uint32), o_thv[hit].astype(np.float32).view(np.uint32))
    d_thv, d_id = oracle.trace(64, mesh.positions, mesh.indices, rays, brute=True)
    agree = np.mean(d_id == g_id)
    assert agree > 0.999, f"FP64 id agreement {agree:.5f}"
    both = hit & (d_id == g_id)
    assert np.allclose(g_thv[both, 0], d_thv[both, 0], rtol=1e-4, atol=1e-5)
    # any-hit: same occlusion answer
    g_any = _gpu_trace(sc, rays, True)[1] != 0xFFFFFFFF
    assert np.array_equal(g_any, hit)


def test_closest_hit_large_scene_vs_oracle_bvh(vgi):
    import oracle
    from violet_amd import scene

    mesh = scene.make_scene(200_000, seed=11)
    sc = scene.Scene(mesh)
    st = sc.stats()
    assert st.depth <= 64
    rays = _rays(50000, 5)
    g_thv, g_id = _gpu_trace(sc, rays, False)
    o_thv, o_id = oracle.trace(32, mesh.positions, mesh.indices, rays, brute=False)
    assert np.array_equal(g_id, o_id), f"hit ids differ on {np.mean(g_id != o_id):.3%} of rays"
    hit = g_id != 0xFFFFFFFF
    assert np.array_equal(g_thv[hit].view(np.uint32), o_thv[hit].astype(np.float32).view(np.uint32))
