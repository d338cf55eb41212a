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


def _onion(n_quads, seed):
    """n_quads unit quads stacked along z (slightly shifted in x and y), so a
    ray along z overlaps every box of the BVH."""
    from violet_amd import scene

    rng = np.random.default_rng(seed)
    z = np.sort(rng.uniform(0.0, 10.0, n_quads))
    j = rng.uniform(-0.2, 0.2, (n_quads, 2))
    c = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float64)
    xy = c[None, :, :] + j[:, None, :]
    pos = np.concatenate([xy, np.broadcast_to(z[:, None, None], (n_quads, 4, 1))], 2).reshape(-1, 3)
    base = 4 * np.arange(n_quads, dtype=np.uint32)[:, None]
    idx = np.concatenate([base + np.array([0, 1, 2], np.uint32), base + np.array([0, 2, 3], np.uint32)], 1)
    n_t = 2 * n_quads
    return scene.Mesh(positions=pos.astype(np.float32), normals=np.tile(np.array([0, 0, 1], np.float32), (4 * n_quads, 1)),
                      indices=idx.reshape(-1, 3), tri_material=np.zeros(n_t, np.uint32), albedo=np.full((1, 3), 0.5),
                      metallic=np.zeros(1), roughness=np.full(1, 0.5))


def test_deep_traversal_stack_spills_exactly(vgi):
    """Rays through 25000 stacked quads hit every box on their way down, so
    their traversal stacks grow past the 16 entries kept in LDS (to 21 on this
    scene, by a host simulation of the traversal) into private memory; hits
    must still equal the brute-force loop bit for bit, also for oblique rays."""
    import oracle
    from violet_amd import scene

    mesh = _onion(25000, 4)
    sc = scene.Scene(mesh)
    rng = np.random.default_rng(4)
    m = 4000
    below = rng.random(m) < 0.5
    o = np.column_stack([rng.uniform(-0.8, 0.8, m), rng.uniform(-0.8, 0.8, m), np.where(below, -1.0, 11.0)])
    d = np.column_stack([rng.normal(0, 0.05, m), rng.normal(0, 0.05, m), np.where(below, 1.0, -1.0)])
    d[: m // 4] = rng.standard_normal((m // 4, 3))  # a quarter in random directions
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((m, 1), 5e-4), d, np.full((m, 1), 1000.0)], 1).astype(np.float32)
    g_thv, g_id = _gpu_trace(sc, rays, False)
    o_thv, o_id = oracle.trace(32, mesh.positions, mesh.indices, rays, brute=True)
    assert np.array_equal(g_id, o_id), f"hit ids differ on {np.mean(g_id != o_id):.3%} of rays"
    hit = g_id != 0xFFFFFFFF
    assert hit.mean() > 0.7
    assert np.array_equal(g_thv[hit].view(np.uint32), o_thv[hit].astype(np.float32).view(np.uint32))
    assert np.array_equal(_gpu_trace(sc, rays, True)[1] != 0xFFFFFFFF, hit)


def test_traversal_statistics_count_exactly_and_change_nothing(vgi):
    """The instrumented traversal (trace statistics) returns the same hits and
    counts what the BVH implies: with a 2-triangle scene the root is a leaf,
    so (wrapped in one interior node) every ray aimed at it visits exactly 1
    node and tests exactly 2 triangles."""
    from violet_amd import _lib, scene

    tiny = scene.Mesh(positions=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32),
                      normals=np.tile(np.array([0, 0, 1], np.float32), (4, 1)),
                      indices=np.array([[0, 1, 2], [0, 1, 3]], np.uint32), tri_material=np.zeros(2, np.uint32),
                      albedo=np.full((1, 3), 0.5), metallic=np.zeros(1), roughness=np.full(1, 0.5))
    big = scene.make_scene(4000, seed=5, extent=8.0)
    rng = np.random.default_rng(9)
    o = rng.standard_normal((5000, 3))
    o = 4.0 * o / np.linalg.norm(o, axis=1, keepdims=True) + 0.3
    d = 0.3 - o  # towards the centre of the tiny mesh's box
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    aimed = np.concatenate([o, np.full((5000, 1), 5e-4), d, np.full((5000, 1), 1000.0)], 1).astype(np.float32)
    for mesh, exact in ((tiny, True), (big, False)):
        sc = scene.Scene(mesh)
        rays = aimed if exact else _rays(5000, 9, extent=8.0)
        ref = [_gpu_trace(sc, rays, a) for a in (False, True)]
        _lib.call("vgi_profiler_reset")
        _lib.call("vgi_trace_stats_enable", 1)
        _lib.call("vgi_profiler_enable", 1)
        try:
            got = [_gpu_trace(sc, rays, a) for a in (False, True)]
        finally:
            _lib.call("vgi_profiler_enable", 0)
            _lib.call("vgi_trace_stats_enable", 0)
        prof = _lib.profiler_read()
        for (rt, ri), (gt, gi) in zip(ref, got):
            assert np.array_equal(ri, gi) and np.array_equal(rt.view(np.uint32), gt.view(np.uint32))
        c, a = prof["trace_rays_closest"], prof["trace_rays_any"]
        assert c["rays_closest"] == len(rays) and c["rays_any"] == 0
        assert a["rays_any"] == len(rays) and a["rays_closest"] == 0
        if exact:
            assert c["nodes"] == len(rays) and c["tris"] == 2 * len(rays)
        else:
            assert c["nodes"] > len(rays) and c["tris"] > 0
            assert a["nodes"] <= c["nodes"]  # any-hit stops early


def test_zero_direction_components_keep_culling(vgi):
    """Rays along the axes (two direction components exactly 0) hit exactly
    what the brute-force loop hits, and the box tests keep culling on every
    axis: they visit no more nodes than the same rays tilted by 1e-6 (an
    infinite inverse direction would make a fused plane distance
    inf - inf = NaN, which drops an axis out of the slab test)."""
    import oracle
    from violet_amd import _lib, scene

    mesh = scene.make_scene(4000, seed=7, extent=8.0)
    sc = scene.Scene(mesh)
    rng = np.random.default_rng(17)
    n = 6000
    o = rng.uniform(-8.0, 8.0, (n, 3))
    o[:, 2] = rng.uniform(0.05, 6.0, n)
    axis = rng.integers(0, 3, n)
    d = np.zeros((n, 3))
    d[np.arange(n), axis] = rng.choice([-1.0, 1.0], n)
    tilt = d + 1e-6 * (d == 0)
    tilt /= np.linalg.norm(tilt, axis=1, keepdims=True)

    def rays(dd):
        return np.concatenate([o, np.full((n, 1), 5e-4), dd, np.full((n, 1), 1000.0)], 1).astype(np.float32)

    counts = []
    for dd in (d, tilt):
        rr = rays(dd)
        _lib.call("vgi_profiler_reset")
        _lib.call("vgi_trace_stats_enable", 1)
        _lib.call("vgi_profiler_enable", 1)
        try:
            thv, ids = _gpu_trace(sc, rr, False)
        finally:
            _lib.call("vgi_profiler_enable", 0)
            _lib.call("vgi_trace_stats_enable", 0)
        counts.append(_lib.profiler_read()["trace_rays_closest"]["nodes"])
        o_thv, o_id = oracle.trace(32, mesh.positions, mesh.indices, rr, brute=True)
        assert np.array_equal(ids, o_id)
        hit = ids != 0xFFFFFFFF
        assert np.array_equal(thv[hit].view(np.uint32), o_thv[hit].astype(np.float32).view(np.uint32))
    assert counts[0] <= 1.2 * counts[1], f"axis-aligned rays visit {counts[0]} nodes vs {counts[1]} tilted"
