"""A glTF scene through the whole path on the GPU: the seeded test scene
(untextured, and textured: texture coordinates, tangents and PNG or JPEG
base colour with PNG metallic-roughness / normal images) written as .glb (Y up),
imported by vgi_gltf_import (Z up, world space, decoded images with
Lanczos3 mips) and rendered with composition; the frame's depth, G-buffer,
sun shadow mask and first-frame GI samples equal the FP32 twin's on the
imported arrays."""
from __future__ import annotations

import numpy as np
import pytest

from gltf_fixture import scene_from_mesh
from restir_harness import make_mesh, run_gpu_frames, seeds_for_frame

pytestmark = pytest.mark.gpu

W, H = 96, 64


@pytest.mark.parametrize("textured", ["none", "png", "jpeg"])
def test_glb_scene_renders_like_the_twin(tmp_path, textured):
    import oracle
    from violet_amd import renderer, scene

    src = make_mesh(textured=textured != "none")
    path = scene_from_mesh(src, jpeg=textured == "jpeg").save_glb(tmp_path / "scene.glb")
    mesh, st = scene.load_gltf(path)
    assert st.triangles == src.num_triangles and st.instances == 1
    if textured != "none":
        assert st.textures >= len(src.textures) and st.images_unsupported == 0 and mesh.tangents is not None
    # Y-up -> Z-up by the f32 rotation of pi/2: the positions come back to rounding
    np.testing.assert_allclose(mesh.positions[mesh.indices], src.positions[src.indices], atol=1e-5)
    cfg = renderer.RestirConfig(compose=True)
    fps, outs, snaps, r, sc = run_gpu_frames(mesh, W, H, 2, config=cfg)
    o32 = oracle.OracleRestir(32, oracle.OracleScene(32, mesh))
    for k in range(2):
        s, fp = snaps[k], fps[k]
        depth, gb = o32.gbuffer(fp, W, H)
        assert np.array_equal(s["depth"].view(np.uint32), depth.view(np.uint32)), f"depth[{k}]"
        assert np.array_equal(s["gbuffer"], gb), f"gbuffer[{k}]"
        assert np.array_equal(s["shadow_mask"], o32.raytraced_shadow(fp, s["depth"])), f"shadow mask[{k}]"
        assert (depth > 0).mean() > 0.3
    q = np.zeros((128 * 1024, 6), np.uint32)
    cnt = np.zeros(1, np.uint32)
    tr = o32.trace_res(np.zeros((65536 * 16, 4), np.uint32), q, cnt)
    hit, rad = o32.diff_sample_gen(fps[0], tr, snaps[0]["gbuffer"], snaps[0]["depth"],
                                   seeds_for_frame(0, False, False)["diff_gen"], False)
    assert np.array_equal(snaps[0]["diff_new_hit_pos_normal"], hit)
    assert np.array_equal(snaps[0]["diff_new_hit_radiance"].view(np.uint32), rad.view(np.uint32))
