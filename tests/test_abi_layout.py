"""Every ctypes mirror of a struct in include/vgi.h has the C layout: a C
program compiled with gcc against the header prints sizeof and the offset of
each field the mirror names, and they must equal ctypes' (a field added on
one side only, or in another order, fails here instead of corrupting
arguments at run time)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _mirrors():
    from violet_amd import _abi, _lib, scene

    m = {
        "vgi_profile_entry": _lib.ProfileEntry, "vgi_frame_params": _lib.FrameParams,
        "vgi_denoise_temporal_io": _lib.DenoiseTemporalIO, "vgi_pc_denoise_temporal": _lib.PCDenoiseTemporal,
        "vgi_denoise_spatial_io": _lib.DenoiseSpatialIO, "vgi_denoise_fast_clamp_io": _lib.DenoiseFastClampIO,
        "vgi_denoise_atrous_io": _lib.DenoiseAtrousIO, "vgi_pc_denoise_atrous": _lib.PCDenoiseAtrous,
        "vgi_denoiser_config": _lib.DenoiserConfig, "vgi_denoiser_input": _lib.DenoiserInput,
        "vgi_denoiser_output": _lib.DenoiserOutput, "vgi_material": scene.Material, "vgi_texture": scene.TextureDesc,
        "vgi_scene_desc": scene.SceneDesc, "vgi_scene_stats": scene.SceneStats,
        "vgi_gltf_config": scene.GltfConfig, "vgi_gltf_stats": scene.GltfStats,
        "vgi_trace_resources": _abi.TraceResources, "vgi_ind_diff_sample_gen_io": _abi.DiffSampleGenIO,
        "vgi_pc_sample_gen": _abi.PCSampleGen, "vgi_ind_diff_sample_validate_io": _abi.DiffSampleValidateIO,
        "vgi_ind_spec_sample_gen_io": _abi.SpecSampleGenIO, "vgi_ind_spec_sample_validate_io": _abi.SpecSampleValidateIO,
        "vgi_ao_gen_io": _abi.AOGenIO, "vgi_pc_ao_gen": _abi.PCAOGen, "vgi_ao_filter_io": _abi.AOFilterIO,
        "vgi_ind_diff_temporal_io": _abi.DiffTemporalIO, "vgi_pc_ind_diff_temporal": _abi.PCDiffTemporal,
        "vgi_ind_diff_spatial_io": _abi.DiffSpatialIO, "vgi_pc_frame_rand": _abi.PCFrameRand,
        "vgi_pc_ind_diff_spatial": _abi.PCDiffSpatial, "vgi_pc_ind_spec_spatial": _abi.PCSpecSpatial,
        "vgi_ind_spec_temporal_io": _abi.SpecTemporalIO, "vgi_pc_ind_spec_temporal": _abi.PCSpecTemporal,
        "vgi_ind_spec_spatial_io": _abi.SpecSpatialIO, "vgi_hash_grid_decay_io": _abi.HGDecayIO,
        "vgi_hash_grid_query_select_io": _abi.HGSelectIO, "vgi_pc_frame_hash": _abi.PCFrameHash,
        "vgi_hash_grid_query_compact_io": _abi.HGCompactIO, "vgi_hash_grid_raygen_io": _abi.HGRaygenIO,
        "vgi_gbuffer_io": _abi.GBufferIO, "vgi_renderer_config": _abi.RendererConfig,
        "vgi_raytraced_shadow_io": _abi.RaytracedShadowIO, "vgi_final_lighting_io": _abi.FinalLightingIO,
        "vgi_temporal_aa_io": _abi.TemporalAAIO, "vgi_pc_temporal_aa": _abi.PCTemporalAA,
        "vgi_renderer_output": _abi.RendererOutput, "vgi_batch_config": _abi.BatchConfig,
        "vgi_resample_params": _abi.ResampleParams, "vgi_denoise_frame_io": _lib.DenoiseFrameIO,
        "vgi_hash_grid_ordered_select_io": _abi.HGOSelectIO, "vgi_hash_grid_ordered_compact_io": _abi.HGOCompactIO,
        "vgi_hash_grid_ordered_raygen_io": _abi.HGORaygenIO, "vgi_tuning": _lib.Tuning,
    }
    return m


# The byte size of every HLSL PushConstants block a pass entry point takes
# (4-byte fields, PushConstantsBuilder shader.rs:1436-1449): a violet caller
# passes exactly these bytes.
HLSL_PUSH_CONSTANTS = {
    "vgi_pc_frame_hash": 4,         # hash_grid_query_select.hlsl:16-20, hash_grid_raygen.hlsl:19-22
    "vgi_pc_sample_gen": 8,         # ind_diff_sample_gen.hlsl:17-23, ind_spec_sample_gen.hlsl:21-26
    "vgi_pc_ao_gen": 12,            # raytraced_ao_gen.hlsl:10-16
    "vgi_pc_ind_diff_temporal": 12,  # ind_diff_temporal_resample.hlsl:31-36
    "vgi_pc_ind_diff_spatial": 4,   # ind_diff_spatial_resample.hlsl:45-49
    "vgi_pc_ind_spec_temporal": 8,  # ind_spec_temporal_resample.hlsl:39-44
    "vgi_pc_ind_spec_spatial": 4,   # ind_spec_spatial_resample.hlsl:31-35
    "vgi_pc_denoise_temporal": 16,  # svgf_temporal_filter.hlsl:41-47
    "vgi_pc_denoise_atrous": 4,     # svgf_atrous_filter.hlsl:17-21
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_push_constants_are_the_hlsl_size(tmp_path):
    lines = ["#include <stdio.h>", '#include "vgi.h"', "int main(void) {"]
    for cname in HLSL_PUSH_CONSTANTS:
        lines.append(f'    printf("%zu\\n", sizeof({cname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "pc.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "pc"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert dict(zip(HLSL_PUSH_CONSTANTS, got)) == HLSL_PUSH_CONSTANTS


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_mirrors_match_c_layout(tmp_path):
    import ctypes as C

    mirrors = _mirrors()
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "vgi.h"', "int main(void) {"]
    expect = []
    for cname, py in mirrors.items():
        lines.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        expect.append(f"{cname} {C.sizeof(py)}")
        for f in py._fields_:
            lines.append(f'    printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
            expect.append(f"{cname}.{f[0]} {getattr(py, f[0]).offset}")
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")[:-1]
    bad = [(e, g) for e, g in zip(expect, got) if e != g]
    assert len(got) == len(expect) and not bad, f"layout mismatches (ctypes, C): {bad[:10]}"
    # every struct the header declares has a mirror
    import re

    declared = set(re.findall(r"typedef struct (vgi_\w+) \{", (ROOT / "include" / "vgi.h").read_text()))
    unmirrored = declared - set(mirrors) - {"vgi_hash_grid_cell", "vgi_hash_grid_query"}
    assert not unmirrored, f"structs without a ctypes mirror: {sorted(unmirrored)}"


def test_tuning_rejects_unknown_shapes():
    """vgi_tuning_set takes 8, 16 or 32 for either shape (0 = the default)
    and rejects anything else without touching the current setting."""
    import ctypes as C

    from violet_amd import _lib

    lib = _lib.lib()
    _lib.set_tuning()
    for ab, tw in ((12, 8), (16, 64), (4, 4)):
        assert lib.vgi_tuning_set(C.byref(_lib.Tuning(ab, tw))) == _lib.VGI_E_INVALID
        assert _lib.get_tuning() == dict(atrous_block=16, trace_tile_width=8)
    _lib.set_tuning(0, 0)
    assert _lib.get_tuning() == dict(atrous_block=16, trace_tile_width=8)
    _lib.set_tuning(32, 16)
    assert _lib.get_tuning() == dict(atrous_block=32, trace_tile_width=16)
    _lib.set_tuning()
