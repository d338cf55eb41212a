"""ctypes mirrors of the ReSTIR / hash-grid / renderer structs of include/vgi.h
and registration of their entry points."""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import FrameParams

vp = C.c_void_p
u32 = C.c_uint32
f32 = C.c_float
P = C.POINTER


def _struct(name, fields):
    return type(name, (C.Structure,), {"_fields_": fields})


WH = [("width", u32), ("height", u32), ("storage", u32)]

TraceResources = _struct("TraceResources", [
    ("scene_tlas", vp), ("sky_radiance", f32 * 3), ("storage", u32),
    ("prev_indirect_diffuse_texture", vp), ("prev_depth_texture", vp), ("prev_width", u32), ("prev_height", u32),
    ("hash_grid_storage_buffer", vp), ("rw_hash_grid_query_buffer", vp), ("rw_hash_grid_query_counter_buffer", vp),
    ("hash_grid_query_threshold", vp)])

DiffSampleGenIO = _struct("DiffSampleGenIO", WH + [
    ("gbuffer_color", vp), ("gbuffer_depth", vp), ("rw_hit_pos_normal_texture", vp), ("rw_hit_radiance_texture", vp)])
PCSampleGen = _struct("PCSampleGen", [("frame_rand", u32), ("has_prev_frame", u32)])
DiffSampleValidateIO = _struct("DiffSampleValidateIO", WH + [
    ("depth_texture", vp), ("rw_reservoir_texture", vp), ("rw_hit_pos_normal_texture", vp),
    ("rw_hit_radiance_texture", vp)])
SpecSampleGenIO = _struct("SpecSampleGenIO", WH + [
    ("gbuffer_color", vp), ("gbuffer_depth", vp), ("rw_hit_pos_texture", vp), ("rw_hit_radiance_texture", vp)])
SpecSampleValidateIO = _struct("SpecSampleValidateIO", WH + [
    ("rw_prev_reservoir_texture", vp), ("rw_prev_hit_pos_texture", vp), ("rw_prev_hit_radiance_texture", vp)])
AOGenIO = _struct("AOGenIO", WH + [
    ("hit_pos_normal_texture", vp), ("depth_buffer", vp), ("prev_ao_texture", vp), ("rw_ao_texture", vp)])
PCAOGen = _struct("PCAOGen", [("has_new_sample", u32), ("has_prev_frame", u32), ("radius_ws", f32)])
AOFilterIO = _struct("AOFilterIO", WH + [("depth_buffer", vp), ("ao_texture", vp), ("rw_filtered_ao_texture", vp)])
DiffTemporalIO = _struct("DiffTemporalIO", WH + [
    ("prev_gbuffer_depth", vp), ("gbuffer_depth", vp), ("new_hit_pos_normal_texture", vp),
    ("new_hit_radiance_texture", vp), ("prev_reservoir_texture", vp), ("prev_hit_pos_normal_texture", vp),
    ("prev_hit_radiance_texture", vp), ("rw_reservoir_texture", vp), ("rw_hit_pos_normal_texture", vp),
    ("rw_hit_radiance_texture", vp)])
PCDiffTemporal = _struct("PCDiffTemporal", [("frame_rand", u32), ("has_prev_frame", u32), ("has_new_sample", u32)])
ResampleParams = _struct("ResampleParams", [("temporal_m_cap", u32), ("diff_spatial_taps", u32),
                                            ("spatial_radius_px", f32)])
DiffSpatialIO = _struct("DiffSpatialIO", WH + [
    ("gbuffer_color", vp), ("gbuffer_depth", vp), ("ao_texture", vp), ("temp_reservoir_texture", vp),
    ("temp_hit_pos_normal_texture", vp), ("temp_hit_radiance_texture", vp), ("rw_reservoir_texture", vp),
    ("rw_hit_pos_normal_texture", vp), ("rw_hit_radiance_texture", vp), ("rw_lighting_texture", vp)])
PCFrameRand = _struct("PCFrameRand", [("frame_rand", u32)])
PCDiffSpatial = PCFrameRand
PCSpecSpatial = PCFrameRand
SpecTemporalIO = _struct("SpecTemporalIO", WH + [
    ("prev_gbuffer_color", vp), ("prev_gbuffer_depth", vp), ("gbuffer_color", vp), ("gbuffer_depth", vp),
    ("prev_reservoir_texture", vp), ("prev_hit_pos_texture", vp), ("prev_hit_radiance_texture", vp),
    ("rw_reservoir_texture", vp), ("rw_hit_pos_texture", vp), ("rw_hit_radiance_texture", vp)])
PCSpecTemporal = _struct("PCSpecTemporal", [("frame_rand", u32), ("has_prev_frame", u32)])
SpecSpatialIO = _struct("SpecSpatialIO", WH + [
    ("gbuffer_color", vp), ("gbuffer_depth", vp), ("reservoir_texture", vp), ("hit_pos_texture", vp),
    ("hit_radiance_texture", vp), ("rw_lighting_texture", vp), ("rw_ray_len_texture", vp)])
HGDecayIO = _struct("HGDecayIO", [("rw_decay_buffer", vp), ("rw_storage_buffer", vp)])
HGSelectIO = _struct("HGSelectIO", [
    ("hash_grid_query_buffer", vp), ("hash_grid_query_counter_buffer", vp), ("rw_hash_grid_storage_buffer", vp),
    ("rw_hash_grid_query_selection_buffer", vp)])
PCFrameHash = _struct("PCFrameHash", [("frame_hash", u32)])
HGCompactIO = _struct("HGCompactIO", [
    ("hash_grid_query_selection_buffer", vp), ("rw_compact_query_index_buffer", vp),
    ("rw_compact_query_cell_addr_buffer", vp), ("rw_compact_query_counter_buffer", vp)])
HGRaygenIO = _struct("HGRaygenIO", [
    ("rw_hash_grid_storage_buffer", vp), ("hash_grid_query_buffer", vp), ("compact_query_counter_buffer", vp),
    ("compact_query_index_buffer", vp), ("compact_query_cell_addr_buffer", vp), ("rw_hash_grid_decay_buffer", vp)])
HGOSelectIO = _struct("HGOSelectIO", [
    ("hash_grid_query_buffer", vp), ("hash_grid_query_counter_buffer", vp), ("rw_hash_grid_storage_buffer", vp),
    ("rw_hash_grid_query_selection_buffer", vp), ("rw_hash_grid_query_winner_buffer", vp),
    ("rw_hash_grid_query_threshold", vp)])
HG_COMPACT_WORKSPACE_WORDS = 257
HGOCompactIO = _struct("HGOCompactIO", [
    ("rw_hash_grid_query_selection_buffer", vp), ("rw_hash_grid_query_winner_buffer", vp),
    ("rw_compact_query_index_buffer", vp), ("rw_compact_query_cell_addr_buffer", vp),
    ("rw_compact_query_counter_buffer", vp), ("rw_workspace_buffer", vp)])
HGORaygenIO = _struct("HGORaygenIO", [("raygen", HGRaygenIO), ("rw_radiance_staging_buffer", vp)])
GBufferIO = _struct("GBufferIO", [("width", u32), ("height", u32), ("rw_depth", vp), ("rw_gbuffer", vp)])
RaytracedShadowIO = _struct("RaytracedShadowIO", [("width", u32), ("height", u32), ("gbuffer_depth", vp),
                                                  ("rw_shadow", vp)])
FinalLightingIO = _struct("FinalLightingIO", WH + [
    ("gbuffer_color", vp), ("shadow_mask_buffer", vp), ("indirect_diffuse_texture", vp),
    ("indirect_specular_texture", vp), ("rw_color_buffer", vp), ("sky_radiance", f32 * 3)])
TemporalAAIO = _struct("TemporalAAIO", WH + [
    ("gbuffer_depth", vp), ("source_texture", vp), ("history_texture", vp), ("rw_target", vp)])
PCTemporalAA = _struct("PCTemporalAA", [("has_history", u32)])

RendererConfig = _struct("RendererConfig", [
    ("storage", u32), ("denoise", u32), ("ind_diff_validate", u32), ("ind_spec_validate", u32),
    ("hash_grid_cache_decay", u32), ("ao_radius", f32), ("sky_radiance", f32 * 3),
    ("temporal_m_cap", u32), ("diff_spatial_taps", u32), ("spatial_radius_px", f32), ("concurrent_lanes", u32),
    ("compose", u32), ("taa", u32), ("denoiser", _lib.DenoiserConfig), ("ordered_cache_update", u32)])
RendererOutput = _struct("RendererOutput", [
    ("indirect_diffuse", vp), ("indirect_specular", vp), ("depth", vp), ("gbuffer", vp), ("denoised", u32),
    ("validation_frame", u32), ("hash_grid_updated", u32), ("color", vp)])

BatchConfig = _struct("BatchConfig", [("max_views", u32), ("pixel_groups", u32), ("trace_groups", u32),
                                     ("atrous_groups", u32), ("use_graph", u32), ("concurrent_lanes", u32)])

FPp = P(FrameParams)
TRp = P(TraceResources)
_SIGS = {
    "vgi_pass_ind_diff_sample_gen": [FPp, TRp, P(DiffSampleGenIO), P(PCSampleGen), vp],
    "vgi_pass_ind_diff_sample_validate": [FPp, TRp, P(DiffSampleValidateIO), vp],
    "vgi_pass_ind_spec_sample_gen": [FPp, TRp, P(SpecSampleGenIO), P(PCSampleGen), vp],
    "vgi_pass_ind_spec_sample_validate": [FPp, TRp, P(SpecSampleValidateIO), vp],
    "vgi_pass_raytraced_ao_gen": [FPp, P(AOGenIO), P(PCAOGen), vp],
    "vgi_pass_raytraced_ao_filter": [P(AOFilterIO), vp],
    "vgi_pass_ind_diff_temporal_resample": [FPp, P(DiffTemporalIO), P(PCDiffTemporal), P(ResampleParams), vp],
    "vgi_pass_ind_diff_spatial_resample": [FPp, P(DiffSpatialIO), P(PCDiffSpatial), P(ResampleParams), vp],
    "vgi_pass_ind_spec_temporal_resample": [FPp, P(SpecTemporalIO), P(PCSpecTemporal), P(ResampleParams), vp],
    "vgi_pass_ind_spec_spatial_resample": [FPp, P(SpecSpatialIO), P(PCSpecSpatial), P(ResampleParams), vp],
    "vgi_pass_hash_grid_decay": [P(HGDecayIO), vp],
    "vgi_pass_hash_grid_query_select": [P(HGSelectIO), P(PCFrameHash), vp],
    "vgi_pass_hash_grid_query_compact": [P(HGCompactIO), vp],
    "vgi_pass_hash_grid_raygen": [FPp, TRp, P(HGRaygenIO), P(PCFrameHash), vp],
    "vgi_pass_hash_grid_query_select_ordered": [P(HGOSelectIO), P(PCFrameHash), vp],
    "vgi_pass_hash_grid_query_compact_ordered": [P(HGOCompactIO), vp],
    "vgi_pass_hash_grid_raygen_ordered": [FPp, TRp, P(HGORaygenIO), P(PCFrameHash), vp],
    "vgi_pass_clear_buffer": [vp, u32, u32, vp],
    "vgi_pass_gbuffer_primary": [FPp, vp, P(GBufferIO), vp],
    "vgi_pass_raytraced_shadow": [FPp, vp, P(RaytracedShadowIO), vp],
    "vgi_pass_final_lighting": [FPp, P(FinalLightingIO), vp],
    "vgi_pass_temporal_aa": [FPp, P(TemporalAAIO), P(PCTemporalAA), vp],
    "vgi_renderer_default_config": [P(RendererConfig)],
    "vgi_renderer_create": [u32, u32, P(RendererConfig), P(vp)],
    "vgi_renderer_destroy": [vp],
    "vgi_renderer_reset": [vp],
    "vgi_renderer_add_passes": [vp, vp, FPp, u32, P(RendererOutput), vp],
    "vgi_renderer_get_image": [vp, C.c_char_p, P(vp), P(C.c_size_t)],
    "vgi_renderer_export_state": [vp, vp, C.c_size_t, vp],
    "vgi_renderer_import_state": [vp, vp, C.c_size_t, vp],
    "vgi_batch_default_config": [P(BatchConfig)],
    "vgi_batch_create": [P(BatchConfig), P(vp)],
    "vgi_batch_destroy": [vp],
    "vgi_batch_add_passes": [vp, vp, u32, P(vp), FPp, P(u32), P(RendererOutput), vp],
}
for _n, _a in _SIGS.items():
    _lib.register(_n, C.c_int32, _a)
_lib.register("vgi_renderer_state_size", C.c_size_t, [vp])
_lib.register("vgi_renderer_device_bytes", C.c_size_t, [vp])
_lib.register("vgi_hash_grid_query_threshold_for", u32, [C.c_uint64])
