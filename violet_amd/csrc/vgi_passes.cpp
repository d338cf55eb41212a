/*
 * vgi_passes.cpp — the direct sink: one pass record -> its C entry point.
 */
#include "vgi_passes.h"

namespace vgi {

vgi_status Lanes::finish() {
    if (!side || !side_work) return VGI_OK;
    side_work = false;
    VGI_TRY_HIP(hipEventRecord(ev_join, side));
    VGI_TRY_HIP(hipStreamWaitEvent(main, ev_join, 0));
    return VGI_OK;
}

vgi_status Lanes::enter(int slot, hipStream_t& st) {
    st = main;
    if (!side) return VGI_OK;
    const int ph = slot_phase(slot);
    if (ph != phase) {
        vgi_status j = finish();
        if (j != VGI_OK) return j;
        VGI_TRY_HIP(hipEventRecord(ev_fork, main));
        VGI_TRY_HIP(hipStreamWaitEvent(side, ev_fork, 0));
        phase = ph;
        cross_pending = false;
    }
    if (slot_lane(slot)) {
        st = side;
        side_work = true;
    }
    if (slot_waits(slot) && cross_pending && cross_from != st) {
        VGI_TRY_HIP(hipStreamWaitEvent(st, ev_cross, 0));
        cross_pending = false;
    }
    return VGI_OK;
}

vgi_status Lanes::leave(int slot, hipStream_t st) {
    if (!side || !ev_cross || !slot_signals(slot)) return VGI_OK;
    VGI_TRY_HIP(hipEventRecord(ev_cross, st));
    cross_from = st;  // may be the null stream: cross_pending says that ev_cross is set
    cross_pending = true;
    return VGI_OK;
}

vgi_status DirectSink::submit(int slot, const PassRec& r) {
    hipStream_t st;
    vgi_status e = lanes.enter(slot, st);
    if (e == VGI_OK) e = launch(r, st);
    if (e == VGI_OK) e = lanes.leave(slot, st);
    return e;
}

vgi_status DirectSink::launch(const PassRec& r, hipStream_t st) {
    void* s = (void*)st;
    switch (r.kind) {
        case kPassClear:
            if (r.clear_value) {
                VGI_REQUIRE(r.clear_bytes % 4 == 0 && ((uintptr_t)r.clear_ptr & 3u) == 0,
                            "internal: word fill of an unaligned buffer");
                return fill_u32((uint32_t*)r.clear_ptr, r.clear_bytes / 4, r.clear_value, st);
            }
            return fill_zero(r.clear_ptr, r.clear_bytes, st);
        case kPassGBuffer: return vgi_pass_gbuffer_primary(&r.fp, r.scene, &r.io.gbuffer, s);
        case kPassHgDecay: return vgi_pass_hash_grid_decay(&r.io.hg_decay, s);
        case kPassHgSelect: return vgi_pass_hash_grid_query_select(&r.io.hg_select, &r.pc.frame_hash, s);
        case kPassHgCompact: return vgi_pass_hash_grid_query_compact(&r.io.hg_compact, s);
        case kPassHgRaygen: return vgi_pass_hash_grid_raygen(&r.fp, &r.tr, &r.io.hg_raygen, &r.pc.frame_hash, s);
        case kPassHgSelectOrdered:
            return vgi_pass_hash_grid_query_select_ordered(&r.io.hg_oselect, &r.pc.frame_hash, s);
        case kPassHgCompactOrdered: return vgi_pass_hash_grid_query_compact_ordered(&r.io.hg_ocompact, s);
        case kPassHgRaygenOrdered:
            return vgi_pass_hash_grid_raygen_ordered(&r.fp, &r.tr, &r.io.hg_oraygen, &r.pc.frame_hash, s);
        case kPassDiffGen: return vgi_pass_ind_diff_sample_gen(&r.fp, &r.tr, &r.io.diff_gen, &r.pc.sample_gen, s);
        case kPassDiffValidate: return vgi_pass_ind_diff_sample_validate(&r.fp, &r.tr, &r.io.diff_validate, s);
        case kPassAoGen: return vgi_pass_raytraced_ao_gen(&r.fp, &r.io.ao_gen, &r.pc.ao_gen, s);
        case kPassAoFilter: return vgi_pass_raytraced_ao_filter(&r.io.ao_filter, s);
        case kPassDiffTemporal:
            return launch_diff_temporal(&r.fp, &r.io.diff_temporal, &r.pc.diff_temporal, &r.resample, r.diff_taps,
                                        r.diff_taps_gbuffer, st);
        case kPassDiffSpatial:
            return launch_diff_spatial(&r.fp, &r.io.diff_spatial, &r.pc.diff_spatial, &r.resample, r.diff_taps, st);
        case kPassSpecGen: return vgi_pass_ind_spec_sample_gen(&r.fp, &r.tr, &r.io.spec_gen, &r.pc.sample_gen, s);
        case kPassSpecValidate: return vgi_pass_ind_spec_sample_validate(&r.fp, &r.tr, &r.io.spec_validate, s);
        case kPassSpecTemporal:
            return vgi_pass_ind_spec_temporal_resample(&r.fp, &r.io.spec_temporal, &r.pc.spec_temporal, &r.resample, s);
        case kPassSpecSpatial:
            return vgi_pass_ind_spec_spatial_resample(&r.fp, &r.io.spec_spatial, &r.pc.spec_spatial, &r.resample, s);
        case kPassDenTemporal: return vgi_pass_denoise_temporal(&r.fp, &r.io.den_temporal, &r.pc.den_temporal, s);
        case kPassDenDisocclusion: return vgi_pass_denoise_disocclusion_fix(&r.io.den_spatial, s);
        case kPassDenFastClamp: return vgi_pass_denoise_fast_history_clamp(&r.io.den_fast_clamp, s);
        case kPassDenFallback: return vgi_pass_denoise_spatial_fallback(&r.io.den_spatial, s);
        case kPassDenAtrous: return vgi_pass_denoise_atrous(&r.io.den_atrous, &r.pc.den_atrous, s);
        case kPassShadow: return vgi_pass_raytraced_shadow(&r.fp, r.scene, &r.io.shadow, s);
        case kPassFinalLighting: return vgi_pass_final_lighting(&r.fp, &r.io.final_lighting, s);
        case kPassTemporalAA: return vgi_pass_temporal_aa(&r.fp, &r.io.taa, &r.pc.taa, s);
        default: return set_error(VGI_E_INVALID, "internal: unknown pass kind %d", r.kind);
    }
}

}  // namespace vgi
