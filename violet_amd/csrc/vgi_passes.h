/*
 * vgi_passes.h — pass records: the arguments of one pass invocation as the
 * sequencers (renderer, denoiser) produce them, and the sinks that consume
 * them.  A DirectSink launches each pass at once through its C entry point
 * (the single-view path); a recording sink keeps them per sequence slot so a
 * batch can fuse the same slot of many views into one persistent launch.
 *
 * Slots follow restir_renderer.rs:258-829 and denoising.rs:66-301; a view's
 * frame issues a subsequence of them, in this order.
 */
#pragma once
#include <string.h>

#include "vgi_internal.h"

namespace vgi {

enum PassKind : int {
    kPassClear,            // fill a device buffer with zeros (util_passes.rs clear_buffer)
    kPassGBuffer,
    kPassHgDecay,
    kPassHgSelect,
    kPassHgCompact,
    kPassHgRaygen,
    kPassDiffGen,
    kPassDiffValidate,
    kPassAoGen,
    kPassAoFilter,
    kPassDiffTemporal,
    kPassDiffSpatial,
    kPassSpecGen,
    kPassSpecValidate,
    kPassSpecTemporal,
    kPassSpecSpatial,
    kPassDenTemporal,
    kPassDenDisocclusion,
    kPassDenFastClamp,
    kPassDenFallback,
    kPassDenAtrous,
    kPassShadow,          // raytraced_shadow.hlsl
    kPassFinalLighting,   // restir/final_lighting.hlsl
    kPassTemporalAA,      // temporal_aa.hlsl
    kPassHgSelectOrdered,   // order-independent cache update (no reference counterpart)
    kPassHgCompactOrdered,
    kPassHgRaygenOrdered,
    kPassKindCount
};

enum Slot : int {
    kSlotGBuffer,
    kSlotClearHgStorage,
    kSlotClearHgDecay,
    kSlotClearHgCounter,
    kSlotFillHgThreshold,
    kSlotHgDecay,
    kSlotClearSelection,
    kSlotHgSelect,
    kSlotClearWinner,
    kSlotHgSelectOrdered,
    kSlotClearCompactCount,
    kSlotHgCompact,
    kSlotHgCompactOrdered,
    kSlotClearNextCounter,
    kSlotHgRaygen,
    kSlotHgRaygenOrdered,
    kSlotSpecValidate,  // issued ahead of the diffuse passes (see slot_lane)
    kSlotDiffGen,
    kSlotDiffValidate,
    kSlotAoGen,
    kSlotAoFilter,
    kSlotDiffTemporal,
    kSlotDiffSpatial,
    kSlotSpecGen,
    kSlotSpecTemporal,
    kSlotSpecSpatial,
    kSlotDenClearFastDiff,
    kSlotDenClearFastSpec,
    kSlotDenTemporal,
    kSlotDenDisocclusion,
    kSlotDenFastClamp,
    kSlotDenFallback,
    kSlotDenAtrous0,  // .. kSlotDenAtrous0 + 4
    kSlotShadow = kSlotDenAtrous0 + 5,
    kSlotFinalLighting,
    kSlotTemporalAA,
    kSlotCount
};

// The pass kind each slot carries.
inline int slot_kind(int slot) {
    static const int k[kSlotCount] = {
        kPassGBuffer,      kPassClear,        kPassClear,         kPassClear,       kPassClear,       kPassHgDecay,
        kPassClear,        kPassHgSelect,     kPassClear,         kPassHgSelectOrdered, kPassClear, kPassHgCompact,
        kPassHgCompactOrdered, kPassClear,    kPassHgRaygen,      kPassHgRaygenOrdered,
        kPassSpecValidate, kPassDiffGen,      kPassDiffValidate,  kPassAoGen,
        kPassAoFilter,     kPassDiffTemporal, kPassDiffSpatial,   kPassSpecGen,     kPassSpecTemporal,
        kPassSpecSpatial,  kPassClear,        kPassClear,         kPassDenTemporal, kPassDenDisocclusion,
        kPassDenFastClamp, kPassDenFallback,  kPassDenAtrous,     kPassDenAtrous,   kPassDenAtrous,
        kPassDenAtrous,    kPassDenAtrous,    kPassShadow,        kPassFinalLighting, kPassTemporalAA};
    return k[slot];
}

struct PassRec {
    int kind;
    vgi_frame_params fp;
    vgi_trace_resources tr;
    const vgi_scene* scene;
    union {
        vgi_gbuffer_io gbuffer;
        vgi_hash_grid_decay_io hg_decay;
        vgi_hash_grid_query_select_io hg_select;
        vgi_hash_grid_query_compact_io hg_compact;
        vgi_hash_grid_raygen_io hg_raygen;
        vgi_hash_grid_ordered_select_io hg_oselect;
        vgi_hash_grid_ordered_compact_io hg_ocompact;
        vgi_hash_grid_ordered_raygen_io hg_oraygen;
        vgi_ind_diff_sample_gen_io diff_gen;
        vgi_ind_diff_sample_validate_io diff_validate;
        vgi_ao_gen_io ao_gen;
        vgi_ao_filter_io ao_filter;
        vgi_ind_diff_temporal_io diff_temporal;
        vgi_ind_diff_spatial_io diff_spatial;
        vgi_ind_spec_sample_gen_io spec_gen;
        vgi_ind_spec_sample_validate_io spec_validate;
        vgi_ind_spec_temporal_io spec_temporal;
        vgi_ind_spec_spatial_io spec_spatial;
        vgi_denoise_temporal_io den_temporal;
        vgi_denoise_spatial_io den_spatial;  // disocclusion fix and spatial fallback
        vgi_denoise_fast_clamp_io den_fast_clamp;
        vgi_denoise_atrous_io den_atrous;
        vgi_raytraced_shadow_io shadow;
        vgi_final_lighting_io final_lighting;
        vgi_temporal_aa_io taa;
    } io;
    union {
        vgi_pc_frame_hash frame_hash;
        vgi_pc_sample_gen sample_gen;
        vgi_pc_ao_gen ao_gen;
        vgi_pc_ind_diff_temporal diff_temporal;
        vgi_pc_ind_diff_spatial diff_spatial;
        vgi_pc_ind_spec_temporal spec_temporal;
        vgi_pc_ind_spec_spatial spec_spatial;
        vgi_pc_denoise_temporal den_temporal;
        vgi_pc_denoise_atrous den_atrous;
        vgi_pc_temporal_aa taa;
    } pc;
    vgi_resample_params resample;  // the four resampling passes (zero = reference constants)
    // diffuse temporal / spatial of the renderer's own frames: the tap
    // records the temporal pass writes (with the G-buffer it reads layer 1
    // from) and the spatial pass reads; NULL for the plain entry points
    DiffTap* diff_taps;
    const uint32_t* diff_taps_gbuffer;
    void* clear_ptr;     // kPassClear
    size_t clear_bytes;  // kPassClear
    uint32_t clear_value;  // kPassClear: the 32-bit word every word of the buffer gets (clear_bytes % 4 == 0 unless 0)
};

inline PassRec make_rec(int kind, const vgi_frame_params* fp = nullptr) {
    PassRec r;
    memset(&r, 0, sizeof(r));
    r.kind = kind;
    if (fp) r.fp = *fp;
    return r;
}
inline PassRec make_fill(void* p, size_t bytes, uint32_t value) {
    PassRec r = make_rec(kPassClear);
    r.clear_ptr = p;
    r.clear_bytes = bytes;
    r.clear_value = value;
    return r;
}
inline PassRec make_clear(void* p, size_t bytes) {
    PassRec r = make_rec(kPassClear);
    r.clear_ptr = p;
    r.clear_bytes = bytes;
    return r;
}

struct PassSink {
    virtual ~PassSink() {}
    virtual vgi_status submit(int slot, const PassRec& r) = 0;
};

// Phases of a frame: the slots of one phase depend only on earlier phases,
// so the two lanes of a phase may run concurrently.  Phase 0: G-buffer
// (lane 0) beside the hash-grid cache update (lane 1: it reads the previous
// frame's depth and queries only); phase 1: diffuse (lane 0) beside specular
// sampling and resampling (lane 1; both read the G-buffer and the updated
// cache, and only append to the shared query buffer); phase 2: denoiser
// (lane 0) beside the sun shadow mask (lane 1: it reads the depth only);
// phase 3: final lighting, then temporal AA.  The specular validation re-
// traces last frame's samples and needs nothing of this frame's specular
// generation, so it runs at the head of lane 0, which balances the two
// lanes' ray tracing; the specular temporal pass waits for it (a cross-lane
// event, slot_signals / slot_waits).
inline int slot_phase(int slot) {
    return slot < kSlotSpecValidate ? 0 : slot < kSlotDenClearFastDiff ? 1 : slot < kSlotFinalLighting ? 2 : 3;
}
inline int slot_lane(int slot) {
    return (slot >= kSlotClearHgStorage && slot <= kSlotHgRaygenOrdered) || (slot >= kSlotSpecGen && slot <= kSlotSpecSpatial) ||
           slot == kSlotShadow;
}
inline bool slot_signals(int slot) { return slot == kSlotSpecValidate; }
inline bool slot_waits(int slot) { return slot == kSlotSpecTemporal; }

// The streams of a frame's slots, issued in slot order.  Without a side
// stream every slot runs on `main`.  With one, lane-1 slots go to `side`:
// entering a phase joins the side stream into the main one (if it has work)
// and forks it again from the main stream, so a lane-1 slot waits for every
// earlier phase and for nothing else; finish() joins it back at the end.
// leave() after a signalling slot records ev_cross on its stream, and a
// waiting slot of the same phase on the other lane waits for it.  Works the
// same under stream capture (the side stream joins the capture).
struct Lanes {
    hipStream_t main = nullptr, side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_cross = nullptr;
    int phase = -1;
    bool side_work = false;
    bool cross_pending = false;        // ev_cross was recorded in this phase
    hipStream_t cross_from = nullptr;  // the stream it was recorded on
    vgi_status enter(int slot, hipStream_t& st);
    vgi_status leave(int slot, hipStream_t st);
    vgi_status finish();
};

// Launches each record at once through its C entry point.
struct DirectSink : PassSink {
    explicit DirectSink(hipStream_t s) { lanes.main = s; }
    DirectSink(hipStream_t s, hipStream_t side, hipEvent_t fork, hipEvent_t join, hipEvent_t cross) {
        lanes.main = s;
        lanes.side = side;
        lanes.ev_fork = fork;
        lanes.ev_join = join;
        lanes.ev_cross = cross;
    }
    vgi_status submit(int slot, const PassRec& r) override;
    vgi_status finish() { return lanes.finish(); }
    Lanes lanes;

   private:
    static vgi_status launch(const PassRec& r, hipStream_t st);
};

// Keeps the records of one view's frame by slot.
struct RecordSink : PassSink {
    RecordSink() { clear(); }
    void clear() {
        for (int s = 0; s < kSlotCount; ++s) has[s] = false;
    }
    vgi_status submit(int slot, const PassRec& r) override {
        if (slot < 0 || slot >= kSlotCount || slot_kind(slot) != r.kind)
            return set_error(VGI_E_INVALID, "internal: pass kind %d recorded in slot %d", r.kind, slot);
        recs[slot] = r;
        has[slot] = true;
        return VGI_OK;
    }
    PassRec recs[kSlotCount];
    bool has[kSlotCount];
};

}  // namespace vgi

// Sequencers over a sink (the C entry points call them with a DirectSink).
uint32_t vgi_renderer_storage(const vgi_renderer* r);
vgi_status vgi_renderer_plan(vgi_renderer* r, const vgi_scene* scene, const vgi_frame_params* fp, uint32_t frame_index,
                             vgi_renderer_output* out, vgi::PassSink& sink);
vgi_status vgi_denoiser_plan(vgi_denoiser* d, const vgi_frame_params* fp, const vgi_denoiser_input* in,
                             vgi_denoiser_output* out, vgi::PassSink& sink);
