/*
 * vgi_renderer.cpp — RestirRenderer::add_passes (restir_renderer.rs:258-1028)
 * over persistent per-view HBM state.
 *
 * Pass order, seeds and flags follow the reference (FrameRand is a jenkins
 * chain started at frame_index; a seed is drawn only by passes that run):
 *   G-buffer -> [clear cache on first frame] -> cache decay ->
 *   [if prev depth & prev denoised diffuse: clear selection, query select
 *    (seed), compact, clear new query counter, cache raygen (seed), swap
 *    query buffers] -> diffuse sample gen (seed) | validate (every 4th frame)
 *   -> AO gen -> AO filter -> diffuse temporal (seed) -> diffuse spatial
 *   (seed) -> specular sample gen (seed) -> [specular validate] ->
 *   specular temporal (seed) -> specular spatial (seed) -> denoiser (if
 *   prev depth & prev G-buffer) -> feedback of temporal state.
 * Images that violet keeps as RGTemporal resources are double-buffered here
 * (index `cur` is written this frame, `1 - cur` holds last frame's).
 */
#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "vgi_denoiser_impl.h"
#include "vgi_passes.h"

using namespace vgi;

struct vgi_renderer {
    uint32_t w = 0, h = 0;
    vgi_renderer_config cfg{};
    // G-buffer + depth, double-buffered (prev G-buffer / prev depth)
    DevBuf depth[2], gbuffer[2];
    // indirect diffuse reservoirs: spatial output (= next frame's prev), 2 sets
    DevBuf dsp_rsv[2], dsp_hit[2], dsp_rad[2];
    DevBuf dtm_rsv, dtm_hit, dtm_rad;  // temporal output (transient)
    DevBuf dtaps;                      // the temporal output as the spatial pass's tap records (transient)
    DevBuf dnew_hit, dnew_rad;         // new diffuse sample (transient)
    DevBuf ao[2], ao_filtered;
    DevBuf diff_light[2];  // raw indirect diffuse (double-buffered: can be next frame's screen cache)
    // indirect specular: new sample / temporal result (in place) = next frame's prev, 2 sets
    DevBuf spec_rsv[2], spec_hitpos[2], spec_rad[2];
    DevBuf spec_light, spec_raylen;
    DevBuf shadow_mask, color;  // compose: sun shadow mask (u8) and final lighting
    DevBuf taa[2];              // taa: resolved colour, double-buffered (= next frame's history)
    // hash grid
    DevBuf hg_storage, hg_decay, hg_query[2], hg_counter[2], hg_selection, hg_cidx, hg_caddr, hg_ccount;
    DevBuf hg_winner, hg_workspace, hg_staging;  // ordered cache update (transient)
    DevBuf hg_threshold;                         // ordered cache update: query threshold word (state)
    int q = 0;  // query buffer holding pending queries
    vgi_denoiser* den = nullptr;
    // history state
    int cur = 0;
    bool has_prev_depth = false, has_prev_gbuffer = false, has_prev_indirect_diffuse = false;
    bool has_prev_diffuse_reservoir = false, has_prev_ind_spec = false, has_prev_ao = false;
    bool hash_grid_ready = false;
    bool hg_scratch_zero = false;  // ordered update: selection and winner buffers hold zeros
    bool has_prev_color = false;  // TAA history (restir_renderer.rs:914-918)
    const void* prev_indirect_diffuse = nullptr;  // last frame's denoised (or raw) diffuse
    int prev_ind_diff_slot = -1;                  // 0/1 = diff_light[k]; 2/3 = denoiser pp[k-2]

    // second lane of add_passes (concurrent_lanes): stream + fork/join events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_cross = nullptr;
    ~vgi_renderer() {
        if (den) vgi_denoiser_destroy(den);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_cross) (void)hipEventDestroy(ev_cross);
        if (side) (void)hipStreamDestroy(side);
    }
    std::vector<std::pair<std::string, DevBuf*>> images() {
        return {{"depth0", &depth[0]},         {"depth1", &depth[1]},
                {"gbuffer0", &gbuffer[0]},     {"gbuffer1", &gbuffer[1]},
                {"dsp_rsv0", &dsp_rsv[0]},     {"dsp_rsv1", &dsp_rsv[1]},
                {"dsp_hit0", &dsp_hit[0]},     {"dsp_hit1", &dsp_hit[1]},
                {"dsp_rad0", &dsp_rad[0]},     {"dsp_rad1", &dsp_rad[1]},
                {"ao0", &ao[0]},               {"ao1", &ao[1]},
                {"diff_light0", &diff_light[0]}, {"diff_light1", &diff_light[1]},
                {"spec_rsv0", &spec_rsv[0]},   {"spec_rsv1", &spec_rsv[1]},
                {"spec_hitpos0", &spec_hitpos[0]}, {"spec_hitpos1", &spec_hitpos[1]},
                {"spec_rad0", &spec_rad[0]},   {"spec_rad1", &spec_rad[1]},
                {"hg_storage", &hg_storage},   {"hg_decay", &hg_decay},
                {"hg_query0", &hg_query[0]},   {"hg_query1", &hg_query[1]},
                {"hg_counter0", &hg_counter[0]}, {"hg_counter1", &hg_counter[1]},
                {"hg_threshold", &hg_threshold},
                {"taa0", &taa[0]},             {"taa1", &taa[1]}};
    }
};

namespace {

uint32_t frame_rand_next(uint32_t& state) {
    state = jenkins_hash(state);
    return state;
}

vgi_trace_resources trace_res(vgi_renderer* r, const vgi_scene* scene, bool with_prev) {
    vgi_trace_resources t{};
    t.scene_tlas = scene;
    t.sky_radiance[0] = r->cfg.sky_radiance[0];
    t.sky_radiance[1] = r->cfg.sky_radiance[1];
    t.sky_radiance[2] = r->cfg.sky_radiance[2];
    t.storage = r->cfg.storage;
    if (with_prev) {
        t.prev_indirect_diffuse_texture = r->prev_indirect_diffuse;
        t.prev_depth_texture = r->depth[1 - r->cur].as<float>();
        t.prev_width = r->w;
        t.prev_height = r->h;
    }
    t.hash_grid_storage_buffer = r->hg_storage.as<vgi_hash_grid_cell>();
    t.rw_hash_grid_query_buffer = r->hg_query[r->q].as<vgi_hash_grid_query>();
    t.rw_hash_grid_query_counter_buffer = r->hg_counter[r->q].as<uint32_t>();
    t.hash_grid_query_threshold = r->cfg.ordered_cache_update ? r->hg_threshold.as<uint32_t>() : nullptr;
    return t;
}

const void* slot_ptr(vgi_renderer* r, int slot) {
    if (slot == 0 || slot == 1) return r->diff_light[slot].p;
    if (slot == 2 || slot == 3) return r->den->pp[slot - 2].diff.p;
    return nullptr;
}

}  // namespace

extern "C" vgi_status vgi_renderer_default_config(vgi_renderer_config* c) {
    VGI_REQUIRE(c, "vgi_renderer_default_config: null");
    memset(c, 0, sizeof(*c));
    c->storage = VGI_STORAGE_F32;
    c->denoise = 1;
    c->ind_diff_validate = 1;
    c->ind_spec_validate = 1;
    c->hash_grid_cache_decay = 1;
    c->concurrent_lanes = 1;
    c->ordered_cache_update = 1;
    c->ao_radius = 2.0f;
    c->sky_radiance[0] = c->sky_radiance[1] = c->sky_radiance[2] = 1.0f;
    c->denoiser.disocclusion_fix = 1;
    c->denoiser.fast_history_clamp = 0;
    c->denoiser.atrous_iterations = 3;
    c->denoiser.compensated_moments = 1;
    return VGI_OK;
}

extern "C" vgi_status vgi_renderer_create(uint32_t width, uint32_t height, const vgi_renderer_config* cfg,
                                          vgi_renderer** out) {
    VGI_REQUIRE(out, "vgi_renderer_create: null out");
    *out = nullptr;
    VGI_REQUIRE(valid_dims(width, height), "vgi_renderer_create: bad size %ux%u", width, height);
    vgi_renderer_config c;
    if (cfg)
        c = *cfg;
    else
        vgi_renderer_default_config(&c);
    VGI_REQUIRE(c.storage == VGI_STORAGE_F32 || c.storage == VGI_STORAGE_F16, "vgi_renderer_create: bad storage");
    VGI_REQUIRE(c.diff_spatial_taps <= 16u, "vgi_renderer_create: diff_spatial_taps %u > 16", c.diff_spatial_taps);
    VGI_REQUIRE(c.temporal_m_cap <= 255u, "vgi_renderer_create: temporal_m_cap %u > 255", c.temporal_m_cap);
    VGI_REQUIRE(c.spatial_radius_px >= 0.0f && c.spatial_radius_px <= 16384.0f, "vgi_renderer_create: bad spatial_radius_px");
    VGI_REQUIRE(!c.taa || c.compose, "vgi_renderer_create: taa needs compose");
    vgi_renderer* r = new (std::nothrow) vgi_renderer();
    if (!r) return set_error(VGI_E_NOMEM, "vgi_renderer_create: out of host memory");
    r->w = width;
    r->h = height;
    r->cfg = c;
    const size_t n = (size_t)width * height;
    const size_t R = radiance_bytes(c.storage), Pb = pair_bytes(c.storage), S = scalar_bytes(c.storage);
    vgi_status st = VGI_OK;
    auto A = [&](DevBuf& b, size_t bytes) {
        if (st == VGI_OK) st = b.alloc(bytes);
    };
    for (int k = 0; k < 2; ++k) {
        A(r->depth[k], 4 * n);
        A(r->gbuffer[k], 16 * n);
        A(r->dsp_rsv[k], 4 * n);
        A(r->dsp_hit[k], 8 * n);
        A(r->dsp_rad[k], R * n);
        A(r->ao[k], Pb * n);
        A(r->diff_light[k], R * n);
        A(r->spec_rsv[k], 8 * n);
        A(r->spec_hitpos[k], 16 * n);
        A(r->spec_rad[k], R * n);
        A(r->hg_query[k], sizeof(vgi_hash_grid_query) * (size_t)VGI_HASH_GRID_MAX_QUERIES);
        A(r->hg_counter[k], 4);
    }
    A(r->dtm_rsv, 4 * n);
    A(r->dtm_hit, 8 * n);
    A(r->dtm_rad, R * n);
    A(r->dtaps, sizeof(DiffTap) * n);
    A(r->dnew_hit, 8 * n);
    A(r->dnew_rad, R * n);
    A(r->ao_filtered, Pb * n);
    A(r->spec_light, R * n);
    A(r->spec_raylen, S * n);
    A(r->hg_storage, sizeof(vgi_hash_grid_cell) * (size_t)VGI_HASH_GRID_NUM_CELLS);
    A(r->hg_decay, VGI_HASH_GRID_NUM_CELLS);
    A(r->hg_selection, 4 * (size_t)VGI_HASH_GRID_NUM_CELLS);
    A(r->hg_cidx, 4 * (size_t)VGI_HASH_GRID_MAX_QUERIES);
    A(r->hg_caddr, 4 * (size_t)VGI_HASH_GRID_MAX_QUERIES);
    A(r->hg_ccount, 4);
    if (c.ordered_cache_update) {
        A(r->hg_winner, 4 * (size_t)VGI_HASH_GRID_NUM_CELLS);
        A(r->hg_workspace, 4 * (size_t)VGI_HASH_GRID_COMPACT_WORKSPACE_WORDS);
        A(r->hg_staging, 16 * (size_t)VGI_HASH_GRID_MAX_QUERIES);
    }
    A(r->hg_threshold, 4);  // read only by the ordered update (kept in every mode: a fixed state layout)
    if (c.compose) {
        A(r->shadow_mask, n);
        A(r->color, R * n);
    }
    if (c.taa) {
        A(r->taa[0], R * n);
        A(r->taa[1], R * n);
    }
    if (st == VGI_OK) st = vgi_denoiser_create(width, height, c.storage, &c.denoiser, &r->den);
    if (st == VGI_OK && c.concurrent_lanes) {
        // non-blocking: no implicit ordering against the legacy null stream,
        // which would serialise the two lanes when the caller passes it
        hipError_t e = hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking);
        // the lane events order two streams of this device only: no
        // system-scope release (a write-back of L2 for the host) at each
        // record (the two joins per frame cost 0.4% of the frame with it)
        const unsigned lane_ev = hipEventDisableTiming | hipEventDisableSystemFence;
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_fork, lane_ev);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_join, lane_ev);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_cross, lane_ev);
        if (e != hipSuccess) st = check_hip(e, "vgi_renderer_create: side stream");
    }
    if (st != VGI_OK) {
        delete r;
        return st;
    }
    *out = r;
    return VGI_OK;
}

extern "C" vgi_status vgi_renderer_destroy(vgi_renderer* r) {
    delete r;
    return VGI_OK;
}

extern "C" vgi_status vgi_renderer_reset(vgi_renderer* r) {
    VGI_REQUIRE(r, "vgi_renderer_reset: null renderer");
    r->has_prev_depth = r->has_prev_gbuffer = r->has_prev_indirect_diffuse = false;
    r->has_prev_diffuse_reservoir = r->has_prev_ind_spec = r->has_prev_ao = false;
    r->hash_grid_ready = false;
    r->hg_scratch_zero = false;
    r->has_prev_color = false;
    r->prev_indirect_diffuse = nullptr;
    r->prev_ind_diff_slot = -1;
    return vgi_denoiser_reset(r->den);
}

#define CALL(x)                           \
    do {                                  \
        vgi_status _s = (x);              \
        if (_s != VGI_OK) return _s;      \
    } while (0)

static vgi_resample_params resample_params(const vgi_renderer* r) {
    return vgi_resample_params{r->cfg.temporal_m_cap, r->cfg.diff_spatial_taps, r->cfg.spatial_radius_px};
}

// One frame of restir_renderer.rs:258-829 issued to `sink` (host state -
// history flags, ping-pong indices, FrameRand chain - advances as the passes
// are issued, so a recorded frame must be executed before the next one).
vgi_status vgi_renderer_plan(vgi_renderer* r, const vgi_scene* scene, const vgi_frame_params* fp, uint32_t frame_index,
                             vgi_renderer_output* out, PassSink& sink) {
    VGI_REQUIRE(r && scene && fp && out, "vgi_renderer_add_passes: null argument");
    const uint32_t w = r->w, h = r->h, S = r->cfg.storage;
    const int c = r->cur, p = 1 - c;
    uint32_t frame_rand = frame_index;  // restir_renderer.rs:274

    // G-buffer (restir_renderer.rs:277-280)
    {
        PassRec g = make_rec(kPassGBuffer, fp);
        g.scene = scene;
        g.io.gbuffer = vgi_gbuffer_io{w, h, r->depth[c].as<float>(), r->gbuffer[c].as<uint32_t>()};
        CALL(sink.submit(kSlotGBuffer, g));
    }

    // Hash grid cache: create/clear on first use (:282-319)
    if (!r->hash_grid_ready) {
        CALL(sink.submit(kSlotClearHgStorage, make_clear(r->hg_storage.p, r->hg_storage.bytes)));
        CALL(sink.submit(kSlotClearHgDecay, make_clear(r->hg_decay.p, r->hg_decay.bytes)));
        CALL(sink.submit(kSlotClearHgCounter, make_clear(r->hg_counter[r->q].p, 4)));
        // the ordered update's query threshold before the first count: at
        // most one diffuse and one specular query per pixel (the first frame
        // runs no cache rays)
        const uint64_t expected = 2ull * w * h;
        CALL(sink.submit(kSlotFillHgThreshold,
                         make_fill(r->hg_threshold.p, 4, vgi_hash_grid_query_threshold_for(expected))));
        r->hash_grid_ready = true;
    }
    if (r->cfg.hash_grid_cache_decay) {
        PassRec d = make_rec(kPassHgDecay);
        d.io.hg_decay = vgi_hash_grid_decay_io{r->hg_decay.as<uint8_t>(), r->hg_storage.as<vgi_hash_grid_cell>()};
        CALL(sink.submit(kSlotHgDecay, d));
    }
    const bool has_prev_depth = r->has_prev_depth, has_prev_gbuffer = r->has_prev_gbuffer;
    const bool has_prev_ind_diff = r->has_prev_indirect_diffuse;

    // Cache update (:361-433)
    bool hg_updated = false;
    if (has_prev_depth && has_prev_ind_diff) {
        const bool ordered = r->cfg.ordered_cache_update != 0;
        // the ordered passes leave the selection and winner buffers zero
        // again, so they are cleared once; the reference's select every frame
        if (!ordered || !r->hg_scratch_zero)
            CALL(sink.submit(kSlotClearSelection, make_clear(r->hg_selection.p, r->hg_selection.bytes)));
        const uint32_t select_hash = frame_rand_next(frame_rand);
        if (ordered) {
            if (!r->hg_scratch_zero)
                CALL(sink.submit(kSlotClearWinner, make_clear(r->hg_winner.p, r->hg_winner.bytes)));
            r->hg_scratch_zero = true;
            PassRec sel = make_rec(kPassHgSelectOrdered);
            sel.io.hg_oselect = vgi_hash_grid_ordered_select_io{
                r->hg_query[r->q].as<vgi_hash_grid_query>(), r->hg_counter[r->q].as<uint32_t>(),
                r->hg_storage.as<vgi_hash_grid_cell>(),      r->hg_selection.as<uint32_t>(),
                r->hg_winner.as<uint32_t>(),                 r->hg_threshold.as<uint32_t>()};
            sel.pc.frame_hash = vgi_pc_frame_hash{select_hash};
            CALL(sink.submit(kSlotHgSelectOrdered, sel));
        } else {
            PassRec sel = make_rec(kPassHgSelect);
            sel.io.hg_select = vgi_hash_grid_query_select_io{
                r->hg_query[r->q].as<vgi_hash_grid_query>(), r->hg_counter[r->q].as<uint32_t>(),
                r->hg_storage.as<vgi_hash_grid_cell>(), r->hg_selection.as<uint32_t>()};
            sel.pc.frame_hash = vgi_pc_frame_hash{select_hash};
            CALL(sink.submit(kSlotHgSelect, sel));
        }
        CALL(sink.submit(kSlotClearCompactCount, make_clear(r->hg_ccount.p, 4)));
        if (ordered) {
            PassRec cmp = make_rec(kPassHgCompactOrdered);
            cmp.io.hg_ocompact = vgi_hash_grid_ordered_compact_io{
                r->hg_selection.as<uint32_t>(), r->hg_winner.as<uint32_t>(), r->hg_cidx.as<uint32_t>(),
                r->hg_caddr.as<uint32_t>(),     r->hg_ccount.as<uint32_t>(), r->hg_workspace.as<uint32_t>()};
            CALL(sink.submit(kSlotHgCompactOrdered, cmp));
        } else {
            PassRec cmp = make_rec(kPassHgCompact);
            cmp.io.hg_compact = vgi_hash_grid_query_compact_io{r->hg_selection.as<uint32_t>(), r->hg_cidx.as<uint32_t>(),
                                                               r->hg_caddr.as<uint32_t>(), r->hg_ccount.as<uint32_t>()};
            CALL(sink.submit(kSlotHgCompact, cmp));
        }
        const int nq = 1 - r->q;
        CALL(sink.submit(kSlotClearNextCounter, make_clear(r->hg_counter[nq].p, 4)));
        PassRec rg = make_rec(ordered ? kPassHgRaygenOrdered : kPassHgRaygen, fp);
        rg.tr = trace_res(r, scene, true);
        rg.tr.rw_hash_grid_query_buffer = r->hg_query[nq].as<vgi_hash_grid_query>();
        rg.tr.rw_hash_grid_query_counter_buffer = r->hg_counter[nq].as<uint32_t>();
        const vgi_hash_grid_raygen_io rio{
            r->hg_storage.as<vgi_hash_grid_cell>(), r->hg_query[r->q].as<vgi_hash_grid_query>(),
            r->hg_ccount.as<uint32_t>(), r->hg_cidx.as<uint32_t>(), r->hg_caddr.as<uint32_t>(), r->hg_decay.as<uint8_t>()};
        if (ordered)
            rg.io.hg_oraygen = vgi_hash_grid_ordered_raygen_io{rio, r->hg_staging.as<float>()};
        else
            rg.io.hg_raygen = rio;
        rg.pc.frame_hash = vgi_pc_frame_hash{frame_rand_next(frame_rand)};
        CALL(sink.submit(ordered ? kSlotHgRaygenOrdered : kSlotHgRaygen, rg));
        r->q = nq;
        hg_updated = true;
    }

    // Specular sample validate (:722-737): issued first, it re-traces last
    // frame's specular samples and is independent of this frame's passes
    // up to the specular temporal resample
    const bool has_prev_spec = r->has_prev_ind_spec;
    if (r->cfg.ind_spec_validate && has_prev_depth && has_prev_ind_diff && has_prev_spec) {
        PassRec v = make_rec(kPassSpecValidate, fp);
        v.tr = trace_res(r, scene, true);
        v.io.spec_validate = vgi_ind_spec_sample_validate_io{w, h, S, r->spec_rsv[p].as<uint32_t>(),
                                                             r->spec_hitpos[p].as<float>(), r->spec_rad[p].p};
        CALL(sink.submit(kSlotSpecValidate, v));
    }

    // Indirect diffuse: sample gen or validation (:436-513)
    const bool has_prev_diff_rsv = r->has_prev_diffuse_reservoir;
    const bool validation = has_prev_depth && has_prev_diff_rsv && ((frame_index & 3u) == 0u) && r->cfg.ind_diff_validate;
    const uint32_t has_new_sample = validation ? 0u : 1u;
    const bool trace_prev = has_prev_depth && has_prev_ind_diff;
    if (!validation) {
        PassRec g = make_rec(kPassDiffGen, fp);
        g.tr = trace_res(r, scene, trace_prev);
        g.io.diff_gen = vgi_ind_diff_sample_gen_io{w, h, S, r->gbuffer[c].as<uint32_t>(), r->depth[c].as<float>(),
                                                   r->dnew_hit.as<uint32_t>(), r->dnew_rad.p};
        g.pc.sample_gen = vgi_pc_sample_gen{frame_rand_next(frame_rand), trace_prev ? 1u : 0u};
        CALL(sink.submit(kSlotDiffGen, g));
    } else {
        PassRec v = make_rec(kPassDiffValidate, fp);
        v.tr = trace_res(r, scene, true);
        v.io.diff_validate = vgi_ind_diff_sample_validate_io{w, h, S, r->depth[p].as<float>(), r->dsp_rsv[p].as<uint32_t>(),
                                                             r->dsp_hit[p].as<uint32_t>(), r->dsp_rad[p].p};
        CALL(sink.submit(kSlotDiffValidate, v));
    }

    // AO gen + filter (:515-563)
    {
        PassRec a = make_rec(kPassAoGen, fp);
        a.io.ao_gen = vgi_ao_gen_io{w, h, S, has_new_sample ? r->dnew_hit.as<uint32_t>() : nullptr, r->depth[c].as<float>(),
                                    r->has_prev_ao ? r->ao[p].p : nullptr, r->ao[c].p};
        a.pc.ao_gen = vgi_pc_ao_gen{has_new_sample, r->has_prev_ao ? 1u : 0u, r->cfg.ao_radius};
        CALL(sink.submit(kSlotAoGen, a));
        PassRec f = make_rec(kPassAoFilter);
        f.io.ao_filter = vgi_ao_filter_io{w, h, S, r->depth[c].as<float>(), r->ao[c].p, r->ao_filtered.p};
        CALL(sink.submit(kSlotAoFilter, f));
    }

    // Diffuse temporal (:589-607) and spatial (:632-659)
    {
        PassRec t = make_rec(kPassDiffTemporal, fp);
        t.io.diff_temporal = vgi_ind_diff_temporal_io{w, h, S, has_prev_diff_rsv ? r->depth[p].as<float>() : nullptr, r->depth[c].as<float>(),
                                    has_new_sample ? r->dnew_hit.as<uint32_t>() : nullptr,
                                    has_new_sample ? r->dnew_rad.p : nullptr,
                                    has_prev_diff_rsv ? r->dsp_rsv[p].as<uint32_t>() : nullptr,
                                    has_prev_diff_rsv ? r->dsp_hit[p].as<uint32_t>() : nullptr,
                                    has_prev_diff_rsv ? r->dsp_rad[p].p : nullptr,
                                    r->dtm_rsv.as<uint32_t>(), r->dtm_hit.as<uint32_t>(), r->dtm_rad.p};
        // has_prev_frame = has_prev_diffuse_reservoir (restir_renderer.rs:605); the
        // previous depth exists whenever the previous reservoir does
        t.pc.diff_temporal = vgi_pc_ind_diff_temporal{frame_rand_next(frame_rand), (has_prev_diff_rsv && has_prev_depth) ? 1u : 0u,
                                    has_new_sample};
        t.resample = resample_params(r);
        t.diff_taps = r->dtaps.as<DiffTap>();
        t.diff_taps_gbuffer = r->gbuffer[c].as<uint32_t>();
        CALL(sink.submit(kSlotDiffTemporal, t));
        PassRec sp = make_rec(kPassDiffSpatial, fp);
        sp.io.diff_spatial = vgi_ind_diff_spatial_io{w, h, S, r->gbuffer[c].as<uint32_t>(), r->depth[c].as<float>(), r->ao_filtered.p,
                                    r->dtm_rsv.as<uint32_t>(), r->dtm_hit.as<uint32_t>(), r->dtm_rad.p,
                                    r->dsp_rsv[c].as<uint32_t>(), r->dsp_hit[c].as<uint32_t>(), r->dsp_rad[c].p,
                                    r->diff_light[c].p};
        sp.pc.diff_spatial = vgi_pc_ind_diff_spatial{frame_rand_next(frame_rand)};
        sp.resample = resample_params(r);
        sp.diff_taps = r->dtaps.as<DiffTap>();
        CALL(sink.submit(kSlotDiffSpatial, sp));
    }

    // Specular sample gen (:685-701), temporal (:743-761), spatial (:770-784)
    {
        PassRec g = make_rec(kPassSpecGen, fp);
        g.tr = trace_res(r, scene, trace_prev);
        g.io.spec_gen = vgi_ind_spec_sample_gen_io{w, h, S, r->gbuffer[c].as<uint32_t>(), r->depth[c].as<float>(),
                                                   r->spec_hitpos[c].as<float>(), r->spec_rad[c].p};
        g.pc.sample_gen = vgi_pc_sample_gen{frame_rand_next(frame_rand), trace_prev ? 1u : 0u};
        CALL(sink.submit(kSlotSpecGen, g));
    }
    {
        const bool hp = has_prev_depth && has_prev_spec && has_prev_gbuffer;
        PassRec t = make_rec(kPassSpecTemporal, fp);
        t.io.spec_temporal = vgi_ind_spec_temporal_io{w, h, S, hp ? r->gbuffer[p].as<uint32_t>() : nullptr,
                                    hp ? r->depth[p].as<float>() : nullptr, r->gbuffer[c].as<uint32_t>(),
                                    r->depth[c].as<float>(), hp ? r->spec_rsv[p].as<uint32_t>() : nullptr,
                                    hp ? r->spec_hitpos[p].as<float>() : nullptr, hp ? r->spec_rad[p].p : nullptr,
                                    r->spec_rsv[c].as<uint32_t>(), r->spec_hitpos[c].as<float>(), r->spec_rad[c].p};
        t.pc.spec_temporal = vgi_pc_ind_spec_temporal{frame_rand_next(frame_rand), hp ? 1u : 0u};
        t.resample = resample_params(r);
        CALL(sink.submit(kSlotSpecTemporal, t));
        PassRec sp = make_rec(kPassSpecSpatial, fp);
        sp.io.spec_spatial = vgi_ind_spec_spatial_io{w, h, S, r->gbuffer[c].as<uint32_t>(), r->depth[c].as<float>(),
                                    r->spec_rsv[c].as<uint32_t>(), r->spec_hitpos[c].as<float>(), r->spec_rad[c].p,
                                    r->spec_light.p, r->spec_raylen.p};
        sp.pc.spec_spatial = vgi_pc_ind_spec_spatial{frame_rand_next(frame_rand)};
        sp.resample = resample_params(r);
        CALL(sink.submit(kSlotSpecSpatial, sp));
    }

    // Denoiser (:790-817)
    const void* den_diff = r->diff_light[c].p;
    const void* den_spec = r->spec_light.p;
    int slot = c;
    bool denoised = false;
    if (r->cfg.denoise && has_prev_depth && has_prev_gbuffer) {
        CALL(vgi_denoiser_set_config(r->den, &r->cfg.denoiser));
        vgi_denoiser_input in{r->depth[c].as<float>(), r->gbuffer[c].as<uint32_t>(), r->depth[p].as<float>(),
                              r->gbuffer[p].as<uint32_t>(), r->diff_light[c].p, r->spec_light.p, r->spec_raylen.p};
        vgi_denoiser_output o{};
        CALL(vgi_denoiser_plan(r->den, fp, &in, &o, sink));
        den_diff = o.diffuse;
        den_spec = o.specular;
        slot = (o.diffuse == r->den->pp[0].diff.p) ? 2 : 3;
        denoised = true;
    } else {
        CALL(vgi_denoiser_reset(r->den));
    }

    // Direct light + composition (:831-911)
    if (r->cfg.compose) {
        PassRec sh = make_rec(kPassShadow, fp);
        sh.scene = scene;
        sh.io.shadow = vgi_raytraced_shadow_io{w, h, r->depth[c].as<float>(), r->shadow_mask.as<uint8_t>()};
        CALL(sink.submit(kSlotShadow, sh));
        PassRec fl = make_rec(kPassFinalLighting, fp);
        fl.io.final_lighting = vgi_final_lighting_io{w, h, S, r->gbuffer[c].as<uint32_t>(), r->shadow_mask.as<uint8_t>(),
                                                     den_diff, den_spec, r->color.p,
                                                     {r->cfg.sky_radiance[0], r->cfg.sky_radiance[1],
                                                      r->cfg.sky_radiance[2]}};
        CALL(sink.submit(kSlotFinalLighting, fl));
        if (r->cfg.taa) {  // :914-944
            PassRec t = make_rec(kPassTemporalAA, fp);
            t.io.taa = vgi_temporal_aa_io{w, h, S, r->depth[c].as<float>(), r->color.p,
                                          r->has_prev_color ? r->taa[p].p : nullptr, r->taa[c].p};
            t.pc.taa = vgi_pc_temporal_aa{r->has_prev_color ? 1u : 0u};
            CALL(sink.submit(kSlotTemporalAA, t));
        }
    }

    // Feedback (:819-829, :1012-1025)
    r->prev_indirect_diffuse = den_diff;
    r->prev_ind_diff_slot = slot;
    r->has_prev_indirect_diffuse = true;
    r->has_prev_diffuse_reservoir = true;
    r->has_prev_ind_spec = true;
    r->has_prev_gbuffer = true;
    r->has_prev_depth = true;
    r->has_prev_ao = true;
    r->has_prev_color = r->cfg.taa != 0;
    r->cur = p;

    out->indirect_diffuse = den_diff;
    out->indirect_specular = den_spec;
    out->depth = r->depth[c].as<float>();
    out->gbuffer = r->gbuffer[c].as<uint32_t>();
    out->denoised = denoised ? 1u : 0u;
    out->validation_frame = validation ? 1u : 0u;
    out->hash_grid_updated = hg_updated ? 1u : 0u;
    out->color = r->cfg.taa ? r->taa[c].p : r->cfg.compose ? r->color.p : nullptr;
    return VGI_OK;
}

uint32_t vgi_renderer_storage(const vgi_renderer* r) { return r->cfg.storage; }

extern "C" vgi_status vgi_renderer_add_passes(vgi_renderer* r, const vgi_scene* scene, const vgi_frame_params* fp,
                                              uint32_t frame_index, vgi_renderer_output* out, void* stream) {
    VGI_REQUIRE(r, "vgi_renderer_add_passes: null renderer");
    DirectSink sink((hipStream_t)stream, r->side, r->ev_fork, r->ev_join, r->ev_cross);
    vgi_status s = vgi_renderer_plan(r, scene, fp, frame_index, out, sink);
    vgi_status j = sink.finish();
    return s != VGI_OK ? s : j;
}

extern "C" vgi_status vgi_renderer_get_image(const vgi_renderer* rc, const char* name, void** ptr, size_t* bytes) {
    VGI_REQUIRE(rc && name && ptr && bytes, "vgi_renderer_get_image: null argument");
    vgi_renderer* r = const_cast<vgi_renderer*>(rc);
    // `cur` was flipped at the end of the last frame: that frame wrote index p
    const int last = 1 - r->cur;
    struct E {
        const char* n;
        DevBuf* b;
    } table[] = {
        {"depth", &r->depth[last]},
        {"gbuffer", &r->gbuffer[last]},
        {"prev_depth", &r->depth[r->cur]},
        {"diff_reservoir", &r->dsp_rsv[last]},
        {"diff_hit_pos_normal", &r->dsp_hit[last]},
        {"diff_hit_radiance", &r->dsp_rad[last]},
        {"diff_temporal_reservoir", &r->dtm_rsv},
        {"diff_temporal_hit_pos_normal", &r->dtm_hit},
        {"diff_temporal_hit_radiance", &r->dtm_rad},
        {"diff_new_hit_pos_normal", &r->dnew_hit},
        {"diff_new_hit_radiance", &r->dnew_rad},
        {"ao", &r->ao[last]},
        {"ao_filtered", &r->ao_filtered},
        {"indirect_diffuse_raw", &r->diff_light[last]},
        {"spec_reservoir", &r->spec_rsv[last]},
        {"spec_hit_pos", &r->spec_hitpos[last]},
        {"spec_hit_radiance", &r->spec_rad[last]},
        {"indirect_specular_raw", &r->spec_light},
        {"spec_ray_len", &r->spec_raylen},
        {"hash_grid_storage", &r->hg_storage},
        {"hash_grid_decay", &r->hg_decay},
        {"hash_grid_queries", &r->hg_query[r->q]},
        {"hash_grid_query_counter", &r->hg_counter[r->q]},
        {"hash_grid_query_threshold", &r->hg_threshold},
        {"shadow_mask", &r->shadow_mask},
        {"color", &r->color},
        {"taa_color", &r->taa[last]},
    };
    for (auto& e : table)
        if (strcmp(e.n, name) == 0) {
            *ptr = e.b->p;
            *bytes = e.b->bytes;
            return VGI_OK;
        }
    return set_error(VGI_E_INVALID, "vgi_renderer_get_image: unknown image '%s'", name);
}

// ---------------------------------------------------------------------------
// state export / import (checkpoint/resume, serving eviction)
// ---------------------------------------------------------------------------
namespace {
struct StateHeader {
    uint32_t magic, w, h, storage;
    int32_t cur, q, slot, den_cur;
    uint8_t flags[8];
    uint8_t den_flags[4];
};
void all_buffers(vgi_renderer* r, std::vector<DevBuf*>& v) {
    for (auto& e : r->images()) v.push_back(e.second);
    denoiser_buffers(r->den, v);
}
}  // namespace

extern "C" size_t vgi_renderer_state_size(const vgi_renderer* rc) {
    if (!rc) return 0;
    vgi_renderer* r = const_cast<vgi_renderer*>(rc);
    std::vector<DevBuf*> v;
    all_buffers(r, v);
    size_t n = sizeof(StateHeader);
    for (auto* b : v) n += b->bytes;
    return n;
}

extern "C" size_t vgi_renderer_device_bytes(const vgi_renderer* rc) {
    if (!rc) return 0;
    vgi_renderer* r = const_cast<vgi_renderer*>(rc);
    std::vector<DevBuf*> v;
    all_buffers(r, v);
    for (DevBuf* b : {&r->dtm_rsv, &r->dtm_hit, &r->dtm_rad, &r->dtaps, &r->dnew_hit, &r->dnew_rad, &r->ao_filtered, &r->spec_light,
                      &r->spec_raylen, &r->hg_selection, &r->hg_cidx, &r->hg_caddr, &r->hg_ccount, &r->hg_winner,
                      &r->hg_workspace, &r->hg_staging, &r->shadow_mask, &r->color})
        if (std::find(v.begin(), v.end(), b) == v.end()) v.push_back(b);
    size_t n = 0;
    for (auto* b : v) n += b->bytes;
    return n;
}

extern "C" vgi_status vgi_renderer_export_state(const vgi_renderer* rc, void* blob, size_t bytes, void* stream) {
    VGI_REQUIRE(rc && blob, "vgi_renderer_export_state: null argument");
    vgi_renderer* r = const_cast<vgi_renderer*>(rc);
    VGI_REQUIRE(bytes >= vgi_renderer_state_size(r), "vgi_renderer_export_state: blob too small");
    StateHeader hd{};
    hd.magic = 0x56474931u;
    hd.w = r->w;
    hd.h = r->h;
    hd.storage = r->cfg.storage;
    hd.cur = r->cur;
    hd.q = r->q;
    hd.slot = r->prev_ind_diff_slot;
    hd.den_cur = r->den->cur;
    bool f[8] = {r->has_prev_depth, r->has_prev_gbuffer, r->has_prev_indirect_diffuse, r->has_prev_diffuse_reservoir,
                 r->has_prev_ind_spec, r->has_prev_ao, r->hash_grid_ready, r->has_prev_color};
    for (int i = 0; i < 8; ++i) hd.flags[i] = f[i];
    hd.den_flags[0] = r->den->has_history;
    hd.den_flags[1] = r->den->has_fast_history;
    hd.den_flags[2] = r->den->hist_comp != nullptr;
    uint8_t* o = (uint8_t*)blob;
    memcpy(o, &hd, sizeof(hd));
    o += sizeof(hd);
    std::vector<DevBuf*> v;
    all_buffers(r, v);
    hipStream_t st = (hipStream_t)stream;
    for (auto* b : v) {
        if (b->bytes) VGI_TRY_HIP(hipMemcpyAsync(o, b->p, b->bytes, hipMemcpyDeviceToHost, st));
        o += b->bytes;
    }
    VGI_TRY_HIP(hipStreamSynchronize(st));
    return VGI_OK;
}

extern "C" vgi_status vgi_renderer_import_state(vgi_renderer* r, const void* blob, size_t bytes, void* stream) {
    VGI_REQUIRE(r && blob, "vgi_renderer_import_state: null argument");
    VGI_REQUIRE(bytes >= vgi_renderer_state_size(r), "vgi_renderer_import_state: blob too small");
    StateHeader hd;
    const uint8_t* s = (const uint8_t*)blob;
    memcpy(&hd, s, sizeof(hd));
    VGI_REQUIRE(hd.magic == 0x56474931u && hd.w == r->w && hd.h == r->h && hd.storage == r->cfg.storage,
                "vgi_renderer_import_state: blob does not match this renderer");
    s += sizeof(hd);
    std::vector<DevBuf*> v;
    all_buffers(r, v);
    hipStream_t st = (hipStream_t)stream;
    for (auto* b : v) {
        if (b->bytes) VGI_TRY_HIP(hipMemcpyAsync(b->p, s, b->bytes, hipMemcpyHostToDevice, st));
        s += b->bytes;
    }
    VGI_TRY_HIP(hipStreamSynchronize(st));
    r->cur = hd.cur;
    r->q = hd.q;
    r->prev_ind_diff_slot = hd.slot;
    r->has_prev_depth = hd.flags[0];
    r->has_prev_gbuffer = hd.flags[1];
    r->has_prev_indirect_diffuse = hd.flags[2];
    r->has_prev_diffuse_reservoir = hd.flags[3];
    r->has_prev_ind_spec = hd.flags[4];
    r->has_prev_ao = hd.flags[5];
    r->hash_grid_ready = hd.flags[6];
    r->hg_scratch_zero = false;
    r->has_prev_color = hd.flags[7] && r->cfg.taa;
    r->prev_indirect_diffuse = slot_ptr(r, hd.slot);
    vgi_denoiser* d = r->den;
    d->cur = hd.den_cur;
    d->has_history = hd.den_flags[0];
    d->has_fast_history = hd.den_flags[1];
    const int hs = 1 - d->cur;
    d->hist_diff = d->cfg.disocclusion_fix ? d->fix[hs].diff.p : d->temporal[hs].diff.p;
    d->hist_spec = d->cfg.disocclusion_fix ? d->fix[hs].spec.p : d->temporal[hs].spec.p;
    d->hist_mom = d->cfg.disocclusion_fix ? d->fix[hs].pair.p : d->temporal[hs].pair.p;
    d->hist_comp = !hd.den_flags[2] ? nullptr : d->cfg.disocclusion_fix ? d->fix[hs].comp.p : d->temporal[hs].comp.p;
    d->hist_len = d->temporal_len[hs].as<uint8_t>();
    return VGI_OK;
}
