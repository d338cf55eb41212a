/*
 * vgi_denoiser.cpp — the Denoiser pass sequencer
 * (crates/violet-restir/src/denoising.rs:41-301) on persistent HBM images.
 *
 * violet's render graph hands out transient textures from a pool and keeps
 * the history as RGTemporal textures; here the denoiser owns fixed images:
 * two alternating sets for the temporal-filter outputs and for the
 * disocclusion-fix outputs (the set written this frame becomes next frame's
 * history, so the one being read is never overwritten), and one ping-pong
 * pair for the spatial fallback and the a-trous iterations.
 */
#include <new>

#include "vgi_denoiser_impl.h"
#include "vgi_passes.h"

using namespace vgi;

static vgi_status alloc_set(ColorSet& s, size_t n, uint32_t storage) {
    size_t r = radiance_bytes(storage) * n, p = pair_bytes(storage) * n;
    vgi_status st;
    if ((st = s.diff.alloc(r)) != VGI_OK) return st;
    if ((st = s.spec.alloc(r)) != VGI_OK) return st;
    return s.pair.alloc(p);
}

extern "C" vgi_status vgi_denoiser_create(uint32_t width, uint32_t height, uint32_t storage,
                                          const vgi_denoiser_config* cfg, vgi_denoiser** out) {
    VGI_REQUIRE(out, "vgi_denoiser_create: null out");
    *out = nullptr;
    VGI_REQUIRE(valid_dims(width, height), "vgi_denoiser_create: bad size %ux%u", width, height);
    VGI_REQUIRE(storage == VGI_STORAGE_F32 || storage == VGI_STORAGE_F16, "vgi_denoiser_create: bad storage %u",
                storage);
    vgi_denoiser* d = new (std::nothrow) vgi_denoiser();
    if (!d) return set_error(VGI_E_NOMEM, "vgi_denoiser_create: out of host memory");
    d->w = width;
    d->h = height;
    d->storage = storage;
    if (cfg) d->cfg = *cfg;
    size_t n = (size_t)width * height;
    vgi_status st = VGI_OK;
    for (int k = 0; k < 2 && st == VGI_OK; ++k) {
        if ((st = alloc_set(d->temporal[k], n, storage)) != VGI_OK) break;
        if ((st = d->temporal_len[k].alloc(n)) != VGI_OK) break;
        if ((st = alloc_set(d->fix[k], n, storage)) != VGI_OK) break;
        if (storage == VGI_STORAGE_F16) {
            if ((st = d->temporal[k].comp.alloc(pair_bytes(storage) * n)) != VGI_OK) break;
            if ((st = d->fix[k].comp.alloc(pair_bytes(storage) * n)) != VGI_OK) break;
        }
        if ((st = alloc_set(d->pp[k], n, storage)) != VGI_OK) break;
        if ((st = d->fast_diff[k].alloc(radiance_bytes(storage) * n)) != VGI_OK) break;
        if ((st = d->fast_spec[k].alloc(radiance_bytes(storage) * n)) != VGI_OK) break;
    }
    if (st == VGI_OK) st = d->fb_var.alloc(pair_bytes(storage) * n);
    if (st != VGI_OK) {
        delete d;
        return st;
    }
    *out = d;
    return VGI_OK;
}

extern "C" vgi_status vgi_denoiser_destroy(vgi_denoiser* d) {
    delete d;
    return VGI_OK;
}

extern "C" vgi_status vgi_denoiser_reset(vgi_denoiser* d) {
    VGI_REQUIRE(d, "vgi_denoiser_reset: null denoiser");
    d->has_history = false;
    d->has_fast_history = false;
    return VGI_OK;
}

extern "C" vgi_status vgi_denoiser_set_config(vgi_denoiser* d, const vgi_denoiser_config* cfg) {
    VGI_REQUIRE(d && cfg, "vgi_denoiser_set_config: null argument");
    d->cfg = *cfg;
    return VGI_OK;
}

vgi_status vgi_denoiser_plan(vgi_denoiser* d, const vgi_frame_params* fp, const vgi_denoiser_input* in,
                             vgi_denoiser_output* out, PassSink& sink) {
    VGI_REQUIRE(d && fp && in && out, "vgi_denoiser_add_passes: null argument");
    VGI_REQUIRE(in->depth && in->gbuffer && in->diffuse && in->specular && in->specular_ray_len,
                "vgi_denoiser_add_passes: missing input");
    const uint32_t iters = d->cfg.atrous_iterations > 5 ? 5 : d->cfg.atrous_iterations;  // denoising.rs:73
    const bool has_history = d->has_history && in->prev_depth && in->prev_gbuffer;
    const bool fast = d->cfg.fast_history_clamp != 0;
    const bool comp = d->storage == VGI_STORAGE_F16 && d->cfg.compensated_moments != 0;
    const int c = d->cur;
    vgi_status s;

    // Pass: Temporal Filter (denoising.rs:136-160)
    vgi_denoise_temporal_io t{};
    t.width = d->w;
    t.height = d->h;
    t.storage = d->storage;
    t.depth_texture = in->depth;
    t.gbuffer_texture = in->gbuffer;
    t.diff_source_texture = in->diffuse;
    t.spec_source_texture = in->specular;
    t.spec_ray_len_texture = in->specular_ray_len;
    if (has_history) {
        t.prev_depth_texture = in->prev_depth;
        t.prev_gbuffer_texture = in->prev_gbuffer;
        t.diff_history_texture = d->hist_diff;
        t.spec_history_texture = d->hist_spec;
        t.moments_history_texture = d->hist_mom;
        t.moments_comp_history_texture = comp ? d->hist_comp : nullptr;  // NULL reads as a zero residual
        t.history_len_texture = d->hist_len;
    }
    const bool fast_hist_in = fast && d->has_fast_history && has_history;
    if (fast) {
        // Without a fast history the reference binds a dummy (zero) texture;
        // the blend then starts from 0 exactly as for a disocclusion.
        if (!fast_hist_in) {
            if ((s = sink.submit(kSlotDenClearFastDiff, make_clear(d->fast_diff[1 - c].p, d->fast_diff[1 - c].bytes))) != VGI_OK)
                return s;
            if ((s = sink.submit(kSlotDenClearFastSpec, make_clear(d->fast_spec[1 - c].p, d->fast_spec[1 - c].bytes))) != VGI_OK)
                return s;
        }
        t.diff_fast_history_texture = d->fast_diff[1 - c].p;
        t.spec_fast_history_texture = d->fast_spec[1 - c].p;
        t.rw_diff_fast_history_texture = d->fast_diff[c].p;
        t.rw_spec_fast_history_texture = d->fast_spec[c].p;
    }
    t.rw_diff_texture = d->temporal[c].diff.p;
    t.rw_spec_texture = d->temporal[c].spec.p;
    t.rw_moments_texture = d->temporal[c].pair.p;
    t.rw_moments_comp_texture = comp ? d->temporal[c].comp.p : nullptr;
    t.rw_history_len_texture = (uint8_t*)d->temporal_len[c].p;
    vgi_pc_denoise_temporal pc{{(float)d->w, (float)d->h}, has_history ? 1u : 0u, fast ? 1u : 0u};
    {
        PassRec rec = make_rec(kPassDenTemporal, fp);
        rec.io.den_temporal = t;
        rec.pc.den_temporal = pc;
        if ((s = sink.submit(kSlotDenTemporal, rec)) != VGI_OK) return s;
    }

    const void* cur_diff = d->temporal[c].diff.p;
    const void* cur_spec = d->temporal[c].spec.p;
    const void* cur_mom = d->temporal[c].pair.p;
    const void* cur_comp = comp ? d->temporal[c].comp.p : nullptr;
    const uint8_t* cur_len = (const uint8_t*)d->temporal_len[c].p;

    // Pass: Disocclusion Fix (denoising.rs:164-188)
    if (d->cfg.disocclusion_fix) {
        vgi_denoise_spatial_io x{};
        x.width = d->w;
        x.height = d->h;
        x.storage = d->storage;
        x.depth_texture = in->depth;
        x.gbuffer_texture = in->gbuffer;
        x.history_len_texture = cur_len;
        x.diff_texture = cur_diff;
        x.spec_texture = cur_spec;
        x.moments_texture = cur_mom;
        x.moments_comp_texture = cur_comp;
        x.rw_diff_texture = d->fix[c].diff.p;
        x.rw_spec_texture = d->fix[c].spec.p;
        x.rw_moments_or_variance_texture = d->fix[c].pair.p;
        x.rw_moments_comp_texture = comp ? d->fix[c].comp.p : nullptr;
        PassRec rec = make_rec(kPassDenDisocclusion);
        rec.io.den_spatial = x;
        if ((s = sink.submit(kSlotDenDisocclusion, rec)) != VGI_OK) return s;
        cur_diff = d->fix[c].diff.p;
        cur_spec = d->fix[c].spec.p;
        cur_mom = d->fix[c].pair.p;
        cur_comp = comp ? d->fix[c].comp.p : nullptr;
    }

    // Pass: Fast History Clamp (denoising.rs:192-207), in place
    if (fast) {
        vgi_denoise_fast_clamp_io fc{};
        fc.width = d->w;
        fc.height = d->h;
        fc.storage = d->storage;
        fc.diff_fast_history_texture = d->fast_diff[c].p;
        fc.spec_fast_history_texture = d->fast_spec[c].p;
        fc.rw_diff_texture = (void*)cur_diff;
        fc.rw_spec_texture = (void*)cur_spec;
        PassRec rec = make_rec(kPassDenFastClamp);
        rec.io.den_fast_clamp = fc;
        if ((s = sink.submit(kSlotDenFastClamp, rec)) != VGI_OK) return s;
    }
    d->has_fast_history = fast;

    // Temporal feedback before the spatial passes (denoising.rs:209-216)
    d->hist_diff = cur_diff;
    d->hist_spec = cur_spec;
    d->hist_mom = cur_mom;
    d->hist_comp = cur_comp;
    d->hist_len = cur_len;
    d->has_history = true;

    // Pass: Spatial Fallback (denoising.rs:218-242)
    vgi_denoise_spatial_io fb{};
    fb.width = d->w;
    fb.height = d->h;
    fb.storage = d->storage;
    fb.depth_texture = in->depth;
    fb.gbuffer_texture = in->gbuffer;
    fb.history_len_texture = cur_len;
    fb.diff_texture = cur_diff;
    fb.spec_texture = cur_spec;
    fb.moments_texture = cur_mom;
    fb.moments_comp_texture = cur_comp;
    fb.rw_diff_texture = d->pp[0].diff.p;
    fb.rw_spec_texture = d->pp[0].spec.p;
    fb.rw_moments_or_variance_texture = d->fb_var.p;
    {
        PassRec rec = make_rec(kPassDenFallback);
        rec.io.den_spatial = fb;
        if ((s = sink.submit(kSlotDenFallback, rec)) != VGI_OK) return s;
    }

    // Passes: A-Trous x N (denoising.rs:244-289)
    int src = 0;
    for (uint32_t it = 0; it < iters; ++it) {
        vgi_denoise_atrous_io a{};
        a.width = d->w;
        a.height = d->h;
        a.storage = d->storage;
        a.depth_texture = in->depth;
        a.gbuffer_texture = in->gbuffer;
        a.diff_texture = d->pp[src].diff.p;
        a.spec_texture = d->pp[src].spec.p;
        a.variance_texture = it == 0 ? d->fb_var.p : d->pp[src].pair.p;
        a.rw_diff_texture = d->pp[1 - src].diff.p;
        a.rw_spec_texture = d->pp[1 - src].spec.p;
        a.rw_variance_texture = d->pp[1 - src].pair.p;
        vgi_pc_denoise_atrous apc{(int32_t)(1u << it)};
        PassRec rec = make_rec(kPassDenAtrous);
        rec.io.den_atrous = a;
        rec.pc.den_atrous = apc;
        if ((s = sink.submit(kSlotDenAtrous0 + (int)it, rec)) != VGI_OK) return s;
        src = 1 - src;
    }
    out->diffuse = d->pp[src].diff.p;
    out->specular = d->pp[src].spec.p;
    out->history_len = cur_len;
    out->variance = d->fb_var.p;  // DebugViews.variance = fallback output (denoising.rs:296)
    d->cur = 1 - c;
    return VGI_OK;
}

extern "C" vgi_status vgi_denoiser_add_passes(vgi_denoiser* d, const vgi_frame_params* fp,
                                              const vgi_denoiser_input* in, vgi_denoiser_output* out,
                                              void* stream) {
    DirectSink sink((hipStream_t)stream);
    return vgi_denoiser_plan(d, fp, in, out, sink);
}
