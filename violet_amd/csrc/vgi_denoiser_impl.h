/*
 * vgi_denoiser_impl.h — the denoiser object (shared with the renderer, which
 * embeds one per view and exports its history with the view state).
 */
#pragma once
#include <vector>

#include "vgi_internal.h"

namespace vgi {

struct ColorSet {
    DevBuf diff, spec, pair;
    DevBuf comp;  // FP16 storage: residual (low f16 halves) of compensated moments
};

}  // namespace vgi

struct vgi_denoiser {
    uint32_t w = 0, h = 0, storage = 0;
    vgi_denoiser_config cfg{1, 0, 3, 1};
    vgi::ColorSet temporal[2];
    vgi::DevBuf temporal_len[2];
    vgi::ColorSet fix[2];
    vgi::DevBuf fast_diff[2], fast_spec[2];
    vgi::ColorSet pp[2];  // fallback colour output + a-trous ping-pong (pair = variance)
    vgi::DevBuf fb_var;   // fallback variance (kept intact: the debug view of denoising.rs:296)
    int cur = 0;           // set index written this frame
    bool has_history = false;
    bool has_fast_history = false;
    // history views (point into the set written last frame)
    const void* hist_diff = nullptr;
    const void* hist_spec = nullptr;
    const void* hist_mom = nullptr;
    const void* hist_comp = nullptr;  // null when the history was written without compensation
    const uint8_t* hist_len = nullptr;
};


namespace vgi {
// Every device image of the denoiser with its size (for state export).
inline void denoiser_buffers(vgi_denoiser* d, std::vector<DevBuf*>& out) {
    for (int k = 0; k < 2; ++k) {
        out.push_back(&d->temporal[k].diff);
        out.push_back(&d->temporal[k].spec);
        out.push_back(&d->temporal[k].pair);
        out.push_back(&d->temporal[k].comp);
        out.push_back(&d->temporal_len[k]);
        out.push_back(&d->fix[k].diff);
        out.push_back(&d->fix[k].spec);
        out.push_back(&d->fix[k].pair);
        out.push_back(&d->fix[k].comp);
        out.push_back(&d->fast_diff[k]);
        out.push_back(&d->fast_spec[k]);
        out.push_back(&d->pp[k].diff);
        out.push_back(&d->pp[k].spec);
        out.push_back(&d->pp[k].pair);
    }
    out.push_back(&d->fb_var);
}
}  // namespace vgi
