/*
 * vgi_denoise_static.hip — one denoiser frame that allocates nothing
 * (vgi_denoise_frame_static).
 *
 * The pass chain is vgi_denoiser_add_passes' (denoising.rs:66-301:
 * temporal filter, disocclusion fix, spatial fallback, a-trous x N), run
 * through the same per-pass entry points.  The caller owns the inputs, the
 * outputs and the history (read from one set of buffers, written to
 * another); the images that live only inside the frame go to `scratch`, a
 * fixed device array that every GPU the library runs on holds until the
 * process ends:
 *
 *   slot A  temporal-filter colours and moments (+ FP16 moment residual);
 *           after the disocclusion fix, an a-trous ping-pong target
 *   slot B  spatial-fallback colours and variance; a-trous ping-pong target
 *
 * Each image starts on a 256-byte boundary.  Bytes per pixel: FP32 64,
 * FP16 36 (32 without compensated moments).  Nothing in the array needs to
 * start at zero: every pass writes each pixel of its outputs before a later
 * pass reads it.  Calls on one GPU share the array, so calls that overlap
 * on one GPU (two streams, or two host threads) race.
 */
#include <mutex>

#include "vgi_internal.h"

// The device array behind the frame's intermediate images.  In host code
// `scratch` only names the symbol: kernels get the device address that
// hipGetSymbolAddress returns for the current device (scratch_base()).
static __device__ __align__(256) unsigned char scratch[VGI_DENOISE_STATIC_SCRATCH_BYTES];

namespace vgi {
namespace {

constexpr size_t kScratchBytes = VGI_DENOISE_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device (null until first use)

// The current device's copy of `scratch`, looked up once per device.
vgi_status scratch_base(unsigned char** out) {
    int dev = -1;
    VGI_TRY_HIP(hipGetDevice(&dev));
    VGI_REQUIRE(dev >= 0 && dev < kMaxDevices, "vgi_denoise_frame_static: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        VGI_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        VGI_REQUIRE(p != nullptr, "vgi_denoise_frame_static: no device address for the scratch array");
        g_scratch_dev[dev] = (unsigned char*)p;
    }
    *out = g_scratch_dev[dev];
    return VGI_OK;
}

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Offsets of the frame's images inside `scratch`.
struct StaticLayout {
    size_t a_diff, a_spec, a_pair, a_comp, b_diff, b_spec, b_pair, total;
    StaticLayout(size_t n, uint32_t storage, bool comp) {
        const size_t R = align256(radiance_bytes(storage) * n), P = align256(pair_bytes(storage) * n);
        size_t o = 0;
        a_diff = o, o += R;
        a_spec = o, o += R;
        a_pair = o, o += P;
        a_comp = o, o += comp ? P : 0;
        b_diff = o, o += R;
        b_spec = o, o += R;
        b_pair = o, o += P;
        total = o;
    }
};

}  // namespace
}  // namespace vgi

using namespace vgi;

extern "C" vgi_status vgi_denoise_frame_static(const vgi_frame_params* fp, const vgi_denoise_frame_io* io,
                                               const vgi_denoiser_config* cfg, void* stream) {
    VGI_REQUIRE(fp && io, "vgi_denoise_frame_static: null argument");
    const vgi_denoiser_config c = cfg ? *cfg : vgi_denoiser_config{1, 0, 3, 1};
    const uint32_t w = io->width, h = io->height, S = io->storage;
    VGI_REQUIRE(valid_dims(w, h), "vgi_denoise_frame_static: bad size %ux%u", w, h);
    VGI_REQUIRE(S == VGI_STORAGE_F32 || S == VGI_STORAGE_F16, "vgi_denoise_frame_static: bad storage %u", S);
    VGI_REQUIRE(!c.fast_history_clamp,
                "vgi_denoise_frame_static: fast_history_clamp is not supported (use vgi_denoiser_add_passes)");
    const bool comp = S == VGI_STORAGE_F16 && c.compensated_moments != 0;
    const StaticLayout L((size_t)w * h, S, comp);
    VGI_REQUIRE(L.total <= kScratchBytes,
                "vgi_denoise_frame_static: %ux%u (%s) needs %zu bytes of scratch, above the maximum of %zu", w, h,
                S == VGI_STORAGE_F16 ? "FP16" : "FP32", L.total, kScratchBytes);
    VGI_REQUIRE(io->depth && io->gbuffer && io->diffuse && io->specular && io->specular_ray_len,
                "vgi_denoise_frame_static: missing input");
    VGI_REQUIRE(io->rw_history_diffuse && io->rw_history_specular && io->rw_history_moments && io->rw_history_len,
                "vgi_denoise_frame_static: missing history output");
    VGI_REQUIRE(!comp || io->rw_history_moments_comp,
                "vgi_denoise_frame_static: compensated FP16 moments need rw_history_moments_comp");
    VGI_REQUIRE(io->rw_diffuse && io->rw_specular, "vgi_denoise_frame_static: missing output");
    const bool has_history = io->history_len && io->prev_depth && io->prev_gbuffer;
    VGI_REQUIRE(!has_history || (io->history_diffuse && io->history_specular && io->history_moments),
                "vgi_denoise_frame_static: incomplete history");
    const uint32_t iters = c.atrous_iterations > 5 ? 5 : c.atrous_iterations;  // denoising.rs:73

    unsigned char* base = nullptr;
    VGI_CHECK(scratch_base(&base));
    void* a_diff = base + L.a_diff;
    void* a_spec = base + L.a_spec;
    void* a_pair = base + L.a_pair;
    void* a_comp = comp ? base + L.a_comp : nullptr;
    void* b_diff = base + L.b_diff;
    void* b_spec = base + L.b_spec;
    void* b_pair = base + L.b_pair;

    // Temporal filter (denoising.rs:136-160): into slot A ahead of the
    // disocclusion fix, else straight into the caller's history
    const bool fix = c.disocclusion_fix != 0;
    vgi_denoise_temporal_io t{};
    t.width = w;
    t.height = h;
    t.storage = S;
    t.depth_texture = io->depth;
    t.gbuffer_texture = io->gbuffer;
    t.diff_source_texture = io->diffuse;
    t.spec_source_texture = io->specular;
    t.spec_ray_len_texture = io->specular_ray_len;
    if (has_history) {
        t.prev_depth_texture = io->prev_depth;
        t.prev_gbuffer_texture = io->prev_gbuffer;
        t.diff_history_texture = io->history_diffuse;
        t.spec_history_texture = io->history_specular;
        t.moments_history_texture = io->history_moments;
        t.moments_comp_history_texture = comp ? io->history_moments_comp : nullptr;  // NULL reads as a zero residual
        t.history_len_texture = io->history_len;
    }
    t.rw_diff_texture = fix ? a_diff : io->rw_history_diffuse;
    t.rw_spec_texture = fix ? a_spec : io->rw_history_specular;
    t.rw_moments_texture = fix ? a_pair : io->rw_history_moments;
    t.rw_moments_comp_texture = comp ? (fix ? a_comp : io->rw_history_moments_comp) : nullptr;
    t.rw_history_len_texture = io->rw_history_len;
    const vgi_pc_denoise_temporal tpc{{(float)w, (float)h}, has_history ? 1u : 0u, 0u};
    VGI_CHECK(vgi_pass_denoise_temporal(fp, &t, &tpc, stream));

    // Disocclusion fix (denoising.rs:164-188): slot A -> the caller's history
    if (fix) {
        vgi_denoise_spatial_io x{};
        x.width = w;
        x.height = h;
        x.storage = S;
        x.depth_texture = io->depth;
        x.gbuffer_texture = io->gbuffer;
        x.history_len_texture = io->rw_history_len;
        x.diff_texture = a_diff;
        x.spec_texture = a_spec;
        x.moments_texture = a_pair;
        x.moments_comp_texture = a_comp;
        x.rw_diff_texture = io->rw_history_diffuse;
        x.rw_spec_texture = io->rw_history_specular;
        x.rw_moments_or_variance_texture = io->rw_history_moments;
        x.rw_moments_comp_texture = comp ? io->rw_history_moments_comp : nullptr;
        VGI_CHECK(vgi_pass_denoise_disocclusion_fix(&x, stream));
    }

    // Spatial fallback (denoising.rs:218-242) on the history: colours into
    // slot B (the output itself without a-trous iterations), variance into
    // the caller's debug view when given
    void* const out_var = io->rw_variance ? io->rw_variance : b_pair;
    vgi_denoise_spatial_io fb{};
    fb.width = w;
    fb.height = h;
    fb.storage = S;
    fb.depth_texture = io->depth;
    fb.gbuffer_texture = io->gbuffer;
    fb.history_len_texture = io->rw_history_len;
    fb.diff_texture = io->rw_history_diffuse;
    fb.spec_texture = io->rw_history_specular;
    fb.moments_texture = io->rw_history_moments;
    fb.moments_comp_texture = comp ? io->rw_history_moments_comp : nullptr;
    fb.rw_diff_texture = iters ? b_diff : io->rw_diffuse;
    fb.rw_spec_texture = iters ? b_spec : io->rw_specular;
    fb.rw_moments_or_variance_texture = out_var;
    VGI_CHECK(vgi_pass_denoise_spatial_fallback(&fb, stream));

    // A-trous x N (denoising.rs:244-289): ping-pong between slots B and A;
    // the last iteration writes the caller's colour outputs
    void* const slot_diff[2] = {a_diff, b_diff};
    void* const slot_spec[2] = {a_spec, b_spec};
    void* const slot_pair[2] = {a_pair, b_pair};
    int src = 1;
    for (uint32_t it = 0; it < iters; ++it) {
        const int dst = 1 - src;
        const bool last = it + 1 == iters;
        vgi_denoise_atrous_io a{};
        a.width = w;
        a.height = h;
        a.storage = S;
        a.depth_texture = io->depth;
        a.gbuffer_texture = io->gbuffer;
        a.diff_texture = slot_diff[src];
        a.spec_texture = slot_spec[src];
        a.variance_texture = it == 0 ? out_var : slot_pair[src];
        a.rw_diff_texture = last ? io->rw_diffuse : slot_diff[dst];
        a.rw_spec_texture = last ? io->rw_specular : slot_spec[dst];
        a.rw_variance_texture = slot_pair[dst];
        const vgi_pc_denoise_atrous apc{(int32_t)(1u << it)};
        VGI_CHECK(vgi_pass_denoise_atrous(&a, &apc, stream));
        src = dst;
    }
    return VGI_OK;
}
