/*
 * vgi_compose.hip — direct light and composition, the passes that follow
 * the GI path in violet's frame (restir_renderer.rs:831-911):
 *
 *   k_raytraced_shadow  raytraced_shadow.hlsl:9-53
 *   k_final_lighting    restir/final_lighting.hlsl:16-62
 *   k_temporal_aa       temporal_aa.hlsl:112-264 (restir_renderer.rs:914-948)
 */
#include "vgi_batch_kernel.h"
#include "vgi_internal.h"
#include "vgi_storage.h"
#include "vgi_traverse.h"

namespace vgi {
const SceneView& scene_view(const vgi_scene* s);

template <bool kCount>
__device__ __forceinline__ void k_raytraced_shadow_body(const FrameParams& fp, const SceneView& s,
                                                        const vgi_raytraced_shadow_io& io, uint32_t bx_, uint32_t by_) {
    VGI_TRAVERSAL_STACK(stack);
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    const float depth = G(io.gbuffer_depth)[i];
    if (depth == 0.0f) {
        G(io.rw_shadow)[i] = 0;
        return;
    }
    // depth_error = 2^-24 (24-bit unorm depth)
    const f3 position = depth_to_position(fp.inv_view_proj, px, py, w, h, depth + 1.0f / 16777216.0f);
    const bool occluded = trace_any<kCount>(s, position, xyz(fp.sun_dir), kRayTMin, kRayTMax, stack);
    G(io.rw_shadow)[i] = occluded ? 0 : 255;
}

template <bool kCount>
__global__ void __launch_bounds__(64, kTraceWaves) k_raytraced_shadow(FrameParams fp, SceneView s,
                                                                      vgi_raytraced_shadow_io io) {
    k_raytraced_shadow_body<kCount>(fp, s, io, blockIdx.x, blockIdx.y);
}

template <class S>
__device__ __forceinline__ void k_final_lighting_body(const FrameParams& fp, const vgi_final_lighting_io& io,
                                                      uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    const size_t plane = (size_t)w * h;
    if (G(io.gbuffer_color)[3 * plane + i] == 0u) {  // has_no_geometry: shading_path 0
        S::st3(io.rw_color_buffer, i, mk3(io.sky_radiance[0], io.sky_radiance[1], io.sky_radiance[2]));
        return;
    }
    const uint32_t l0 = G(io.gbuffer_color)[i];
    const f3 color = decode_gbuffer_color(l0);
    const float metallic = decode_gbuffer_metallic(l0);
    const GBufferNormal gn = decode_gbuffer_normal(G(io.gbuffer_color)[plane + i]);
    // direct light (final_lighting.hlsl:29-43)
    const float direct_atten = (float)G(io.shadow_mask_buffer)[i] / 255.0f;
    // cs_view_ray_direction(dispatch_id, buffer_size, 0): lerp(-0.5, 0.5, 0)
    // shifts the pixel centre back to its corner (frame_bindings.hlsl:71-79)
    const float ndx = (((float)px + 0.5f) + -0.5f) / (float)w * 2.0f - 1.0f;
    const float ndy = (((float)py + 0.5f) + -0.5f) / (float)h * 2.0f - 1.0f;
    const f4 e = mul44(fp.inv_view_proj, mk4(ndx, ndy, 1.0f, 1.0f));
    const f3 view_dir = -normalize3(mk3(e.x / e.w, e.y / e.w, e.z / e.w) - xyz(fp.view_pos));
    const float perceptual_roughness = fmaxf(gn.perceptual_roughness, 0.045f);  // DIRECT_SPECULAR_SUPRESSION
    const f3 diffuse_rho = get_diffuse_rho(color, metallic);
    const f3 specular_f0 = get_specular_f0(color, metallic);
    const f3 sun_dir = xyz(fp.sun_dir);
    const float NoL = saturate(dot3(gn.normal, sun_dir));
    const f3 direct = (eval_GGX_Lambertian(view_dir, sun_dir, gn.normal, perceptual_roughness, diffuse_rho,
                                           specular_f0) *
                       (NoL * direct_atten)) *
                      xyz(fp.sun_inten);
    // indirect specular re-modulation (:45-54)
    const float NoV = saturate(dot3(gn.normal, view_dir));
    const float roughness = fmaxf(gn.perceptual_roughness * gn.perceptual_roughness, 0.002f);
    const f3 ri = ggx_brdf_integral_approx(NoV, roughness, specular_f0);
    const float rmin = 1.0f / 65536.0f;
    const f3 specular_reflectance = mk3(fmaxf(ri.x, rmin), fmaxf(ri.y, rmin), fmaxf(ri.z, rmin));
    // indirect lighting (:56-59)
    const f3 ind = S::ld3(io.indirect_diffuse_texture, i) * diffuse_rho +
                   S::ld3(io.indirect_specular_texture, i) * specular_reflectance;
    S::st3(io.rw_color_buffer, i, direct + ind);
}

template <class S>
__global__ void __launch_bounds__(64) k_final_lighting(FrameParams fp, vgi_final_lighting_io io) {
    k_final_lighting_body<S>(fp, io, blockIdx.x, blockIdx.y);
}

// Mitchell-Netravali cubic with B = C = 1/3 (temporal_aa.hlsl:17-29).
__device__ __forceinline__ float mitchell(float x) {
    const float B = 1.0f / 3.0f, C = 1.0f / 3.0f;
    float y = 0.0f;
    const float x2 = x * x, x3 = x * x * x;
    if (x < 1.0f)
        y = (12.0f - 9.0f * B - 6.0f * C) * x3 + (-18.0f + 12.0f * B + 6.0f * C) * x2 + (6.0f - 2.0f * B);
    else if (x <= 2.0f)
        y = (-B - 6.0f * C) * x3 + (6.0f * B + 30.0f * C) * x2 + (-12.0f * B - 48.0f * C) * x + (8.0f * B + 24.0f * C);
    return y / 6.0f;
}

__device__ __forceinline__ float sat_nan0(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }  // NaN -> 0

// One pixel of the temporal AA resolve.  The reference also tracks the
// nearest-depth neighbour but then reprojects the pixel's own depth
// (temporal_aa.hlsl:186-187), so only the latter is computed here.
template <class S>
__device__ __forceinline__ void k_temporal_aa_body(const FrameParams& fp, const vgi_temporal_aa_io& io,
                                                   const vgi_pc_temporal_aa& pc, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    // the jittered samples sit at -jitter.xy * 0.5 * size pixels from the centre
    const float sox = -fp.jitter[0] * 0.5f * (float)w;
    const float soy = -fp.jitter[1] * 0.5f * (float)h;
    f3 accu = mk3s(0.0f), nmin = mk3s(1e5f), nmax = mk3s(-1e5f), m1 = mk3s(0.0f), m2 = mk3s(0.0f);
    float wsum = 0.0f;
#pragma unroll
    for (int x = -1; x <= 1; ++x) {
#pragma unroll
        for (int y = -1; y <= 1; ++y) {
            const int cx = clampi((int)px + x, 0, (int)w - 1), cy = clampi((int)py + y, 0, (int)h - 1);
            const float ox = (float)(cx - (int)px) + sox, oy = (float)(cy - (int)py) + soy;
            const f3 nb = S::ld3(io.source_texture, (size_t)cy * w + cx);
            const float wt = mitchell(sqrtf(ox * ox + oy * oy));
            accu = accu + nb * wt;
            wsum += wt;
            nmin = mk3(fminf(nmin.x, nb.x), fminf(nmin.y, nb.y), fminf(nmin.z, nb.z));
            nmax = mk3(fmaxf(nmax.x, nb.x), fmaxf(nmax.y, nb.y), fmaxf(nmax.z, nb.z));
            m1 = m1 + nb;
            m2 = m2 + nb * nb;
        }
    }
    const f3 source = mk3(accu.x / wsum, accu.y / wsum, accu.z / wsum);
    // reprojection of this pixel's depth (static scene, :183-193)
    const f3 pos = depth_to_position(fp.inv_view_proj, px, py, w, h, G(io.gbuffer_depth)[i]);
    const f4 hc = mul44(fp.view_proj, mk4(pos.x, pos.y, pos.z, 1.0f));
    const f4 hr = mul44(fp.prev_view_proj, mk4(pos.x, pos.y, pos.z, 1.0f));
    const float mvx = ((hc.x / hc.w - fp.jitter[0]) - (hr.x / hr.w - fp.jitter[2])) * 0.5f;
    const float mvy = ((hc.y / hc.w - fp.jitter[1]) - (hr.y / hr.w - fp.jitter[3])) * 0.5f;
    const float hu = ((float)px + 0.5f) / (float)w - mvx;
    const float hv = ((float)py + 0.5f) / (float)h - mvy;
    const bool not_in_last_view = hu != sat_nan0(hu) || hv != sat_nan0(hv);
    if (not_in_last_view || !pc.has_history) {
        S::st3(io.rw_target, i, source);
        return;
    }
    f3 hist = sample_bilinear3<S>(io.history_texture, w, h, hu, hv);
    // neighbourhood clamp (:217) and variance clip, gamma 1 (:225-232)
    hist = mk3(fminf(fmaxf(hist.x, nmin.x), nmax.x), fminf(fmaxf(hist.y, nmin.y), nmax.y),
               fminf(fmaxf(hist.z, nmin.z), nmax.z));
    const float inv_n = 1.0f / 9.0f;
    const f3 m1u = m1 * inv_n, m2u = m2 * inv_n;
    const f3 d2 = m2u - m1u * m1u;
    const f3 e = mk3(sqrtf(fabsf(d2.x)) + 1e-8f, sqrtf(fabsf(d2.y)) + 1e-8f, sqrtf(fabsf(d2.z)) + 1e-8f);
    const f3 v = hist - m1u;
    const float ma = fmaxf(fabsf(v.x / e.x), fmaxf(fabsf(v.y / e.y), fabsf(v.z / e.z)));
    if (ma > 1.0f) hist = m1u + mk3(v.x / ma, v.y / ma, v.z / ma);
    // anti-flicker weighted blend (:238-255)
    float sw = 1.0f / 32.0f, hw = 1.0f - 1.0f / 32.0f;
    const float rs = 1.0f / (fmaxf(fmaxf(source.x, source.y), source.z) + 1.0f);
    const float rh = 1.0f / (fmaxf(fmaxf(hist.x, hist.y), hist.z) + 1.0f);
    sw *= 1.0f / (1.0f + luminance(source * rs));
    hw *= 1.0f / (1.0f + luminance(hist * rh));
    const float den = fmaxf(sw + hw, 1e-5f);
    const f3 num = source * sw + hist * hw;
    S::st3(io.rw_target, i, mk3(num.x / den, num.y / den, num.z / den));
}

template <class S>
__global__ void __launch_bounds__(64) k_temporal_aa(FrameParams fp, vgi_temporal_aa_io io, vgi_pc_temporal_aa pc) {
    k_temporal_aa_body<S>(fp, io, pc, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Fused multi-view launches (vgi_batch.h)
// ---------------------------------------------------------------------------
template <class>
struct JobShadow {
    FrameParams fp;
    SceneView s;
    vgi_raytraced_shadow_io io;
    __device__ static void run(const JobShadow& j, uint32_t bx, uint32_t by) {
        k_raytraced_shadow_body<false>(j.fp, j.s, j.io, bx, by);
    }
};
template <class S>
struct BatchMinWaves<JobShadow<S>> {
    static constexpr int value = kTraceWaves;
};
template <class S>
struct JobFinalLighting {
    FrameParams fp;
    vgi_final_lighting_io io;
    __device__ static void run(const JobFinalLighting& j, uint32_t bx, uint32_t by) {
        k_final_lighting_body<S>(j.fp, j.io, bx, by);
    }
};

template <class S>
struct JobTemporalAA {
    FrameParams fp;
    vgi_temporal_aa_io io;
    vgi_pc_temporal_aa pc;
    __device__ static void run(const JobTemporalAA& j, uint32_t bx, uint32_t by) {
        k_temporal_aa_body<S>(j.fp, j.io, j.pc, bx, by);
    }
};

static vgi_status fill_shadow(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobShadow>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        VGI_REQUIRE(r.scene && valid_dims(r.io.shadow.width, r.io.shadow.height), "batched shadow: bad job");
        auto& j = t->jobs[v];
        memcpy(&j.fp, &r.fp, sizeof(j.fp));
        j.s = scene_view(r.scene);
        j.s.counters = nullptr;
        j.io = r.io.shadow;
        batch_add_pixel_tiles(t->hd, v, j.io.width, j.io.height);
    }
    return VGI_OK;
}
static vgi_status fill_final_lighting(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobFinalLighting>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        memcpy(&j.fp, &r.fp, sizeof(j.fp));
        j.io = r.io.final_lighting;
        batch_add_pixel_tiles(t->hd, v, j.io.width, j.io.height);
    }
    return VGI_OK;
}

static vgi_status fill_temporal_aa(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobTemporalAA>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        memcpy(&j.fp, &r.fp, sizeof(j.fp));
        j.io = r.io.taa;
        j.pc = r.pc.taa;
        batch_add_pixel_tiles(t->hd, v, j.io.width, j.io.height);
    }
    return VGI_OK;
}

const BatchOps* batch_ops_compose(int kind) {
    static const BatchOps shadow{sizeof(BatchTable<JobShadow<StorageF32>>), fill_shadow, batch_launch<JobShadow>, 64,
                                 true};
    static const BatchOps lighting{sizeof(BatchTable<JobFinalLighting<StorageF32>>), fill_final_lighting,
                                   batch_launch<JobFinalLighting>, 64, false};
    static const BatchOps taa{sizeof(BatchTable<JobTemporalAA<StorageF32>>), fill_temporal_aa,
                              batch_launch<JobTemporalAA>, 64, false};
    switch (kind) {
        case kPassShadow: return &shadow;
        case kPassFinalLighting: return &lighting;
        case kPassTemporalAA: return &taa;
        default: return nullptr;
    }
}

}  // namespace vgi

using namespace vgi;

extern "C" vgi_status vgi_pass_raytraced_shadow(const vgi_frame_params* fp, const vgi_scene* scene,
                                                const vgi_raytraced_shadow_io* io, void* stream) {
    VGI_REQUIRE(fp && scene && io, "vgi_pass_raytraced_shadow: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_raytraced_shadow: bad size");
    VGI_REQUIRE(io->gbuffer_depth && io->rw_shadow, "vgi_pass_raytraced_shadow: missing binding");
    FrameParams f;
    memcpy(&f, fp, sizeof(f));
    hipStream_t st = (hipStream_t)stream;
    // depth 4 + mask 1 per pixel (plus the traversal)
    ProfScope prof("raytraced_shadow", st, (double)io->width * io->height * 5.0);
    SceneView sv = scene_view(scene);
    sv.counters = trace_counters();
    const dim3 g = tile_grid(io->width, io->height);
    if (sv.counters)
        hipLaunchKernelGGL(k_raytraced_shadow<true>, g, dim3(64), traversal_lds_bytes(sv), st, f, sv, *io);
    else
        hipLaunchKernelGGL(k_raytraced_shadow<false>, g, dim3(64), traversal_lds_bytes(sv), st, f, sv, *io);
    trace_stats_collect("raytraced_shadow", st);
    return check_launch("k_raytraced_shadow");
}

extern "C" vgi_status vgi_pass_final_lighting(const vgi_frame_params* fp, const vgi_final_lighting_io* io,
                                              void* stream) {
    VGI_REQUIRE(fp && io, "vgi_pass_final_lighting: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_final_lighting: bad size");
    VGI_REQUIRE(io->storage == VGI_STORAGE_F32 || io->storage == VGI_STORAGE_F16, "vgi_pass_final_lighting: bad storage");
    VGI_REQUIRE(io->gbuffer_color && io->shadow_mask_buffer && io->indirect_diffuse_texture &&
                    io->indirect_specular_texture && io->rw_color_buffer,
                "vgi_pass_final_lighting: missing binding");
    FrameParams f;
    memcpy(&f, fp, sizeof(f));
    hipStream_t st = (hipStream_t)stream;
    // G-buffer layers 0, 1, 3 + mask + two radiance inputs + one output
    ProfScope prof("final_lighting", st, (double)io->width * io->height * (12.0 + 1.0 + 3.0 * radiance_bytes(io->storage)));
    const dim3 g = tile_grid(io->width, io->height);
    if (io->storage == VGI_STORAGE_F32)
        hipLaunchKernelGGL(k_final_lighting<StorageF32>, g, dim3(64), 0, st, f, *io);
    else
        hipLaunchKernelGGL(k_final_lighting<StorageF16>, g, dim3(64), 0, st, f, *io);
    return check_launch("k_final_lighting");
}

extern "C" vgi_status vgi_pass_temporal_aa(const vgi_frame_params* fp, const vgi_temporal_aa_io* io,
                                          const vgi_pc_temporal_aa* pc, void* stream) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_temporal_aa: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_temporal_aa: bad size");
    VGI_REQUIRE(io->storage == VGI_STORAGE_F32 || io->storage == VGI_STORAGE_F16, "vgi_pass_temporal_aa: bad storage");
    VGI_REQUIRE(io->gbuffer_depth && io->source_texture && io->rw_target && (io->history_texture || !pc->has_history),
                "vgi_pass_temporal_aa: missing binding");
    VGI_REQUIRE(io->source_texture != io->rw_target && io->history_texture != io->rw_target,
                "vgi_pass_temporal_aa: the target must not alias an input");
    FrameParams f;
    memcpy(&f, fp, sizeof(f));
    hipStream_t st = (hipStream_t)stream;
    // depth + source + history + target, each once per pixel
    ProfScope prof("temporal_aa", st, (double)io->width * io->height * (4.0 + 3.0 * radiance_bytes(io->storage)));
    const dim3 g = tile_grid(io->width, io->height);
    if (io->storage == VGI_STORAGE_F32)
        hipLaunchKernelGGL(k_temporal_aa<StorageF32>, g, dim3(64), 0, st, f, *io, *pc);
    else
        hipLaunchKernelGGL(k_temporal_aa<StorageF16>, g, dim3(64), 0, st, f, *io, *pc);
    return check_launch("k_temporal_aa");
}
