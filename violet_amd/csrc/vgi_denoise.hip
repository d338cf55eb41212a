/*
 * vgi_denoise.hip — ReLAX/SVGF denoiser passes for gfx950.
 *
 * One wave64 per 8x8 pixel tile (lanes grouped in 2x2 quads, see
 * tile_pixel()).  All passes are HBM-bandwidth-bound gathers; neighbour
 * taps of the stencil passes hit L2/MALL because a tile's footprint is
 * shared with its neighbours' waves.
 *
 * Reference passes (violet, shader/denoise/):
 *   k_temporal        svgf_temporal_filter.hlsl:96-362
 *   k_disocclusion    relax_disocclusion_fix.hlsl:27-122
 *   k_fast_clamp      relax_fast_history_clamp.hlsl:12-75
 *   k_fallback        svgf_spatial_fallback_filter.hlsl:146-246
 *   k_atrous          svgf_atrous_filter.hlsl:24-88 (+ svgf.inc.hlsl:6-84)
 *
 * Pixels without geometry (depth == 0): the reference leaves the colour
 * outputs of the temporal, disocclusion and fallback passes unwritten
 * (undefined pooled-texture contents); these kernels write zeros there.
 */
#include <algorithm>

#include "vgi_batch_kernel.h"
#include "vgi_runtime.h"
#include "vgi_storage.h"

namespace vgi {

// ---------------------------------------------------------------------------
// Shared helpers
// ---------------------------------------------------------------------------

// pow(x, 2^k) by k squarings (x in [0,1]); sigma_n is 128 or 32 here.
__device__ __forceinline__ float pow_pow2(float x, int k) {
    for (int i = 0; i < k; ++i) x = x * x;
    return x;
}

constexpr float kLog2e = 1.4426950408889634f;

// G-buffer oct 12:12 normal with reciprocal multiplies and a hardware rsqrt.
__device__ __forceinline__ f3 decode_normal_fast(uint32_t enc) {
    const float k = 2.0f / 4095.0f;
    float fx = (float)(enc & 0xFFFu) * k - 1.0f;
    float fy = (float)((enc >> 12) & 0xFFFu) * k - 1.0f;
    float fz = (1.0f - fabsf(fx)) - fabsf(fy);
    float t = saturate(-fz);
    fx += (fx >= 0.0f) ? -t : t;
    fy += (fy >= 0.0f) ? -t : t;
    float inv = __builtin_amdgcn_rsqf(fx * fx + fy * fy + fz * fz);
    return mk3(fx * inv, fy * inv, fz * inv);
}

// EdgeStoppingFunc (svgf.inc.hlsl:6-84) is evaluated inline by the spatial
// passes: w = pow(sat(n.n_c), sigma_n) * exp(-|z - z_c| / (sigma_z fwidth_z
// dist + 1e-8) - |l - l_c| / (sigma_l std + 1e-8)), with exp(a) as
// exp2(a log2 e) and the divisions as reciprocal multiplies.

__device__ __forceinline__ f3 load_normal(const uint32_t* gb, size_t plane, size_t i) {
    return decode_gbuffer_normal(G(gb)[plane + i]).normal;
}

// Second moments.  With FP16 storage and a residual texture `lo`, a moment
// is kept compensated as hi + lo: hi = f16(m) and lo = f16(m - hi), which
// carries the bits the f16 rounding of the EMA drops (about 22 significant
// bits instead of 11 for values well inside the f16 normal range).
template <class S>
__device__ __forceinline__ f2 ld_moments(const void* hi, const void* lo, size_t i) {
    f2 m = S::ld2(hi, i);
    if (S::kMode == StorageF16::kMode && lo) {
        f2 c = StorageF16::ld2(lo, i);
        m = mk2(m.x + c.x, m.y + c.y);
    }
    return m;
}
template <class S>
__device__ __forceinline__ void st_moments(void* hi, void* lo, size_t i, f2 m) {
    S::st2(hi, i, m);
    if (S::kMode == StorageF16::kMode && lo) {
        const float hx = f16tof32(f32tof16(m.x)), hy = f16tof32(f32tof16(m.y));  // the stored high half
        StorageF16::st2(lo, i, mk2(m.x - hx, m.y - hy));
    }
}

// ---------------------------------------------------------------------------
// Temporal filter (svgf_temporal_filter.hlsl)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float dsr_weight(f3 color, f3 mean, f3 deviation) {
    float dist = luminance(abs3(color - mean) / deviation);
    return exp2f(-10.0f * dist);
}

template <class S>
__device__ __forceinline__ void k_temporal_body(const FrameParams& fp, const vgi_denoise_temporal_io& io, const vgi_pc_denoise_temporal& pc, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    const size_t plane = (size_t)w * h;

    float depth = G(io.depth_texture)[i];
    if (depth == 0.0f) {
        // (the reference leaves these texels unwritten; zeros keep the fast
        // history clamp's 3x3 box at the sky's edge defined)
        G(io.rw_history_len_texture)[i] = 0;
        S::st3(io.rw_diff_texture, i, mk3s(0.0f));
        S::st3(io.rw_spec_texture, i, mk3s(0.0f));
        st_moments<S>(io.rw_moments_texture, io.rw_moments_comp_texture, i, mk2(0.0f, 0.0f));
        if (pc.fast_history) {
            S::st3(io.rw_diff_fast_history_texture, i, mk3s(0.0f));
            S::st3(io.rw_spec_fast_history_texture, i, mk3s(0.0f));
        }
        return;
    }

    const float bsx = pc.buffer_size[0], bsy = pc.buffer_size[1];
    const int bix = (int)bsx, biy = (int)bsy;
    const float suv_x = ((float)px + 0.5f) / bsx;
    const float suv_y = ((float)py + 0.5f) / bsy;

    GBufferNormal gbn = decode_gbuffer_normal(G(io.gbuffer_texture)[plane + i]);
    f2 fwidth = decode_gbuffer_fwidth(G(io.gbuffer_texture)[2 * plane + i]);
    const bool has_history = pc.has_history != 0;
    const bool fast = pc.fast_history != 0;

    // Neighbourhood moments of the specular source (:118-132)
    f3 m1 = mk3s(0.0f), m2 = mk3s(0.0f);
    for (int x = -1; x <= 1; x++) {
        for (int y = -1; y <= 1; y++) {
            int cx = clampi((int)px + x, 0, bix - 1);
            int cy = clampi((int)py + y, 0, biy - 1);
            f3 s = S::ld3(io.spec_source_texture, (size_t)cy * w + cx);
            m1 = m1 + s;
            m2 = m2 + s * s;
        }
    }
    m1 = m1 * (1.0f / 9.0f);
    m2 = m2 * (1.0f / 9.0f);

    // Virtual (reflected-point) reprojection (:138-167)
    f3 virtual_spec_history = mk3s(0.0f);
    f3 virtual_spec_fast_history = mk3s(0.0f);
    f3 position;
    {
        f4 ph = mul44(fp.inv_view_proj, mk4(suv_x * 2.0f - 1.0f, suv_y * 2.0f - 1.0f, depth, 1.0f));
        position = mk3(ph.x / ph.w, ph.y / ph.w, ph.z / ph.w);
        f3 view_ray_dir = normalize3(position - xyz(fp.view_pos));
        float ray_len = S::ld1(io.spec_ray_len_texture, i);
        f3 vpos = position + view_ray_dir * ray_len;
        f4 hr = mul44(fp.prev_view_proj, mk4(vpos.x, vpos.y, vpos.z, 1.0f));
        float nx = hr.x / hr.w - fp.jitter[2];
        float ny = hr.y / hr.w - fp.jitter[3];
        float vu = (nx + fp.jitter[0]) * 0.5f + 0.5f;
        float vv = (ny + fp.jitter[1]) * 0.5f + 0.5f;
        bool valid = (vu == saturate(vu)) && (vv == saturate(vv)) && has_history;
        if (valid) {
            virtual_spec_history = sample_bilinear3<S>(io.spec_history_texture, w, h, vu, vv);
            if (fast) virtual_spec_fast_history = sample_bilinear3<S>(io.spec_fast_history_texture, w, h, vu, vv);
        }
    }

    // Reprojection (:170-189)
    float mvx, mvy, reproj_depth;
    {
        f4 hp = mul44(fp.view_proj, mk4(position.x, position.y, position.z, 1.0f));
        float nx = hp.x / hp.w - fp.jitter[0];
        float ny = hp.y / hp.w - fp.jitter[1];
        f4 hr = mul44(fp.prev_view_proj, mk4(position.x, position.y, position.z, 1.0f));
        float rx = hr.x / hr.w - fp.jitter[2];
        float ry = hr.y / hr.w - fp.jitter[3];
        mvx = (nx - rx) * 0.5f;
        mvy = (ny - ry) * 0.5f;
        reproj_depth = hr.z / hr.w;
    }
    float hu = suv_x - mvx;
    float hv = suv_y - mvy;

    f3 diff_history = mk3s(0.0f), spec_history = mk3s(0.0f);
    f2 sec_moments_history = mk2(0.0f, 0.0f);
    float history_len = 0.0f;
    f3 diff_fast_history = mk3s(0.0f), spec_fast_history = mk3s(0.0f);
    bool valid = (hu == saturate(hu)) && (hv == saturate(hv)) && has_history;
    if (valid) {
        float pcx = hu * bsx - 0.5f;
        float pcy = hv * bsy - 0.5f;
        float frx = pcx - floorf(pcx);
        float fry = pcy - floorf(pcy);
        int bx = (int)pcx, by = (int)pcy;  // truncation, as int2() in the reference
        float depth_tol = fwidth.x * 4.0f + 1e-4f;
        float normal_tol = fwidth.y * 16.0f + 1e-2f;
        float normal_tol_sq = normal_tol * normal_tol;
        float b = fry, r = frx, t = 1.0f - b, l = 1.0f - r;
        float wts[2][2] = {{t * l, t * r}, {b * l, b * r}};
        f3 dsum = mk3s(0.0f), ssum = mk3s(0.0f), dfsum = mk3s(0.0f), sfsum = mk3s(0.0f);
        f2 msum = mk2(0.0f, 0.0f);
        float lsum = 0.0f, wsum = 0.0f;
        for (int x = 0; x < 2; ++x) {
            for (int y = 0; y < 2; ++y) {
                int cx = bx + x, cy = by + y;
                bool coord_valid = cx >= 0 && cy >= 0 && cx <= bix - 1 && cy <= biy - 1;
                if (!coord_valid) continue;
                size_t j = (size_t)cy * w + cx;
                f3 n = load_normal(io.prev_gbuffer_texture, plane, j);
                f3 nd = n - gbn.normal;
                bool ok = (fabsf(G(io.prev_depth_texture)[j] - reproj_depth) < depth_tol) &&
                          (dot3(nd, nd) < normal_tol_sq);
                if (ok) {
                    float wgt = wts[y][x];
                    wsum += wgt;
                    dsum = dsum + S::ld3(io.diff_history_texture, j) * wgt;
                    ssum = ssum + S::ld3(io.spec_history_texture, j) * wgt;
                    f2 mh = ld_moments<S>(io.moments_history_texture, io.moments_comp_history_texture, j);
                    msum = mk2(msum.x + wgt * mh.x, msum.y + wgt * mh.y);
                    lsum += wgt * history_len_load(G(io.history_len_texture)[j]);
                    if (fast) {
                        dfsum = dfsum + S::ld3(io.diff_fast_history_texture, j) * wgt;
                        sfsum = sfsum + S::ld3(io.spec_fast_history_texture, j) * wgt;
                    }
                }
            }
        }
        valid = wsum > 0.00390625f;
        if (valid) {
            float inv = rcp(wsum);
            diff_history = dsum * inv;
            spec_history = ssum * inv;
            sec_moments_history = mk2(msum.x * inv, msum.y * inv);
            history_len = lsum * inv;
            diff_fast_history = dfsum * inv;
            spec_fast_history = sfsum * inv;
        }
    }
    if (!valid) {
        diff_history = mk3s(0.0f);
        spec_history = mk3s(0.0f);
        sec_moments_history = mk2(0.0f, 0.0f);
        history_len = 0.0f;
        diff_fast_history = mk3s(0.0f);
        spec_fast_history = mk3s(0.0f);
    }
    history_len += kHistoryLenUnit;

    f3 diff_source = S::ld3(io.diff_source_texture, i);
    f3 spec_source = S::ld3(io.spec_source_texture, i);
    float lx = luminance(diff_source), ly = luminance(spec_source);

    const float BF_DIFF = 1.0f / 20.0f;
    const float BF_FAST = 1.0f / 3.0f;
    float blurness = sqrtf(saturate(gbn.perceptual_roughness * 6.0f));
    float BF_SPEC = lerp(7.0f / 8.0f, BF_DIFF, blurness);
    float BF_SPEC_FAST = lerp(1.0f, BF_FAST, blurness);
    float len_alpha = kHistoryLenUnit / history_len;
    float ax = fmaxf(BF_DIFF, len_alpha), ay = fmaxf(BF_SPEC, len_alpha);
    f3 diff_filtered = lerp3(diff_history, diff_source, ax);
    f2 mom = mk2(lerp(sec_moments_history.x, lx * lx, ax), lerp(sec_moments_history.y, ly * ly, ay));
    float afx = fmaxf(BF_FAST, len_alpha), afy = fmaxf(BF_SPEC_FAST, len_alpha);
    f3 diff_filtered_fast = lerp3(diff_fast_history, diff_source, afx);

    // Dual-source blend [Stachowiak 2018] (:303-331)
    f3 spec_dev = max3s(sqrt3(abs3(m2 - m1 * m1)), 1e-19f);
    float wv = dsr_weight(virtual_spec_history, m1, spec_dev);
    float wh = dsr_weight(spec_history, m1, spec_dev);
    f3 dual = virtual_spec_history * wv + spec_history * wh;
    dual = dual / fmaxf(wv + wh, 1e-37f);
    f3 spec_filtered = lerp3(dual, spec_source, ay);
    f3 spec_filtered_fast = mk3s(0.0f);
    if (fast) {
        float wvf = dsr_weight(virtual_spec_fast_history, m1, spec_dev);
        float whf = dsr_weight(spec_fast_history, m1, spec_dev);
        f3 dualf = virtual_spec_fast_history * wvf + spec_fast_history * whf;
        dualf = dualf / fmaxf(wvf + whf, 1e-37f);
        spec_filtered_fast = lerp3(dualf, spec_source, afy);
    }

    S::st3(io.rw_diff_texture, i, diff_filtered);
    S::st3(io.rw_spec_texture, i, spec_filtered);
    st_moments<S>(io.rw_moments_texture, io.rw_moments_comp_texture, i, mom);
    G(io.rw_history_len_texture)[i] = history_len_store(history_len);
    if (fast) {
        S::st3(io.rw_diff_fast_history_texture, i, diff_filtered_fast);
        S::st3(io.rw_spec_fast_history_texture, i, spec_filtered_fast);
    }
}

template <class S>
__global__ void __launch_bounds__(64) k_temporal(FrameParams fp, vgi_denoise_temporal_io io, vgi_pc_denoise_temporal pc) {
    k_temporal_body<S>(fp, io, pc, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Quad max of the history length (QuadReadAcrossX/Y), lanes 4q..4q+3.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float quad_max(float v) {
    v = fmaxf(v, __shfl_xor(v, 1));
    v = fmaxf(v, __shfl_xor(v, 2));
    return v;
}

// ---------------------------------------------------------------------------
// Disocclusion fix (relax_disocclusion_fix.hlsl)
// ---------------------------------------------------------------------------
template <class S>
__device__ __forceinline__ void k_disocclusion_body(const vgi_denoise_spatial_io& io, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    const bool inside = px < w && py < h;
    const size_t i = (size_t)py * w + px;
    const size_t plane = (size_t)w * h;
    float len = inside ? history_len_load(G(io.history_len_texture)[i]) : 0.0f;
    float qmax = quad_max(len);
    if (!inside) return;
    if (len == 0.0f) {
        S::st3(io.rw_diff_texture, i, mk3s(0.0f));
        S::st3(io.rw_spec_texture, i, mk3s(0.0f));
        st_moments<S>(io.rw_moments_or_variance_texture, io.rw_moments_comp_texture, i, mk2(0.0f, 0.0f));
        return;
    }
    len = qmax;
    if (len < 4.0f * kHistoryLenUnit) {
        // EdgeStoppingFunc with sigma_z = 5, sigma_n = 32 and no luminance
        // term: w = pow(sat(n.n_c), 32) * exp(-|z - z_c| / (5 fwidth_z dist + 1e-8))
        const float depth_center = G(io.depth_texture)[i];
        const f3 nc = decode_normal_fast(G(io.gbuffer_texture)[plane + i]);
        const float fwz = 5.0f * decode_gbuffer_fwidth(G(io.gbuffer_texture)[2 * plane + i]).x;
        f2 wsum = mk2(0.0f, 0.0f);
        f3 dsum = mk3s(0.0f), ssum = mk3s(0.0f);
        f2 msum = mk2(0.0f, 0.0f);
        const int stride = (int)((14.0f * kHistoryLenUnit) / len);
        const float fs = (float)stride;
        // A row of 5 taps at a time: the row's depth and normal loads are
        // issued together, then its colour and moment loads, so a pixel's
        // 25 taps cost 10 dependent round trips instead of up to 50.  The
        // loads are unconditional (a tap outside the image reads this
        // pixel and gets weight 0); the accumulation keeps the taps' order
        // and adds a tap only when its weight is positive, so a zero-weight
        // tap (whose colours may be anything) adds exactly nothing.
#pragma unroll
        for (int y = -2; y <= 2; y++) {
            size_t jt[5];
            bool in[5];
            float zt[5];
            uint32_t nt[5];
#pragma unroll
            for (int x = -2; x <= 2; x++) {
                const int cx = (int)px + x * stride, cy = (int)py + y * stride;
                in[x + 2] = cx >= 0 && cy >= 0 && cx <= (int)w - 1 && cy <= (int)h - 1;
                jt[x + 2] = in[x + 2] ? (size_t)cy * w + cx : i;
                zt[x + 2] = G(io.depth_texture)[jt[x + 2]];
                nt[x + 2] = G(io.gbuffer_texture)[plane + jt[x + 2]];
            }
            float wt[5];
#pragma unroll
            for (int x = -2; x <= 2; x++) {
                const float dist = fs * sqrtf((float)(x * x + y * y));  // folded per tap
                const float dz = fabsf(zt[x + 2] - depth_center);
                const f3 n = decode_normal_fast(nt[x + 2]);
                const float wn = pow_pow2(saturate(dot3(nc, n)), 5);
                const float wk =
                    wn * __builtin_amdgcn_exp2f(-(dz * (kLog2e * __builtin_amdgcn_rcpf(fwz * dist + 1e-8f))));
                wt[x + 2] = in[x + 2] ? wk : 0.0f;
            }
            f3 dv[5], sv[5];
            f2 mv[5];
#pragma unroll
            for (int x = 0; x < 5; x++) {
                dv[x] = S::ld3(io.diff_texture, jt[x]);
                sv[x] = S::ld3(io.spec_texture, jt[x]);
                mv[x] = ld_moments<S>(io.moments_texture, io.moments_comp_texture, jt[x]);
            }
            // left alone the compiler sinks each tap's loads into its
            // weight test below, one round trip per tap; using the values
            // here (the empty asm emits nothing) keeps the row's loads together
#pragma unroll
            for (int x = 0; x < 5; x++)
                asm volatile("" ::"v"(dv[x].x), "v"(dv[x].y), "v"(dv[x].z), "v"(sv[x].x), "v"(sv[x].y), "v"(sv[x].z),
                             "v"(mv[x].x), "v"(mv[x].y));
#pragma unroll
            for (int x = 0; x < 5; x++) {
                if (wt[x] > 0.0f) {
                    wsum = mk2(wsum.x + wt[x], wsum.y + wt[x]);
                    dsum = dsum + dv[x] * wt[x];
                    ssum = ssum + sv[x] * wt[x];
                    msum = mk2(msum.x + mv[x].x * wt[x], msum.y + mv[x].y * wt[x]);
                }
            }
        }
        const float iw = __builtin_amdgcn_rcpf(wsum.x);
        S::st3(io.rw_diff_texture, i, dsum * iw);
        S::st3(io.rw_spec_texture, i, ssum * iw);
        st_moments<S>(io.rw_moments_or_variance_texture, io.rw_moments_comp_texture, i,
                      mk2(msum.x * iw, msum.y * iw));
    } else {
        S::st3(io.rw_diff_texture, i, S::ld3(io.diff_texture, i));
        S::st3(io.rw_spec_texture, i, S::ld3(io.spec_texture, i));
        st_moments<S>(io.rw_moments_or_variance_texture, io.rw_moments_comp_texture, i,
                      ld_moments<S>(io.moments_texture, io.moments_comp_texture, i));
    }
}

template <class S>
__global__ void __launch_bounds__(64) k_disocclusion(vgi_denoise_spatial_io io) {
    k_disocclusion_body<S>(io, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Spatial fallback / variance estimation (svgf_spatial_fallback_filter.hlsl)
//
// A 256-thread workgroup handles a 16x16 block (each wave an 8x8 quarter).
// Pixels whose quad history length is below 4 frames take a 7x7 edge-
// stopping filter; when a block holds any, its 22x22 footprint (depth,
// decoded normal, colours, moments) is staged in LDS once and the 49 taps
// read LDS.  Out-of-image footprint entries are zero: a zero normal gives
// weight 0 and zero colours, so such a tap adds exactly nothing, as the
// skipped tap of the reference's bounds test does.
// ---------------------------------------------------------------------------
constexpr int kFbB = 16;               // block edge
constexpr int kFbF = kFbB + 6;         // footprint edge (3 each side)
constexpr int kFbN = kFbF * kFbF;      // footprint entries
constexpr int kFbP = 24;               // footprint row pitch (conflict-free reads, as the a-trous pass)
constexpr int kFbRow = kFbF * kFbP;    // dwords per field
enum FallbackField { kGDepth, kGNx, kGNy, kGNz, kGDr, kGDg, kGDb, kGSr, kGSg, kGSb, kGMx, kGMy, kGCount };

template <class S>
__device__ __forceinline__ void k_fallback_body(const vgi_denoise_spatial_io& io, uint32_t bx_, uint32_t by_) {
    __shared__ float lds[kGCount][kFbRow];
    const int w = (int)io.width, h = (int)io.height;
    const size_t plane = (size_t)w * h;
    const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
    const int li = (wv & 1) * 8 + (lane & 7), lj = (wv >> 1) * 8 + (lane >> 3);
    const int bx = (int)bx_, by = (int)by_;
    const int px = bx * kFbB + li, py = by * kFbB + lj;
    const bool inside = px < w && py < h;
    const size_t i = (size_t)py * w + px;
    // this pixel's history length and the max over its 2x2 quad
    // (QuadReadAcrossX/Y; a partner outside the image counts 0).  A wave's
    // 8x8 pixels are row-major in its lanes, so the x partner is lane ^ 1
    // and the y partner lane ^ 8: two lane swaps instead of three loads.
    const float len0 = inside ? history_len_load(G(io.history_len_texture)[i]) : 0.0f;
    const float mx = fmaxf(len0, __shfl_xor(len0, 1));
    float len = len0;
    const float qmax = inside ? fmaxf(mx, __shfl_xor(mx, 8)) : 0.0f;
    const bool filtered = inside && len != 0.0f && qmax < 4.0f * kHistoryLenUnit;
    if (__syncthreads_or(filtered)) {
#pragma unroll 1
        for (int e = t; e < kFbN; e += 256) {
            const int gx = bx * kFbB + e % kFbF - 3, gy = by * kFbB + e / kFbF - 3;
            float v[kGCount];
#pragma unroll
            for (int k = 0; k < kGCount; ++k) v[k] = 0.0f;
            if (gx >= 0 && gy >= 0 && gx < w && gy < h) {
                const size_t j = (size_t)gy * w + gx;
                const f3 n = decode_normal_fast(G(io.gbuffer_texture)[plane + j]);
                const f3 dv = S::ld3(io.diff_texture, j);
                const f3 sv = S::ld3(io.spec_texture, j);
                const f2 mv = ld_moments<S>(io.moments_texture, io.moments_comp_texture, j);
                v[kGDepth] = G(io.depth_texture)[j];
                v[kGNx] = n.x, v[kGNy] = n.y, v[kGNz] = n.z;
                v[kGDr] = dv.x, v[kGDg] = dv.y, v[kGDb] = dv.z;
                v[kGSr] = sv.x, v[kGSg] = sv.y, v[kGSb] = sv.z;
                v[kGMx] = mv.x, v[kGMy] = mv.y;
            }
            const int le = (e / kFbF) * kFbP + e % kFbF;
#pragma unroll
            for (int k = 0; k < kGCount; ++k) lds[k][le] = v[k];
        }
        __syncthreads();
    }
    if (!inside) return;
    if (len == 0.0f) {
        S::st3(io.rw_diff_texture, i, mk3s(0.0f));
        S::st3(io.rw_spec_texture, i, mk3s(0.0f));
        S::st2(io.rw_moments_or_variance_texture, i, mk2(0.0f, 0.0f));
        return;
    }
    len = qmax;
    f3 diff, spec;
    f2 sm;
    if (filtered) {
        // EdgeStoppingFunc with fwidth_z * 3, sigma_n = 128 and no luminance
        // term, over a 7x7 window
        const int c = (lj + 3) * kFbP + (li + 3);
        const float depth_center = lds[kGDepth][c];
        const f3 nc = mk3(lds[kGNx][c], lds[kGNy][c], lds[kGNz][c]);
        const float fwz = decode_gbuffer_fwidth(G(io.gbuffer_texture)[2 * plane + i]).x * 3.0f;
        float wsum = 0.0f;
        f3 dsum = mk3s(0.0f), ssum = mk3s(0.0f);
        f2 msum = mk2(0.0f, 0.0f);
#pragma unroll
        for (int y = -3; y <= 3; y++) {
#pragma unroll
            for (int x = -3; x <= 3; x++) {
                const int j = c + y * kFbP + x;
                const float dist = sqrtf((float)(x * x + y * y));  // folded per tap
                const float dz = fabsf(lds[kGDepth][j] - depth_center);
                const f3 n = mk3(lds[kGNx][j], lds[kGNy][j], lds[kGNz][j]);
                const float wn = pow_pow2(saturate(dot3(nc, n)), 7);
                const float wt = wn * __builtin_amdgcn_exp2f(-(dz * (kLog2e * __builtin_amdgcn_rcpf(fwz * dist + 1e-8f))));
                wsum += wt;
                dsum = dsum + mk3(lds[kGDr][j], lds[kGDg][j], lds[kGDb][j]) * wt;
                ssum = ssum + mk3(lds[kGSr][j], lds[kGSg][j], lds[kGSb][j]) * wt;
                msum = mk2(msum.x + lds[kGMx][j] * wt, msum.y + lds[kGMy][j] * wt);
            }
        }
        const float iw = __builtin_amdgcn_rcpf(wsum);
        diff = dsum * iw;
        spec = ssum * iw;
        sm = mk2(msum.x * iw, msum.y * iw);
    } else {
        diff = S::ld3(io.diff_texture, i);
        spec = S::ld3(io.spec_texture, i);
        sm = ld_moments<S>(io.moments_texture, io.moments_comp_texture, i);
    }
    float l1x = luminance(diff), l1y = luminance(spec);
    f2 var = mk2(fmaxf(sm.x - l1x * l1x, 0.0f), fmaxf(sm.y - l1y * l1y, 0.0f));
    float boost = fmaxf(1.0f, (4.0f * kHistoryLenUnit) / len);
    var = mk2(var.x * boost, var.y * boost);
    S::st3(io.rw_diff_texture, i, diff);
    S::st3(io.rw_spec_texture, i, spec);
    S::st2(io.rw_moments_or_variance_texture, i, var);
}

// 5 waves per SIMD (a register budget under the LDS limit of six 25.6 KB
// workgroups per CU; at 6 the unrolled 7x7 window spills).
constexpr int kFallbackWaves = 5;
template <class S>
__global__ void __launch_bounds__(256, kFallbackWaves) k_fallback(vgi_denoise_spatial_io io) {
    k_fallback_body<S>(io, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Fast history clamp (relax_fast_history_clamp.hlsl), in place on rw_*
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 rgb_to_ycocg(f3 c) {
    float y = (c.x * 0.25f + c.y * 0.5f) + c.z * 0.25f;
    float co = c.x * 0.5f - c.z * 0.5f;
    float cg = (c.x * -0.25f + c.y * 0.5f) - c.z * 0.25f;
    return mk3(y, co, cg);
}
__device__ __forceinline__ f3 ycocg_to_rgb(f3 c) {
    float tmp = c.x - c.z;
    return mk3(tmp + c.y, c.x + c.z, tmp - c.y);
}
__device__ __forceinline__ f3 clamp3(f3 v, f3 lo, f3 hi) { return min3v(max3v(v, lo), hi); }

template <class S>
__device__ __forceinline__ void k_fast_clamp_body(const vgi_denoise_fast_clamp_io& io, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    f3 dm1 = mk3s(0.0f), dm2 = mk3s(0.0f), sm1 = mk3s(0.0f), sm2 = mk3s(0.0f);
    for (int y = -1; y <= 1; y++) {
        for (int x = -1; x <= 1; x++) {
            int cx = clampi((int)px + x, 0, (int)w - 1), cy = clampi((int)py + y, 0, (int)h - 1);
            size_t j = (size_t)cy * w + cx;
            f3 d = rgb_to_ycocg(S::ld3(io.diff_fast_history_texture, j));
            f3 s = rgb_to_ycocg(S::ld3(io.spec_fast_history_texture, j));
            dm1 = dm1 + d;
            dm2 = dm2 + d * d;
            sm1 = sm1 + s;
            sm2 = sm2 + s * s;
        }
    }
    const float ws = 9.0f;
    dm1 = dm1 / ws;
    dm2 = dm2 / ws;
    sm1 = sm1 / ws;
    sm2 = sm2 / ws;
    f3 ddev = sqrt3(dm2 - dm1 * dm1) * 0.75f;
    f3 sdev = sqrt3(sm2 - sm1 * sm1) * 0.75f;
    f3 d = rgb_to_ycocg(S::ld3(io.rw_diff_texture, i));
    f3 s = rgb_to_ycocg(S::ld3(io.rw_spec_texture, i));
    d = clamp3(d, dm1 - ddev, dm1 + ddev);
    s = clamp3(s, sm1 - sdev, sm1 + sdev);
    S::st3(io.rw_diff_texture, i, ycocg_to_rgb(d));
    S::st3(io.rw_spec_texture, i, ycocg_to_rgb(s));
}

template <class S>
__global__ void __launch_bounds__(64) k_fast_clamp(vgi_denoise_fast_clamp_io io) {
    k_fast_clamp_body<S>(io, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// A-trous iteration (svgf_atrous_filter.hlsl).  Taps with x != 0 AND y != 0
// plus the explicit centre (the reference's `not_center` test drops the
// whole centre row and column).
//
// Lattice tiling: a 256-thread workgroup filters a 16x16 block of pixels
// that share the residue (px mod step, py mod step).  The 5x5 taps at
// spacing `step` are then the block's lattice neighbours, so the 20x20
// lattice footprint is staged once in LDS (normal decoded, luminances
// computed) and every tap is an LDS read.  Out-of-image footprint entries
// hold a zero normal and zero colour: their edge weight is exactly 0, so
// they add nothing, as a skipped tap adds nothing.
//
// Arithmetic is FP32 with hardware transcendentals (exp2, rsqrt, rcp): the
// pass's contract is the FP64-oracle tolerance, not bit equality.
// ---------------------------------------------------------------------------
// Lattice block shapes: a workgroup of B x B threads filters a B x B block
// of lattice points (one residue class of the step) from its (B + 4)^2
// footprint staged in LDS.  B is a launch shape (vgi_tuning): 8, 16 (the
// default) or 32.  A footprint entry is three 16-byte records (diffuse
// colour and variance, specular colour and variance, normal and depth) and
// one 8-byte record (the two luminances), each record kind in its own array
// with rows P entries apart, so a tap is four LDS reads (ds_read_b128 x 3,
// ds_read_b64) instead of fourteen dword reads.  Each wave filters an 8x8
// part of the block; lanes 8r..8r+7 read 8 consecutive entries of row r.
// With P = 24 or 40, one row starts 4P = 32 (mod 64) banks after the one
// above for the 16-byte records (16 lanes, 64 distinct banks) and 2P = 48
// or 16 (mod 64) banks after it for the 8-byte records (32 lanes, 64
// distinct banks).
template <int B>
struct AtrousShape {
    static constexpr int F = B + 4;                // footprint edge (2 each side)
    static constexpr int N = F * F;                // footprint entries
    static constexpr int P = B == 32 ? 40 : 24;    // footprint row pitch
    static constexpr int Row = F * P;              // entries per record array
    static constexpr int T = B * B;                // threads per workgroup
    static_assert(P >= F, "pitch holds a footprint row");
};
typedef float AtrousV2 __attribute__((ext_vector_type(2)));
typedef float AtrousV4 __attribute__((ext_vector_type(4)));
enum AtrousRecord { kRDiff, kRSpec, kRNormal, kRCount };  // .xyz colour or normal, .w variance or depth

template <class S, int B>
__device__ __forceinline__ void k_atrous_body(const vgi_denoise_atrous_io& io, int step, uint32_t nbx, uint32_t nby, uint32_t bx_, uint32_t by_) {
    typedef AtrousShape<B> Sh;
    constexpr int kAtrousF = Sh::F, kAtrousN = Sh::N, kAtrousP = Sh::P, kAtrousRow = Sh::Row;
    __shared__ AtrousV4 lq[kRCount][kAtrousRow];
    __shared__ AtrousV2 llum[kAtrousRow];
    const int w = (int)io.width, h = (int)io.height;
    const size_t plane = (size_t)w * h;
    // residue innermost: the s x s residue classes of one block region are
    // consecutive tiles (the same XCD's L2 serves their shared cache lines)
    const int rx = (int)(bx_ % (uint32_t)step), bx = (int)(bx_ / (uint32_t)step);
    const int ry = (int)(by_ % (uint32_t)step), by = (int)(by_ / (uint32_t)step);
    const int t = (int)threadIdx.x;

    // Staging: a thread's footprint entries (at most kIter) are all loaded
    // before any is stored, so their memory round trips overlap.  The loads
    // are unconditional (an entry outside the image or past the footprint
    // reads pixel 0 and is then replaced by zeros): a load under a branch
    // would have to complete before the branch joins.
    constexpr int kIter = (kAtrousN + Sh::T - 1) / Sh::T;
    AtrousV4 rd[kIter], rs[kIter], rn[kIter];
#pragma unroll
    for (int it = 0; it < kIter; ++it) {
        const int e = t + it * Sh::T;
        const int gx = bx * B + e % kAtrousF - 2, gy = by * B + e / kAtrousF - 2;
        const int cx = rx + step * gx, cy = ry + step * gy;
        const bool inside = e < kAtrousN && gx >= 0 && gy >= 0 && cx < w && cy < h;
        const size_t j = inside ? (size_t)cy * w + cx : 0;
        const f3 n = decode_normal_fast(G(io.gbuffer_texture)[plane + j]);
        const f3 dv = S::ld3(io.diff_texture, j);
        const f3 sv = S::ld3(io.spec_texture, j);
        const f2 vv = S::ld2(io.variance_texture, j);
        const float z = G(io.depth_texture)[j];
        const AtrousV4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
        rd[it] = inside ? AtrousV4{dv.x, dv.y, dv.z, vv.x} : zero;
        rs[it] = inside ? AtrousV4{sv.x, sv.y, sv.z, vv.y} : zero;
        rn[it] = inside ? AtrousV4{n.x, n.y, n.z, z} : zero;
    }
#pragma unroll
    for (int it = 0; it < kIter; ++it) {
        const int e = t + it * Sh::T;
        if (e < kAtrousN) {
            const int le = (e / kAtrousF) * kAtrousP + e % kAtrousF;
            lq[kRDiff][le] = rd[it];
            lq[kRSpec][le] = rs[it];
            lq[kRNormal][le] = rn[it];
            llum[le] = AtrousV2{luminance(mk3(rd[it].x, rd[it].y, rd[it].z)), luminance(mk3(rs[it].x, rs[it].y, rs[it].z))};
        }
    }
    __syncthreads();

    const int lane = t & 63, wv = t >> 6;
    constexpr int kWavesPerRow = B / 8;
    const int li = (wv % kWavesPerRow) * 8 + (lane & 7), lj = (wv / kWavesPerRow) * 8 + (lane >> 3);
    const int px = rx + step * (bx * B + li), py = ry + step * (by * B + lj);
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    const int c = (lj + 2) * kAtrousP + (li + 2);

    const AtrousV4 cd = lq[kRDiff][c], cs = lq[kRSpec][c], cn = lq[kRNormal][c];
    const AtrousV2 cl = llum[c];
    const float depth_center = cn.w;
    const f3 nc = mk3(cn.x, cn.y, cn.z);
    const f2 lc = mk2(cl.x, cl.y);
    f3 dsum = mk3(cd.x, cd.y, cd.z);
    f3 ssum = mk3(cs.x, cs.y, cs.z);
    f2 vsum = mk2(cd.w, cs.w);
    const float fwz = decode_gbuffer_fwidth(G(io.gbuffer_texture)[2 * plane + i]).x;
    // exp(-|dl| / (sigma_l*std + eps) - dz / (fwidth_z*dist + eps)) as exp2 of
    // pre-scaled terms (sigma_l = 4, sigma_z = 1, eps = 1e-8)
    const float kl_d = kLog2e * __builtin_amdgcn_rcpf(4.0f * sqrtf(vsum.x) + 1e-8f);
    const float kl_s = kLog2e * __builtin_amdgcn_rcpf(4.0f * sqrtf(vsum.y) + 1e-8f);
    float kz[3];  // by tap class |x|+|y| - 2: distances sqrt(2), sqrt(5), sqrt(8) times step
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float a = (float)((q == 2 ? 2 : 1) * step), b = (float)((q == 0 ? 1 : 2) * step);
        kz[q] = kLog2e * __builtin_amdgcn_rcpf(fwz * sqrtf(a * a + b * b) + 1e-8f);
    }
    // Tap weight w = kw[|x|] kw[|y|] * ndot^128 * exp(-dl/(4 std) - dz/(fwz dist))
    // evaluated as one exp2 per colour of an exponent that carries every
    // factor: log2 of the kernel weight (a constant per tap), 128 log2(ndot)
    // (ndot = 0 gives -inf, hence w = 0), the depth term and the luminance
    // term.  Contracted to fused multiply-adds.
    const float lkw[3] = {0.0f, -0.5849625007211562f, -2.584962500721156f};  // log2 of 1, 2/3, 1/6
    f2 wsum = mk2(1.0f, 1.0f);
    // every tap read is at a constant offset from the footprint corner of
    // this pixel's 5x5 window: one base register per record array, kept
    // opaque so the compiler addresses the taps as immediate offsets from it
    // instead of adding an array offset per tap (the empty asm emits nothing)
    typedef __attribute__((address_space(3))) const AtrousV4* Lds4;
    typedef __attribute__((address_space(3))) const AtrousV2* Lds2;
    const int corner = c - 2 * kAtrousP - 2;
    Lds4 Q[kRCount];
#pragma unroll
    for (int k = 0; k < kRCount; ++k) {
        Q[k] = (Lds4)(&lq[k][0] + corner);
        asm volatile("" : "+v"(Q[k]));
    }
    Lds2 QL = (Lds2)(&llum[0] + corner);
    asm volatile("" : "+v"(QL));
#pragma unroll
    for (int y = -2; y <= 2; y++) {
#pragma unroll
        for (int x = -2; x <= 2; x++) {
            if (x == 0 || y == 0) continue;
            const int ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
            const int j = (y + 2) * kAtrousP + (x + 2);
            const AtrousV4 td = Q[kRDiff][j], ts = Q[kRSpec][j], tn = Q[kRNormal][j];
            const AtrousV2 tl = QL[j];
            const float dz = fabsf(tn.w - depth_center);
            const float ndot = saturate(__builtin_fmaf(nc.z, tn.z, __builtin_fmaf(nc.y, tn.y, nc.x * tn.x)));
            // sigma_n = 128
            const float e = __builtin_fmaf(__builtin_amdgcn_logf(ndot), 128.0f,
                                           __builtin_fmaf(dz, -kz[ax + ay - 2], lkw[ay] + lkw[ax]));
            const float wx = __builtin_amdgcn_exp2f(__builtin_fmaf(-fabsf(tl.x - lc.x), kl_d, e));
            const float wy = __builtin_amdgcn_exp2f(__builtin_fmaf(-fabsf(tl.y - lc.y), kl_s, e));
            dsum = mk3(__builtin_fmaf(td.x, wx, dsum.x), __builtin_fmaf(td.y, wx, dsum.y),
                       __builtin_fmaf(td.z, wx, dsum.z));
            ssum = mk3(__builtin_fmaf(ts.x, wy, ssum.x), __builtin_fmaf(ts.y, wy, ssum.y),
                       __builtin_fmaf(ts.z, wy, ssum.z));
            vsum = mk2(__builtin_fmaf(wx * wx, td.w, vsum.x), __builtin_fmaf(wy * wy, ts.w, vsum.y));
            wsum = mk2(wsum.x + wx, wsum.y + wy);
        }
    }
    // the output addresses are formed after the taps: redefining the pixel
    // index here (the empty asm emits nothing) keeps the compiler from
    // computing them up front and holding them through the loop, which
    // spills in FP16 mode (three 64-bit addresses per image)
    uint32_t io_px = (uint32_t)i;
    asm volatile("" : "+v"(io_px));
    const float id = __builtin_amdgcn_rcpf(wsum.x), is = __builtin_amdgcn_rcpf(wsum.y);
    S::st3(io.rw_diff_texture, io_px, dsum * id);
    S::st3(io.rw_spec_texture, io_px, ssum * is);
    S::st2(io.rw_variance_texture, io_px, mk2(vsum.x * (id * id), vsum.y * (is * is)));
}

// Register budget of the a-trous kernels in waves per SIMD: the LDS limit of
// the block shape (16: 26.9 KB workgroups of 4 waves, 6 per SIMD; 8: 16.1 KB
// single-wave workgroups; 32: 80.6 KB workgroups of 16 waves, one per CU half).
template <int B>
struct AtrousWaves {
    static constexpr int value = B == 8 ? 8 : B == 16 ? 6 : 4;
};

template <class S, int B>
__global__ void __launch_bounds__(B * B, AtrousWaves<B>::value) k_atrous(vgi_denoise_atrous_io io, int step, uint32_t nbx, uint32_t nby) {
    const uint32_t gx = gridDim.x;
    const uint32_t l = xcd_remap(blockIdx.y * gx + blockIdx.x, gx * gridDim.y);
    k_atrous_body<S, B>(io, step, nbx, nby, l % gx, l / gx);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static bool storage_ok(uint32_t s) { return s == VGI_STORAGE_F32 || s == VGI_STORAGE_F16; }

// Algorithmic bytes per pixel (each used field read once, each output
// written once): radiance R, pair P, scalar S bytes by storage mode.
static double px_bytes(uint32_t storage, int nR, int nP, int nS, int fixed) {
    return (double)(nR * (storage == VGI_STORAGE_F16 ? 6 : 12) + nP * (storage == VGI_STORAGE_F16 ? 4 : 8) +
                    nS * (storage == VGI_STORAGE_F16 ? 2 : 4) + fixed);
}

vgi_status launch_denoise_temporal(const FrameParams& fp, const vgi_denoise_temporal_io& io,
                                   const vgi_pc_denoise_temporal& pc, hipStream_t st) {
    // reads: depth 4, L1+L2 8, diff/spec src, ray len, prev depth 4, prev L1 4,
    // diff/spec/moments history, len 1; writes diff, spec, moments, len 1
    ProfScope prof("denoise_temporal", st,
                   px_bytes(io.storage, 6, 2, 1, 4 + 8 + 4 + 4 + 1 + 1) * io.width * io.height);
    dim3 g = tile_grid(io.width, io.height);
    if (io.storage == VGI_STORAGE_F32)
        hipLaunchKernelGGL(k_temporal<StorageF32>, g, dim3(64), 0, st, fp, io, pc);
    else
        hipLaunchKernelGGL(k_temporal<StorageF16>, g, dim3(64), 0, st, fp, io, pc);
    return check_launch("k_temporal");
}
vgi_status launch_denoise_disocclusion(const vgi_denoise_spatial_io& io, hipStream_t st) {
    ProfScope prof("denoise_disocclusion_fix", st, px_bytes(io.storage, 4, 2, 0, 1) * io.width * io.height);
    dim3 g = tile_grid(io.width, io.height);
    if (io.storage == VGI_STORAGE_F32)
        hipLaunchKernelGGL(k_disocclusion<StorageF32>, g, dim3(64), 0, st, io);
    else
        hipLaunchKernelGGL(k_disocclusion<StorageF16>, g, dim3(64), 0, st, io);
    return check_launch("k_disocclusion");
}
vgi_status launch_denoise_fallback(const vgi_denoise_spatial_io& io, hipStream_t st) {
    ProfScope prof("denoise_spatial_fallback", st, px_bytes(io.storage, 4, 2, 0, 1) * io.width * io.height);
    const dim3 g((io.width + kFbB - 1) / kFbB, (io.height + kFbB - 1) / kFbB);
    if (io.storage == VGI_STORAGE_F32)
        hipLaunchKernelGGL(k_fallback<StorageF32>, g, dim3(256), 0, st, io);
    else
        hipLaunchKernelGGL(k_fallback<StorageF16>, g, dim3(256), 0, st, io);
    return check_launch("k_fallback");
}
vgi_status launch_denoise_fast_clamp(const vgi_denoise_fast_clamp_io& io, hipStream_t st) {
    ProfScope prof("denoise_fast_history_clamp", st, px_bytes(io.storage, 6, 0, 0, 0) * io.width * io.height);
    dim3 g = tile_grid(io.width, io.height);
    if (io.storage == VGI_STORAGE_F32)
        hipLaunchKernelGGL(k_fast_clamp<StorageF32>, g, dim3(64), 0, st, io);
    else
        hipLaunchKernelGGL(k_fast_clamp<StorageF16>, g, dim3(64), 0, st, io);
    return check_launch("k_fast_clamp");
}
// Lattice blocks of one a-trous launch.  A step at or beyond the larger image
// edge puts every tap outside the image, exactly as that edge does: it is
// clamped there to bound the residue count.
struct AtrousGrid {
    uint32_t s, nbx, nby;
};
static AtrousGrid atrous_grid(uint32_t w, uint32_t h, uint32_t step, uint32_t block) {
    AtrousGrid g;
    g.s = step < std::max(w, h) ? step : std::max(w, h);
    const uint32_t lw = (w + g.s - 1) / g.s, lh = (h + g.s - 1) / g.s;  // lattice extent of residue 0
    g.nbx = (lw + block - 1) / block;
    g.nby = (lh + block - 1) / block;
    return g;
}

template <int B>
static void launch_atrous_shape(const vgi_denoise_atrous_io& io, const AtrousGrid& ag, hipStream_t st) {
    const dim3 g(ag.nbx * ag.s, ag.nby * ag.s, 1);
    if (io.storage == VGI_STORAGE_F32)
        hipLaunchKernelGGL((k_atrous<StorageF32, B>), g, dim3(B * B), 0, st, io, (int)ag.s, ag.nbx, ag.nby);
    else
        hipLaunchKernelGGL((k_atrous<StorageF16, B>), g, dim3(B * B), 0, st, io, (int)ag.s, ag.nbx, ag.nby);
}

vgi_status launch_denoise_atrous(const vgi_denoise_atrous_io& io, int step, hipStream_t st) {
    // reads depth 4, L1+L2 8, diff, spec, variance; writes diff, spec, variance
    ProfScope prof("denoise_atrous", st, px_bytes(io.storage, 4, 2, 0, 12) * io.width * io.height);
    const uint32_t block = tuning_atrous_block();
    const AtrousGrid ag = atrous_grid(io.width, io.height, (uint32_t)step, block);
    if (block == 8)
        launch_atrous_shape<8>(io, ag, st);
    else if (block == 32)
        launch_atrous_shape<32>(io, ag, st);
    else
        launch_atrous_shape<16>(io, ag, st);
    return check_launch("k_atrous");
}

// ---------------------------------------------------------------------------
// Fused multi-view launches (vgi_batch.h)
// ---------------------------------------------------------------------------
template <class S>
struct JobDenTemporal {
    FrameParams fp;
    vgi_denoise_temporal_io io;
    vgi_pc_denoise_temporal pc;
    __device__ static void run(const JobDenTemporal& j, uint32_t bx, uint32_t by) {
        k_temporal_body<S>(j.fp, j.io, j.pc, bx, by);
    }
};
template <class S>
struct JobDenDisocclusion {
    vgi_denoise_spatial_io io;
    __device__ static void run(const JobDenDisocclusion& j, uint32_t bx, uint32_t by) {
        k_disocclusion_body<S>(j.io, bx, by);
    }
};
template <class S>
struct JobDenFallback {
    vgi_denoise_spatial_io io;
    __device__ static void run(const JobDenFallback& j, uint32_t bx, uint32_t by) { k_fallback_body<S>(j.io, bx, by); }
};
template <class S>
struct BatchMinWaves<JobDenFallback<S>> {
    static constexpr int value = kFallbackWaves;
};
template <class S>
struct JobDenFastClamp {
    vgi_denoise_fast_clamp_io io;
    __device__ static void run(const JobDenFastClamp& j, uint32_t bx, uint32_t by) {
        k_fast_clamp_body<S>(j.io, bx, by);
    }
};
struct AtrousArgs {
    vgi_denoise_atrous_io io;
    int step;
    uint32_t nbx, nby;
};
template <int B>
struct AtrousJobs {
    template <class S>
    struct J : AtrousArgs {
        __device__ static void run(const J& j, uint32_t bx, uint32_t by) {
            k_atrous_body<S, B>(j.io, j.step, j.nbx, j.nby, bx, by);
        }
    };
};
template <class S>
struct BatchMinWaves<AtrousJobs<8>::J<S>> {
    static constexpr int value = AtrousWaves<8>::value;
};
template <class S>
struct BatchMinWaves<AtrousJobs<16>::J<S>> {
    static constexpr int value = AtrousWaves<16>::value;
};
template <class S>
struct BatchMinWaves<AtrousJobs<32>::J<S>> {
    static constexpr int value = AtrousWaves<32>::value;
};

static vgi_status fill_den_temporal(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobDenTemporal>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        memcpy(&j.fp, &r.fp, sizeof(j.fp));
        j.io = r.io.den_temporal;
        j.pc = r.pc.den_temporal;
        batch_add_pixel_tiles(t->hd, v, j.io.width, j.io.height);
    }
    return VGI_OK;
}
template <template <class> class J>
static vgi_status fill_den_spatial(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<J>(host);
    for (int v = 0; v < n; ++v) {
        t->jobs[v].io = recs[v]->io.den_spatial;
        batch_add_pixel_tiles(t->hd, v, recs[v]->io.den_spatial.width, recs[v]->io.den_spatial.height);
    }
    return VGI_OK;
}
static vgi_status fill_den_fallback(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobDenFallback>(host);
    for (int v = 0; v < n; ++v) {
        const vgi_denoise_spatial_io& io = recs[v]->io.den_spatial;
        t->jobs[v].io = io;
        batch_add_tiles(t->hd, v, (io.width + kFbB - 1) / kFbB, (io.height + kFbB - 1) / kFbB);
    }
    return VGI_OK;
}
static vgi_status fill_den_fast_clamp(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobDenFastClamp>(host);
    for (int v = 0; v < n; ++v) {
        t->jobs[v].io = recs[v]->io.den_fast_clamp;
        batch_add_pixel_tiles(t->hd, v, recs[v]->io.den_fast_clamp.width, recs[v]->io.den_fast_clamp.height);
    }
    return VGI_OK;
}
static vgi_status fill_den_atrous(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<AtrousJobs<16>::J>(host);
    const uint32_t block = tuning_atrous_block();
    for (int v = 0; v < n; ++v) {
        const vgi_denoise_atrous_io& io = recs[v]->io.den_atrous;
        const int step = recs[v]->pc.den_atrous.step_size;
        VGI_REQUIRE(step >= 1, "batched a-trous: bad step %d", step);
        const AtrousGrid g = atrous_grid(io.width, io.height, (uint32_t)step, block);
        AtrousArgs& a = t->jobs[v];
        a.io = io;
        a.step = (int)g.s;
        a.nbx = g.nbx;
        a.nby = g.nby;
        batch_add_tiles(t->hd, v, g.nbx * g.s, g.nby * g.s);
    }
    return VGI_OK;
}
// The table was filled for the current block shape; the launch uses it too.
static vgi_status launch_den_atrous(const void* dev, uint32_t storage, uint32_t grid, size_t lds, hipStream_t st) {
    switch (tuning_atrous_block()) {
        case 8: return batch_launch<AtrousJobs<8>::J, 64, true, true>(dev, storage, grid, lds, st);
        case 32: return batch_launch<AtrousJobs<32>::J, 1024, true, true>(dev, storage, grid, lds, st);
        default: return batch_launch<AtrousJobs<16>::J, 256, true, true>(dev, storage, grid, lds, st);
    }
}

const BatchOps* batch_ops_denoise(int kind) {
    static const BatchOps temporal{sizeof(BatchTable<JobDenTemporal<StorageF32>>), fill_den_temporal,
                                   batch_launch<JobDenTemporal>, 64, false};
    static const BatchOps disocclusion{sizeof(BatchTable<JobDenDisocclusion<StorageF32>>),
                                       fill_den_spatial<JobDenDisocclusion>, batch_launch<JobDenDisocclusion>, 64, false};
    static const BatchOps fallback{sizeof(BatchTable<JobDenFallback<StorageF32>>), fill_den_fallback,
                                   batch_launch<JobDenFallback, 256, true>, 256, false};
    static const BatchOps fast_clamp{sizeof(BatchTable<JobDenFastClamp<StorageF32>>), fill_den_fast_clamp,
                                     batch_launch<JobDenFastClamp>, 64, false};
    // (threads 256 marks the a-trous class of persistent group counts; the
    // launch's workgroup size follows the block shape)
    static const BatchOps atrous{sizeof(BatchTable<AtrousJobs<16>::J<StorageF32>>), fill_den_atrous,
                                 launch_den_atrous, 256, false};
    switch (kind) {
        case kPassDenTemporal: return &temporal;
        case kPassDenDisocclusion: return &disocclusion;
        case kPassDenFallback: return &fallback;
        case kPassDenFastClamp: return &fast_clamp;
        case kPassDenAtrous: return &atrous;
        default: return nullptr;
    }
}

}  // namespace vgi

// ---------------------------------------------------------------------------
// C ABI: per-pass entry points
// ---------------------------------------------------------------------------
using namespace vgi;

extern "C" vgi_status vgi_pass_denoise_temporal(const vgi_frame_params* fp, const vgi_denoise_temporal_io* io,
                                                const vgi_pc_denoise_temporal* pc, void* stream) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_denoise_temporal: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_denoise_temporal: bad size %ux%u", io->width,
                io->height);
    VGI_REQUIRE(storage_ok(io->storage), "vgi_pass_denoise_temporal: bad storage mode %u", io->storage);
    VGI_REQUIRE((uint32_t)pc->buffer_size[0] == io->width && (uint32_t)pc->buffer_size[1] == io->height,
                "vgi_pass_denoise_temporal: buffer_size does not match the image size");
    VGI_REQUIRE(io->depth_texture && io->gbuffer_texture && io->diff_source_texture && io->spec_source_texture &&
                    io->spec_ray_len_texture && io->rw_diff_texture && io->rw_spec_texture &&
                    io->rw_moments_texture && io->rw_history_len_texture,
                "vgi_pass_denoise_temporal: missing binding");
    if (pc->has_history)
        VGI_REQUIRE(io->prev_depth_texture && io->prev_gbuffer_texture && io->diff_history_texture &&
                        io->spec_history_texture && io->moments_history_texture && io->history_len_texture,
                    "vgi_pass_denoise_temporal: has_history=1 but a history binding is NULL");
    if (pc->fast_history)
        VGI_REQUIRE(io->rw_diff_fast_history_texture && io->rw_spec_fast_history_texture &&
                        (!pc->has_history || (io->diff_fast_history_texture && io->spec_fast_history_texture)),
                    "vgi_pass_denoise_temporal: fast_history=1 but a fast-history binding is NULL");
    FrameParams f;
    static_assert(sizeof(vgi_frame_params) == sizeof(FrameParams), "layout");
    memcpy(&f, fp, sizeof(f));
    return launch_denoise_temporal(f, *io, *pc, (hipStream_t)stream);
}

static vgi_status check_spatial_io(const vgi_denoise_spatial_io* io, const char* name) {
    VGI_REQUIRE(io, "%s: null argument", name);
    VGI_REQUIRE(valid_dims(io->width, io->height), "%s: bad size %ux%u", name, io->width, io->height);
    VGI_REQUIRE(storage_ok(io->storage), "%s: bad storage mode %u", name, io->storage);
    VGI_REQUIRE(io->depth_texture && io->gbuffer_texture && io->history_len_texture && io->diff_texture &&
                    io->spec_texture && io->moments_texture && io->rw_diff_texture && io->rw_spec_texture &&
                    io->rw_moments_or_variance_texture,
                "%s: missing binding", name);
    return VGI_OK;
}

extern "C" vgi_status vgi_pass_denoise_disocclusion_fix(const vgi_denoise_spatial_io* io, void* stream) {
    vgi_status s = check_spatial_io(io, "vgi_pass_denoise_disocclusion_fix");
    if (s != VGI_OK) return s;
    return launch_denoise_disocclusion(*io, (hipStream_t)stream);
}

extern "C" vgi_status vgi_pass_denoise_spatial_fallback(const vgi_denoise_spatial_io* io, void* stream) {
    vgi_status s = check_spatial_io(io, "vgi_pass_denoise_spatial_fallback");
    if (s != VGI_OK) return s;
    return launch_denoise_fallback(*io, (hipStream_t)stream);
}

extern "C" vgi_status vgi_pass_denoise_fast_history_clamp(const vgi_denoise_fast_clamp_io* io, void* stream) {
    VGI_REQUIRE(io, "vgi_pass_denoise_fast_history_clamp: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_denoise_fast_history_clamp: bad size");
    VGI_REQUIRE(storage_ok(io->storage), "vgi_pass_denoise_fast_history_clamp: bad storage mode");
    VGI_REQUIRE(io->diff_fast_history_texture && io->spec_fast_history_texture && io->rw_diff_texture &&
                    io->rw_spec_texture,
                "vgi_pass_denoise_fast_history_clamp: missing binding");
    return launch_denoise_fast_clamp(*io, (hipStream_t)stream);
}

extern "C" vgi_status vgi_pass_denoise_atrous(const vgi_denoise_atrous_io* io, const vgi_pc_denoise_atrous* pc,
                                              void* stream) {
    VGI_REQUIRE(io && pc, "vgi_pass_denoise_atrous: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_denoise_atrous: bad size %ux%u", io->width,
                io->height);
    VGI_REQUIRE(storage_ok(io->storage), "vgi_pass_denoise_atrous: bad storage mode %u", io->storage);
    VGI_REQUIRE(pc->step_size >= 1 && pc->step_size <= 4096, "vgi_pass_denoise_atrous: bad step_size %d",
                pc->step_size);
    VGI_REQUIRE(io->depth_texture && io->gbuffer_texture && io->diff_texture && io->spec_texture &&
                    io->variance_texture && io->rw_diff_texture && io->rw_spec_texture && io->rw_variance_texture,
                "vgi_pass_denoise_atrous: missing binding");
    return launch_denoise_atrous(*io, pc->step_size, (hipStream_t)stream);
}
