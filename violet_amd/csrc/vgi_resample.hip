/*
 * vgi_resample.hip — reservoir resampling and AO passes of the ReSTIR GI path:
 *
 *   k_ao_gen          restir/raytraced_ao_gen.hlsl:18-81
 *   k_ao_filter       restir/raytraced_ao_filter.hlsl:13-70
 *   k_diff_temporal   restir/ind_diff_temporal_resample.hlsl:49-200
 *   k_diff_spatial    restir/ind_diff_spatial_resample.hlsl:94-337
 *   k_spec_temporal   restir/ind_spec_temporal_resample.hlsl:146-362
 *   k_spec_spatial    restir/ind_spec_spatial_resample.hlsl:169-448
 *
 * One lane per pixel on 8x8 tiles.  Every float expression that feeds a
 * reservoir selection is written in the reference's evaluation order with
 * IEEE operations only (no fused multiply-add, deterministic sine/cosine),
 * so the selections equal the oracle's single-precision build bit for bit.
 * Geometry-free pixels also clear the outputs the reference leaves stale.
 */
#include <cmath>

#include "vgi_batch_kernel.h"
#include "vgi_internal.h"
#include "vgi_storage.h"

namespace vgi {

// Bilinear fetch of a scalar f32 image (sampler_linear_clamp), exact f32.
__device__ __forceinline__ float sample_bilinear1(const float* tex, uint32_t w, uint32_t h, float u, float v) {
    float tx = u * (float)w - 0.5f, ty = v * (float)h - 0.5f;
    float x0f = floorf(tx), y0f = floorf(ty);
    float fx = tx - x0f, fy = ty - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int xa = clampi(x0, 0, (int)w - 1), xb = clampi(x0 + 1, 0, (int)w - 1);
    int ya = clampi(y0, 0, (int)h - 1), yb = clampi(y0 + 1, 0, (int)h - 1);
    float c00 = G(tex)[(size_t)ya * w + xa], c10 = G(tex)[(size_t)ya * w + xb];
    float c01 = G(tex)[(size_t)yb * w + xa], c11 = G(tex)[(size_t)yb * w + xb];
    float top = c00 * (1.0f - fx) + c10 * fx;
    float bot = c01 * (1.0f - fx) + c11 * fx;
    return top * (1.0f - fy) + bot * fy;
}
template <class S>
__device__ __forceinline__ f2 sample_bilinear2(const void* tex, uint32_t w, uint32_t h, float u, float v) {
    float tx = u * (float)w - 0.5f, ty = v * (float)h - 0.5f;
    float x0f = floorf(tx), y0f = floorf(ty);
    float fx = tx - x0f, fy = ty - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int xa = clampi(x0, 0, (int)w - 1), xb = clampi(x0 + 1, 0, (int)w - 1);
    int ya = clampi(y0, 0, (int)h - 1), yb = clampi(y0 + 1, 0, (int)h - 1);
    f2 c00 = S::ld2(tex, (size_t)ya * w + xa), c10 = S::ld2(tex, (size_t)ya * w + xb);
    f2 c01 = S::ld2(tex, (size_t)yb * w + xa), c11 = S::ld2(tex, (size_t)yb * w + xb);
    f2 top = mk2(c00.x * (1.0f - fx) + c10.x * fx, c00.y * (1.0f - fx) + c10.y * fx);
    f2 bot = mk2(c01.x * (1.0f - fx) + c11.x * fx, c01.y * (1.0f - fx) + c11.y * fx);
    return mk2(top.x * (1.0f - fy) + bot.x * fy, top.y * (1.0f - fy) + bot.y * fy);
}

__device__ __forceinline__ uint32_t ld_u32(const uint32_t* p, size_t i) { return p[i]; }

// ---------------------------------------------------------------------------
// AO (raytraced_ao_gen.hlsl / raytraced_ao_filter.hlsl).  The reference
// stores AO in an RGBA16F image and uses .x (AO) and .a (valid); the pair
// (x, a) is stored here.
// ---------------------------------------------------------------------------
template <class S>
__device__ __forceinline__ void k_ao_gen_body(const FrameParams& fp, const vgi_ao_gen_io& io, const vgi_pc_ao_gen& pc, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    float depth = G(io.depth_buffer)[i];
    if (depth == 0.0f) {
        S::st2(io.rw_ao_texture, i, mk2(0.0f, 0.0f));
        return;
    }
    f3 position = depth_to_position(fp.inv_view_proj, px, py, w, h, depth);
    f4 hp = mul44(fp.view_proj, mk4(position.x, position.y, position.z, 1.0f));
    float nx = hp.x / hp.w - fp.jitter[0], ny = hp.y / hp.w - fp.jitter[1];
    f4 hr = mul44(fp.prev_view_proj, mk4(position.x, position.y, position.z, 1.0f));
    float rx = hr.x / hr.w - fp.jitter[2], ry = hr.y / hr.w - fp.jitter[3];
    float mvx = (nx - rx) * 0.5f, mvy = (ny - ry) * 0.5f;
    float hu = ((float)px + 0.5f) / (float)w - mvx;
    float hv = ((float)py + 0.5f) / (float)h - mvy;
    f2 hist = mk2(0.0f, 0.0f);
    if (hu == saturate(hu) && hv == saturate(hv) && pc.has_prev_frame)
        hist = sample_bilinear2<S>(io.prev_ao_texture, w, h, hu, hv);
    if (pc.has_new_sample) {
        HitPosNormal hit = hit_decode(G(io.hit_pos_normal_texture)[2 * i], G(io.hit_pos_normal_texture)[2 * i + 1]);
        f3 pos2 = depth_to_position(fp.inv_view_proj, px, py, w, h, depth);
        float dist = length3(hit.pos - pos2);
        float ao = dist < pc.radius_ws ? 0.0f : 1.0f;
        hist.x = lerp(ao, hist.x, ((32.0f - 1.0f) / 32.0f) * hist.y);
        hist.y = 1.0f;
    } else {
        hist.x = hist.y > 0.0f ? hist.x : 1.0f;
    }
    S::st2(io.rw_ao_texture, i, hist);
}

template <class S>
__global__ void __launch_bounds__(64) k_ao_gen(FrameParams fp, vgi_ao_gen_io io, vgi_pc_ao_gen pc) {
    k_ao_gen_body<S>(fp, io, pc, blockIdx.x, blockIdx.y);
}

__device__ __forceinline__ f2 min2v(f2 a, f2 b) { return mk2(fminf(a.x, b.x), fminf(a.y, b.y)); }
__device__ __forceinline__ f2 max2v(f2 a, f2 b) { return mk2(fmaxf(a.x, b.x), fmaxf(a.y, b.y)); }
__device__ __forceinline__ f2 clamp2v(f2 v, f2 lo, f2 hi) { return min2v(max2v(v, lo), hi); }

template <class S>
__device__ __forceinline__ void k_ao_filter_body(const vgi_ao_filter_io& io, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    float depth = G(io.depth_buffer)[i];
    if (depth == 0.0f) {
        S::st2(io.rw_filtered_ao_texture, i, mk2(0.0f, 0.0f));
        return;
    }
    f2 center = S::ld2(io.ao_texture, i);
    // the nine taps' depths and AO values are loaded together (every tap's
    // AO, then a select), one round trip instead of one or two per tap
    float sd[9];
    f2 av[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        // the reference reads min(id + (x,y), size) as unsigned (one texel out of
        // bounds at the far edges, wrapping at the near edges); clamped here
        const int x = k % 3 - 1, y = k / 3 - 1;
        const int cx = clampi((int)px + x, 0, (int)w - 1), cy = clampi((int)py + y, 0, (int)h - 1);
        const size_t j = (size_t)cy * w + cx;
        sd[k] = G(io.depth_buffer)[j];
        av[k] = S::ld2(io.ao_texture, j);
    }
    // using the values here (the empty asm emits nothing) keeps the compiler
    // from sinking each AO load into its depth test
#pragma unroll
    for (int k = 0; k < 9; k++) asm volatile("" ::"v"(av[k].x), "v"(av[k].y));
    f2 col_med[3], col_min[3], col_max[3];
    for (int y = -1; y <= 1; y++) {
        f2 row[3];
        for (int x = -1; x <= 1; x++) {
            const int k = (y + 1) * 3 + (x + 1);
            row[x + 1] = fabsf(depth - sd[k]) > 0.0005f ? center : av[k];
        }
        // the reference's "median" is clamp(a, b, c) = min(max(a, b), c)
        col_med[y + 1] = clamp2v(row[0], row[1], row[2]);
        col_min[y + 1] = min2v(min2v(row[0], row[1]), row[2]);
        col_max[y + 1] = max2v(max2v(row[0], row[1]), row[2]);
    }
    f2 A = min2v(min2v(col_max[0], col_max[1]), col_max[2]);
    f2 B = max2v(max2v(col_min[0], col_min[1]), col_min[2]);
    f2 C = clamp2v(col_med[0], col_med[1], col_med[2]);
    f2 med = clamp2v(C, A, B);
    f2 mn = min2v(min2v(col_min[0], col_min[1]), col_min[2]);
    S::st2(io.rw_filtered_ao_texture, i, mk2(lerp(med.x, mn.x, 0.5f), lerp(med.y, mn.y, 0.5f)));
}

template <class S>
__global__ void __launch_bounds__(64) k_ao_filter(vgi_ao_filter_io io) {
    k_ao_filter_body<S>(io, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Indirect diffuse temporal resampling
// ---------------------------------------------------------------------------

// Tap records as two 16-byte words.
typedef uint32_t U4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ U4 mk_u4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    U4 r;
    r.x = x;
    r.y = y;
    r.z = z;
    r.w = w;
    return r;
}
struct TapWords {
    U4 a, b;  // {depth, layer 1, reservoir, hit.x}, {hit.y, radiance rgb}
};
__device__ __forceinline__ void st_tap(DiffTap* p, const TapWords& t) {
    G((U4*)p)[0] = t.a;
    G((U4*)p)[1] = t.b;
}
__device__ __forceinline__ TapWords ld_tap(const DiffTap* p) {
    TapWords t;
    t.a = G((const U4*)p)[0];
    t.b = G((const U4*)p)[1];
    return t;
}

// Kernel arguments of the two temporal passes: the HLSL push constants plus
// the history M clamp, resolved on the host (M_MAX 16 unless configured).
// taps / gbuffer_color: when set (the renderer's own frames), the diffuse
// temporal pass also writes each pixel's tap record for the spatial pass.
struct DiffTemporalArgs {
    uint32_t frame_rand, has_prev_frame, has_new_sample, m_cap;
    DiffTap* taps;
    const uint32_t* gbuffer_color;
};
struct SpecTemporalArgs {
    uint32_t frame_rand, has_prev_frame, m_cap;
};

template <class S>
__device__ __forceinline__ void k_diff_temporal_body(const FrameParams& fp, const vgi_ind_diff_temporal_io& io, const DiffTemporalArgs& pc, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    const float depth = G(io.gbuffer_depth)[i];
    // (a pixel without geometry writes reservoir 0 here but carries on; the
    // final store overwrites it, as in the reference)
    f3 position = depth_to_position(fp.inv_view_proj, px, py, w, h, depth);
    f4 hr = mul44(fp.prev_view_proj, mk4(position.x, position.y, position.z, 1.0f));
    float nx = hr.x / hr.w - fp.jitter[2], ny = hr.y / hr.w - fp.jitter[3], nz = hr.z / hr.w - 0.0f;
    float u = nx * 0.5f + 0.5f, v = ny * 0.5f + 0.5f;
    bool in_view = fabsf(nx) < 1.0f && fabsf(ny) < 1.0f;
    bool sample_prev = pc.has_prev_frame && in_view;
    if (sample_prev) {
        float prev_depth = sample_bilinear1(io.prev_gbuffer_depth, w, h, u, v);
        if (fabsf(prev_depth - nz) > 0.001f) sample_prev = false;
    }
    uint32_t M = 0;
    float W = 0.0f;
    f3 rad = mk3s(0.0f);
    uint32_t ex = 0, ey = 0;
    if (sample_prev) {
        uint32_t ppx = minu(f2u(u * (float)w), w - 1u), ppy = minu(f2u(v * (float)h), h - 1u);
        size_t j = (size_t)ppy * w + ppx;
        ReservoirSimple rs = reservoir_decode_32b(G(io.prev_reservoir_texture)[j]);
        M = rs.M;
        W = rs.W;
        rad = S::ld3(io.prev_hit_radiance_texture, j);
        ex = G(io.prev_hit_pos_normal_texture)[2 * j];
        ey = G(io.prev_hit_pos_normal_texture)[2 * j + 1];
        if (M == 0u) {
            rad = mk3s(0.0f);
            ex = ey = 0u;
        }
        M = M < pc.m_cap ? M : pc.m_cap;
    }
    if (pc.has_new_sample) {
        float target = luminance(rad);
        float prev_w = (W * target) * (float)M;
        f3 new_rad = S::ld3(io.new_hit_radiance_texture, i);
        float new_target = luminance(new_rad);
        float new_w = new_target * kTwoPi;
        uint32_t rng = lcg_init_with_seed(px, py, pc.frame_rand);
        float w_sum = prev_w + new_w;
        float chance = new_w / w_sum;
        if (lcg_rand(rng) < chance) {
            ex = G(io.new_hit_pos_normal_texture)[2 * i];
            ey = G(io.new_hit_pos_normal_texture)[2 * i + 1];
            rad = new_rad;
            target = new_target;
        }
        M += 1u;
        W = target > 0.0f ? w_sum / (target * (float)M) : 0.0f;
    }
    ReservoirSimple rs;
    rs.M = M;
    rs.W = W;
    const uint32_t rsv = reservoir_encode_32b(rs);
    G(io.rw_reservoir_texture)[i] = rsv;
    G(io.rw_hit_pos_normal_texture)[2 * i] = ex;
    G(io.rw_hit_pos_normal_texture)[2 * i + 1] = ey;
    S::st3(io.rw_hit_radiance_texture, i, rad);
    if (pc.taps) {
        // the radiance as the spatial pass would read it back from storage
        const f3 rs3 = S::round3(rad);
        TapWords t;
        t.a = mk_u4(__float_as_uint(depth), G(pc.gbuffer_color)[(size_t)w * h + i], rsv, ex);
        t.b = mk_u4(ey, __float_as_uint(rs3.x), __float_as_uint(rs3.y), __float_as_uint(rs3.z));
        st_tap(pc.taps + i, t);
    }
}

template <class S>
__global__ void __launch_bounds__(64) k_diff_temporal(FrameParams fp, vgi_ind_diff_temporal_io io, DiffTemporalArgs pc) {
    k_diff_temporal_body<S>(fp, io, pc, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Indirect diffuse spatial resampling (golden-spiral taps)
// ---------------------------------------------------------------------------
constexpr int kMaxSpiralTaps = 16;
struct SpiralTable {
    float theta[kMaxSpiralTaps];  // TWO_PI * (i+1)/n
    float r[kMaxSpiralTaps];      // A * exp(B * theta_i)
    float T;                      // 0.6186 * n
    float cos_tol;                // cos(25 deg)
    int n;                        // taps (MAX_ITERATION)
};

struct DiffWR {
    uint32_t M;
    float W;
    f3 rad, pos, nrm;
    float target;
};

__device__ __forceinline__ DiffWR diff_make(uint32_t rsv, uint32_t ex, uint32_t ey, f3 rad, f3 pix_pos, f3 pix_n) {
    DiffWR r;
    ReservoirSimple rs = reservoir_decode_32b(rsv);
    r.M = rs.M;
    r.W = rs.W;
    HitPosNormal hit = hit_decode(ex, ey);
    r.pos = hit.pos;
    r.nrm = hit.normal;
    r.rad = rad;
    float nol = saturate(dot3(pix_n, normalize3(hit.pos - pix_pos)));
    r.target = nol * luminance(r.rad);
    return r;
}

template <class S>
__device__ __forceinline__ DiffWR diff_load(const vgi_ind_diff_spatial_io& io, size_t j, f3 pix_pos, f3 pix_n) {
    return diff_make(G(io.temp_reservoir_texture)[j], G(io.temp_hit_pos_normal_texture)[2 * j],
                     G(io.temp_hit_pos_normal_texture)[2 * j + 1], S::ld3(io.temp_hit_radiance_texture, j), pix_pos,
                     pix_n);
}
__device__ __forceinline__ DiffWR diff_make_tap(const TapWords& t, f3 pix_pos, f3 pix_n) {
    return diff_make(t.a.z, t.a.w, t.b.x, mk3(__uint_as_float(t.b.y), __uint_as_float(t.b.z), __uint_as_float(t.b.w)),
                     pix_pos, pix_n);
}

// kTaps: the pixels' depth, normal layer and temporal reservoirs come from
// the tap records (same values as the separate images, one load pair each).
template <class S, bool kTaps>
__device__ __forceinline__ void k_diff_spatial_body(const FrameParams& fp, const vgi_ind_diff_spatial_io& io, const DiffTap* taps, uint32_t frame_rand, float radius0, const SpiralTable& tab, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    const size_t plane = (size_t)w * h;
    TapWords ct;
    if constexpr (kTaps) ct = ld_tap(taps + i);
    float depth = kTaps ? __uint_as_float(ct.a.x) : G(io.gbuffer_depth)[i];
    if (depth == 0.0f) {
        G(io.rw_reservoir_texture)[i] = 0u;
        G(io.rw_hit_pos_normal_texture)[2 * i] = 0u;
        G(io.rw_hit_pos_normal_texture)[2 * i + 1] = 0u;
        S::st3(io.rw_hit_radiance_texture, i, mk3s(0.0f));
        S::st3(io.rw_lighting_texture, i, mk3s(0.0f));
        return;
    }
    const float rw = 1.0f / (float)w, rh = 1.0f / (float)h;
    const f3 position = depth_to_position_r(fp.inv_view_proj, px, py, w, h, rw, rh, depth);
    f3 nrm = decode_gbuffer_normal(kTaps ? ct.a.y : G(io.gbuffer_color)[plane + i]).normal;
    float ao = S::ld2(io.ao_texture, i).x;
    DiffWR res = kTaps ? diff_make_tap(ct, position, nrm) : diff_load<S>(io, i, position, nrm);
    float w_sum = (res.W * res.target) * (float)res.M;
    float radius = fmaxf(radius0 * ao, 1.5f);
    uint32_t rng = lcg_init_with_seed(px, py, frame_rand);
    float rand_angle = lcg_rand(rng) * kTwoPi;
    for (int k = 0; k < tab.n; k++) {
        float sv, cv;
        vgi_sincosf(tab.theta[k] * tab.T + rand_angle, &sv, &cv);
        float x = tab.r[k] * cv;
        float y = tab.r[k] * sv;
        int ox = (int)(x * radius), oy = (int)(y * radius);
        int cx = clampi((int)px + ox, 0, (int)w - 1), cy = clampi((int)py + oy, 0, (int)h - 1);
        size_t j = (size_t)cy * w + cx;
        TapWords tt;
        if constexpr (kTaps) tt = ld_tap(taps + j);
        f3 nn = decode_gbuffer_normal(kTaps ? tt.a.y : G(io.gbuffer_color)[plane + j]).normal;
        float depth_n = kTaps ? __uint_as_float(tt.a.x) : G(io.gbuffer_depth)[j];
        bool geometrical_diff = dot3(nrm, nn) < tab.cos_tol;
        f3 position_n = depth_to_position_r(fp.inv_view_proj, (uint32_t)cx, (uint32_t)cy, w, h, rw, rh, depth_n);
        float dist_to_plane = dot3(position_n - position, nrm);
        geometrical_diff = geometrical_diff || (fabsf(dist_to_plane) > 0.05f);
        if (geometrical_diff) {
            radius = fmaxf(radius * 0.5f, 1.5f);
            continue;
        }
        DiffWR nb = kTaps ? diff_make_tap(tt, position, nrm) : diff_load<S>(io, j, position, nrm);
        f3 off_qq = position_n - nb.pos;
        f3 off_rq = position - nb.pos;
        float L2_qq = dot3(off_qq, off_qq);
        float L2_rq = dot3(off_rq, off_rq);
        float cos_rq = dot3(nb.nrm, off_rq * rsqrt(L2_rq));
        float cos_qq = dot3(nb.nrm, off_qq * rsqrt(L2_qq));
        float jac = (cos_rq / cos_qq) * (L2_qq / L2_rq);
        jac = fminf(fmaxf(jac, 0.0f), 10.0f);
        jac = fminf(fmaxf(jac, 0.0f), 1.0f);
        float target_n = nb.target * jac;
        if (target_n <= 0.0f) continue;
        uint32_t nM = nb.M < 255u ? nb.M : 255u;
        float w_sum_n = (nb.W * target_n) * (float)nM;
        w_sum += w_sum_n;
        float chance = w_sum_n / w_sum;
        if (lcg_rand(rng) < chance) {
            res.nrm = nb.nrm;
            res.pos = nb.pos;
            res.rad = nb.rad;
            res.target = nb.target;
        }
        res.M += nM;
    }
    if (res.target > 0.0f) res.W = w_sum / (res.target * (float)res.M);
    ReservoirSimple rs;
    rs.M = res.M;
    rs.W = res.W;
    G(io.rw_reservoir_texture)[i] = reservoir_encode_32b(rs);
    uint32_t ex, ey;
    hit_encode(res.pos, res.nrm, ex, ey);
    G(io.rw_hit_pos_normal_texture)[2 * i] = ex;
    G(io.rw_hit_pos_normal_texture)[2 * i + 1] = ey;
    S::st3(io.rw_hit_radiance_texture, i, res.rad);
    f3 dir = normalize3(res.pos - position);
    float NoL = saturate(dot3(nrm, dir));
    S::st3(io.rw_lighting_texture, i, res.rad * ((kOneOverPi * NoL) * res.W));
}

template <class S>
__global__ void __launch_bounds__(64) k_diff_spatial(FrameParams fp, vgi_ind_diff_spatial_io io, uint32_t frame_rand, float radius0, SpiralTable tab) {
    k_diff_spatial_body<S, false>(fp, io, nullptr, frame_rand, radius0, tab, blockIdx.x, blockIdx.y);
}
template <class S>
__global__ void __launch_bounds__(64) k_diff_spatial_taps(FrameParams fp, vgi_ind_diff_spatial_io io, const DiffTap* taps, uint32_t frame_rand, float radius0, SpiralTable tab) {
    k_diff_spatial_body<S, true>(fp, io, taps, frame_rand, radius0, tab, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Specular target function (ind_spec_temporal_resample.hlsl:63-144,
// ind_spec_spatial_resample.hlsl:62-151)
// ---------------------------------------------------------------------------
struct SpecSampling {
    float source_pdf;
    f3 brdf_NoL_over_source_pdf;
};

__device__ __forceinline__ SpecSampling calc_things(f3 view_pos, f3 position, f3 normal, float perceptual_roughness,
                                                    f3 f0, f3 hit_pos) {
    SpecSampling r;
    float roughness = fmaxf(perceptual_roughness * perceptual_roughness, 0.002f);
    f3 view_dir = normalize3(view_pos - position);
    f3 light_dir = normalize3(hit_pos - position);
    f3 H = normalize3(view_dir + light_dir);
    float NoV = saturate(dot3(normal, view_dir));
    float NoH = saturate(dot3(normal, H));
    float D = D_GGX(NoH, roughness);
    float lambda_pomc = smith_lambda_GGX_plus_one_mul_cos(NoV, roughness);
    r.source_pdf = (0.25f * D) / lambda_pomc;
    float NoL = saturate(dot3(normal, light_dir));
    float LoH = saturate(dot3(light_dir, H));
    f3 F = F_Schlick(LoH, f0);
    float g = smith_G2_over_G1_height_correlated_GGX(NoL, NoV, roughness);
    if (NoL <= 0.0f) g = 0.0f;
    if (NoV <= 0.0f) g = 0.0f;
    r.brdf_NoL_over_source_pdf = F * g;
    return r;
}

struct GB {
    f3 color;
    float metallic;
    f3 normal;
    float roughness;
};
__device__ __forceinline__ GB load_gb(const uint32_t* g, size_t plane, size_t i) {
    GB r;
    uint32_t l0 = g[i], l1 = g[plane + i];
    r.color = decode_gbuffer_color(l0);
    r.metallic = decode_gbuffer_metallic(l0);
    GBufferNormal n = decode_gbuffer_normal(l1);
    r.normal = n.normal;
    r.roughness = n.perceptual_roughness;
    return r;
}

template <class S>
__device__ __forceinline__ void k_spec_temporal_body(const FrameParams& fp, const vgi_ind_spec_temporal_io& io, const SpecTemporalArgs& pc, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    const size_t plane = (size_t)w * h;
    float depth = G(io.gbuffer_depth)[i];
    if (depth == 0.0f) {
        G(io.rw_reservoir_texture)[2 * i] = 0u;
        G(io.rw_reservoir_texture)[2 * i + 1] = 0u;
        return;
    }
    const f3 position = depth_to_position(fp.inv_view_proj, px, py, w, h, depth);
    f4 hr = mul44(fp.prev_view_proj, mk4(position.x, position.y, position.z, 1.0f));
    float nx = hr.x / hr.w - fp.jitter[2], ny = hr.y / hr.w - fp.jitter[3], nz = hr.z / hr.w - 0.0f;
    float u = nx * 0.5f + 0.5f, v = ny * 0.5f + 0.5f;
    bool sample_prev = (u == saturate(u)) && (v == saturate(v)) && pc.has_prev_frame;
    if (sample_prev) {
        float prev_depth = sample_bilinear1(io.prev_gbuffer_depth, w, h, u, v);
        if (fabsf(prev_depth - nz) > 0.001f) sample_prev = false;
    }
    uint32_t M = 0;
    float W = 0.0f;
    f3 res_pos = mk3s(0.0f), res_rad = mk3s(0.0f);
    float res_tf = 0.0f;
    GB pg;
    pg.color = mk3s(0.0f);
    pg.metallic = 0.0f;
    pg.normal = mk3s(0.0f);
    pg.roughness = 0.0f;
    if (sample_prev) {
        uint32_t ppx = minu(f2u(u * (float)w), w - 1u), ppy = minu(f2u(v * (float)h), h - 1u);
        size_t j = (size_t)ppy * w + ppx;
        M = G(io.prev_reservoir_texture)[2 * j];
        W = asfloat(G(io.prev_reservoir_texture)[2 * j + 1]);
        const float* hp = io.prev_hit_pos_texture + 4 * j;
        res_pos = mk3(hp[0], hp[1], hp[2]);
        res_tf = hp[3];
        res_rad = S::ld3(io.prev_hit_radiance_texture, j);
        pg = load_gb(io.prev_gbuffer_color, plane, j);
    }
    const float* nhp = io.rw_hit_pos_texture + 4 * i;
    f3 new_pos = mk3(nhp[0], nhp[1], nhp[2]);
    f3 new_rad = S::ld3(io.rw_hit_radiance_texture, i);
    GB g = load_gb(io.gbuffer_color, plane, i);
    uint32_t M_max = f2u((float)pc.m_cap * g.roughness);
    M = M < M_max ? M : M_max;
    f3 f0 = get_specular_f0(g.color, g.metallic);
    SpecSampling cur = calc_things(xyz(fp.view_pos), position, g.normal, g.roughness, f0, new_pos);
    float new_w = luminance(new_rad * cur.brdf_NoL_over_source_pdf);
    float new_tf = new_w * cur.source_pdf;
    float prev_w = W * (res_tf * (float)M);
    if (M > 0u) {
        SpecSampling poc = calc_things(xyz(fp.prev_view_pos), position, pg.normal, pg.roughness,
                                       get_specular_f0(pg.color, pg.metallic), new_pos);
        float prev_tf_on_curr = luminance(new_rad * poc.brdf_NoL_over_source_pdf) * poc.source_pdf;
        float curr_m = new_tf / (new_tf + prev_tf_on_curr * (float)M);
        if (new_tf == 0.0f) curr_m = 0.0f;
        new_w = curr_m * new_w;
        SpecSampling cop = calc_things(xyz(fp.view_pos), position, g.normal, g.roughness, f0, res_pos);
        float curr_tf_on_prev = luminance(res_rad * cop.brdf_NoL_over_source_pdf) * cop.source_pdf;
        float m_num = res_tf * (float)M;
        float prev_m = m_num / (m_num + curr_tf_on_prev);
        if (m_num == 0.0f) prev_m = 0.0f;
        prev_w = prev_m * (curr_tf_on_prev * W);
        res_tf = curr_tf_on_prev;
    }
    uint32_t rng = lcg_init_with_seed(px, py, pc.frame_rand);
    const float w_sum = prev_w + new_w;
    const float chance_prev = prev_w / w_sum;
    const bool use_prev = lcg_rand(rng) < chance_prev;
    if (!use_prev) {
        res_pos = new_pos;
        res_rad = new_rad;
        res_tf = new_tf;
    }
    M += 1u;
    if (res_tf > 0.0f)
        W = w_sum / res_tf;  // the 1/M of the MIS weights is already in w_sum
    else
        W = 0.0f;
    G(io.rw_reservoir_texture)[2 * i] = M;
    G(io.rw_reservoir_texture)[2 * i + 1] = asuint(W);
    float* o = io.rw_hit_pos_texture + 4 * i;
    o[0] = res_pos.x;
    o[1] = res_pos.y;
    o[2] = res_pos.z;
    o[3] = res_tf;
    if (use_prev) S::st3(io.rw_hit_radiance_texture, i, res_rad);
}

template <class S>
__global__ void __launch_bounds__(64) k_spec_temporal(FrameParams fp, vgi_ind_spec_temporal_io io, SpecTemporalArgs pc) {
    k_spec_temporal_body<S>(fp, io, pc, blockIdx.x, blockIdx.y);
}

struct SpecTF {
    f3 view_pos, position, normal, f0;
    float pr;
    __device__ __forceinline__ float target(f3 hit_pos, f3 hit_rad) const {
        SpecSampling s = calc_things(view_pos, position, normal, pr, f0, hit_pos);
        return luminance(hit_rad * s.brdf_NoL_over_source_pdf) * s.source_pdf;
    }
};

template <class S>
__device__ __forceinline__ void k_spec_spatial_body(const FrameParams& fp, const vgi_ind_spec_spatial_io& io, uint32_t frame_rand, float radius0, float cos_tol, uint32_t bx_, uint32_t by_) {
    uint32_t px, py;
    tile_pixel_at(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    if (px >= w || py >= h) return;
    const size_t i = (size_t)py * w + px;
    const size_t plane = (size_t)w * h;
    float depth = G(io.gbuffer_depth)[i];
    if (depth == 0.0f) {
        S::st3(io.rw_lighting_texture, i, mk3s(0.0f));
        S::st1(io.rw_ray_len_texture, i, 0.0f);
        return;
    }
    const float rw = 1.0f / (float)w, rh = 1.0f / (float)h;
    const f3 position = depth_to_position_r(fp.inv_view_proj, px, py, w, h, rw, rh, depth);
    GB g = load_gb(io.gbuffer_color, plane, i);
    uint32_t M = G(io.reservoir_texture)[2 * i];
    float W = asfloat(G(io.reservoir_texture)[2 * i + 1]);
    const float* hp = io.hit_pos_texture + 4 * i;
    f3 hit_pos = mk3(hp[0], hp[1], hp[2]);
    f3 hit_rad = S::ld3(io.hit_radiance_texture, i);
    f3 f0 = get_specular_f0(g.color, g.metallic);
    SpecTF center{xyz(fp.view_pos), position, g.normal, f0, g.roughness};
    float res_tf = hp[3];
    const uint32_t num_iteration = f2u(0.0f + g.roughness * (4.0f - 0.0f));
    float radius = radius0;
    uint32_t rng = lcg_init_with_seed(px, py, frame_rand);
    for (uint32_t k = 0; k < num_iteration; k++) {
        float ux = lcg_rand(rng);
        float uy = lcg_rand(rng);
        float r = sqrtf(ux);
        float theta = kTwoPi * uy;
        float sv, cv;
        vgi_sincosf(theta, &sv, &cv);
        float x = r * cv, y = r * sv;
        int ox = (int)(x * radius), oy = (int)(y * radius);
        int cx = clampi((int)px + ox, 0, (int)w - 1), cy = clampi((int)py + oy, 0, (int)h - 1);
        if (cx == (int)px && cy == (int)py) continue;
        size_t j = (size_t)cy * w + cx;
        GB gn = load_gb(io.gbuffer_color, plane, j);
        float depth_n = G(io.gbuffer_depth)[j];
        bool diff = dot3(g.normal, gn.normal) < cos_tol;
        f3 position_n = depth_to_position_r(fp.inv_view_proj, (uint32_t)cx, (uint32_t)cy, w, h, rw, rh, depth_n);
        float dist_to_plane = dot3(position_n - position, g.normal);
        diff = diff || (fabsf(dist_to_plane) > 0.05f);
        if (diff) {
            radius = fmaxf(radius * 0.5f, 1.5f);
            continue;
        }
        uint32_t Mn = G(io.reservoir_texture)[2 * j];
        float Wn = asfloat(G(io.reservoir_texture)[2 * j + 1]);
        const float* hpn = io.hit_pos_texture + 4 * j;
        f3 hit_pos_n = mk3(hpn[0], hpn[1], hpn[2]);
        f3 hit_rad_n = S::ld3(io.hit_radiance_texture, j);
        float tf_n = hpn[3];
        SpecTF neighb{xyz(fp.view_pos), position_n, gn.normal, get_specular_f0(gn.color, gn.metallic), gn.roughness};
        float c_on_c = res_tf;
        float c_on_n = center.target(hit_pos_n, hit_rad_n);
        float m_neighb = (tf_n * (float)Mn) / (tf_n * (float)Mn + c_on_n * (float)M);
        if (tf_n * (float)Mn == 0.0f) m_neighb = 0.0f;
        float w_neighb = (m_neighb * tf_n) * Wn;
        float n_on_c = neighb.target(hit_pos, hit_rad);
        float m_center = (c_on_c * (float)M) / (c_on_c * (float)M + n_on_c * (float)Mn);
        if (c_on_c * (float)M == 0.0f) m_center = 0.0f;
        float w_center = (m_center * c_on_c) * W;
        float w_sum = w_center + w_neighb;
        float chance = w_neighb / w_sum;
        if (lcg_rand(rng) < chance) {
            hit_pos = hit_pos_n;
            hit_rad = hit_rad_n;
            res_tf = tf_n;
        }
        M += Mn;
        W = res_tf > 0.0f ? w_sum / res_tf : 0.0f;
    }
    float ray_len = length3(hit_pos - position);
    float roughness = fmaxf(g.roughness * g.roughness, 0.002f);
    f3 sel = normalize3(hit_pos - position);
    f3 view_dir = normalize3(xyz(fp.view_pos) - position);
    f3 H = normalize3(view_dir + sel);
    float NoL = saturate(dot3(g.normal, sel));
    float NoV = saturate(dot3(g.normal, view_dir));
    float NoH = saturate(dot3(g.normal, H));
    float LoH = saturate(dot3(sel, H));
    f3 F = F_Schlick(LoH, f0);
    float D = D_GGX(NoH, roughness);
    float lambda_pomc = smith_lambda_GGX_plus_one_mul_cos(NoV, roughness);
    float vndf = (0.25f * D) / lambda_pomc;
    float g21 = smith_G2_over_G1_height_correlated_GGX(NoL, NoV, roughness);
    if (NoL <= 0.0f) g21 = 0.0f;
    if (NoV <= 0.0f) g21 = 0.0f;
    f3 lighting = hit_rad * ((F * g21) * (vndf * W));
    f3 refl = max3s(ggx_brdf_integral_approx(NoV, roughness, f0), 1.0f / 65536.0f);
    lighting = lighting / refl;
    S::st3(io.rw_lighting_texture, i, lighting);
    S::st1(io.rw_ray_len_texture, i, ray_len);
}

template <class S>
__global__ void __launch_bounds__(64) k_spec_spatial(FrameParams fp, vgi_ind_spec_spatial_io io, uint32_t frame_rand, float radius0, float cos_tol) {
    k_spec_spatial_body<S>(fp, io, frame_rand, radius0, cos_tol, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static FrameParams to_fp(const vgi_frame_params* fp) {
    FrameParams f;
    memcpy(&f, fp, sizeof(f));
    return f;
}

// Constants of the reference spiral (ind_diff_spatial_resample.hlsl:148-151,
// 186) for n taps, folded in double and rounded to f32 once.
SpiralTable make_spiral_table(int n) {
    SpiralTable t;
    const double PI = (double)(float)3.14159265359;
    const float B = (float)(std::log(1.61803398) / (0.5 * PI));
    const float A = (float)(1.0 / std::exp((double)B * (double)(2.0f * (float)PI)));
    for (int i = 0; i < kMaxSpiralTaps; ++i) {
        float theta = i < n ? kTwoPi * (((float)i + 1.0f) / (float)n) : 0.0f;
        t.theta[i] = theta;
        t.r[i] = i < n ? (float)((double)A * std::exp((double)B * (double)theta)) : 0.0f;
    }
    t.T = (float)(0.6186 * (double)n);
    t.cos_tol = (float)std::cos((double)(float)(PI * (25.0 / 180.0)));
    t.n = n;
    return t;
}

// Initial spatial search radius: INIT_RADIUS_FACTOR * min(w, h) unless set.
static float spatial_radius(const vgi_resample_params* p, uint32_t w, uint32_t h) {
    return p && p->spatial_radius_px > 0.0f ? p->spatial_radius_px : (float)(w < h ? w : h) * 0.05f;
}
static uint32_t temporal_m_cap(const vgi_resample_params* p) { return p && p->temporal_m_cap ? p->temporal_m_cap : 16u; }
static int spiral_taps(const vgi_resample_params* p) { return p && p->diff_spatial_taps ? (int)p->diff_spatial_taps : 9; }
static vgi_status check_resample_params(const vgi_resample_params* p, const char* who) {
    if (!p) return VGI_OK;
    VGI_REQUIRE(p->diff_spatial_taps <= (uint32_t)kMaxSpiralTaps, "%s: diff_spatial_taps %u > %d", who,
                p->diff_spatial_taps, kMaxSpiralTaps);
    VGI_REQUIRE(p->spatial_radius_px >= 0.0f && p->spatial_radius_px <= 16384.0f, "%s: bad spatial_radius_px", who);
    VGI_REQUIRE(p->temporal_m_cap <= 255u, "%s: temporal_m_cap %u > 255", who, p->temporal_m_cap);
    return VGI_OK;
}

#define VGI_LAUNCH_S(mode, kernel, grid, ...)                                             \
    do {                                                                                  \
        if ((mode) == VGI_STORAGE_F32)                                                    \
            hipLaunchKernelGGL(kernel<StorageF32>, grid, dim3(64), 0, st, __VA_ARGS__);   \
        else                                                                              \
            hipLaunchKernelGGL(kernel<StorageF16>, grid, dim3(64), 0, st, __VA_ARGS__);   \
    } while (0)

static bool storage_ok(uint32_t s) { return s == VGI_STORAGE_F32 || s == VGI_STORAGE_F16; }

// ---------------------------------------------------------------------------
// Fused multi-view launches (vgi_batch.h)
// ---------------------------------------------------------------------------
template <class S>
struct JobAoGen {
    FrameParams fp;
    vgi_ao_gen_io io;
    vgi_pc_ao_gen pc;
    __device__ static void run(const JobAoGen& j, uint32_t bx, uint32_t by) { k_ao_gen_body<S>(j.fp, j.io, j.pc, bx, by); }
};
template <class S>
struct JobAoFilter {
    vgi_ao_filter_io io;
    __device__ static void run(const JobAoFilter& j, uint32_t bx, uint32_t by) { k_ao_filter_body<S>(j.io, bx, by); }
};
template <class S>
struct JobDiffTemporal {
    FrameParams fp;
    vgi_ind_diff_temporal_io io;
    DiffTemporalArgs pc;
    __device__ static void run(const JobDiffTemporal& j, uint32_t bx, uint32_t by) {
        k_diff_temporal_body<S>(j.fp, j.io, j.pc, bx, by);
    }
};
template <class S>
struct JobDiffSpatial {
    FrameParams fp;
    vgi_ind_diff_spatial_io io;
    const DiffTap* taps;
    uint32_t frame_rand;
    float radius0;
    SpiralTable tab;
    __device__ static void run(const JobDiffSpatial& j, uint32_t bx, uint32_t by) {
        if (j.taps)
            k_diff_spatial_body<S, true>(j.fp, j.io, j.taps, j.frame_rand, j.radius0, j.tab, bx, by);
        else
            k_diff_spatial_body<S, false>(j.fp, j.io, nullptr, j.frame_rand, j.radius0, j.tab, bx, by);
    }
};
template <class S>
struct JobSpecTemporal {
    FrameParams fp;
    vgi_ind_spec_temporal_io io;
    SpecTemporalArgs pc;
    __device__ static void run(const JobSpecTemporal& j, uint32_t bx, uint32_t by) {
        k_spec_temporal_body<S>(j.fp, j.io, j.pc, bx, by);
    }
};
template <class S>
struct JobSpecSpatial {
    FrameParams fp;
    vgi_ind_spec_spatial_io io;
    uint32_t frame_rand;
    float radius0;
    float cos_tol;
    __device__ static void run(const JobSpecSpatial& j, uint32_t bx, uint32_t by) {
        k_spec_spatial_body<S>(j.fp, j.io, j.frame_rand, j.radius0, j.cos_tol, bx, by);
    }
};

static vgi_status fill_ao_gen(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobAoGen>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        t->jobs[v] = {to_fp(&r.fp), r.io.ao_gen, r.pc.ao_gen};
        batch_add_pixel_tiles(t->hd, v, r.io.ao_gen.width, r.io.ao_gen.height);
    }
    return VGI_OK;
}
static vgi_status fill_ao_filter(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobAoFilter>(host);
    for (int v = 0; v < n; ++v) {
        t->jobs[v] = {recs[v]->io.ao_filter};
        batch_add_pixel_tiles(t->hd, v, recs[v]->io.ao_filter.width, recs[v]->io.ao_filter.height);
    }
    return VGI_OK;
}
static vgi_status fill_diff_temporal(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobDiffTemporal>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        const vgi_pc_ind_diff_temporal& pc = r.pc.diff_temporal;
        t->jobs[v] = {to_fp(&r.fp), r.io.diff_temporal,
                      {pc.frame_rand, pc.has_prev_frame, pc.has_new_sample, temporal_m_cap(&r.resample), r.diff_taps,
                       r.diff_taps_gbuffer}};
        batch_add_pixel_tiles(t->hd, v, r.io.diff_temporal.width, r.io.diff_temporal.height);
    }
    return VGI_OK;
}
static vgi_status fill_diff_spatial(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobDiffSpatial>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        VGI_CHECK(check_resample_params(&r.resample, "batched diffuse spatial"));
        t->jobs[v] = {to_fp(&r.fp), r.io.diff_spatial, r.diff_taps, r.pc.diff_spatial.frame_rand,
                      spatial_radius(&r.resample, r.io.diff_spatial.width, r.io.diff_spatial.height),
                      make_spiral_table(spiral_taps(&r.resample))};
        batch_add_pixel_tiles(t->hd, v, r.io.diff_spatial.width, r.io.diff_spatial.height);
    }
    return VGI_OK;
}
static vgi_status fill_spec_temporal(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobSpecTemporal>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        const vgi_pc_ind_spec_temporal& pc = r.pc.spec_temporal;
        t->jobs[v] = {to_fp(&r.fp), r.io.spec_temporal,
                      {pc.frame_rand, pc.has_prev_frame, temporal_m_cap(&r.resample)}};
        batch_add_pixel_tiles(t->hd, v, r.io.spec_temporal.width, r.io.spec_temporal.height);
    }
    return VGI_OK;
}
static vgi_status fill_spec_spatial(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobSpecSpatial>(host);
    static const SpiralTable tab = make_spiral_table(9);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        t->jobs[v] = {to_fp(&r.fp), r.io.spec_spatial, r.pc.spec_spatial.frame_rand,
                      spatial_radius(&r.resample, r.io.spec_spatial.width, r.io.spec_spatial.height),
                      tab.cos_tol};
        batch_add_pixel_tiles(t->hd, v, r.io.spec_spatial.width, r.io.spec_spatial.height);
    }
    return VGI_OK;
}

const BatchOps* batch_ops_resample(int kind) {
    static const BatchOps ao_gen{sizeof(BatchTable<JobAoGen<StorageF32>>), fill_ao_gen, batch_launch<JobAoGen>, 64, false};
    static const BatchOps ao_filter{sizeof(BatchTable<JobAoFilter<StorageF32>>), fill_ao_filter,
                                    batch_launch<JobAoFilter>, 64, false};
    static const BatchOps diff_temporal{sizeof(BatchTable<JobDiffTemporal<StorageF32>>), fill_diff_temporal,
                                        batch_launch<JobDiffTemporal>, 64, false};
    static const BatchOps diff_spatial{sizeof(BatchTable<JobDiffSpatial<StorageF32>>), fill_diff_spatial,
                                       batch_launch<JobDiffSpatial>, 64, false};
    static const BatchOps spec_temporal{sizeof(BatchTable<JobSpecTemporal<StorageF32>>), fill_spec_temporal,
                                        batch_launch<JobSpecTemporal>, 64, false};
    static const BatchOps spec_spatial{sizeof(BatchTable<JobSpecSpatial<StorageF32>>), fill_spec_spatial,
                                       batch_launch<JobSpecSpatial>, 64, false};
    switch (kind) {
        case kPassAoGen: return &ao_gen;
        case kPassAoFilter: return &ao_filter;
        case kPassDiffTemporal: return &diff_temporal;
        case kPassDiffSpatial: return &diff_spatial;
        case kPassSpecTemporal: return &spec_temporal;
        case kPassSpecSpatial: return &spec_spatial;
        default: return nullptr;
    }
}

}  // namespace vgi

using namespace vgi;

extern "C" vgi_status vgi_pass_raytraced_ao_gen(const vgi_frame_params* fp, const vgi_ao_gen_io* io,
                                                const vgi_pc_ao_gen* pc, void* stream) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_raytraced_ao_gen: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height) && storage_ok(io->storage), "vgi_pass_raytraced_ao_gen: bad size/storage");
    VGI_REQUIRE(io->depth_buffer && io->rw_ao_texture, "vgi_pass_raytraced_ao_gen: missing binding");
    VGI_REQUIRE(!pc->has_new_sample || io->hit_pos_normal_texture, "vgi_pass_raytraced_ao_gen: missing hit texture");
    VGI_REQUIRE(!pc->has_prev_frame || io->prev_ao_texture, "vgi_pass_raytraced_ao_gen: missing prev AO");
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("raytraced_ao_gen", st,
                   (double)io->width * io->height * (4 + 8 + 2 * pair_bytes(io->storage)));
    VGI_LAUNCH_S(io->storage, k_ao_gen, tile_grid(io->width, io->height), to_fp(fp), *io, *pc);
    return check_launch("k_ao_gen");
}

extern "C" vgi_status vgi_pass_raytraced_ao_filter(const vgi_ao_filter_io* io, void* stream) {
    VGI_REQUIRE(io, "vgi_pass_raytraced_ao_filter: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height) && storage_ok(io->storage), "vgi_pass_raytraced_ao_filter: bad size/storage");
    VGI_REQUIRE(io->depth_buffer && io->ao_texture && io->rw_filtered_ao_texture,
                "vgi_pass_raytraced_ao_filter: missing binding");
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("raytraced_ao_filter", st, (double)io->width * io->height * (4 + 2 * pair_bytes(io->storage)));
    VGI_LAUNCH_S(io->storage, k_ao_filter, tile_grid(io->width, io->height), *io);
    return check_launch("k_ao_filter");
}

vgi_status vgi::launch_diff_temporal(const vgi_frame_params* fp, const vgi_ind_diff_temporal_io* io,
                                     const vgi_pc_ind_diff_temporal* pc, const vgi_resample_params* params,
                                     DiffTap* taps, const uint32_t* gbuffer_color, hipStream_t st) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_ind_diff_temporal_resample: null argument");
    VGI_REQUIRE(!taps || gbuffer_color, "internal: diffuse tap records without the G-buffer");
    VGI_CHECK(check_resample_params(params, "vgi_pass_ind_diff_temporal_resample"));
    VGI_REQUIRE(valid_dims(io->width, io->height) && storage_ok(io->storage),
                "vgi_pass_ind_diff_temporal_resample: bad size/storage");
    VGI_REQUIRE(io->gbuffer_depth && io->rw_reservoir_texture && io->rw_hit_pos_normal_texture && io->rw_hit_radiance_texture,
                "vgi_pass_ind_diff_temporal_resample: missing binding");
    VGI_REQUIRE(!pc->has_new_sample || (io->new_hit_pos_normal_texture && io->new_hit_radiance_texture),
                "vgi_pass_ind_diff_temporal_resample: missing new sample");
    VGI_REQUIRE(!pc->has_prev_frame || (io->prev_gbuffer_depth && io->prev_reservoir_texture &&
                                        io->prev_hit_pos_normal_texture && io->prev_hit_radiance_texture),
                "vgi_pass_ind_diff_temporal_resample: missing previous reservoir");
    size_t R = radiance_bytes(io->storage);
    ProfScope prof("ind_diff_temporal_resample", st,
                   (double)io->width * io->height * (4 + 4 + 3 * (4 + 8 + R) - 4 + (taps ? 4 + sizeof(DiffTap) : 0)));
    const DiffTemporalArgs args{pc->frame_rand, pc->has_prev_frame, pc->has_new_sample, temporal_m_cap(params), taps,
                                gbuffer_color};
    VGI_LAUNCH_S(io->storage, k_diff_temporal, tile_grid(io->width, io->height), to_fp(fp), *io, args);
    return check_launch("k_diff_temporal");
}

extern "C" vgi_status vgi_pass_ind_diff_temporal_resample(const vgi_frame_params* fp, const vgi_ind_diff_temporal_io* io,
                                                          const vgi_pc_ind_diff_temporal* pc,
                                                          const vgi_resample_params* params, void* stream) {
    return launch_diff_temporal(fp, io, pc, params, nullptr, nullptr, (hipStream_t)stream);
}

vgi_status vgi::launch_diff_spatial(const vgi_frame_params* fp, const vgi_ind_diff_spatial_io* io,
                                    const vgi_pc_ind_diff_spatial* pc, const vgi_resample_params* params,
                                    const DiffTap* taps, hipStream_t st) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_ind_diff_spatial_resample: null argument");
    VGI_CHECK(check_resample_params(params, "vgi_pass_ind_diff_spatial_resample"));
    VGI_REQUIRE(valid_dims(io->width, io->height) && storage_ok(io->storage),
                "vgi_pass_ind_diff_spatial_resample: bad size/storage");
    VGI_REQUIRE(io->gbuffer_color && io->gbuffer_depth && io->ao_texture && io->temp_reservoir_texture &&
                    io->temp_hit_pos_normal_texture && io->temp_hit_radiance_texture && io->rw_reservoir_texture &&
                    io->rw_hit_pos_normal_texture && io->rw_hit_radiance_texture && io->rw_lighting_texture,
                "vgi_pass_ind_diff_spatial_resample: missing binding");
    const SpiralTable tab = make_spiral_table(spiral_taps(params));
    size_t R = radiance_bytes(io->storage);
    // algorithmic bytes of the pass's fields (a tap record carries the same ones)
    ProfScope prof("ind_diff_spatial_resample", st,
                   (double)io->width * io->height * (4 + 4 + pair_bytes(io->storage) / 2 + 2 * (4 + 8 + R) + R));
    if (taps) {
        VGI_LAUNCH_S(io->storage, k_diff_spatial_taps, tile_grid(io->width, io->height), to_fp(fp), *io, taps,
                     pc->frame_rand, spatial_radius(params, io->width, io->height), tab);
        return check_launch("k_diff_spatial_taps");
    }
    VGI_LAUNCH_S(io->storage, k_diff_spatial, tile_grid(io->width, io->height), to_fp(fp), *io, pc->frame_rand,
                 spatial_radius(params, io->width, io->height), tab);
    return check_launch("k_diff_spatial");
}

extern "C" vgi_status vgi_pass_ind_diff_spatial_resample(const vgi_frame_params* fp, const vgi_ind_diff_spatial_io* io,
                                                         const vgi_pc_ind_diff_spatial* pc,
                                                         const vgi_resample_params* params, void* stream) {
    return launch_diff_spatial(fp, io, pc, params, nullptr, (hipStream_t)stream);
}

extern "C" vgi_status vgi_pass_ind_spec_temporal_resample(const vgi_frame_params* fp, const vgi_ind_spec_temporal_io* io,
                                                          const vgi_pc_ind_spec_temporal* pc,
                                                          const vgi_resample_params* params, void* stream) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_ind_spec_temporal_resample: null argument");
    VGI_CHECK(check_resample_params(params, "vgi_pass_ind_spec_temporal_resample"));
    VGI_REQUIRE(valid_dims(io->width, io->height) && storage_ok(io->storage),
                "vgi_pass_ind_spec_temporal_resample: bad size/storage");
    VGI_REQUIRE(io->gbuffer_color && io->gbuffer_depth && io->rw_reservoir_texture && io->rw_hit_pos_texture &&
                    io->rw_hit_radiance_texture,
                "vgi_pass_ind_spec_temporal_resample: missing binding");
    VGI_REQUIRE(!pc->has_prev_frame || (io->prev_gbuffer_color && io->prev_gbuffer_depth && io->prev_reservoir_texture &&
                                        io->prev_hit_pos_texture && io->prev_hit_radiance_texture),
                "vgi_pass_ind_spec_temporal_resample: missing previous reservoir");
    hipStream_t st = (hipStream_t)stream;
    size_t R = radiance_bytes(io->storage);
    ProfScope prof("ind_spec_temporal_resample", st,
                   (double)io->width * io->height * (4 + 4 + 8 + 4 + 8 + 8 + 16 + R + 16 + R + 8 + 16 + R));
    const SpecTemporalArgs args{pc->frame_rand, pc->has_prev_frame, temporal_m_cap(params)};
    VGI_LAUNCH_S(io->storage, k_spec_temporal, tile_grid(io->width, io->height), to_fp(fp), *io, args);
    return check_launch("k_spec_temporal");
}

extern "C" vgi_status vgi_pass_ind_spec_spatial_resample(const vgi_frame_params* fp, const vgi_ind_spec_spatial_io* io,
                                                         const vgi_pc_ind_spec_spatial* pc,
                                                         const vgi_resample_params* params, void* stream) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_ind_spec_spatial_resample: null argument");
    VGI_CHECK(check_resample_params(params, "vgi_pass_ind_spec_spatial_resample"));
    VGI_REQUIRE(valid_dims(io->width, io->height) && storage_ok(io->storage),
                "vgi_pass_ind_spec_spatial_resample: bad size/storage");
    VGI_REQUIRE(io->gbuffer_color && io->gbuffer_depth && io->reservoir_texture && io->hit_pos_texture &&
                    io->hit_radiance_texture && io->rw_lighting_texture && io->rw_ray_len_texture,
                "vgi_pass_ind_spec_spatial_resample: missing binding");
    static const SpiralTable tab = make_spiral_table(9);
    hipStream_t st = (hipStream_t)stream;
    size_t R = radiance_bytes(io->storage);
    ProfScope prof("ind_spec_spatial_resample", st,
                   (double)io->width * io->height * (4 + 8 + 8 + 16 + R + R + scalar_bytes(io->storage)));
    VGI_LAUNCH_S(io->storage, k_spec_spatial, tile_grid(io->width, io->height), to_fp(fp), *io, pc->frame_rand,
                 spatial_radius(params, io->width, io->height), tab.cos_tol);
    return check_launch("k_spec_spatial");
}
