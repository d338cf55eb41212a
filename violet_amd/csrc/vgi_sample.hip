/*
 * vgi_sample.hip — the radiance-tracing passes of the ReSTIR GI path:
 *
 *   k_diff_sample_gen      restir/ind_diff_sample_gen.hlsl:25-67
 *   k_diff_sample_validate restir/ind_diff_sample_validate.hlsl:90-145
 *   k_spec_sample_gen      restir/ind_spec_sample_gen.hlsl:28-128
 *   k_spec_sample_validate restir/ind_spec_sample_validate.hlsl:144-191
 *   k_hash_grid_raygen     restir/hash_grid_raygen.hlsl:141-204
 *
 * One lane per pixel (or per compacted cache query, 16 per wave).  Each lane traces one
 * GI ray plus one sun shadow ray through the LDS-stack traversal.  Pixels
 * without geometry leave the wave early, so the remaining lanes are the
 * active rays.
 */
#include "vgi_batch_kernel.h"
#include "vgi_internal.h"
#include "vgi_radiance.h"

namespace vgi {
const SceneView& scene_view(const vgi_scene* s);

constexpr float kDepthError = 1.0f / 16777216.0f;  // 24-bit depth bias of the sample passes

__device__ __forceinline__ f3 position_biased(const float* m, uint32_t px, uint32_t py, uint32_t w, uint32_t h,
                                              float depth) {
    return depth_to_position(m, px, py, w, h, depth + kDepthError);
}

// Each sample pass is a pair of functions around trace_radiance: gen_px()
// turns a pixel into its GI ray (or writes the outputs of a pixel without
// one and returns false) and finish_px() writes the outputs of the traced
// sample; sample_pixel() runs the pair for one lane per pixel.
template <class S>
struct DiffGenPass {
    const FrameParams& fp;
    const vgi_ind_diff_sample_gen_io& io;
    const vgi_pc_sample_gen& pc;
    __device__ bool gen_px(uint32_t px, uint32_t py, f3& o, f3& d) const {
        const uint32_t w = io.width, h = io.height;
        if (px >= w || py >= h) return false;
        const size_t i = (size_t)py * w + px;
        float depth = G(io.gbuffer_depth)[i];
        if (depth == 0.0f) {
            G(io.rw_hit_pos_normal_texture)[2 * i] = 0u;
            G(io.rw_hit_pos_normal_texture)[2 * i + 1] = 0u;
            S::st3(io.rw_hit_radiance_texture, i, mk3s(0.0f));
            return false;
        }
        GBufferNormal g = decode_gbuffer_normal(G(io.gbuffer_color)[(size_t)w * h + i]);
        o = position_biased(fp.inv_view_proj, px, py, w, h, depth);
        uint32_t rng = lcg_init_with_seed(px, py, pc.frame_rand);
        float u0 = lcg_rand(rng);
        float u1 = lcg_rand(rng);
        d = sample_hemisphere_uniform_with_normal(u0, u1, g.normal);
        return true;
    }
    __device__ void finish_px(uint32_t px, uint32_t py, f3, f3, const RadianceTraceResult& r) const {
        const size_t i = (size_t)py * io.width + px;
        uint32_t ex, ey;
        hit_encode(r.position_ws, r.normal_ws, ex, ey);
        G(io.rw_hit_pos_normal_texture)[2 * i] = ex;
        G(io.rw_hit_pos_normal_texture)[2 * i + 1] = ey;
        S::st3(io.rw_hit_radiance_texture, i, r.radiance);
    }
};

// One lane per pixel of tile (bx_, by_): 8x8 pixels in 2x2 quads, or rows of
// rt.tile_w pixels (16 x 4, 32 x 2); the shape is a launch parameter
// (vgi_tuning) and does not change any pixel's result.
template <class P, class S, bool kCount>
__device__ __forceinline__ void sample_pixel(const P& p, const TraceRes& rt, const FrameParams& fp, bool has_prev,
                                             uint32_t bx_, uint32_t by_) {
    VGI_TRAVERSAL_STACK(stack);
    uint32_t px, py;
    if (rt.tile_w == 8u) {
        quad_tile_pixel_at<8>(bx_, by_, px, py);
    } else {
        const uint32_t lane = threadIdx.x & 63u, sh = rt.tile_w == 16u ? 4u : 5u;
        px = (bx_ << sh) + (lane & (rt.tile_w - 1u));
        py = by_ * (64u >> sh) + (lane >> sh);
    }
    f3 o, d;
    if (!p.gen_px(px, py, o, d)) return;
    RadianceTraceResult r = trace_radiance<S, kCount>(rt, fp, o, d, has_prev, stack);
    p.finish_px(px, py, o, d, r);
}

template <class S, bool kCount>
__device__ __forceinline__ void k_diff_sample_gen_body(const FrameParams& fp, const TraceRes& rt, const vgi_ind_diff_sample_gen_io& io, const vgi_pc_sample_gen& pc, uint32_t bx_, uint32_t by_) {
    const DiffGenPass<S> p{fp, io, pc};
    sample_pixel<DiffGenPass<S>, S, kCount>(p, rt, fp, pc.has_prev_frame != 0, bx_, by_);
}

template <class S, bool kCount>
__global__ void __launch_bounds__(64, kTraceWaves) k_diff_sample_gen(FrameParams fp, TraceRes rt, vgi_ind_diff_sample_gen_io io, vgi_pc_sample_gen pc) {
    k_diff_sample_gen_body<S, kCount>(fp, rt, io, pc, blockIdx.x, blockIdx.y);
}

template <class S>
struct DiffValidatePass {
    const FrameParams& fp;
    const vgi_ind_diff_sample_validate_io& io;
    __device__ bool gen_px(uint32_t px, uint32_t py, f3& o, f3& d) const {
        const uint32_t w = io.width, h = io.height;
        if (px >= w || py >= h) return false;
        const size_t i = (size_t)py * w + px;
        float depth = G(io.depth_texture)[i];
        if (depth == 0.0f) return false;
        HitPosNormal hit =
            hit_decode(G(io.rw_hit_pos_normal_texture)[2 * i], G(io.rw_hit_pos_normal_texture)[2 * i + 1]);
        o = depth_to_position(fp.prev_inv_view_proj, px, py, w, h, depth);
        d = normalize3(hit.pos - o);
        return true;
    }
    __device__ void finish_px(uint32_t px, uint32_t py, f3 origin, f3, const RadianceTraceResult& r) const {
        const size_t i = (size_t)py * io.width + px;
        HitPosNormal hit =
            hit_decode(G(io.rw_hit_pos_normal_texture)[2 * i], G(io.rw_hit_pos_normal_texture)[2 * i + 1]);
        f3 prev_rad = S::ld3(io.rw_hit_radiance_texture, i);
        float lumi_prev = luminance(prev_rad);
        float lumi_diff = luminance(abs3(r.radiance - prev_rad));
        float prev_ray_len = length3(hit.pos - origin);
        bool pos_changed = component_max(abs3(hit.pos - r.position_ws)) > fmaxf(prev_ray_len / 128.0f, 0.01f);
        float m_scale = saturate(1.0f - (1.0f * lumi_diff) / lumi_prev);
        m_scale = pos_changed ? 0.0f : m_scale;
        if ((lumi_prev > 0.0f) && (m_scale <= (1.0f - 0.5f / 32.0f))) {
            ReservoirSimple rs = reservoir_decode_32b(G(io.rw_reservoir_texture)[i]);
            rs.M = f2u(floorf((float)rs.M * m_scale));
            G(io.rw_reservoir_texture)[i] = reservoir_encode_32b(rs);
        }
        uint32_t ex, ey;
        hit_encode(r.position_ws, r.normal_ws, ex, ey);
        G(io.rw_hit_pos_normal_texture)[2 * i] = ex;
        G(io.rw_hit_pos_normal_texture)[2 * i + 1] = ey;
        S::st3(io.rw_hit_radiance_texture, i, r.radiance);
    }
};

template <class S, bool kCount>
__device__ __forceinline__ void k_diff_sample_validate_body(const FrameParams& fp, const TraceRes& rt, const vgi_ind_diff_sample_validate_io& io, uint32_t bx_, uint32_t by_) {
    const DiffValidatePass<S> p{fp, io};
    sample_pixel<DiffValidatePass<S>, S, kCount>(p, rt, fp, true, bx_, by_);
}

template <class S, bool kCount>
__global__ void __launch_bounds__(64, kTraceWaves) k_diff_sample_validate(FrameParams fp, TraceRes rt, vgi_ind_diff_sample_validate_io io) {
    k_diff_sample_validate_body<S, kCount>(fp, rt, io, blockIdx.x, blockIdx.y);
}

// GGX VNDF sample [Heitz 2014] of ind_spec_sample_gen.hlsl:67-120; returns the
// reflected world-space direction.
__device__ __forceinline__ f3 sample_spec_dir(f3 view_dir, f3 normal, float roughness, uint32_t& rng) {
    f4 rot = get_rotation_to_z_from(normal);
    f3 V = rotate_point(rot, view_dir);
    f3 H;
    if (roughness != 0.0f) {
        float ux = lcg_rand(rng);
        float uy = lcg_rand(rng);
        f3 v_h = normalize3(mk3(roughness * V.x, roughness * V.y, V.z));
        float lensq = v_h.x * v_h.x + v_h.y * v_h.y;
        f3 T1 = lensq > 0.0f ? mk3(-v_h.y, v_h.x, 0.0f) * rsqrt(lensq) : mk3(1.0f, 0.0f, 0.0f);
        f3 T2 = cross3(v_h, T1);
        float r = sqrtf(ux);
        float phi = kTwoPi * uy;
        float sp, cp;
        vgi_sincosf(phi, &sp, &cp);
        float t1 = r * cp;
        float t2 = r * sp;
        float s = 0.5f * (1.0f + v_h.z);
        t2 = lerp(sqrtf(1.0f - t1 * t1), t2, s);
        f3 n_h = (T1 * t1 + T2 * t2) + v_h * sqrtf(fmaxf(0.0f, (1.0f - t1 * t1) - t2 * t2));
        H = normalize3(mk3(roughness * n_h.x, roughness * n_h.y, fmaxf(0.0f, n_h.z)));
    } else {
        H = mk3(0.0f, 0.0f, 1.0f);
    }
    f3 I = -V;
    f3 L = I - H * (2.0f * dot3(H, I));  // reflect(-V, H)
    return rotate_point(invert_rotation(rot), L);
}

template <class S>
struct SpecGenPass {
    const FrameParams& fp;
    const vgi_ind_spec_sample_gen_io& io;
    const vgi_pc_sample_gen& pc;
    __device__ bool gen_px(uint32_t px, uint32_t py, f3& o, f3& d) const {
        const uint32_t w = io.width, h = io.height;
        if (px >= w || py >= h) return false;
        const size_t i = (size_t)py * w + px;
        const size_t plane = (size_t)w * h;
        float depth = G(io.gbuffer_depth)[i];
        if (depth == 0.0f) {
            // the reference leaves the hit position of a sky pixel unwritten;
            // the specular temporal pass reads it back (times M = 0) when a
            // reprojection lands on such a pixel, so it must not be stale
            G(io.rw_hit_pos_texture)[4 * i] = 0.0f;
            G(io.rw_hit_pos_texture)[4 * i + 1] = 0.0f;
            G(io.rw_hit_pos_texture)[4 * i + 2] = 0.0f;
            G(io.rw_hit_pos_texture)[4 * i + 3] = 0.0f;
            S::st3(io.rw_hit_radiance_texture, i, mk3s(0.0f));
            return false;
        }
        GBufferNormal g = decode_gbuffer_normal(G(io.gbuffer_color)[plane + i]);
        o = position_biased(fp.inv_view_proj, px, py, w, h, depth);
        uint32_t rng = lcg_init_with_seed(px, py, pc.frame_rand);
        f3 view_dir = normalize3(xyz(fp.view_pos) - o);
        float roughness = fmaxf(g.perceptual_roughness * g.perceptual_roughness, 0.002f);
        d = sample_spec_dir(view_dir, g.normal, roughness, rng);
        return true;
    }
    __device__ void finish_px(uint32_t px, uint32_t py, f3, f3, const RadianceTraceResult& r) const {
        const size_t i = (size_t)py * io.width + px;
        float* hp = io.rw_hit_pos_texture + 4 * i;
        hp[0] = r.position_ws.x;
        hp[1] = r.position_ws.y;
        hp[2] = r.position_ws.z;
        hp[3] = 1.0f;
        S::st3(io.rw_hit_radiance_texture, i, r.radiance);
    }
};

template <class S, bool kCount>
__device__ __forceinline__ void k_spec_sample_gen_body(const FrameParams& fp, const TraceRes& rt, const vgi_ind_spec_sample_gen_io& io, const vgi_pc_sample_gen& pc, uint32_t bx_, uint32_t by_) {
    const SpecGenPass<S> p{fp, io, pc};
    sample_pixel<SpecGenPass<S>, S, kCount>(p, rt, fp, pc.has_prev_frame != 0, bx_, by_);
}

template <class S, bool kCount>
__global__ void __launch_bounds__(64, kTraceWaves) k_spec_sample_gen(FrameParams fp, TraceRes rt, vgi_ind_spec_sample_gen_io io, vgi_pc_sample_gen pc) {
    k_spec_sample_gen_body<S, kCount>(fp, rt, io, pc, blockIdx.x, blockIdx.y);
}

template <class S>
struct SpecValidatePass {
    const FrameParams& fp;
    const TraceRes& rt;
    const vgi_ind_spec_sample_validate_io& io;
    __device__ bool gen_px(uint32_t px, uint32_t py, f3& o, f3& d) const {
        const uint32_t w = io.width, h = io.height;
        if (px >= w || py >= h) return false;
        const size_t i = (size_t)py * w + px;
        float prev_depth = G(rt.prev_depth)[i];
        if (prev_depth == 0.0f) return false;
        o = position_biased(fp.prev_inv_view_proj, px, py, w, h, prev_depth);
        const float* hp = io.rw_prev_hit_pos_texture + 4 * i;
        f3 prev_hit_pos = mk3(hp[0], hp[1], hp[2]);
        d = normalize3(prev_hit_pos - o);
        return true;
    }
    __device__ void finish_px(uint32_t px, uint32_t py, f3, f3, const RadianceTraceResult& r) const {
        const size_t i = (size_t)py * io.width + px;
        uint32_t Wbits = G(io.rw_prev_reservoir_texture)[2 * i + 1];
        float prev_target = io.rw_prev_hit_pos_texture[4 * i + 3];
        f3 prev_rad = S::ld3(io.rw_prev_hit_radiance_texture, i);
        float max_rad = fmaxf(component_max(r.radiance), component_max(prev_rad));
        f3 dr = abs3(r.radiance - prev_rad);
        float lim = 0.5f * max_rad;
        bool changed = dr.x > lim || dr.y > lim || dr.z > lim;
        if (changed) {
            G(io.rw_prev_reservoir_texture)[2 * i] = 1u;
            G(io.rw_prev_reservoir_texture)[2 * i + 1] = Wbits;
            float* out = io.rw_prev_hit_pos_texture + 4 * i;
            out[0] = r.position_ws.x;
            out[1] = r.position_ws.y;
            out[2] = r.position_ws.z;
            out[3] = prev_target;
            S::st3(io.rw_prev_hit_radiance_texture, i, r.radiance);
        }
    }
};

template <class S, bool kCount>
__device__ __forceinline__ void k_spec_sample_validate_body(const FrameParams& fp, const TraceRes& rt, const vgi_ind_spec_sample_validate_io& io, uint32_t bx_, uint32_t by_) {
    const SpecValidatePass<S> p{fp, rt, io};
    sample_pixel<SpecValidatePass<S>, S, kCount>(p, rt, fp, true, bx_, by_);
}

template <class S, bool kCount>
__global__ void __launch_bounds__(64, kTraceWaves) k_spec_sample_validate(FrameParams fp, TraceRes rt, vgi_ind_spec_sample_validate_io io) {
    k_spec_sample_validate_body<S, kCount>(fp, rt, io, blockIdx.x, blockIdx.y);
}

// One cache ray: hash_grid_raygen.hlsl:141-204 for compacted query `id`.
// With `staging` the blended radiance goes to staging[id] (4 floats) and
// k_hgo_apply writes it into the cell after every ray has traced, so each
// ray's cache lookups read the cache as it was before the pass.
template <class S, bool kCount>
__device__ __forceinline__ void hash_grid_ray(const FrameParams& fp, const TraceRes& rt, const vgi_hash_grid_raygen_io& io,
                                              float* staging, uint32_t frame_hash, uint32_t id, LdsStack stack) {
    uint32_t qi = G(io.compact_query_index_buffer)[id];
    uint32_t cell = G(io.compact_query_cell_addr_buffer)[id];
    uint32_t rng = jenkins_hash(id ^ frame_hash);
    const HashQuery q = reinterpret_cast<const HashQuery*>(io.hash_grid_query_buffer)[qi];
    f3 normal = normal_decode_oct_u32(q.normal);
    float u0 = lcg_rand(rng);
    float u1 = lcg_rand(rng);
    f3 local = sample_hemisphere_cosine(u0, u1);
    f3 dir = rotate_point(invert_rotation(get_rotation_to_z_from(normal)), local);
    RadianceTraceResult r = trace_radiance<S, kCount>(rt, fp, mk3(q.px, q.py, q.pz), dir, true, stack);
    HashCell* storage = reinterpret_cast<HashCell*>(io.rw_hash_grid_storage_buffer);
    f3 prev = mk3(storage[cell].r, storage[cell].g, storage[cell].b);
    const float bf = 1.0f / 16.0f;
    f3 blend = r.radiance * bf + prev * (1.0f - bf);
    if (staging) {
        typedef float F4 __attribute__((ext_vector_type(4)));
        F4 v = {blend.x, blend.y, blend.z, 0.0f};
        G(reinterpret_cast<F4*>(staging))[id] = v;
    } else {
        storage[cell].r = blend.x;
        storage[cell].g = blend.y;
        storage[cell].b = blend.z;
    }
    G(io.rw_hash_grid_decay_buffer)[cell] = 10;
}

// Writes the staged blends of an ordered cache-ray pass into their cells.
__device__ __forceinline__ void k_hgo_apply_body(const vgi_hash_grid_raygen_io& io, const float* staging,
                                                 uint32_t bx_) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    const uint32_t id = bx_ * 64u + threadIdx.x;
    if (id >= kHashGridMaxQueries || id >= G(io.compact_query_counter_buffer)[0]) return;
    const uint32_t cell = G(io.compact_query_cell_addr_buffer)[id];
    const F4 v = G(reinterpret_cast<const F4*>(staging))[id];
    HashCell* c = reinterpret_cast<HashCell*>(io.rw_hash_grid_storage_buffer) + cell;
    G(&c->r)[0] = v.x;
    G(&c->g)[0] = v.y;
    G(&c->b)[0] = v.z;
}

__global__ void __launch_bounds__(64) k_hgo_apply(vgi_hash_grid_raygen_io io, const float* staging) {
    k_hgo_apply_body(io, staging, blockIdx.x);
}

// Cache rays start at scattered surface points in random directions, so a
// wave's lanes share no traversal path and a wave lasts as long as the union
// of its rays' paths.  Few rays per wave (in more waves) shortens that union;
// there are only ~1e4 such rays per frame, against ~1e6 sample rays.  The
// groups stride over the compacted queries (at most kHashGridMaxQueries).
constexpr uint32_t kRaygenRaysPerWave = 16;
constexpr uint32_t kRaygenGroups = kHashGridMaxQueries / 64;  // groups per launch (per view in a batch)

template <class S, bool kCount>
__device__ __forceinline__ void k_hash_grid_raygen_body(const FrameParams& fp, const TraceRes& rt, const vgi_hash_grid_raygen_io& io, float* staging, uint32_t frame_hash, uint32_t bx_, uint32_t by_) {
    VGI_TRAVERSAL_STACK(stack);
    const uint32_t lane = threadIdx.x & 63u;
    if (lane >= kRaygenRaysPerWave) return;
    const uint32_t count = minu(G(io.compact_query_counter_buffer)[0], kHashGridMaxQueries);
    for (uint32_t id = bx_ * kRaygenRaysPerWave + lane; id < count; id += kRaygenGroups * kRaygenRaysPerWave)
        hash_grid_ray<S, kCount>(fp, rt, io, staging, frame_hash, id, stack);
}

template <class S, bool kCount>
__global__ void __launch_bounds__(64) k_hash_grid_raygen(FrameParams fp, TraceRes rt, vgi_hash_grid_raygen_io io, float* staging, uint32_t frame_hash) {
    k_hash_grid_raygen_body<S, kCount>(fp, rt, io, staging, frame_hash, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
static vgi_status make_trace_res(const vgi_trace_resources* r, bool need_prev, TraceRes& t) {
    VGI_REQUIRE(r && r->scene_tlas, "trace resources: no scene");
    VGI_REQUIRE(r->hash_grid_storage_buffer && r->rw_hash_grid_query_buffer && r->rw_hash_grid_query_counter_buffer,
                "trace resources: hash-grid buffers missing");
    VGI_REQUIRE(r->storage == VGI_STORAGE_F32 || r->storage == VGI_STORAGE_F16, "trace resources: bad storage");
    if (need_prev)
        VGI_REQUIRE(r->prev_indirect_diffuse_texture && r->prev_depth_texture,
                    "trace resources: has_prev_frame=1 but prev_* is NULL");
    t.scene = scene_view(r->scene_tlas);
    t.scene.counters = trace_counters();
    t.sky = mk3(r->sky_radiance[0], r->sky_radiance[1], r->sky_radiance[2]);
    t.prev_indirect_diffuse = r->prev_indirect_diffuse_texture;
    t.prev_depth = r->prev_depth_texture;
    t.pw = r->prev_width;
    t.ph = r->prev_height;
    if (need_prev) VGI_REQUIRE(valid_dims(t.pw, t.ph), "trace resources: bad prev_width/prev_height");
    t.hg_storage = reinterpret_cast<const HashCell*>(r->hash_grid_storage_buffer);
    t.queries = reinterpret_cast<HashQuery*>(r->rw_hash_grid_query_buffer);
    t.query_counter = r->rw_hash_grid_query_counter_buffer;
    t.query_threshold = r->hash_grid_query_threshold;
    t.tile_w = tuning_trace_tile_width();
    return VGI_OK;
}

static FrameParams to_fp(const vgi_frame_params* fp) {
    FrameParams f;
    memcpy(&f, fp, sizeof(f));
    return f;
}

// Every kernel here traces rays: launch with the scene's LDS traversal stack,
// and the instrumented build when traversal statistics are on.
#define VGI_DISPATCH_STORAGE(mode, kernel, grid, ...)                                                      \
    do {                                                                                                   \
        const size_t lds_ = traversal_lds_bytes(t.scene);                                                  \
        const bool f32_ = (mode) == VGI_STORAGE_F32;                                                       \
        if (t.scene.counters) {                                                                            \
            if (f32_)                                                                                      \
                hipLaunchKernelGGL((kernel<StorageF32, true>), grid, dim3(64), lds_, st, __VA_ARGS__);     \
            else                                                                                           \
                hipLaunchKernelGGL((kernel<StorageF16, true>), grid, dim3(64), lds_, st, __VA_ARGS__);     \
        } else {                                                                                           \
            if (f32_)                                                                                      \
                hipLaunchKernelGGL((kernel<StorageF32, false>), grid, dim3(64), lds_, st, __VA_ARGS__);    \
            else                                                                                           \
                hipLaunchKernelGGL((kernel<StorageF16, false>), grid, dim3(64), lds_, st, __VA_ARGS__);    \
        }                                                                                                  \
    } while (0)

// ---------------------------------------------------------------------------
// Fused multi-view launches (vgi_batch.h); the scene (and so the LDS stack
// size) is shared by every view of a batch.
// ---------------------------------------------------------------------------
template <class S>
struct JobDiffGen {
    FrameParams fp;
    TraceRes rt;
    vgi_ind_diff_sample_gen_io io;
    vgi_pc_sample_gen pc;
    __device__ static void run(const JobDiffGen& j, uint32_t bx, uint32_t by) {
        k_diff_sample_gen_body<S, false>(j.fp, j.rt, j.io, j.pc, bx, by);
    }
};
template <class S>
struct JobDiffValidate {
    FrameParams fp;
    TraceRes rt;
    vgi_ind_diff_sample_validate_io io;
    __device__ static void run(const JobDiffValidate& j, uint32_t bx, uint32_t by) {
        k_diff_sample_validate_body<S, false>(j.fp, j.rt, j.io, bx, by);
    }
};
template <class S>
struct JobSpecGen {
    FrameParams fp;
    TraceRes rt;
    vgi_ind_spec_sample_gen_io io;
    vgi_pc_sample_gen pc;
    __device__ static void run(const JobSpecGen& j, uint32_t bx, uint32_t by) {
        k_spec_sample_gen_body<S, false>(j.fp, j.rt, j.io, j.pc, bx, by);
    }
};
template <class S>
struct JobSpecValidate {
    FrameParams fp;
    TraceRes rt;
    vgi_ind_spec_sample_validate_io io;
    __device__ static void run(const JobSpecValidate& j, uint32_t bx, uint32_t by) {
        k_spec_sample_validate_body<S, false>(j.fp, j.rt, j.io, bx, by);
    }
};
struct HgRaygenArgs {
    FrameParams fp;
    TraceRes rt;
    vgi_hash_grid_raygen_io io;
    float* staging;  // null: blends go straight into the cells
    uint32_t frame_hash;
};
template <class S>
struct JobHgRaygen : HgRaygenArgs {
    __device__ static void run(const JobHgRaygen& j, uint32_t bx, uint32_t by) {
        k_hash_grid_raygen_body<S, false>(j.fp, j.rt, j.io, j.staging, j.frame_hash, bx, by);
    }
};
template <class S>
struct JobHgOApply : HgRaygenArgs {
    __device__ static void run(const JobHgOApply& j, uint32_t bx, uint32_t) { k_hgo_apply_body(j.io, j.staging, bx); }
};
static_assert(sizeof(JobHgOApply<StorageF32>) == sizeof(JobHgRaygen<StorageF32>), "one table, two launches");
static_assert(kRaygenGroups * 64 >= kHashGridMaxQueries, "the cache-ray tiles cover the compact list");
template <class S>
struct BatchMinWaves<JobDiffGen<S>> {
    static constexpr int value = kTraceWaves;
};
template <class S>
struct BatchMinWaves<JobDiffValidate<S>> {
    static constexpr int value = kTraceWaves;
};
template <class S>
struct BatchMinWaves<JobSpecGen<S>> {
    static constexpr int value = kTraceWaves;
};
template <class S>
struct BatchMinWaves<JobSpecValidate<S>> {
    static constexpr int value = kTraceWaves;
};

static vgi_status fill_diff_gen(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobDiffGen>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        vgi_status s = make_trace_res(&r.tr, r.pc.sample_gen.has_prev_frame != 0, j.rt);
        if (s != VGI_OK) return s;
        j.fp = to_fp(&r.fp);
        j.io = r.io.diff_gen;
        j.pc = r.pc.sample_gen;
        {
            const dim3 tg = trace_tile_grid(j.io.width, j.io.height, j.rt.tile_w);
            batch_add_tiles(t->hd, v, tg.x, tg.y);
        }
    }
    return VGI_OK;
}
static vgi_status fill_diff_validate(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobDiffValidate>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        vgi_status s = make_trace_res(&r.tr, true, j.rt);
        if (s != VGI_OK) return s;
        j.fp = to_fp(&r.fp);
        j.io = r.io.diff_validate;
        {
            const dim3 tg = trace_tile_grid(j.io.width, j.io.height, j.rt.tile_w);
            batch_add_tiles(t->hd, v, tg.x, tg.y);
        }
    }
    return VGI_OK;
}
static vgi_status fill_spec_gen(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobSpecGen>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        vgi_status s = make_trace_res(&r.tr, r.pc.sample_gen.has_prev_frame != 0, j.rt);
        if (s != VGI_OK) return s;
        j.fp = to_fp(&r.fp);
        j.io = r.io.spec_gen;
        j.pc = r.pc.sample_gen;
        {
            const dim3 tg = trace_tile_grid(j.io.width, j.io.height, j.rt.tile_w);
            batch_add_tiles(t->hd, v, tg.x, tg.y);
        }
    }
    return VGI_OK;
}
static vgi_status fill_spec_validate(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobSpecValidate>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        vgi_status s = make_trace_res(&r.tr, true, j.rt);
        if (s != VGI_OK) return s;
        j.fp = to_fp(&r.fp);
        j.io = r.io.spec_validate;
        {
            const dim3 tg = trace_tile_grid(j.io.width, j.io.height, j.rt.tile_w);
            batch_add_tiles(t->hd, v, tg.x, tg.y);
        }
    }
    return VGI_OK;
}
static vgi_status fill_hg_raygen(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobHgRaygen>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        vgi_status s = make_trace_res(&r.tr, true, j.rt);
        if (s != VGI_OK) return s;
        j.fp = to_fp(&r.fp);
        j.io = r.io.hg_raygen;
        j.staging = nullptr;
        j.frame_hash = r.pc.frame_hash.frame_hash;
        batch_add_tiles(t->hd, v, kRaygenGroups, 1);
    }
    return VGI_OK;
}
static vgi_status fill_hg_oraygen(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobHgRaygen>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        auto& j = t->jobs[v];
        vgi_status s = make_trace_res(&r.tr, true, j.rt);
        if (s != VGI_OK) return s;
        VGI_REQUIRE(r.io.hg_oraygen.rw_radiance_staging_buffer &&
                        ((uintptr_t)r.io.hg_oraygen.rw_radiance_staging_buffer & 15u) == 0,
                    "batched ordered cache rays: staging buffer missing or not 16-byte aligned");
        j.fp = to_fp(&r.fp);
        j.io = r.io.hg_oraygen.raygen;
        j.staging = r.io.hg_oraygen.rw_radiance_staging_buffer;
        j.frame_hash = r.pc.frame_hash.frame_hash;
        batch_add_tiles(t->hd, v, kRaygenGroups, 1);
    }
    return VGI_OK;
}
static vgi_status launch_hg_oraygen(const void* dev, uint32_t storage, uint32_t grid, size_t lds, hipStream_t st) {
    VGI_CHECK(batch_launch<JobHgRaygen>(dev, storage, grid, lds, st));
    return batch_launch<JobHgOApply>(dev, storage, grid, 0, st);
}

const BatchOps* batch_ops_sample(int kind) {
    static const BatchOps diff_gen{sizeof(BatchTable<JobDiffGen<StorageF32>>), fill_diff_gen, batch_launch<JobDiffGen>,
                                   64, true};
    static const BatchOps diff_validate{sizeof(BatchTable<JobDiffValidate<StorageF32>>), fill_diff_validate,
                                        batch_launch<JobDiffValidate>, 64, true};
    static const BatchOps spec_gen{sizeof(BatchTable<JobSpecGen<StorageF32>>), fill_spec_gen, batch_launch<JobSpecGen>,
                                   64, true};
    static const BatchOps spec_validate{sizeof(BatchTable<JobSpecValidate<StorageF32>>), fill_spec_validate,
                                        batch_launch<JobSpecValidate>, 64, true};
    static const BatchOps hg_raygen{sizeof(BatchTable<JobHgRaygen<StorageF32>>), fill_hg_raygen,
                                    batch_launch<JobHgRaygen>, 64, true};
    static const BatchOps hg_oraygen{sizeof(BatchTable<JobHgRaygen<StorageF32>>), fill_hg_oraygen, launch_hg_oraygen,
                                     64, true};
    switch (kind) {
        case kPassDiffGen: return &diff_gen;
        case kPassDiffValidate: return &diff_validate;
        case kPassSpecGen: return &spec_gen;
        case kPassSpecValidate: return &spec_validate;
        case kPassHgRaygen: return &hg_raygen;
        case kPassHgRaygenOrdered: return &hg_oraygen;
        default: return nullptr;
    }
}

}  // namespace vgi

using namespace vgi;

extern "C" vgi_status vgi_pass_ind_diff_sample_gen(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                                   const vgi_ind_diff_sample_gen_io* io, const vgi_pc_sample_gen* pc,
                                                   void* stream) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_ind_diff_sample_gen: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_ind_diff_sample_gen: bad size");
    VGI_REQUIRE(io->storage == rt->storage, "vgi_pass_ind_diff_sample_gen: storage mismatch");
    VGI_REQUIRE(io->gbuffer_color && io->gbuffer_depth && io->rw_hit_pos_normal_texture && io->rw_hit_radiance_texture,
                "vgi_pass_ind_diff_sample_gen: missing binding");
    TraceRes t;
    vgi_status s = make_trace_res(rt, pc->has_prev_frame != 0, t);
    if (s != VGI_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("ind_diff_sample_gen", st, (double)io->width * io->height * (8 + 4 + 8 + radiance_bytes(io->storage)));
    VGI_DISPATCH_STORAGE(io->storage, k_diff_sample_gen, trace_tile_grid(io->width, io->height, t.tile_w), to_fp(fp), t, *io, *pc);
    trace_stats_collect("ind_diff_sample_gen", st);
    return check_launch("k_diff_sample_gen");
}

extern "C" vgi_status vgi_pass_ind_diff_sample_validate(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                                        const vgi_ind_diff_sample_validate_io* io, void* stream) {
    VGI_REQUIRE(fp && io, "vgi_pass_ind_diff_sample_validate: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_ind_diff_sample_validate: bad size");
    VGI_REQUIRE(io->storage == rt->storage, "vgi_pass_ind_diff_sample_validate: storage mismatch");
    VGI_REQUIRE(io->depth_texture && io->rw_reservoir_texture && io->rw_hit_pos_normal_texture &&
                    io->rw_hit_radiance_texture,
                "vgi_pass_ind_diff_sample_validate: missing binding");
    TraceRes t;
    vgi_status s = make_trace_res(rt, true, t);
    if (s != VGI_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("ind_diff_sample_validate", st,
                   (double)io->width * io->height * (4 + 2 * (4 + 8 + radiance_bytes(io->storage))));
    VGI_DISPATCH_STORAGE(io->storage, k_diff_sample_validate, trace_tile_grid(io->width, io->height, t.tile_w), to_fp(fp), t, *io);
    trace_stats_collect("ind_diff_sample_validate", st);
    return check_launch("k_diff_sample_validate");
}

extern "C" vgi_status vgi_pass_ind_spec_sample_gen(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                                   const vgi_ind_spec_sample_gen_io* io, const vgi_pc_sample_gen* pc,
                                                   void* stream) {
    VGI_REQUIRE(fp && io && pc, "vgi_pass_ind_spec_sample_gen: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_ind_spec_sample_gen: bad size");
    VGI_REQUIRE(io->storage == rt->storage, "vgi_pass_ind_spec_sample_gen: storage mismatch");
    VGI_REQUIRE(io->gbuffer_color && io->gbuffer_depth && io->rw_hit_pos_texture && io->rw_hit_radiance_texture,
                "vgi_pass_ind_spec_sample_gen: missing binding");
    TraceRes t;
    vgi_status s = make_trace_res(rt, pc->has_prev_frame != 0, t);
    if (s != VGI_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("ind_spec_sample_gen", st, (double)io->width * io->height * (4 + 4 + 16 + radiance_bytes(io->storage)));
    VGI_DISPATCH_STORAGE(io->storage, k_spec_sample_gen, trace_tile_grid(io->width, io->height, t.tile_w), to_fp(fp), t, *io, *pc);
    trace_stats_collect("ind_spec_sample_gen", st);
    return check_launch("k_spec_sample_gen");
}

extern "C" vgi_status vgi_pass_ind_spec_sample_validate(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                                        const vgi_ind_spec_sample_validate_io* io, void* stream) {
    VGI_REQUIRE(fp && io, "vgi_pass_ind_spec_sample_validate: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_ind_spec_sample_validate: bad size");
    VGI_REQUIRE(io->storage == rt->storage, "vgi_pass_ind_spec_sample_validate: storage mismatch");
    VGI_REQUIRE(io->rw_prev_reservoir_texture && io->rw_prev_hit_pos_texture && io->rw_prev_hit_radiance_texture,
                "vgi_pass_ind_spec_sample_validate: missing binding");
    TraceRes t;
    vgi_status s = make_trace_res(rt, true, t);
    if (s != VGI_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("ind_spec_sample_validate", st,
                   (double)io->width * io->height * (4 + 2 * (8 + 16 + radiance_bytes(io->storage))));
    VGI_DISPATCH_STORAGE(io->storage, k_spec_sample_validate, trace_tile_grid(io->width, io->height, t.tile_w), to_fp(fp), t, *io);
    trace_stats_collect("ind_spec_sample_validate", st);
    return check_launch("k_spec_sample_validate");
}

extern "C" vgi_status vgi_pass_hash_grid_raygen(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                                const vgi_hash_grid_raygen_io* io, const vgi_pc_frame_hash* pc,
                                                void* stream) {
    VGI_REQUIRE(fp && io && pc && rt, "vgi_pass_hash_grid_raygen: null argument");
    VGI_REQUIRE(io->rw_hash_grid_storage_buffer && io->hash_grid_query_buffer && io->compact_query_counter_buffer &&
                    io->compact_query_index_buffer && io->compact_query_cell_addr_buffer && io->rw_hash_grid_decay_buffer,
                "vgi_pass_hash_grid_raygen: missing binding");
    TraceRes t;
    vgi_status s = make_trace_res(rt, true, t);
    if (s != VGI_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    // ray per compacted query: query 24 B + index 4 + addr 4 + cell rmw 32 + decay 1
    ProfScope prof("hash_grid_raygen", st, 0.0);
    VGI_DISPATCH_STORAGE(rt->storage, k_hash_grid_raygen, dim3(kRaygenGroups), to_fp(fp), t, *io, (float*)nullptr,
                         pc->frame_hash);
    trace_stats_collect("hash_grid_raygen", st);
    return check_launch("k_hash_grid_raygen");
}

extern "C" vgi_status vgi_pass_hash_grid_raygen_ordered(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                                        const vgi_hash_grid_ordered_raygen_io* oio,
                                                        const vgi_pc_frame_hash* pc, void* stream) {
    VGI_REQUIRE(fp && oio && pc && rt, "vgi_pass_hash_grid_raygen_ordered: null argument");
    const vgi_hash_grid_raygen_io* io = &oio->raygen;
    VGI_REQUIRE(io->rw_hash_grid_storage_buffer && io->hash_grid_query_buffer && io->compact_query_counter_buffer &&
                    io->compact_query_index_buffer && io->compact_query_cell_addr_buffer &&
                    io->rw_hash_grid_decay_buffer && oio->rw_radiance_staging_buffer,
                "vgi_pass_hash_grid_raygen_ordered: missing binding");
    VGI_REQUIRE(((uintptr_t)oio->rw_radiance_staging_buffer & 15u) == 0,
                "vgi_pass_hash_grid_raygen_ordered: staging buffer not 16-byte aligned");
    TraceRes t;
    vgi_status s = make_trace_res(rt, true, t);
    if (s != VGI_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("hash_grid_raygen_ordered", st, 0.0);
    VGI_DISPATCH_STORAGE(rt->storage, k_hash_grid_raygen, dim3(kRaygenGroups), to_fp(fp), t, *io,
                         oio->rw_radiance_staging_buffer, pc->frame_hash);
    trace_stats_collect("hash_grid_raygen_ordered", st);
    VGI_CHECK(check_launch("k_hash_grid_raygen"));
    hipLaunchKernelGGL(k_hgo_apply, dim3(kHashGridMaxQueries / 64), dim3(64), 0, st, *io,
                       (const float*)oio->rw_radiance_staging_buffer);
    return check_launch("k_hgo_apply");
}
