/*
 * vgi_radiance.h — device-side `trace_radiance` and `trace_shadow`
 * (restir/raytrace.inc.hlsl:30-208) and the closest-hit shading of
 * raytrace/geometry_rchit.inc.hlsl:376-472, plus the hash-grid probe
 * (restir/hash_grid.inc.hlsl:124-147).
 */
#pragma once
#include "vgi_storage.h"
#include "vgi_traverse.h"

namespace vgi {

struct HashCell {
    uint32_t checksum;
    float r, g, b;
};
struct HashQuery {
    float px, py, pz;
    uint32_t normal;
    uint32_t hash;
    uint32_t checksum;
};
static_assert(sizeof(HashCell) == 16, "HashCell");
static_assert(sizeof(HashQuery) == 24, "HashQuery");

// Device copy of vgi_trace_resources.
struct TraceRes {
    SceneView scene;
    f3 sky;
    const void* prev_indirect_diffuse;
    const float* prev_depth;
    uint32_t pw, ph;  // size of the previous-frame images
    const HashCell* hg_storage;
    HashQuery* queries;
    uint32_t* query_counter;
    const uint32_t* query_threshold;  // null: append every query
    uint32_t tile_w;                  // sample passes: a wave's pixel tile is tile_w x (64 / tile_w)
};

// Hash of a query's 24 bytes (a jenkins chain over its words): the thinning
// of the appends and the ordered select's random key are functions of it,
// so equal queries are treated alike wherever they were appended.
__device__ __forceinline__ uint32_t query_content_hash(const HashQuery& q) {
    uint32_t h = jenkins_hash(__float_as_uint(q.px));
    h = jenkins_hash(h ^ __float_as_uint(q.py));
    h = jenkins_hash(h ^ __float_as_uint(q.pz));
    h = jenkins_hash(h ^ q.normal);
    h = jenkins_hash(h ^ q.hash);
    return jenkins_hash(h ^ q.checksum);
}

struct RadianceTraceResult {
    f3 position_ws;
    f3 normal_ws;
    f3 radiance;
};

// hash_grid.inc.hlsl:124-147
__device__ __forceinline__ bool hash_grid_find(const HashCell* buf, uint32_t hash, uint32_t checksum, f3& radiance,
                                               uint32_t& addr) {
    uint32_t beg = hash << kHashGridBucketShift;
    for (addr = beg; addr < beg + kHashGridBucketSize; ++addr) {
        HashCell c = gld(buf + addr);
        if (c.checksum == checksum) {
            radiance = mk3(c.r, c.g, c.b);
            return true;
        }
        if (c.checksum == 0u) break;
    }
    return false;
}

struct SurfaceHit {
    bool missed;
    f3 position;
    f3 normal;      // shading normal (flipped to face the ray)
    f3 normal_geo;  // geometric normal (flipped to face the ray)
    f3 base_color;
    float metallic;
    float roughness;  // perceptual
};

// geometry_rchit.inc.hlsl:376-472 for an untextured, pre-transformed scene:
// barycentric interpolation of position and normal (normal map = flat),
// geometric normal from the edges, two-sided flip.
__device__ __forceinline__ SurfaceHit shade_hit(const SceneView& s, const Hit& h, f3 ray_dir) {
    SurfaceHit r;
    r.missed = h.tri == 0xFFFFFFFFu;
    if (r.missed) return r;
    TriShade sh = gld(s.shade + h.tri);
    const float* P = s.positions;
    f3 p0 = mk3(gld(P + 3 * sh.i0), gld(P + 3 * sh.i0 + 1), gld(P + 3 * sh.i0 + 2));
    f3 p1 = mk3(gld(P + 3 * sh.i1), gld(P + 3 * sh.i1 + 1), gld(P + 3 * sh.i1 + 2));
    f3 p2 = mk3(gld(P + 3 * sh.i2), gld(P + 3 * sh.i2 + 1), gld(P + 3 * sh.i2 + 2));
    float w0 = (1.0f - h.u) - h.v;
    r.position = (p0 * w0 + p1 * h.u) + p2 * h.v;
    f3 ng = normalize3(cross3(p1 - p0, p2 - p0));
    f3 n;
    if (s.normals) {
        const float* N = s.normals;
        f3 n0 = mk3(gld(N + 3 * sh.i0), gld(N + 3 * sh.i0 + 1), gld(N + 3 * sh.i0 + 2));
        f3 n1 = mk3(gld(N + 3 * sh.i1), gld(N + 3 * sh.i1 + 1), gld(N + 3 * sh.i1 + 2));
        f3 n2 = mk3(gld(N + 3 * sh.i2), gld(N + 3 * sh.i2 + 1), gld(N + 3 * sh.i2 + 2));
        n = normalize3((n0 * w0 + n1 * h.u) + n2 * h.v);
    } else {
        n = ng;
    }
    if (dot3(ng, ray_dir) > 0.0f) {
        ng = -ng;
        n = -n;
    }
    r.normal = n;
    r.normal_geo = ng;
    const Material m = gld(s.materials + sh.material);
    r.base_color = mk3(m.base_color[0], m.base_color[1], m.base_color[2]);
    r.metallic = m.metallic;
    r.roughness = m.roughness;
    return r;
}

// raytrace.inc.hlsl:65-208 after its two TraceRay calls: the radiance of
// the ray (o, d) given its closest hit `sh` (shade_hit) and whether the sun
// shadow ray from the hit was unoccluded.  `S` is the storage policy of the
// previous frame's denoised diffuse (the screen-space radiance cache).
template <class S>
__device__ __forceinline__ RadianceTraceResult radiance_at(const TraceRes& rt, const FrameParams& fp, f3 o, f3 d,
                                                           const SurfaceHit& sh, bool sun_visible,
                                                           bool has_prev_frame) {
    RadianceTraceResult ret;
    if (sh.missed) {
        ret.radiance = rt.sky;
        ret.position_ws = o + d * kRayTMax;
        ret.normal_ws = -d;
        return ret;
    }
    ret.position_ws = sh.position;
    ret.normal_ws = sh.normal;
    f3 diffuse_rho = get_diffuse_rho(sh.base_color, sh.metallic);
    f3 sun_dir = xyz(fp.sun_dir);
    f3 sun_inten = xyz(fp.sun_inten);
    f3 radiance = mk3s(0.0f);
    if (sun_visible) {
        float pr = fmaxf(sh.roughness, 0.045f);
        pr = fmaxf(pr, 0.5f);  // RAYTRACE_INC_FIREFLY_SUPPRESS
        float NoL = saturate(dot3(sh.normal, sun_dir));
        f3 f0 = get_specular_f0(sh.base_color, sh.metallic);
        radiance = (eval_GGX_Lambertian(-d, sun_dir, sh.normal, pr, diffuse_rho, f0) * NoL) * sun_inten;
    }
    bool sample_world_radiance_cache = true;
    if (has_prev_frame) {
        f4 ph = mul44(fp.prev_view_proj, mk4(sh.position.x, sh.position.y, sh.position.z, 1.0f));
        float nx = ph.x / ph.w - fp.jitter[2];
        float ny = ph.y / ph.w - fp.jitter[3];
        float reproj_depth = ph.z / ph.w;
        if (fabsf(nx) < 1.0f && fabsf(ny) < 1.0f) {
            float u = nx * 0.5f + 0.5f, v = ny * 0.5f + 0.5f;
            // uint2(floor(uv * size)); clamped to the image (the product can
            // round up to `size` for uv just below 1)
            uint32_t ppx = minu(f2u(floorf(u * (float)rt.pw)), rt.pw - 1u);
            uint32_t ppy = minu(f2u(floorf(v * (float)rt.ph)), rt.ph - 1u);
            size_t pi = (size_t)ppy * rt.pw + ppx;
            float prev_depth_value = gld(rt.prev_depth + pi);
            if (fabsf(reproj_depth - prev_depth_value) < 0.0001f) {
                radiance = radiance + diffuse_rho * S::ld3(rt.prev_indirect_diffuse, pi);
                sample_world_radiance_cache = false;
            }
        }
    }
    if (sample_world_radiance_cache) {
        VertexDescriptor vd = vertex_descriptor_create(sh.position, xyz(fp.view_pos));
        uint32_t hsh = vertex_descriptor_hash(vd), chk = vertex_descriptor_checksum(vd);
        f3 cell_radiance;
        uint32_t cell_addr;
        if (hash_grid_find(rt.hg_storage, hsh, chk, cell_radiance, cell_addr))
            radiance = radiance + diffuse_rho * cell_radiance;
        HashQuery q;
        q.px = sh.position.x;
        q.py = sh.position.y;
        q.pz = sh.position.z;
        q.normal = normal_encode_oct_u32(sh.normal_geo);
        q.hash = hsh;
        q.checksum = chk;
        if (!rt.query_threshold || query_content_hash(q) >= G(rt.query_threshold)[0]) {
            uint32_t qi = g_atomic_add(rt.query_counter, 1u);
            if (qi < kHashGridMaxQueries) gst(rt.queries + qi, q);
        }
    }
    ret.radiance = radiance;
    return ret;
}

// raytrace.inc.hlsl:65-208: closest hit of (o, d), then the sun shadow ray
// from the hit (any hit), then radiance_at.
template <class S, bool kCount = false>
__device__ __forceinline__ RadianceTraceResult trace_radiance(const TraceRes& rt, const FrameParams& fp, f3 o, f3 d,
                                                              bool has_prev_frame, LdsStack stack) {
    Hit hit = trace_closest<kCount>(rt.scene, o, d, kRayTMin, kRayTMax, stack);
    SurfaceHit sh = shade_hit(rt.scene, hit, d);
    bool sun_visible = false;
    if (!sh.missed) sun_visible = !trace_any<kCount>(rt.scene, sh.position, xyz(fp.sun_dir), kRayTMin, kRayTMax, stack);
    return radiance_at<S>(rt, fp, o, d, sh, sun_visible, has_prev_frame);
}

}  // namespace vgi
