/*
 * vgi_radiance.h — device-side `trace_radiance` and `trace_shadow`
 * (restir/raytrace.inc.hlsl:30-208) and the closest-hit shading of
 * raytrace/geometry_rchit.inc.hlsl:376-472, plus the hash-grid probe
 * (restir/hash_grid.inc.hlsl:124-147).
 */
#pragma once
#include "vgi_storage.h"
#include "vgi_texture.h"
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

// The G-buffer's screen-space footprint of a primary ray: the camera
// position and the rays through the pixel's 2x2-quad partners, (px ^ 1, py)
// and (px, py ^ 1).  A raster pixel shader's fine derivatives difference an
// attribute between the two pixels of a quad row (column), both evaluated on
// the pixel's own triangle plane; the partner's barycentrics come from the
// same ray/plane intersection as the hit's, without the edge tests.
struct PrimaryFootprint {
    f3 o;
    f3 dir_x, dir_y;
    bool x_odd, y_odd;
};

constexpr float kNormalUpXY = (float)(128.0 / 255.0);  // render_scene.rs:808 texel, UNORM

__device__ __forceinline__ float2_ tri_uv(const float* uv, uint32_t i) {
    return float2_{gld(uv + 2 * i), gld(uv + 2 * i + 1)};
}

// Barycentric interpolation of a 2-vector, weights (w0, u, v).
__device__ __forceinline__ float2_ bary2(float2_ a, float2_ b, float2_ c, float w0, float u, float v) {
    return float2_{(a.x * w0 + b.x * u) + c.x * v, (a.y * w0 + b.y * u) + c.y * v};
}

// geometry_rchit.inc.hlsl:54-150 (and mesh_gbuffer.hlsl:62-97 for the
// G-buffer, kPrimary) on a pre-transformed scene: barycentric interpolation
// of position, normal, texture coordinates and tangent; material textures
// sampled at LOD 2 (indirect) or with the quad derivatives (primary);
// factors applied; normal mapping with the MikkTSpace bitangent; geometric
// normal from the edges; two-sided flip.  Texture indices 0 and 1 are
// violet's built-in white and flat-normal textures, taken as constants.  A
// scene without tangents and a material without a normal map keep the
// interpolated normal (the tangent terms vanish and the map's z is 1).
template <bool kPrimary>
__device__ __forceinline__ SurfaceHit shade_hit_t(const SceneView& s, const Hit& h, f3 ray_dir,
                                                  const PrimaryFootprint* pf) {
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
    f3 nl = ng;  // interpolated (unnormalised) vertex normal
    if (s.normals) {
        const float* N = s.normals;
        f3 n0 = mk3(gld(N + 3 * sh.i0), gld(N + 3 * sh.i0 + 1), gld(N + 3 * sh.i0 + 2));
        f3 n1 = mk3(gld(N + 3 * sh.i1), gld(N + 3 * sh.i1 + 1), gld(N + 3 * sh.i1 + 2));
        f3 n2 = mk3(gld(N + 3 * sh.i2), gld(N + 3 * sh.i2 + 1), gld(N + 3 * sh.i2 + 2));
        nl = (n0 * w0 + n1 * h.u) + n2 * h.v;
    }
    const Material m = gld(s.materials + sh.material);
    const uint32_t bc_i = m.color_metalrough_texture & 0xFFFFu, mr_i = m.color_metalrough_texture >> 16,
                   nm_i = m.normal_texture & 0xFFFFu;
    Tex4 bc{1.0f, 1.0f, 1.0f, 1.0f}, mr{1.0f, 1.0f, 1.0f, 1.0f}, nm{1.0f, 1.0f, 1.0f, 1.0f};
    if (bc_i == VGI_TEXTURE_NORMAL_UP) bc = Tex4{kNormalUpXY, kNormalUpXY, 1.0f, 1.0f};
    if (mr_i == VGI_TEXTURE_NORMAL_UP) mr = Tex4{kNormalUpXY, kNormalUpXY, 1.0f, 1.0f};
    if (nm_i == VGI_TEXTURE_NORMAL_UP) nm = Tex4{kNormalUpXY, kNormalUpXY, 1.0f, 1.0f};
    if (bc_i >= VGI_TEXTURE_FIRST || mr_i >= VGI_TEXTURE_FIRST || nm_i >= VGI_TEXTURE_FIRST) {
        float2_ t0{0.0f, 0.0f}, t1{0.0f, 0.0f}, t2{0.0f, 0.0f};
        if (s.texcoords) {
            t0 = tri_uv(s.texcoords, sh.i0);
            t1 = tri_uv(s.texcoords, sh.i1);
            t2 = tri_uv(s.texcoords, sh.i2);
        }
        const float2_ uv = bary2(t0, t1, t2, w0, h.u, h.v);
        TexGrad g{0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (kPrimary) {
            const f3 e1 = p1 - p0, e2 = p2 - p0, tvec = pf->o - p0, qvec = cross3(tvec, e1);
            auto uv_on_plane = [&](f3 d) {
                const f3 pvec = cross3(d, e2);
                const float inv_det = 1.0f / dot3(e1, pvec);
                const float bu = dot3(tvec, pvec) * inv_det;
                const float bv = dot3(d, qvec) * inv_det;
                return bary2(t0, t1, t2, (1.0f - bu) - bv, bu, bv);
            };
            const float2_ ux = uv_on_plane(pf->dir_x), uy = uv_on_plane(pf->dir_y);
            g.dsdx = pf->x_odd ? uv.x - ux.x : ux.x - uv.x;
            g.dtdx = pf->x_odd ? uv.y - ux.y : ux.y - uv.y;
            g.dsdy = pf->y_odd ? uv.x - uy.x : uy.x - uv.x;
            g.dtdy = pf->y_odd ? uv.y - uy.y : uy.y - uv.y;
        }
        auto sample = [&](uint32_t idx) {
            const TexDesc& t = s.textures[idx - VGI_TEXTURE_FIRST];
            if constexpr (kPrimary)
                return tex_sample_grad(s, t, uv.x, uv.y, g);
            else
                return tex_trilinear(s, t, kIndirectTexLod, uv.x, uv.y);
        };
        if (bc_i >= VGI_TEXTURE_FIRST) bc = sample(bc_i);
        if (mr_i >= VGI_TEXTURE_FIRST) mr = sample(mr_i);
        if (nm_i >= VGI_TEXTURE_FIRST) nm = sample(nm_i);
    }
    f3 n;
    if (!s.tangents && nm_i < VGI_TEXTURE_FIRST) {
        n = s.normals ? normalize3(nl) : ng;
    } else {
        const f3 ts = mk3(nm.r * 2.0f - 1.0f, nm.g * 2.0f - 1.0f, nm.b * 2.0f - 1.0f);
        f3 tg = mk3s(0.0f), bt = mk3s(0.0f);
        if (s.tangents) {
            const float* T = s.tangents;
            const f4 a = mk4(gld(T + 4 * sh.i0), gld(T + 4 * sh.i0 + 1), gld(T + 4 * sh.i0 + 2), gld(T + 4 * sh.i0 + 3));
            const f4 b = mk4(gld(T + 4 * sh.i1), gld(T + 4 * sh.i1 + 1), gld(T + 4 * sh.i1 + 2), gld(T + 4 * sh.i1 + 3));
            const f4 c = mk4(gld(T + 4 * sh.i2), gld(T + 4 * sh.i2 + 1), gld(T + 4 * sh.i2 + 2), gld(T + 4 * sh.i2 + 3));
            tg = (mk3(a.x, a.y, a.z) * w0 + mk3(b.x, b.y, b.z) * h.u) + mk3(c.x, c.y, c.z) * h.v;
            const float tw = (a.w * w0 + b.w * h.u) + c.w * h.v;
            bt = tw * cross3(nl, tg);
        }
        n = normalize3((ts.x * tg + ts.y * bt) + ts.z * nl);
    }
    if (dot3(ng, ray_dir) > 0.0f) {
        ng = -ng;
        n = -n;
    }
    r.normal = n;
    r.normal_geo = ng;
    r.base_color = mk3(bc.r * m.base_color[0], bc.g * m.base_color[1], bc.b * m.base_color[2]);
    r.metallic = mr.r * m.metallic;
    r.roughness = mr.g * m.roughness;
    return r;
}

// Indirect rays (geometry_ind.rchit.hlsl): textures at LOD 2.
__device__ __forceinline__ SurfaceHit shade_hit(const SceneView& s, const Hit& h, f3 ray_dir) {
    return shade_hit_t<false>(s, h, ray_dir, nullptr);
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
