/*
 * oracle_restir.c — ORACLE (test infrastructure only).  CPU restatement of
 * violet's ReSTIR GI passes, one function per HLSL pass, pixel loops in
 * dispatch order:
 *
 *   oracle_diff_sample_gen      restir/ind_diff_sample_gen.hlsl:25-67
 *   oracle_diff_sample_validate restir/ind_diff_sample_validate.hlsl:90-145
 *   oracle_spec_sample_gen      restir/ind_spec_sample_gen.hlsl:28-128
 *   oracle_spec_sample_validate restir/ind_spec_sample_validate.hlsl:144-191
 *   oracle_ao_gen / _filter     restir/raytraced_ao_gen.hlsl:18-81, raytraced_ao_filter.hlsl:13-70
 *   oracle_diff_temporal        restir/ind_diff_temporal_resample.hlsl:49-200
 *   oracle_diff_spatial         restir/ind_diff_spatial_resample.hlsl:94-337
 *   oracle_spec_temporal        restir/ind_spec_temporal_resample.hlsl:146-362
 *   oracle_spec_spatial         restir/ind_spec_spatial_resample.hlsl:169-448
 *   trace_radiance / shading    restir/raytrace.inc.hlsl:65-208,
 *                               raytrace/geometry_rchit.inc.hlsl:376-472
 *   hash-grid passes            restir/hash_grid_{decay,query_select,query_compact,raygen}.hlsl
 *   oracle_raytraced_shadow     raytraced_shadow.hlsl:9-53
 *   oracle_final_lighting       restir/final_lighting.hlsl:16-62
 *   oracle_temporal_aa          temporal_aa.hlsl:112-264
 *
 * Scene: world-space f32 positions/normals, u32 indices, per-triangle
 * material {base_color rgba, metallic, roughness, pad, pad}; queried through
 * the oracle's own median-split BVH (oracle_trace.inc).
 * Hash-grid queries are appended in pixel order (the GPU appends in
 * scheduling order; tests compare them as sets).
 */
#include "oracle_trace.inc"

static inline v3 vtx(const float* A, uint32_t i) { return V3(A[3 * i], A[3 * i + 1], A[3 * i + 2]); }

typedef struct {
    uint32_t w, h, levels, srgb;
    const uint32_t* level[16]; /* RGBA8 words, R in the low byte */
} otex;

typedef struct {
    obvh* bvh;
    const float* P;
    const float* N; /* may be NULL */
    const uint32_t* I;
    const uint32_t* TM;
    const float* mats; /* 8 words per material: base colour rgba, metallic, roughness, texture words */
    uint32_t nt;
    const float* UV; /* 2 per vertex or NULL */
    const float* TG; /* 4 per vertex or NULL */
    otex* tex;
    uint32_t ntex;
    real lut[512]; /* byte -> UNORM [0,256), sRGB-decoded [256,512) */
} oscene;

oscene* oracle_scene_create(const float* P, const float* N, const uint32_t* I, uint32_t nt, const uint32_t* TM,
                            const float* mats) {
    oscene* s = (oscene*)calloc(1, sizeof(oscene));
    s->bvh = obvh_create(P, I, nt);
    s->P = P;
    s->N = N;
    s->I = I;
    s->TM = TM;
    s->mats = mats;
    s->nt = nt;
    for (int i = 0; i < 256; ++i) {
        double c = i / 255.0;
        s->lut[i] = (real)c;
        s->lut[256 + i] = (real)(c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4));
    }
    return s;
}
/* Texturing inputs: texcoords (2 per vertex) and tangents (4 per vertex),
 * either NULL, and ntex textures whose texels are `texels` at the word
 * offsets of `meta` (per texture: width, height, levels, srgb, then one
 * offset per level; 20 words). */
void oracle_scene_set_textures(oscene* s, const float* UV, const float* TG, const uint32_t* texels,
                               const uint32_t* meta, uint32_t ntex) {
    s->UV = UV;
    s->TG = TG;
    free(s->tex);
    s->tex = ntex ? (otex*)calloc(ntex, sizeof(otex)) : NULL;
    s->ntex = ntex;
    for (uint32_t t = 0; t < ntex; ++t) {
        const uint32_t* m = meta + 20 * t;
        s->tex[t].w = m[0], s->tex[t].h = m[1], s->tex[t].levels = m[2], s->tex[t].srgb = m[3];
        for (uint32_t l = 0; l < m[2] && l < 16; ++l) s->tex[t].level[l] = texels + m[4 + l];
    }
}
void oracle_scene_destroy(oscene* s) {
    if (!s) return;
    obvh_destroy(s->bvh);
    free(s->tex);
    free(s);
}

/* ---- material textures: violet's sampler_linear_wrap (render_loop.rs:199-212)
 * through the Vulkan texel filtering equations (bilinear within a level,
 * linear between levels, REPEAT, anisotropy up to 16 in the specification's
 * reference form), sRGB decoded per texel before filtering ---- */
typedef struct { real r, g, b, a; } o4;
static inline o4 o4lerp(o4 x, o4 y, real t) {
    o4 o = {x.r + (y.r - x.r) * t, x.g + (y.g - x.g) * t, x.b + (y.b - x.b) * t, x.a + (y.a - x.a) * t};
    return o;
}
static inline uint32_t o_wrap(real f, uint32_t n) {
    int i = R_FABS(f) < RC(16777216.0) ? (int)f : 0;
    int r = i % (int)n;
    return (uint32_t)(r < 0 ? r + (int)n : r);
}
static inline o4 o_texel(const oscene* s, const otex* t, uint32_t l, uint32_t x, uint32_t y) {
    uint32_t wl = t->w >> l ? t->w >> l : 1u;
    uint32_t w = t->level[l][(size_t)y * wl + x];
    const real* lut = s->lut + (t->srgb ? 256 : 0);
    o4 o = {lut[w & 255u], lut[(w >> 8) & 255u], lut[(w >> 16) & 255u], s->lut[w >> 24]};
    return o;
}
static o4 o_tex_bilinear(const oscene* s, const otex* t, uint32_t l, real u, real v) {
    uint32_t wl = t->w >> l ? t->w >> l : 1u, hl = t->h >> l ? t->h >> l : 1u;
    real fu = u * (real)wl - RC(0.5), fv = v * (real)hl - RC(0.5);
    real x0f = R_FLOOR(fu), y0f = R_FLOOR(fv);
    real a = fu - x0f, b = fv - y0f;
    uint32_t x0 = o_wrap(x0f, wl), y0 = o_wrap(y0f, hl);
    uint32_t x1 = x0 + 1 == wl ? 0 : x0 + 1, y1 = y0 + 1 == hl ? 0 : y0 + 1;
    return o4lerp(o4lerp(o_texel(s, t, l, x0, y0), o_texel(s, t, l, x1, y0), a),
                  o4lerp(o_texel(s, t, l, x0, y1), o_texel(s, t, l, x1, y1), a), b);
}
static o4 o_tex_trilinear(const oscene* s, const otex* t, real lod, real u, real v) {
    uint32_t last = t->levels - 1;
    real d = lod > RC(0.0) ? lod : RC(0.0);
    if (d >= (real)last) return o_tex_bilinear(s, t, last, u, v);
    real dl = R_FLOOR(d), fr = d - dl;
    o4 lo = o_tex_bilinear(s, t, (uint32_t)dl, u, v);
    if (fr == RC(0.0)) return lo;
    return o4lerp(lo, o_tex_bilinear(s, t, (uint32_t)dl + 1, u, v), fr);
}
static o4 o_tex_grad(const oscene* s, const otex* t, real u, real v, real dsdx, real dtdx, real dsdy, real dtdy) {
    real W = (real)t->w, H = (real)t->h;
    real ax = dsdx * W, bx = dtdx * H, ay = dsdy * W, by = dtdy * H;
    real rx = R_SQRT(ax * ax + bx * bx), ry = R_SQRT(ay * ay + by * by);
    real rmax = R_FMAX(rx, ry), rmin = R_FMIN(rx, ry), lod = RC(0.0);
    uint32_t n = 1;
    if (!(rmax > RC(0.0))) {
        lod = RC(0.0);
    } else if (!(rmax <= RC(3.0e38))) {
        lod = RC(1.0e30);
    } else {
        real eta = rmin > RC(0.0) ? R_FMIN(rmax / rmin, RC(16.0)) : RC(16.0);
        n = (uint32_t)R_CEIL(eta);
        lod = R_LOG2(rmax / eta);
    }
    if (n <= 1) return o_tex_trilinear(s, t, lod, u, v);
    int ax_major = rx > ry;
    real du = ax_major ? dsdx : dsdy, dv = ax_major ? dtdx : dtdy;
    o4 sum = {0, 0, 0, 0};
    for (uint32_t i = 1; i <= n; ++i) {
        real o = (real)i / (real)(n + 1) - RC(0.5);
        o4 x = o_tex_trilinear(s, t, lod, u + du * o, v + dv * o);
        sum.r = sum.r + x.r, sum.g = sum.g + x.g, sum.b = sum.b + x.b, sum.a = sum.a + x.a;
    }
    o4 r = {sum.r / (real)n, sum.g / (real)n, sum.b / (real)n, sum.a / (real)n};
    return r;
}

/* Test access to the sampler: texture `t` (0-based in the scene's list) at
 * (u, v), either at an explicit level of detail (grad == NULL) or with the
 * four derivatives grad = {ds/dx, dt/dx, ds/dy, dt/dy}; out = rgba. */
void oracle_tex_sample(const oscene* s, uint32_t t, real u, real v, real lod, const real* grad, real* out) {
    o4 r = grad ? o_tex_grad(s, &s->tex[t], u, v, grad[0], grad[1], grad[2], grad[3])
                : o_tex_trilinear(s, &s->tex[t], lod, u, v);
    out[0] = r.r, out[1] = r.g, out[2] = r.b, out[3] = r.a;
}

/* A primary ray's quad partners (mesh_gbuffer.hlsl's implicit derivatives). */
typedef struct {
    v3 o, dir_x, dir_y;
    int x_odd, y_odd;
} ofoot;

typedef struct {
    int missed;
    v3 pos, n, ng, base;
    real metal, rough;
} osurf;

static inline v2 o_uv(const float* UV, uint32_t i) { return V2(UV[2 * i], UV[2 * i + 1]); }
static inline v2 o_bary2(v2 a, v2 b, v2 c, real w0, real u, real v) {
    return V2((a.x * w0 + b.x * u) + c.x * v, (a.y * w0 + b.y * u) + c.y * v);
}

/* geometry_rchit.inc.hlsl:54-150 (indirect rays: textures at LOD 2,
 * geometry_ind.rchit.hlsl:3) or mesh_gbuffer.hlsl:62-97 (foot != NULL:
 * implicit derivatives) on a pre-transformed scene.  Texture indices 0 and
 * 1 are violet's white and flat-normal textures (render_scene.rs:344-346). */
static osurf o_surface(const oscene* s, ohit h, v3 d, const ofoot* foot) {
    osurf r;
    r.missed = h.id == 0xFFFFFFFFu;
    if (r.missed) return r;
    uint32_t i0 = s->I[3 * h.id], i1 = s->I[3 * h.id + 1], i2 = s->I[3 * h.id + 2];
    v3 p0 = vtx(s->P, i0), p1 = vtx(s->P, i1), p2 = vtx(s->P, i2);
    real w0 = (RC(1.0) - h.u) - h.v;
    r.pos = add3(add3(scale3(p0, w0), scale3(p1, h.u)), scale3(p2, h.v));
    v3 ng = norm3(cross3(sub3(p1, p0), sub3(p2, p0)));
    v3 nl = ng;
    if (s->N) nl = add3(add3(scale3(vtx(s->N, i0), w0), scale3(vtx(s->N, i1), h.u)), scale3(vtx(s->N, i2), h.v));
    const float* m = s->mats + 8 * (s->TM ? s->TM[h.id] : 0);
    uint32_t cm = f_bits(m[6]), nw = f_bits(m[7]);
    uint32_t bci = cm & 0xFFFFu, mri = cm >> 16, nmi = nw & 0xFFFFu;
    const real up = (real)(128.0 / 255.0);
    o4 white = {1, 1, 1, 1}, flat = {up, up, 1, 1};
    o4 bc = bci == 1 ? flat : white, mr = mri == 1 ? flat : white, nm = nmi == 1 ? flat : white;
    if (bci >= 2 || mri >= 2 || nmi >= 2) {
        v2 t0 = V2(0, 0), t1 = V2(0, 0), t2 = V2(0, 0);
        if (s->UV) t0 = o_uv(s->UV, i0), t1 = o_uv(s->UV, i1), t2 = o_uv(s->UV, i2);
        v2 uv = o_bary2(t0, t1, t2, w0, h.u, h.v);
        real g[4] = {0, 0, 0, 0};
        if (foot) {
            v3 e1 = sub3(p1, p0), e2 = sub3(p2, p0), tv = sub3(foot->o, p0), qv = cross3(tv, e1);
            v2 q[2];
            for (int k = 0; k < 2; ++k) {
                v3 dd = k ? foot->dir_y : foot->dir_x;
                v3 pv = cross3(dd, e2);
                real inv = RC(1.0) / dot3(e1, pv);
                real bu = dot3(tv, pv) * inv, bv = dot3(dd, qv) * inv;
                q[k] = o_bary2(t0, t1, t2, (RC(1.0) - bu) - bv, bu, bv);
            }
            g[0] = foot->x_odd ? uv.x - q[0].x : q[0].x - uv.x;
            g[1] = foot->x_odd ? uv.y - q[0].y : q[0].y - uv.y;
            g[2] = foot->y_odd ? uv.x - q[1].x : q[1].x - uv.x;
            g[3] = foot->y_odd ? uv.y - q[1].y : q[1].y - uv.y;
        }
        uint32_t idx[3] = {bci, mri, nmi};
        o4* dst[3] = {&bc, &mr, &nm};
        for (int k = 0; k < 3; ++k) {
            if (idx[k] < 2) continue;
            const otex* t = &s->tex[idx[k] - 2];
            *dst[k] = foot ? o_tex_grad(s, t, uv.x, uv.y, g[0], g[1], g[2], g[3])
                           : o_tex_trilinear(s, t, RC(2.0), uv.x, uv.y);
        }
    }
    v3 n;
    if (!s->TG && nmi < 2) {
        n = s->N ? norm3(nl) : ng;
    } else {
        v3 ts = V3(nm.r * RC(2.0) - RC(1.0), nm.g * RC(2.0) - RC(1.0), nm.b * RC(2.0) - RC(1.0));
        v3 tg = V3s(0), bt = V3s(0);
        if (s->TG) {
            const float* a = s->TG + 4 * i0;
            const float* b = s->TG + 4 * i1;
            const float* c = s->TG + 4 * i2;
            tg = add3(add3(scale3(V3(a[0], a[1], a[2]), w0), scale3(V3(b[0], b[1], b[2]), h.u)),
                      scale3(V3(c[0], c[1], c[2]), h.v));
            real tw = ((real)a[3] * w0 + (real)b[3] * h.u) + (real)c[3] * h.v;
            bt = scale3(cross3(nl, tg), tw);
        }
        n = norm3(add3(add3(scale3(tg, ts.x), scale3(bt, ts.y)), scale3(nl, ts.z)));
    }
    if (dot3(ng, d) > RC(0.0)) {
        ng = neg3(ng);
        n = neg3(n);
    }
    r.n = n;
    r.ng = ng;
    r.base = V3(bc.r * m[0], bc.g * m[1], bc.b * m[2]);
    r.metal = mr.r * m[4];
    r.rough = mr.g * m[5];
    return r;
}

typedef struct {
    oscene* scene;
    float sky[3];
    const real* prev_ind_diff; /* radiance (3 reals per pixel) */
    const real* prev_depth;
    int32_t pw, ph;
    const uint32_t* hg_cells; /* 4 words per cell: checksum, rgb (f32 bits) */
    uint32_t* queries;        /* 6 words per query */
    uint32_t* query_counter;
    const uint32_t* query_threshold; /* NULL: append every query (the reference) */
} otrace_res;

/* Hash of a query's six words (jenkins chain); see oracle_hg_select_ordered. */
static uint32_t o_query_hash(const uint32_t* q) {
    uint32_t h = o_jenkins(q[0]);
    for (int k = 1; k < 6; ++k) h = o_jenkins(h ^ q[k]);
    return h;
}

/* Threshold word keeping about HG_MAX_QUERIES / 2 of `expected` query
 * candidates (a candidate is kept if its hash is >= the word). */
#define HG_QUERY_TARGET (128u * 1024u / 2u)
uint32_t oracle_hg_query_threshold_for(uint64_t expected) {
    if (expected <= HG_QUERY_TARGET) return 0u;
    uint64_t keep = ((uint64_t)HG_QUERY_TARGET << 32) / expected;
    if (keep == 0) keep = 1;
    return (uint32_t)((1ull << 32) - keep);
}

typedef struct {
    v3 pos, normal, radiance;
} otrace_out;

#define HG_MAX_QUERIES (128u * 1024u)

/* hash_grid.inc.hlsl:32-105 VertexDescriptor keys */
static inline void o_hg_key(v3 pos, v3 cam, uint32_t* hash, uint32_t* checksum) {
    v3 dd = sub3(pos, cam);
    real d2 = dot3(dd, dd);
    double l = 0.5 * log2((double)d2);
    if (!(l > 0.0)) l = 0.0;
    uint32_t lod = (uint32_t)l;
    if (lod > 31) lod = 31;
    real cell = (RC(1.0) / RC(16.0)) * (real)(1u << lod);
    real inv = RC(1.0) / cell;
    /* quantise the single-precision position (the stored value) */
    float fx = (float)pos.x, fy = (float)pos.y, fz = (float)pos.z;
    int32_t qx = (int32_t)floor((double)fx * (double)(float)inv);
    int32_t qy = (int32_t)floor((double)fy * (double)(float)inv);
    int32_t qz = (int32_t)floor((double)fz * (double)(float)inv);
    uint32_t ux = (uint32_t)qx, uy = (uint32_t)qy, uz = (uint32_t)qz;
    uint32_t px_ = ux ^ (uy << 22);
    uint32_t py_ = (uy >> 10) ^ (uz << 22);
    *hash = o_xxh_eval(o_xxh_add(o_xxh_add(o_xxh_init(px_), py_), lod)) & 0xFFFFu;
    *checksum = o_xxh_eval(o_xxh_add(o_xxh_add(o_xxh_add(o_xxh_init(uy), uz), ux), lod)) | 1u;
}

static inline int o_hg_find(const uint32_t* cells, uint32_t hash, uint32_t checksum, v3* rad, uint32_t* addr) {
    for (uint32_t a = hash << 4; a < (hash << 4) + 16; ++a) {
        const uint32_t* c = cells + 4 * (size_t)a;
        if (c[0] == checksum) {
            *rad = V3(bits_f(c[1]), bits_f(c[2]), bits_f(c[3]));
            *addr = a;
            return 1;
        }
        if (c[0] == 0) break;
    }
    return 0;
}

/* brdf.hlsl */
static inline real o_pow5(real x) {
#if defined(ORACLE_F32)
    real x2 = x * x, x4 = x2 * x2;
    return x4 * x;
#else
    return pow(x, 5.0);
#endif
}
static inline v3 o_F_schlick(real u, v3 f0) {
    real f = o_pow5(RC(1.0) - u);
    return V3(f + f0.x * (RC(1.0) - f), f + f0.y * (RC(1.0) - f), f + f0.z * (RC(1.0) - f));
}
static inline real o_D_ggx(real NoH, real r) {
    real a = NoH * r;
    real k = r / ((RC(1.0) - NoH * NoH) + a * a);
    return (k * k) * O_ONE_OVER_PI;
}
static inline real o_vis_smith(real NoL, real NoV, real r) {
    real r2 = r * r;
    real uL = NoV * R_SQRT((-r2 * NoL + NoL) * NoL + r2);
    real uV = NoL * R_SQRT((-r2 * NoV + NoV) * NoV + r2);
    return RC(0.5) / (uL + uV);
}
static inline real o_g2_over_g1(real NoL, real NoV, real r) {
    real r2 = r * r;
    real kV = R_SQRT((-r2 * NoV + NoV) * NoV + r2) / NoV;
    real kL = R_SQRT((-r2 * NoL + NoL) * NoL + r2) / NoL;
    return (RC(1.0) + kV) / (kV + kL);
}
static inline real o_lambda_pomc(real c, real r) {
    real r2 = r * r;
    return (c + R_SQRT((-r2 * c + c) * c + r2)) / RC(2.0);
}
static inline v3 o_f0(v3 base, real metal) {
    return V3(lerp1(RC(0.04), base.x, metal), lerp1(RC(0.04), base.y, metal), lerp1(RC(0.04), base.z, metal));
}
static inline v3 o_eval_ggx_lambert(v3 v, v3 l, v3 n, real pr, v3 rho, v3 f0) {
    v3 h = norm3(add3(v, l));
    real NoL = sat(dot3(n, l)), NoV = sat(dot3(n, v)), NoH = sat(dot3(n, h)), LoH = sat(dot3(l, h));
    real r = pr * pr;
    real D = o_D_ggx(NoH, r);
    v3 F = o_F_schlick(LoH, f0);
    real V = o_vis_smith(NoL, NoV, r);
    real dv = D * V;
    if (dv > RC(1e19)) dv = RC(1e19);
    v3 Fr = scale3(F, dv);
    v3 Fd = scale3(rho, O_ONE_OVER_PI);
    return add3(Fd, Fr);
}
static inline v3 o_ggx_integral(real c, real r, v3 f0) {
    real a = r * r, a3 = a * (a * a), c2 = c * c, c3 = c * c2;
    real k0 = RC(1.0) * (RC(0.99044) * RC(1.0) + RC(-1.28514) * c) + a * (RC(1.29678) * RC(1.0) + RC(-0.755907) * c);
    k0 = k0 / ((RC(1.0) * ((RC(1.0) * RC(1.0) + RC(2.92338) * c) + RC(59.4188) * c3) +
                a * ((RC(20.3225) * RC(1.0) + RC(-27.0302) * c) + RC(222.592) * c3)) +
               a3 * ((RC(121.563) * RC(1.0) + RC(626.13) * c) + RC(316.627) * c3));
    real k1 = RC(1.0) * (RC(0.0365463) * RC(1.0) + RC(3.32707) * c) + a * (RC(9.0632) * RC(1.0) + RC(-9.04756) * c);
    k1 = k1 / ((RC(1.0) * ((RC(1.0) * RC(1.0) + RC(3.59685) * c2) + RC(-1.36772) * c3) +
                a * ((RC(9.04401) * RC(1.0) + RC(-16.3174) * c2) + RC(9.22949) * c3)) +
               a3 * ((RC(5.56589) * RC(1.0) + RC(19.7886) * c2) + RC(-20.2123) * c3));
    return V3(k0 + k1 * f0.x, k0 + k1 * f0.y, k0 + k1 * f0.z);
}

/* sampling.hlsl */
static inline v4 o_rot_to_z(v3 v) {
    if (v.z < RC(-0.99999)) return V4(1, 0, 0, 0);
    v4 q = V4(v.y, -v.x, RC(0.0), RC(1.0) + v.z);
    real inv = RC(1.0) / R_SQRT(((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w);
    return V4(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
}
static inline v4 o_inv_rot(v4 q) { return V4(-q.x, -q.y, -q.z, q.w); }
static inline v3 o_rotate(v4 q, v3 v) {
    v3 ax = V3(q.x, q.y, q.z);
    v3 a = scale3(ax, RC(2.0) * dot3(ax, v));
    v3 b = scale3(v, q.w * q.w - dot3(ax, ax));
    v3 c = scale3(cross3(ax, v), RC(2.0) * q.w);
    return add3(add3(a, b), c);
}

/* TraceRay + geometry_rchit.inc.hlsl + raytrace.inc.hlsl:65-208 */
static otrace_out o_trace_radiance(const otrace_res* rt, const ofp* fp, v3 o, v3 d, int has_prev) {
    otrace_out r;
    const oscene* s = rt->scene;
    ohit h = obvh_trace(s->bvh, o, d, RC(0.0005), RC(1000.0), 0);
    if (h.id == 0xFFFFFFFFu) {
        r.radiance = V3(rt->sky[0], rt->sky[1], rt->sky[2]);
        r.pos = add3(o, scale3(d, RC(1000.0)));
        r.normal = neg3(d);
        return r;
    }
    osurf sf = o_surface(s, h, d, NULL);
    v3 pos = sf.pos, n = sf.n, ng = sf.ng, base = sf.base;
    real metal = sf.metal, rough = sf.rough;
    v3 rho = scale3(base, RC(1.0) - metal);
    v3 sun_dir = o_xyz(fp->sun_dir), sun = o_xyz(fp->sun_inten);
    v3 rad = V3s(0);
    ohit sh = obvh_trace(s->bvh, pos, sun_dir, RC(0.0005), RC(1000.0), 1);
    if (sh.id == 0xFFFFFFFFu) {
        real pr = R_FMAX(rough, RC(0.045));
        pr = R_FMAX(pr, RC(0.5));
        real NoL = sat(dot3(n, sun_dir));
        rad = mul3(scale3(o_eval_ggx_lambert(neg3(d), sun_dir, n, pr, rho, o_f0(base, metal)), NoL), sun);
    }
    int world = 1;
    if (has_prev) {
        v4 ph = o_mul44(fp->prev_view_proj, V4(pos.x, pos.y, pos.z, RC(1.0)));
        real nx = ph.x / ph.w - fp->jitter[2], ny = ph.y / ph.w - fp->jitter[3];
        real rd = ph.z / ph.w;
        if (R_FABS(nx) < RC(1.0) && R_FABS(ny) < RC(1.0)) {
            real u = nx * RC(0.5) + RC(0.5), v = ny * RC(0.5) + RC(0.5);
            int ppx = (int)R_FLOOR(u * (real)rt->pw), ppy = (int)R_FLOOR(v * (real)rt->ph);
            if (ppx > rt->pw - 1) ppx = rt->pw - 1;
            if (ppy > rt->ph - 1) ppy = rt->ph - 1;
            size_t pi = (size_t)ppy * rt->pw + ppx;
            if (R_FABS(rd - rt->prev_depth[pi]) < RC(0.0001)) {
                const real* c = rt->prev_ind_diff + 3 * pi;
                rad = add3(rad, mul3(rho, V3(c[0], c[1], c[2])));
                world = 0;
            }
        }
    }
    if (world) {
        uint32_t hh, ck, addr;
        o_hg_key(pos, o_xyz(fp->view_pos), &hh, &ck);
        v3 cr;
        if (o_hg_find(rt->hg_cells, hh, ck, &cr, &addr)) rad = add3(rad, mul3(rho, cr));
        uint32_t qw[6] = {f_bits((float)pos.x), f_bits((float)pos.y), f_bits((float)pos.z), o_oct_enc_u32(ng), hh, ck};
        if (!rt->query_threshold || o_query_hash(qw) >= *rt->query_threshold) {
            uint32_t qi;
#pragma omp atomic capture
            qi = (*rt->query_counter)++;
            if (qi < HG_MAX_QUERIES) memcpy(rt->queries + 6 * (size_t)qi, qw, sizeof(qw));
        }
    }
    r.pos = pos;
    r.normal = n;
    r.radiance = rad;
    return r;
}

static inline v3 ld3r(const real* p, size_t i) { return V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st3r(real* p, size_t i, v3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

#define DEPTH_ERR (RC(1.0) / RC(16777216.0))

/* Batch trace_radiance (tests): rays {o, -, d, -}, out 9 reals per ray. */
void oracle_trace_radiance(const float* fpf, const otrace_res* rt, const float* rays, uint32_t n, int has_prev,
                           real* out) {
    ofp fp;
    o_load_fp(fpf, &fp);
    for (uint32_t k = 0; k < n; ++k) {
        const float* r = rays + 8 * (size_t)k;
        otrace_out t = o_trace_radiance(rt, &fp, V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), has_prev);
        real* o = out + 9 * (size_t)k;
        o[0] = t.pos.x, o[1] = t.pos.y, o[2] = t.pos.z;
        o[3] = t.normal.x, o[4] = t.normal.y, o[5] = t.normal.z;
        o[6] = t.radiance.x, o[7] = t.radiance.y, o[8] = t.radiance.z;
    }
}

void oracle_diff_sample_gen(const float* fpf, const otrace_res* rt, int w, int h, const uint32_t* gb, const real* depth,
                            uint32_t frame_rand, int has_prev, uint32_t* o_hit, real* o_rad) {
    ofp fp;
    o_load_fp(fpf, &fp);
    size_t plane = (size_t)w * h;
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real dpt = depth[i];
            if (dpt == RC(0.0)) {
                o_hit[2 * i] = o_hit[2 * i + 1] = 0;
                st3r(o_rad, i, V3s(0));
                continue;
            }
            v3 nrm = o_gb_normal(gb[plane + i]);
            v3 pos = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, dpt + DEPTH_ERR);
            uint32_t rng = o_lcg_init(px, py, frame_rand);
            real u0 = o_rand(&rng), u1 = o_rand(&rng);
            real cos_t = RC(1.0) - u0, sin_t = R_SQRT(RC(1.0) - cos_t * cos_t), sp, cp;
            r_sincos(O_TWO_PI * u1, &sp, &cp);
            v3 l = V3(cp * sin_t, sp * sin_t, cos_t);
            v3 dir = o_rotate(o_inv_rot(o_rot_to_z(nrm)), l);
            otrace_out t = o_trace_radiance(rt, &fp, pos, dir, has_prev);
            o_hit_enc(t.pos, t.normal, &o_hit[2 * i], &o_hit[2 * i + 1]);
            st3r(o_rad, i, t.radiance);
        }
}

void oracle_diff_sample_validate(const float* fpf, const otrace_res* rt, int w, int h, const real* prev_depth,
                                 uint32_t* rw_rsv, uint32_t* rw_hit, real* rw_rad) {
    ofp fp;
    o_load_fp(fpf, &fp);
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real dpt = prev_depth[i];
            if (dpt == RC(0.0)) continue;
            v3 hp, hn;
            o_hit_dec(rw_hit[2 * i], rw_hit[2 * i + 1], &hp, &hn);
            v3 prev_rad = ld3r(rw_rad, i);
            v3 org = o_depth_to_pos(fp.prev_inv_view_proj, px, py, w, h, dpt);
            v3 dir = norm3(sub3(hp, org));
            otrace_out t = o_trace_radiance(rt, &fp, org, dir, 1);
            real lp = lum(prev_rad);
            real ld = lum(abs3(sub3(t.radiance, prev_rad)));
            real len = len3(sub3(hp, org));
            int moved = cmax3(abs3(sub3(hp, t.pos))) > R_FMAX(len / RC(128.0), RC(0.01));
            real ms = sat(RC(1.0) - (RC(1.0) * ld) / lp);
            if (moved) ms = RC(0.0);
            if (lp > RC(0.0) && ms <= (RC(1.0) - RC(0.5) / RC(32.0))) {
                uint32_t M;
                float W;
                o_rsv_dec(rw_rsv[i], &M, &W);
                M = r2u(R_FLOOR((real)M * ms));
                rw_rsv[i] = o_rsv_enc(M, W);
            }
            o_hit_enc(t.pos, t.normal, &rw_hit[2 * i], &rw_hit[2 * i + 1]);
            st3r(rw_rad, i, t.radiance);
        }
}

static v3 o_spec_dir(v3 view_dir, v3 n, real a, uint32_t* rng) {
    v4 rot = o_rot_to_z(n);
    v3 V = o_rotate(rot, view_dir);
    v3 H;
    if (a != RC(0.0)) {
        real ux = o_rand(rng), uy = o_rand(rng);
        v3 vh = norm3(V3(a * V.x, a * V.y, V.z));
        real lensq = vh.x * vh.x + vh.y * vh.y;
        v3 T1 = lensq > RC(0.0) ? scale3(V3(-vh.y, vh.x, RC(0.0)), RC(1.0) / R_SQRT(lensq)) : V3(1, 0, 0);
        v3 T2 = cross3(vh, T1);
        real r = R_SQRT(ux), sp, cp;
        r_sincos(O_TWO_PI * uy, &sp, &cp);
        real t1 = r * cp, t2 = r * sp;
        real s = RC(0.5) * (RC(1.0) + vh.z);
        t2 = lerp1(R_SQRT(RC(1.0) - t1 * t1), t2, s);
        v3 nh = add3(add3(scale3(T1, t1), scale3(T2, t2)), scale3(vh, R_SQRT(R_FMAX(RC(0.0), (RC(1.0) - t1 * t1) - t2 * t2))));
        H = norm3(V3(a * nh.x, a * nh.y, R_FMAX(RC(0.0), nh.z)));
    } else {
        H = V3(0, 0, 1);
    }
    v3 I = neg3(V);
    v3 L = sub3(I, scale3(H, RC(2.0) * dot3(H, I)));
    return o_rotate(o_inv_rot(rot), L);
}

void oracle_spec_sample_gen(const float* fpf, const otrace_res* rt, int w, int h, const uint32_t* gb, const real* depth,
                            uint32_t frame_rand, int has_prev, real* o_hitpos, real* o_rad) {
    ofp fp;
    o_load_fp(fpf, &fp);
    size_t plane = (size_t)w * h;
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real dpt = depth[i];
            if (dpt == RC(0.0)) {
                /* unwritten in the reference; zero here and in the library */
                for (int k = 0; k < 4; ++k) o_hitpos[4 * i + k] = RC(0.0);
                st3r(o_rad, i, V3s(0));
                continue;
            }
            v3 nrm = o_gb_normal(gb[plane + i]);
            real pr = o_gb_rough(gb[plane + i]);
            v3 pos = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, dpt + DEPTH_ERR);
            uint32_t rng = o_lcg_init(px, py, frame_rand);
            v3 vd = norm3(sub3(o_xyz(fp.view_pos), pos));
            real a = R_FMAX(pr * pr, RC(0.002));
            v3 dir = o_spec_dir(vd, nrm, a, &rng);
            otrace_out t = o_trace_radiance(rt, &fp, pos, dir, has_prev);
            real* hp = o_hitpos + 4 * i;
            hp[0] = t.pos.x, hp[1] = t.pos.y, hp[2] = t.pos.z, hp[3] = RC(1.0);
            st3r(o_rad, i, t.radiance);
        }
}

void oracle_spec_sample_validate(const float* fpf, const otrace_res* rt, int w, int h, uint32_t* rw_rsv, real* rw_hitpos,
                                 real* rw_rad) {
    ofp fp;
    o_load_fp(fpf, &fp);
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real pd = rt->prev_depth[i];
            if (pd == RC(0.0)) continue;
            v3 pos = o_depth_to_pos(fp.prev_inv_view_proj, px, py, w, h, pd + DEPTH_ERR);
            real* hp = rw_hitpos + 4 * i;
            v3 php = V3(hp[0], hp[1], hp[2]);
            real ptf = hp[3];
            v3 prad = ld3r(rw_rad, i);
            otrace_out t = o_trace_radiance(rt, &fp, pos, norm3(sub3(php, pos)), 1);
            real mr = R_FMAX(cmax3(t.radiance), cmax3(prad));
            v3 dr = abs3(sub3(t.radiance, prad));
            real lim = RC(0.5) * mr;
            if (dr.x > lim || dr.y > lim || dr.z > lim) {
                rw_rsv[2 * i] = 1;
                hp[0] = t.pos.x, hp[1] = t.pos.y, hp[2] = t.pos.z, hp[3] = ptf;
                st3r(rw_rad, i, t.radiance);
            }
        }
}

/* ---- AO ---- */
static real bil1(const real* t, int w, int h, real u, real v) {
    real tx = u * (real)w - RC(0.5), ty = v * (real)h - RC(0.5);
    real x0 = R_FLOOR(tx), y0 = R_FLOOR(ty), fx = tx - x0, fy = ty - y0;
    int xa = clampi((int)x0, 0, w - 1), xb = clampi((int)x0 + 1, 0, w - 1);
    int ya = clampi((int)y0, 0, h - 1), yb = clampi((int)y0 + 1, 0, h - 1);
    real top = t[(size_t)ya * w + xa] * (RC(1.0) - fx) + t[(size_t)ya * w + xb] * fx;
    real bot = t[(size_t)yb * w + xa] * (RC(1.0) - fx) + t[(size_t)yb * w + xb] * fx;
    return top * (RC(1.0) - fy) + bot * fy;
}
static v2 bil2(const real* t, int w, int h, real u, real v) {
    real tx = u * (real)w - RC(0.5), ty = v * (real)h - RC(0.5);
    real x0 = R_FLOOR(tx), y0 = R_FLOOR(ty), fx = tx - x0, fy = ty - y0;
    int xa = clampi((int)x0, 0, w - 1), xb = clampi((int)x0 + 1, 0, w - 1);
    int ya = clampi((int)y0, 0, h - 1), yb = clampi((int)y0 + 1, 0, h - 1);
    v2 r;
    for (int c = 0; c < 2; ++c) {
        real top = t[2 * ((size_t)ya * w + xa) + c] * (RC(1.0) - fx) + t[2 * ((size_t)ya * w + xb) + c] * fx;
        real bot = t[2 * ((size_t)yb * w + xa) + c] * (RC(1.0) - fx) + t[2 * ((size_t)yb * w + xb) + c] * fx;
        real val = top * (RC(1.0) - fy) + bot * fy;
        if (c == 0) r.x = val; else r.y = val;
    }
    return r;
}

void oracle_ao_gen(const float* fpf, int w, int h, const uint32_t* hit, const real* depth, const real* prev_ao,
                   int has_new, int has_prev, real radius, real* o_ao) {
    ofp fp;
    o_load_fp(fpf, &fp);
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real d = depth[i];
            if (d == RC(0.0)) {
                o_ao[2 * i] = o_ao[2 * i + 1] = 0;
                continue;
            }
            v3 p = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, d);
            v4 hp = o_mul44(fp.view_proj, V4(p.x, p.y, p.z, RC(1.0)));
            v4 hr = o_mul44(fp.prev_view_proj, V4(p.x, p.y, p.z, RC(1.0)));
            real mvx = ((hp.x / hp.w - fp.jitter[0]) - (hr.x / hr.w - fp.jitter[2])) * RC(0.5);
            real mvy = ((hp.y / hp.w - fp.jitter[1]) - (hr.y / hr.w - fp.jitter[3])) * RC(0.5);
            real hu = ((real)px + RC(0.5)) / (real)w - mvx, hv = ((real)py + RC(0.5)) / (real)h - mvy;
            v2 hist = V2(0, 0);
            if (hu == sat(hu) && hv == sat(hv) && has_prev) hist = bil2(prev_ao, w, h, hu, hv);
            if (has_new) {
                v3 hpos, hn;
                o_hit_dec(hit[2 * i], hit[2 * i + 1], &hpos, &hn);
                real dist = len3(sub3(hpos, p));
                real ao = dist < radius ? RC(0.0) : RC(1.0);
                hist.x = lerp1(ao, hist.x, ((RC(32.0) - RC(1.0)) / RC(32.0)) * hist.y);
                hist.y = RC(1.0);
            } else {
                hist.x = hist.y > RC(0.0) ? hist.x : RC(1.0);
            }
            o_ao[2 * i] = hist.x;
            o_ao[2 * i + 1] = hist.y;
        }
}

/* HLSL clamp(a, b, c) = min(max(a, b), c): the filter's "median"
 * (raytraced_ao_filter.hlsl:52-60) is this clamp chain, a median only for
 * some orderings of its arguments. */
static inline real med3c(real a, real b, real c) { return R_FMIN(R_FMAX(a, b), c); }

void oracle_ao_filter(int w, int h, const real* depth, const real* ao, real* o_ao) {
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real d = depth[i];
            if (d == RC(0.0)) {
                o_ao[2 * i] = o_ao[2 * i + 1] = 0;
                continue;
            }
            for (int c = 0; c < 2; ++c) {
                real center = ao[2 * i + c];
                real cmed[3], cmin[3], cmax[3];
                for (int y = -1; y <= 1; y++) {
                    real row[3];
                    for (int x = -1; x <= 1; x++) {
                        size_t j = (size_t)clampi(py + y, 0, h - 1) * w + clampi(px + x, 0, w - 1);
                        row[x + 1] = R_FABS(d - depth[j]) > RC(0.0005) ? center : ao[2 * j + c];
                    }
                    cmed[y + 1] = med3c(row[0], row[1], row[2]);
                    cmin[y + 1] = R_FMIN(R_FMIN(row[0], row[1]), row[2]);
                    cmax[y + 1] = R_FMAX(R_FMAX(row[0], row[1]), row[2]);
                }
                real A = R_FMIN(R_FMIN(cmax[0], cmax[1]), cmax[2]);
                real B = R_FMAX(R_FMAX(cmin[0], cmin[1]), cmin[2]);
                real C = med3c(cmed[0], cmed[1], cmed[2]);
                real med = med3c(C, A, B);
                real mn = R_FMIN(R_FMIN(cmin[0], cmin[1]), cmin[2]);
                o_ao[2 * i + c] = lerp1(med, mn, RC(0.5));
            }
        }
}

/* ---- diffuse temporal ---- */
void oracle_diff_temporal(const float* fpf, int w, int h, const real* prev_depth, const real* depth,
                          const uint32_t* new_hit, const real* new_rad, const uint32_t* prev_rsv, const uint32_t* prev_hit,
                          const real* prev_rad, uint32_t frame_rand, int has_prev, int has_new, uint32_t* o_rsv,
                          uint32_t* o_hit, real* o_rad, uint32_t m_cap) {
    /* M_MAX (ind_diff_temporal_resample.hlsl:136) unless set */
    const uint32_t Mcap = m_cap ? m_cap : 16u;
    ofp fp;
    o_load_fp(fpf, &fp);
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real d = depth[i];
            v3 p = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, d);
            v4 hr = o_mul44(fp.prev_view_proj, V4(p.x, p.y, p.z, RC(1.0)));
            real nx = hr.x / hr.w - fp.jitter[2], ny = hr.y / hr.w - fp.jitter[3], nz = hr.z / hr.w - RC(0.0);
            real u = nx * RC(0.5) + RC(0.5), v = ny * RC(0.5) + RC(0.5);
            int sample_prev = has_prev && R_FABS(nx) < RC(1.0) && R_FABS(ny) < RC(1.0);
            if (sample_prev && R_FABS(bil1(prev_depth, w, h, u, v) - nz) > RC(0.001)) sample_prev = 0;
            uint32_t M = 0, ex = 0, ey = 0;
            float Wf = 0.0f;
            real W = 0;
            v3 rad = V3s(0);
            if (sample_prev) {
                int ppx = (int)r2u(u * (real)w), ppy = (int)r2u(v * (real)h);
                if (ppx > w - 1) ppx = w - 1;
                if (ppy > h - 1) ppy = h - 1;
                size_t j = (size_t)ppy * w + ppx;
                o_rsv_dec(prev_rsv[j], &M, &Wf);
                W = Wf;
                rad = ld3r(prev_rad, j);
                ex = prev_hit[2 * j];
                ey = prev_hit[2 * j + 1];
                if (M == 0) {
                    rad = V3s(0);
                    ex = ey = 0;
                }
                if (M > Mcap) M = Mcap;
            }
            if (has_new) {
                real tp = lum(rad);
                real pw = (W * tp) * (real)M;
                v3 nr = ld3r(new_rad, i);
                real nt = lum(nr);
                real nw = nt * O_TWO_PI;
                uint32_t rng = o_lcg_init(px, py, frame_rand);
                real ws = pw + nw;
                real chance = nw / ws;
                if (o_rand(&rng) < chance) {
                    ex = new_hit[2 * i];
                    ey = new_hit[2 * i + 1];
                    rad = nr;
                    tp = nt;
                }
                M += 1;
                W = tp > RC(0.0) ? ws / (tp * (real)M) : RC(0.0);
            }
            o_rsv[i] = o_rsv_enc(M, (float)W);
            o_hit[2 * i] = ex;
            o_hit[2 * i + 1] = ey;
            st3r(o_rad, i, rad);
        }
}

/* ---- diffuse spatial ---- */
typedef struct {
    uint32_t M;
    real W, target;
    v3 rad, pos, nrm;
} owr;

static owr o_diff_load(const uint32_t* rsv, const uint32_t* hit, const real* rad, size_t j, v3 pp, v3 pn) {
    owr r;
    float Wf;
    o_rsv_dec(rsv[j], &r.M, &Wf);
    r.W = Wf;
    o_hit_dec(hit[2 * j], hit[2 * j + 1], &r.pos, &r.nrm);
    r.rad = ld3r(rad, j);
    r.target = sat(dot3(pn, norm3(sub3(r.pos, pp)))) * lum(r.rad);
    return r;
}

void oracle_diff_spatial(const float* fpf, int w, int h, const uint32_t* gb, const real* depth, const real* ao,
                         const uint32_t* t_rsv, const uint32_t* t_hit, const real* t_rad, uint32_t frame_rand,
                         uint32_t* o_rsv, uint32_t* o_hit, real* o_rad, real* o_light, int num_taps, real radius_px) {
    /* MAX_ITERATION (ind_diff_spatial_resample.hlsl) and the initial radius
     * INIT_RADIUS_FACTOR * min(w, h) unless set */
    const int N = num_taps > 0 ? num_taps : 9;
    const real radius0 = radius_px > RC(0.0) ? radius_px : (real)(w < h ? w : h) * RC(0.05);
    ofp fp;
    o_load_fp(fpf, &fp);
    size_t plane = (size_t)w * h;
    /* spiral constants (ind_diff_spatial_resample.hlsl:148-151, 186), folded in double */
    const double PI_ = (double)(float)3.14159265359;
    const float Bf = (float)(log(1.61803398) / (0.5 * PI_));
    const float Af = (float)(1.0 / exp((double)Bf * (double)(2.0f * (float)PI_)));
    real theta[16], rr[16];
    for (int k = 0; k < N && k < 16; ++k) {
        theta[k] = O_TWO_PI * (((real)k + RC(1.0)) / (real)N);
#if defined(ORACLE_F32)
        rr[k] = (float)((double)Af * exp((double)Bf * (double)theta[k]));
#else
        rr[k] = (double)Af * exp((double)Bf * theta[k]);
#endif
    }
    const real T = (real)(float)(0.6186 * (double)N);
    const real cos_tol = (real)(float)cos((double)(float)(PI_ * (25.0 / 180.0)));
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real d = depth[i];
            if (d == RC(0.0)) {
                o_rsv[i] = 0;
                o_hit[2 * i] = o_hit[2 * i + 1] = 0;
                st3r(o_rad, i, V3s(0));
                st3r(o_light, i, V3s(0));
                continue;
            }
            v3 p = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, d);
            v3 n = o_gb_normal(gb[plane + i]);
            real a = ao[2 * i];
            owr res = o_diff_load(t_rsv, t_hit, t_rad, i, p, n);
            real ws = (res.W * res.target) * (real)res.M;
            real radius = R_FMAX(radius0 * a, RC(1.5));
            uint32_t rng = o_lcg_init(px, py, frame_rand);
            real ra = o_rand(&rng) * O_TWO_PI;
            for (int k = 0; k < N; ++k) {
                real s, c;
                r_sincos(theta[k] * T + ra, &s, &c);
                real x = rr[k] * c, y = rr[k] * s;
                int ox = (int)(x * radius), oy = (int)(y * radius);
                int cx = clampi(px + ox, 0, w - 1), cy = clampi(py + oy, 0, h - 1);
                size_t j = (size_t)cy * w + cx;
                v3 nn = o_gb_normal(gb[plane + j]);
                real dn = depth[j];
                int diff = dot3(n, nn) < cos_tol;
                v3 pn = o_depth_to_pos(fp.inv_view_proj, cx, cy, w, h, dn);
                diff = diff || R_FABS(dot3(sub3(pn, p), n)) > RC(0.05);
                if (diff) {
                    radius = R_FMAX(radius * RC(0.5), RC(1.5));
                    continue;
                }
                owr nb = o_diff_load(t_rsv, t_hit, t_rad, j, p, n);
                v3 oqq = sub3(pn, nb.pos), orq = sub3(p, nb.pos);
                real L2qq = dot3(oqq, oqq), L2rq = dot3(orq, orq);
                real crq = dot3(nb.nrm, scale3(orq, RC(1.0) / R_SQRT(L2rq)));
                real cqq = dot3(nb.nrm, scale3(oqq, RC(1.0) / R_SQRT(L2qq)));
                real jac = (crq / cqq) * (L2qq / L2rq);
                jac = R_FMIN(R_FMAX(jac, RC(0.0)), RC(10.0));
                jac = R_FMIN(R_FMAX(jac, RC(0.0)), RC(1.0));
                real tn = nb.target * jac;
                if (tn <= RC(0.0)) continue;
                uint32_t nM = nb.M < 255 ? nb.M : 255;
                real wn = (nb.W * tn) * (real)nM;
                ws += wn;
                real chance = wn / ws;
                if (o_rand(&rng) < chance) {
                    res.nrm = nb.nrm;
                    res.pos = nb.pos;
                    res.rad = nb.rad;
                    res.target = nb.target;
                }
                res.M += nM;
            }
            if (res.target > RC(0.0)) res.W = ws / (res.target * (real)res.M);
            o_rsv[i] = o_rsv_enc(res.M, (float)res.W);
            o_hit_enc(res.pos, res.nrm, &o_hit[2 * i], &o_hit[2 * i + 1]);
            st3r(o_rad, i, res.rad);
            real NoL = sat(dot3(n, norm3(sub3(res.pos, p))));
            st3r(o_light, i, scale3(res.rad, (O_ONE_OVER_PI * NoL) * res.W));
        }
}

/* ---- specular ---- */
typedef struct {
    real src_pdf;
    v3 f;
} ospec;

static ospec o_calc_things(v3 view_pos, v3 p, v3 n, real pr, v3 f0, v3 hit) {
    ospec r;
    real a = R_FMAX(pr * pr, RC(0.002));
    v3 vd = norm3(sub3(view_pos, p)), ld = norm3(sub3(hit, p)), H = norm3(add3(vd, ld));
    real NoV = sat(dot3(n, vd)), NoH = sat(dot3(n, H));
    real D = o_D_ggx(NoH, a);
    r.src_pdf = (RC(0.25) * D) / o_lambda_pomc(NoV, a);
    real NoL = sat(dot3(n, ld)), LoH = sat(dot3(ld, H));
    v3 F = o_F_schlick(LoH, f0);
    real g = o_g2_over_g1(NoL, NoV, a);
    if (NoL <= RC(0.0)) g = RC(0.0);
    if (NoV <= RC(0.0)) g = RC(0.0);
    r.f = scale3(F, g);
    return r;
}
static inline real o_spec_target(v3 view_pos, v3 p, v3 n, real pr, v3 f0, v3 hit, v3 rad) {
    ospec s = o_calc_things(view_pos, p, n, pr, f0, hit);
    return lum(mul3(rad, s.f)) * s.src_pdf;
}

void oracle_spec_temporal(const float* fpf, int w, int h, const uint32_t* prev_gb, const real* prev_depth,
                          const uint32_t* gb, const real* depth, const uint32_t* prev_rsv, const real* prev_hitpos,
                          const real* prev_rad, uint32_t frame_rand, int has_prev, uint32_t* rw_rsv, real* rw_hitpos,
                          real* rw_rad, uint32_t m_cap) {
    /* M_MAX (ind_spec_temporal_resample.hlsl:252) unless set */
    const real Mcap = (real)(m_cap ? m_cap : 16u);
    ofp fp;
    o_load_fp(fpf, &fp);
    size_t plane = (size_t)w * h;
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real d = depth[i];
            if (d == RC(0.0)) {
                rw_rsv[2 * i] = rw_rsv[2 * i + 1] = 0;
                continue;
            }
            v3 p = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, d);
            v4 hr = o_mul44(fp.prev_view_proj, V4(p.x, p.y, p.z, RC(1.0)));
            real nx = hr.x / hr.w - fp.jitter[2], ny = hr.y / hr.w - fp.jitter[3], nz = hr.z / hr.w - RC(0.0);
            real u = nx * RC(0.5) + RC(0.5), v = ny * RC(0.5) + RC(0.5);
            int sp_ = (u == sat(u)) && (v == sat(v)) && has_prev;
            if (sp_ && R_FABS(bil1(prev_depth, w, h, u, v) - nz) > RC(0.001)) sp_ = 0;
            uint32_t M = 0;
            real W = 0, tf = 0;
            v3 rpos = V3s(0), rrad = V3s(0);
            v3 pg_col = V3s(0), pg_n = V3s(0);
            real pg_metal = 0, pg_rough = 0;
            if (sp_) {
                int ppx = (int)r2u(u * (real)w), ppy = (int)r2u(v * (real)h);
                if (ppx > w - 1) ppx = w - 1;
                if (ppy > h - 1) ppy = h - 1;
                size_t j = (size_t)ppy * w + ppx;
                M = prev_rsv[2 * j];
                W = bits_f(prev_rsv[2 * j + 1]);
                const real* hp = prev_hitpos + 4 * j;
                rpos = V3(hp[0], hp[1], hp[2]);
                tf = hp[3];
                rrad = ld3r(prev_rad, j);
                pg_col = o_gb_color(prev_gb[j]);
                pg_metal = o_gb_metal(prev_gb[j]);
                pg_n = o_gb_normal(prev_gb[plane + j]);
                pg_rough = o_gb_rough(prev_gb[plane + j]);
            }
            const real* nh = rw_hitpos + 4 * i;
            v3 npos = V3(nh[0], nh[1], nh[2]);
            v3 nrad = ld3r(rw_rad, i);
            v3 col = o_gb_color(gb[i]);
            real metal = o_gb_metal(gb[i]);
            v3 n = o_gb_normal(gb[plane + i]);
            real pr = o_gb_rough(gb[plane + i]);
            uint32_t Mmax = r2u(Mcap * pr);
            if (M > Mmax) M = Mmax;
            v3 f0 = o_f0(col, metal);
            ospec cur = o_calc_things(o_xyz(fp.view_pos), p, n, pr, f0, npos);
            real nw = lum(mul3(nrad, cur.f));
            real ntf = nw * cur.src_pdf;
            real pw = W * (tf * (real)M);
            if (M > 0) {
                real ptc = o_spec_target(o_xyz(fp.prev_view_pos), p, pg_n, pg_rough, o_f0(pg_col, pg_metal), npos, nrad);
                real cm = ntf / (ntf + ptc * (real)M);
                if (ntf == RC(0.0)) cm = RC(0.0);
                nw = cm * nw;
                real ctp = o_spec_target(o_xyz(fp.view_pos), p, n, pr, f0, rpos, rrad);
                real mn = tf * (real)M;
                real pm = mn / (mn + ctp);
                if (mn == RC(0.0)) pm = RC(0.0);
                pw = pm * (ctp * W);
                tf = ctp;
            }
            uint32_t rng = o_lcg_init(px, py, frame_rand);
            real ws = pw + nw;
            real cp = pw / ws;
            int use_prev = o_rand(&rng) < cp;
            if (!use_prev) {
                rpos = npos;
                rrad = nrad;
                tf = ntf;
            }
            M += 1;
            W = tf > RC(0.0) ? ws / tf : RC(0.0);
            rw_rsv[2 * i] = M;
            rw_rsv[2 * i + 1] = f_bits((float)W);
            real* o = rw_hitpos + 4 * i;
            o[0] = rpos.x, o[1] = rpos.y, o[2] = rpos.z, o[3] = tf;
            if (use_prev) st3r(rw_rad, i, rrad);
        }
}

void oracle_spec_spatial(const float* fpf, int w, int h, const uint32_t* gb, const real* depth, const uint32_t* rsv,
                         const real* hitpos, const real* rad, uint32_t frame_rand, real* o_light, real* o_raylen,
                         real radius_px) {
    /* INIT_RADIUS_FACTOR * min(w, h) (ind_spec_spatial_resample.hlsl:13,213) unless set */
    const real radius0 = radius_px > RC(0.0) ? radius_px : (real)(w < h ? w : h) * RC(0.05);
    ofp fp;
    o_load_fp(fpf, &fp);
    size_t plane = (size_t)w * h;
    const double PI_ = (double)(float)3.14159265359;
    const real cos_tol = (real)(float)cos((double)(float)(PI_ * (25.0 / 180.0)));
    v3 vp = o_xyz(fp.view_pos);
#pragma omp parallel for schedule(dynamic, 2)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real d = depth[i];
            if (d == RC(0.0)) {
                st3r(o_light, i, V3s(0));
                o_raylen[i] = 0;
                continue;
            }
            v3 p = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, d);
            v3 col = o_gb_color(gb[i]);
            real metal = o_gb_metal(gb[i]);
            v3 n = o_gb_normal(gb[plane + i]);
            real pr = o_gb_rough(gb[plane + i]);
            uint32_t M = rsv[2 * i];
            real W = bits_f(rsv[2 * i + 1]);
            const real* hp = hitpos + 4 * i;
            v3 hpos = V3(hp[0], hp[1], hp[2]);
            v3 hrad = ld3r(rad, i);
            v3 f0 = o_f0(col, metal);
            real tf = hp[3];
            uint32_t iters = r2u(RC(0.0) + pr * (RC(4.0) - RC(0.0)));
            real radius = radius0;
            uint32_t rng = o_lcg_init(px, py, frame_rand);
            for (uint32_t k = 0; k < iters; ++k) {
                real ux = o_rand(&rng), uy = o_rand(&rng);
                real r = R_SQRT(ux), s, c;
                r_sincos(O_TWO_PI * uy, &s, &c);
                real x = r * c, y = r * s;
                int ox = (int)(x * radius), oy = (int)(y * radius);
                int cx = clampi(px + ox, 0, w - 1), cy = clampi(py + oy, 0, h - 1);
                if (cx == px && cy == py) continue;
                size_t j = (size_t)cy * w + cx;
                v3 ncol = o_gb_color(gb[j]);
                real nmetal = o_gb_metal(gb[j]);
                v3 nn = o_gb_normal(gb[plane + j]);
                real npr = o_gb_rough(gb[plane + j]);
                real dn = depth[j];
                int diff = dot3(n, nn) < cos_tol;
                v3 pn = o_depth_to_pos(fp.inv_view_proj, cx, cy, w, h, dn);
                diff = diff || R_FABS(dot3(sub3(pn, p), n)) > RC(0.05);
                if (diff) {
                    radius = R_FMAX(radius * RC(0.5), RC(1.5));
                    continue;
                }
                uint32_t Mn = rsv[2 * j];
                real Wn = bits_f(rsv[2 * j + 1]);
                const real* hpn = hitpos + 4 * j;
                v3 hposn = V3(hpn[0], hpn[1], hpn[2]);
                v3 hradn = ld3r(rad, j);
                real tfn = hpn[3];
                real c_on_n = o_spec_target(vp, p, n, pr, f0, hposn, hradn);
                real mnb = (tfn * (real)Mn) / (tfn * (real)Mn + c_on_n * (real)M);
                if (tfn * (real)Mn == RC(0.0)) mnb = RC(0.0);
                real wnb = (mnb * tfn) * Wn;
                real n_on_c = o_spec_target(vp, pn, nn, npr, o_f0(ncol, nmetal), hpos, hrad);
                real mc = (tf * (real)M) / (tf * (real)M + n_on_c * (real)Mn);
                if (tf * (real)M == RC(0.0)) mc = RC(0.0);
                real wc = (mc * tf) * W;
                real ws = wc + wnb;
                real chance = wnb / ws;
                if (o_rand(&rng) < chance) {
                    hpos = hposn;
                    hrad = hradn;
                    tf = tfn;
                }
                M += Mn;
                W = tf > RC(0.0) ? ws / tf : RC(0.0);
            }
            o_raylen[i] = len3(sub3(hpos, p));
            real a = R_FMAX(pr * pr, RC(0.002));
            v3 sel = norm3(sub3(hpos, p)), vd = norm3(sub3(vp, p)), H = norm3(add3(vd, sel));
            real NoL = sat(dot3(n, sel)), NoV = sat(dot3(n, vd)), NoH = sat(dot3(n, H)), LoH = sat(dot3(sel, H));
            v3 F = o_F_schlick(LoH, f0);
            real D = o_D_ggx(NoH, a);
            real vndf = (RC(0.25) * D) / o_lambda_pomc(NoV, a);
            real g = o_g2_over_g1(NoL, NoV, a);
            if (NoL <= RC(0.0)) g = RC(0.0);
            if (NoV <= RC(0.0)) g = RC(0.0);
            v3 L = mul3(hrad, scale3(scale3(F, g), vndf * W));
            v3 refl = o_ggx_integral(NoV, a, f0);
            refl = V3(R_FMAX(refl.x, RC(1.0) / RC(65536.0)), R_FMAX(refl.y, RC(1.0) / RC(65536.0)),
                      R_FMAX(refl.z, RC(1.0) / RC(65536.0)));
            st3r(o_light, i, div3(L, refl));
        }
}

/* ---- primary-visibility G-buffer (the producer that stands in for the
 * raster pass gbuffer_pass.rs:45-137 / mesh_gbuffer.hlsl:62-126) ---- */
static inline uint32_t enc_unorm(real x, real s) { return r2u(x * s); }

static v3 o_primary_dir(const ofp* fp, v3 o, int px, int py, int w, int h) {
    real ndx = ((real)px + RC(0.5)) / (real)w * RC(2.0) - RC(1.0);
    real ndy = ((real)py + RC(0.5)) / (real)h * RC(2.0) - RC(1.0);
    v4 e = o_mul44(fp->inv_view_proj, V4(ndx, ndy, RC(1.0), RC(1.0)));
    return norm3(sub3(V3(e.x / e.w, e.y / e.w, e.z / e.w), o));
}

void oracle_gbuffer(const float* fpf, const oscene* s, int w, int h, real* o_depth, uint32_t* o_gb) {
    ofp fp;
    o_load_fp(fpf, &fp);
    size_t plane = (size_t)w * h;
    real* nrm = (real*)calloc(3 * plane, sizeof(real));
    uint8_t* geo = (uint8_t*)calloc(plane, 1);
#pragma omp parallel for schedule(dynamic, 16)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            v3 o = o_xyz(fp.view_pos);
            v3 d = o_primary_dir(&fp, o, px, py, w, h);
            ohit hh = obvh_trace(s->bvh, o, d, RC(0.05), RC(100000.0), 0);
            o_depth[i] = 0;
            o_gb[i] = o_gb[plane + i] = o_gb[2 * plane + i] = o_gb[3 * plane + i] = 0;
            if (hh.id == 0xFFFFFFFFu) continue;
            ofoot ft;
            ft.o = o;
            ft.dir_x = o_primary_dir(&fp, o, px ^ 1, py, w, h);
            ft.dir_y = o_primary_dir(&fp, o, px, py ^ 1, w, h);
            ft.x_odd = px & 1, ft.y_odd = py & 1;
            osurf sf = o_surface(s, hh, d, &ft);
            v3 pos = sf.pos, n = sf.n;
            v4 c = o_mul44(fp.view_proj, V4(pos.x, pos.y, pos.z, RC(1.0)));
            real dep = c.z / c.w;
            if (!(dep > RC(0.0))) continue;
            o_depth[i] = dep;
            geo[i] = 1;
            nrm[3 * i] = n.x, nrm[3 * i + 1] = n.y, nrm[3 * i + 2] = n.z;
            o_gb[i] = enc_unorm(sf.base.x, RC(255.0)) | (enc_unorm(sf.base.y, RC(255.0)) << 8) |
                      (enc_unorm(sf.base.z, RC(255.0)) << 16) | (enc_unorm(sf.metal, RC(255.0)) << 24);
            v2 oe = o_oct_enc(n);
            o_gb[plane + i] = enc_unorm(oe.x, RC(4095.0)) | (enc_unorm(oe.y, RC(4095.0)) << 12) |
                              (enc_unorm(sf.rough, RC(255.0)) << 24);
            o_gb[3 * plane + i] = 1;
        }
    /* fwidth from fine differences inside aligned 2x2 quads */
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            if (!geo[i]) continue;
            int qx = px ^ 1, qy = py ^ 1;
            int okx = qx < w && geo[(size_t)py * w + qx];
            int oky = qy < h && geo[(size_t)qy * w + px];
            size_t jx = (size_t)py * w + (qx < w ? qx : px), jy = (size_t)(qy < h ? qy : py) * w + px;
            real fz = (okx ? R_FABS(o_depth[jx] - o_depth[i]) : RC(0.0)) + (oky ? R_FABS(o_depth[jy] - o_depth[i]) : RC(0.0));
            v3 ni = V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
            v3 ax = okx ? abs3(sub3(V3(nrm[3 * jx], nrm[3 * jx + 1], nrm[3 * jx + 2]), ni)) : V3s(0);
            v3 ay = oky ? abs3(sub3(V3(nrm[3 * jy], nrm[3 * jy + 1], nrm[3 * jy + 2]), ni)) : V3s(0);
            real fn = len3(add3(ax, ay));
            o_gb[2 * plane + i] = o_double_to_half((double)(float)fz) | (o_double_to_half((double)(float)fn) << 16);
        }
    free(nrm);
    free(geo);
}

/* ---- hash grid maintenance ---- */
void oracle_hg_decay(uint8_t* decay, uint32_t* cells) {
    for (uint32_t b = 0; b < 65536u; ++b) {
        uint32_t off = b * 16u, occ = 0, set = 0, slice[4] = {0, 0, 0, 0};
        while (set < 4) {
            uint32_t base = off + set * 4, packed = 0;
            memcpy(&packed, decay + base, 4);
            if (packed == 0) break;
            uint32_t upd = 0, dv = 0;
            for (uint32_t i = 0; i < 4; ++i) {
                dv = (packed >> (8 * i)) & 0xFFu;
                if (dv == 0) break;
                dv -= 1;
                if (dv == 0) {
                    memset(cells + 4 * (size_t)(base + i), 0, 16);
                } else {
                    upd |= dv << (8 * i);
                    occ |= 1u << (set * 4 + i);
                }
            }
            slice[set++] = upd;
            if (dv == 0) break;
        }
        while (occ) {
            uint32_t last = 31u - (uint32_t)__builtin_clz(occ);
            uint32_t fs = (uint32_t)__builtin_ctz(~occ);
            if (!(fs < last)) break;
            uint32_t dv = slice[last >> 2] >> ((last & 3) * 8);
            slice[last >> 2] ^= dv << ((last & 3) * 8);
            slice[fs >> 2] |= dv << ((fs & 3) * 8);
            memcpy(cells + 4 * (size_t)(off + fs), cells + 4 * (size_t)(off + last), 16);
            memset(cells + 4 * (size_t)(off + last), 0, 16);
            occ |= 1u << fs;
            occ ^= 1u << last;
        }
        for (uint32_t i = 0; i < set; ++i) memcpy(decay + off + 4 * i, &slice[i], 4);
    }
}

/* Sequential restatement: queries processed in index order (one valid
 * schedule of the reference's concurrent CAS inserts / atomicMax). */
void oracle_hg_select(const uint32_t* queries, uint32_t counter, uint32_t* cells, uint32_t* selection,
                      uint32_t frame_hash) {
    uint32_t n = counter < HG_MAX_QUERIES ? counter : HG_MAX_QUERIES;
    for (uint32_t id = 0; id < n; ++id) {
        uint32_t hash = queries[6 * (size_t)id + 4], ck = queries[6 * (size_t)id + 5];
        int found = 0;
        uint32_t addr = 0;
        for (uint32_t a = hash << 4; a < (hash << 4) + 16; ++a) {
            uint32_t slot = cells[4 * (size_t)a];
            if (slot == ck) { found = 1; addr = a; break; }
            if (slot == 0) { cells[4 * (size_t)a] = ck; found = 1; addr = a; break; }
        }
        if (!found) continue;
        uint32_t st = o_jenkins(id ^ frame_hash);
        uint32_t r = o_xorshift(&st);
        uint32_t word = (r & (0xFFFFFFFFu << 18)) | (1u << 31) | id;
        if (word > selection[addr]) selection[addr] = word;
    }
}

uint32_t oracle_hg_compact(const uint32_t* selection, uint32_t* out_index, uint32_t* out_addr) {
    uint32_t n = 0;
    for (uint32_t c = 0; c < 65536u * 16u; ++c)
        if (selection[c] != 0) {
            out_index[n] = selection[c] & 0x1FFFFu;
            out_addr[n] = c;
            ++n;
        }
    return n;
}

/* Order-independent cache update (the library's *_ordered passes; no
 * reference counterpart).  The keys that the reference's probe would not
 * find in front of their bucket's first free slot go in one at a time, in
 * ascending checksum order, through find-or-insert (the reference's result
 * for the schedule that runs them in that order).  Each query draws a
 * selection word from its 24 bytes (the reference's jenkins + xorshift draw
 * seeded by a hash of the contents instead of the index, bit 31 set); a cell
 * keeps the largest word and the index of a query that drew it. */
static uint32_t o_query_select_word(const uint32_t* q, uint32_t frame_hash) {
    uint32_t st = o_jenkins(o_query_hash(q) ^ frame_hash);
    return o_xorshift(&st) | 0x80000000u;
}

/* First slot of the bucket holding `checksum`, or -1. */
static int64_t o_key_cell(const uint32_t* cells, uint32_t hash, uint32_t checksum) {
    for (uint32_t a = hash << 4; a < (hash << 4) + 16; ++a)
        if (cells[4 * (size_t)a] == checksum) return a;
    return -1;
}

static int o_cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}

/* The `n` queries kept under the word `thr` sample about n * 2^32 / (2^32 -
 * thr) candidates: the word that keeps the target of as many. */
uint32_t oracle_hg_next_query_threshold(uint32_t n, uint32_t thr) {
    uint64_t keep = (1ull << 32) - thr;
    return oracle_hg_query_threshold_for(((uint64_t)n << 32) / keep);
}

void oracle_hg_select_ordered(const uint32_t* queries, uint32_t counter, uint32_t* cells, uint32_t* selection,
                              uint32_t* winner, uint32_t* threshold, uint32_t frame_hash) {
    uint32_t n = counter < HG_MAX_QUERIES ? counter : HG_MAX_QUERIES;
    if (threshold) *threshold = oracle_hg_next_query_threshold(counter, *threshold);
    /* candidate keys per bucket (at most 16, as the library's table) */
    uint32_t* cand = (uint32_t*)calloc(65536u * 16u, 4);
    for (uint32_t id = 0; id < n; ++id) {
        uint32_t hash = queries[6 * (size_t)id + 4], ck = queries[6 * (size_t)id + 5], a = hash << 4;
        for (; a < (hash << 4) + 16; ++a) {
            uint32_t c = cells[4 * (size_t)a];
            if (c == ck || c == 0) break;
        }
        if (a == (hash << 4) + 16 || cells[4 * (size_t)a] == ck) continue;
        for (uint32_t t = hash << 4; t < (hash << 4) + 16; ++t) {
            if (cand[t] == ck) break;
            if (cand[t] == 0) {
                cand[t] = ck;
                break;
            }
        }
    }
    for (uint32_t b = 0; b < 65536u; ++b) {
        uint32_t* key = cand + 16 * (size_t)b;
        uint32_t m = 0;
        while (m < 16 && key[m]) ++m;
        qsort(key, m, 4, o_cmp_u32);
        for (uint32_t i = 0; i < m; ++i)
            for (uint32_t a = b * 16; a < b * 16 + 16; ++a) {
                uint32_t* c = cells + 4 * (size_t)a;
                if (c[0] == key[i]) break;
                if (c[0] == 0) {
                    c[0] = key[i];
                    break;
                }
            }
    }
    free(cand);
    for (uint32_t id = 0; id < n; ++id) {
        const uint32_t* q = queries + 6 * (size_t)id;
        int64_t a = o_key_cell(cells, q[4], q[5]);
        if (a < 0) continue;
        uint32_t w = o_query_select_word(q, frame_hash);
        if (w > selection[a]) selection[a] = w;
    }
    for (uint32_t id = 0; id < n; ++id) {
        const uint32_t* q = queries + 6 * (size_t)id;
        int64_t a = o_key_cell(cells, q[4], q[5]);
        if (a >= 0 && selection[a] == o_query_select_word(q, frame_hash)) winner[a] = id;
    }
}

uint32_t oracle_hg_compact_ordered(const uint32_t* selection, const uint32_t* winner, uint32_t* out_index,
                                   uint32_t* out_addr) {
    uint32_t n = 0;
    for (uint32_t c = 0; c < 65536u * 16u; ++c)
        if (selection[c] >> 31) {
            if (n < HG_MAX_QUERIES) {
                out_index[n] = winner[c] & (HG_MAX_QUERIES - 1u);
                out_addr[n] = c;
            }
            ++n;
        }
    return n;
}

/* One cosine ray per compacted query, in compacted order; the cache lookup
 * of the trace reads rt->hg_cells (`cells` as updated so far when they are
 * the same buffer: one sequential schedule of the reference's pass; a copy
 * of the cache from before the pass: the ordered pass). */
void oracle_hg_raygen(const float* fpf, const otrace_res* rt, uint32_t* cells, const uint32_t* queries, uint32_t count,
                      const uint32_t* cidx, const uint32_t* caddr, uint8_t* decay, uint32_t frame_hash) {
    ofp fp;
    o_load_fp(fpf, &fp);
    for (uint32_t id = 0; id < count && id < HG_MAX_QUERIES; ++id) {
        const uint32_t* q = queries + 6 * (size_t)cidx[id];
        uint32_t cell = caddr[id];
        uint32_t rng = o_jenkins(id ^ frame_hash);
        v3 n = o_oct_dec_u32(q[3]);
        real u0 = o_rand(&rng), u1 = o_rand(&rng);
        real st_, sc;
        r_sincos(O_TWO_PI * u1, &st_, &sc);
        real sin_t = R_SQRT(u0), cos_t = R_SQRT(RC(1.0) - u0);
        v3 l = V3(sin_t * sc, sin_t * st_, cos_t);
        v3 dir = o_rotate(o_inv_rot(o_rot_to_z(n)), l);
        otrace_out t = o_trace_radiance(rt, &fp, V3(bits_f(q[0]), bits_f(q[1]), bits_f(q[2])), dir, 1);
        uint32_t* c = cells + 4 * (size_t)cell;
        const real bf = RC(1.0) / RC(16.0);
        v3 prev = V3(bits_f(c[1]), bits_f(c[2]), bits_f(c[3]));
        v3 b = add3(scale3(t.radiance, bf), scale3(prev, RC(1.0) - bf));
        c[1] = f_bits((float)b.x);
        c[2] = f_bits((float)b.y);
        c[3] = f_bits((float)b.z);
        decay[cell] = 10;
    }
}

/* raytraced_shadow.hlsl:9-53: one any-hit ray toward the sun per pixel with
 * geometry; 255 (R8_UNORM 1.0) where it misses, 0 where it hits.  Pixels
 * without geometry hold 0 (the reference leaves them unwritten; the final
 * lighting never reads them). */
void oracle_raytraced_shadow(const float* fpf, const oscene* s, int w, int h, const real* depth, uint8_t* o_mask) {
    ofp fp;
    o_load_fp(fpf, &fp);
#pragma omp parallel for schedule(dynamic, 16)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real d = depth[i];
            if (d == RC(0.0)) {
                o_mask[i] = 0;
                continue;
            }
            v3 pos = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, d + DEPTH_ERR);
            ohit sh = obvh_trace(s->bvh, pos, o_xyz(fp.sun_dir), RC(0.0005), RC(1000.0), 1);
            o_mask[i] = sh.id == 0xFFFFFFFFu ? 255 : 0;
        }
}

/* restir/final_lighting.hlsl:16-62 (DIRECT_SPECULAR_SUPRESSION 1,
 * DEMODULATE_INDIRECT_SPECULAR_FOR_DENOISER 1, IND_SPEC_R_MIN 0.002,
 * integral floor 1/65536: restir/config.inc.hlsl:5-10).  Pixels without
 * geometry take the constant sky radiance (the sky pass's colour). */
void oracle_final_lighting(const float* fpf, int w, int h, const uint32_t* gb, const uint8_t* mask,
                           const real* ind_diff, const real* ind_spec, const float* sky, real* o_color) {
    ofp fp;
    o_load_fp(fpf, &fp);
    size_t plane = (size_t)w * h;
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            if (gb[3 * plane + i] == 0u) {
                st3r(o_color, i, V3(sky[0], sky[1], sky[2]));
                continue;
            }
            v3 col = o_gb_color(gb[i]);
            real metal = o_gb_metal(gb[i]);
            v3 n = o_gb_normal(gb[plane + i]);
            real perc = o_gb_rough(gb[plane + i]);
            real atten = (real)mask[i] / RC(255.0);
            /* cs_view_ray_direction with zero jitter: the pixel corner */
            real ndx = (((real)px + RC(0.5)) + RC(-0.5)) / (real)w * RC(2.0) - RC(1.0);
            real ndy = (((real)py + RC(0.5)) + RC(-0.5)) / (real)h * RC(2.0) - RC(1.0);
            v4 e = o_mul44(fp.inv_view_proj, V4(ndx, ndy, RC(1.0), RC(1.0)));
            v3 vd = neg3(norm3(sub3(V3(e.x / e.w, e.y / e.w, e.z / e.w), o_xyz(fp.view_pos))));
            real pr = R_FMAX(perc, RC(0.045));
            v3 rho = scale3(col, RC(1.0) - metal), f0 = o_f0(col, metal);
            v3 sun = o_xyz(fp.sun_dir);
            real NoL = sat(dot3(n, sun));
            v3 direct = mul3(scale3(o_eval_ggx_lambert(vd, sun, n, pr, rho, f0), NoL * atten), o_xyz(fp.sun_inten));
            real NoV = sat(dot3(n, vd));
            v3 ri = o_ggx_integral(NoV, R_FMAX(perc * perc, RC(0.002)), f0);
            real rmin = RC(1.0) / RC(65536.0);
            v3 sr = V3(R_FMAX(ri.x, rmin), R_FMAX(ri.y, rmin), R_FMAX(ri.z, rmin));
            v3 ind = add3(mul3(ld3r(ind_diff, i), rho), mul3(ld3r(ind_spec, i), sr));
            st3r(o_color, i, add3(direct, ind));
        }
}

/* temporal_aa.hlsl:17-29 */
static real o_mitchell(real x) {
    const real B = RC(1.0) / RC(3.0), C = RC(1.0) / RC(3.0);
    real y = RC(0.0), x2 = x * x, x3 = x * x * x;
    if (x < RC(1.0))
        y = (RC(12.0) - RC(9.0) * B - RC(6.0) * C) * x3 + (RC(-18.0) + RC(12.0) * B + RC(6.0) * C) * x2 +
            (RC(6.0) - RC(2.0) * B);
    else if (x <= RC(2.0))
        y = (-B - RC(6.0) * C) * x3 + (RC(6.0) * B + RC(30.0) * C) * x2 + (RC(-12.0) * B - RC(48.0) * C) * x +
            (RC(8.0) * B + RC(24.0) * C);
    return y / RC(6.0);
}

/* sampler_linear_clamp: texel = uv*size - 0.5, clamp-to-edge */
static v3 o_bilinear3r(const real* tex, int w, int h, real u, real v) {
    real tx = u * (real)w - RC(0.5), ty = v * (real)h - RC(0.5);
    real x0f = R_FLOOR(tx), y0f = R_FLOOR(ty);
    real fx = tx - x0f, fy = ty - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    v3 top = add3(scale3(ld3r(tex, (size_t)ya * w + xa), RC(1.0) - fx), scale3(ld3r(tex, (size_t)ya * w + xb), fx));
    v3 bot = add3(scale3(ld3r(tex, (size_t)yb * w + xa), RC(1.0) - fx), scale3(ld3r(tex, (size_t)yb * w + xb), fx));
    return add3(scale3(top, RC(1.0) - fy), scale3(bot, fy));
}

/* NaN-safe saturate: fmin/fmax return the other operand for a NaN */
static inline real o_sat_nan0(real x) { return R_FMIN(R_FMAX(x, RC(0.0)), RC(1.0)); }

/* temporal_aa.hlsl:112-264.  The nearest-depth neighbour the reference
 * tracks is overwritten before use (:186-187), so it is not computed. */
void oracle_temporal_aa(const float* fpf, int w, int h, const real* depth, const real* source, const real* history,
                        int has_history, real* o_target) {
    ofp fp;
    o_load_fp(fpf, &fp);
    real sox = -fp.jitter[0] * RC(0.5) * (real)w, soy = -fp.jitter[1] * RC(0.5) * (real)h;
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            v3 accu = V3s(0), nmin = V3s(RC(1e5)), nmax = V3s(RC(-1e5)), m1 = V3s(0), m2 = V3s(0);
            real wsum = RC(0.0);
            for (int x = -1; x <= 1; ++x)
                for (int y = -1; y <= 1; ++y) {
                    int cx = clampi(px + x, 0, w - 1), cy = clampi(py + y, 0, h - 1);
                    real ox = (real)(cx - px) + sox, oy = (real)(cy - py) + soy;
                    v3 nb = ld3r(source, (size_t)cy * w + cx);
                    real wt = o_mitchell(R_SQRT(ox * ox + oy * oy));
                    accu = add3(accu, scale3(nb, wt));
                    wsum += wt;
                    nmin = V3(R_FMIN(nmin.x, nb.x), R_FMIN(nmin.y, nb.y), R_FMIN(nmin.z, nb.z));
                    nmax = V3(R_FMAX(nmax.x, nb.x), R_FMAX(nmax.y, nb.y), R_FMAX(nmax.z, nb.z));
                    m1 = add3(m1, nb);
                    m2 = add3(m2, mul3(nb, nb));
                }
            v3 src = V3(accu.x / wsum, accu.y / wsum, accu.z / wsum);
            v3 pos = o_depth_to_pos(fp.inv_view_proj, px, py, w, h, depth[i]);
            v4 hc = o_mul44(fp.view_proj, V4(pos.x, pos.y, pos.z, RC(1.0)));
            v4 hr = o_mul44(fp.prev_view_proj, V4(pos.x, pos.y, pos.z, RC(1.0)));
            real mvx = ((hc.x / hc.w - fp.jitter[0]) - (hr.x / hr.w - fp.jitter[2])) * RC(0.5);
            real mvy = ((hc.y / hc.w - fp.jitter[1]) - (hr.y / hr.w - fp.jitter[3])) * RC(0.5);
            real hu = ((real)px + RC(0.5)) / (real)w - mvx, hv = ((real)py + RC(0.5)) / (real)h - mvy;
            int out_of_view = hu != o_sat_nan0(hu) || hv != o_sat_nan0(hv);
            if (out_of_view || !has_history) {
                st3r(o_target, i, src);
                continue;
            }
            v3 hs = o_bilinear3r(history, w, h, hu, hv);
            hs = V3(R_FMIN(R_FMAX(hs.x, nmin.x), nmax.x), R_FMIN(R_FMAX(hs.y, nmin.y), nmax.y),
                    R_FMIN(R_FMAX(hs.z, nmin.z), nmax.z));
            real inv_n = RC(1.0) / RC(9.0);
            v3 m1u = scale3(m1, inv_n), m2u = scale3(m2, inv_n);
            v3 d2 = sub3(m2u, mul3(m1u, m1u));
            v3 e = V3(R_SQRT(R_FABS(d2.x)) + RC(1e-8), R_SQRT(R_FABS(d2.y)) + RC(1e-8), R_SQRT(R_FABS(d2.z)) + RC(1e-8));
            v3 v = sub3(hs, m1u);
            real ma = R_FMAX(R_FABS(v.x / e.x), R_FMAX(R_FABS(v.y / e.y), R_FABS(v.z / e.z)));
            if (ma > RC(1.0)) hs = add3(m1u, V3(v.x / ma, v.y / ma, v.z / ma));
            real sw = RC(1.0) / RC(32.0), hw = RC(1.0) - RC(1.0) / RC(32.0);
            real rs = RC(1.0) / (R_FMAX(R_FMAX(src.x, src.y), src.z) + RC(1.0));
            real rh = RC(1.0) / (R_FMAX(R_FMAX(hs.x, hs.y), hs.z) + RC(1.0));
            sw *= RC(1.0) / (RC(1.0) + lum(scale3(src, rs)));
            hw *= RC(1.0) / (RC(1.0) + lum(scale3(hs, rh)));
            real den = R_FMAX(sw + hw, RC(1e-5));
            v3 num = add3(scale3(src, sw), scale3(hs, hw));
            st3r(o_target, i, V3(num.x / den, num.y / den, num.z / den));
        }
}
