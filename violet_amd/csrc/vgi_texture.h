/*
 * vgi_texture.h — material texture sampling on the device, standing in for
 * the bindless Texture2D fetches of violet's hit and G-buffer shaders
 * (geometry_rchit.inc.hlsl:79-81 SampleLevel(sampler_linear_wrap, uv,
 * TEX_LOD_BIAS) with TEX_LOD_BIAS 2 for indirect rays,
 * geometry_ind.rchit.hlsl:3; mesh_gbuffer.hlsl:79-81 Sample with implicit
 * derivatives).  The sampler is violet's sampler_linear_wrap
 * (render_loop.rs:199-212): linear min/mag filter, linear between mips,
 * REPEAT addressing, no LOD clamp, anisotropy at the device maximum (16).
 *
 * Filtering follows the Vulkan specification's texel filtering equations in
 * float arithmetic (the hardware's fixed-point sub-texel weights are not
 * reproduced):
 *   level l size   w_l = max(1, w >> l), h_l = max(1, h >> l)
 *   bilinear       u' = s * w_l - 0.5, i0 = floor(u'), alpha = u' - i0,
 *                  i1 = i0 + 1, both wrapped mod w_l (likewise t);
 *                  lerp(lerp(t00, t10, alpha), lerp(t01, t11, alpha), beta)
 *   trilinear      d = clamp(lambda, 0, levels - 1); lerp(level floor(d),
 *                  level floor(d) + 1, frac(d))
 *   implicit LOD   rho_x = |(ds/dx * w, dt/dx * h)|, rho_y likewise;
 *                  eta = min(rho_max / rho_min, 16), N = ceil(eta),
 *                  lambda = log2(rho_max / eta); N trilinear samples at
 *                  uv + (i / (N + 1) - 1/2) * (duv along the major axis),
 *                  i = 1..N, averaged (the specification's reference form
 *                  of anisotropic filtering)
 * sRGB texels are decoded to linear before filtering, alpha is linear.
 */
#pragma once
#include "vgi_bvh.h"
#include "vgi_mathf.h"

namespace vgi {

constexpr float kMaxAnisotropy = 16.0f;
constexpr float kIndirectTexLod = 2.0f;  // geometry_ind.rchit.hlsl:3

struct Tex4 {
    float r, g, b, a;
};

__device__ __forceinline__ Tex4 tex_lerp(Tex4 x, Tex4 y, float t) {
    Tex4 o;
    o.r = x.r + (y.r - x.r) * t;
    o.g = x.g + (y.g - x.g) * t;
    o.b = x.b + (y.b - x.b) * t;
    o.a = x.a + (y.a - x.a) * t;
    return o;
}

// REPEAT: an integral coordinate (as float) reduced mod n.  Coordinates of
// 2^24 and beyond (and NaN) fold to 0.
__device__ __forceinline__ uint32_t tex_wrap(float f, uint32_t n) {
    const int i = fabsf(f) < 16777216.0f ? (int)f : 0;
    int r = i % (int)n;
    return (uint32_t)(r < 0 ? r + (int)n : r);
}

__device__ __forceinline__ Tex4 tex_decode(const SceneView& s, uint32_t word, uint32_t srgb) {
    const float* lut = s.texel_lut + (srgb ? 256u : 0u);
    Tex4 t;
    t.r = gld(lut + (word & 0xFFu));
    t.g = gld(lut + ((word >> 8) & 0xFFu));
    t.b = gld(lut + ((word >> 16) & 0xFFu));
    t.a = gld(s.texel_lut + (word >> 24));
    return t;
}

__device__ __forceinline__ Tex4 tex_bilinear(const SceneView& s, const TexDesc& t, uint32_t level, float u, float v) {
    const uint32_t wl = max(1u, t.width >> level), hl = max(1u, t.height >> level);
    const float fu = u * (float)wl - 0.5f, fv = v * (float)hl - 0.5f;
    const float x0f = floorf(fu), y0f = floorf(fv);
    const float a = fu - x0f, b = fv - y0f;
    const uint32_t x0 = tex_wrap(x0f, wl), y0 = tex_wrap(y0f, hl);
    const uint32_t x1 = x0 + 1u == wl ? 0u : x0 + 1u, y1 = y0 + 1u == hl ? 0u : y0 + 1u;
    const uint32_t* L = s.texels + t.level_offset[level];
    const Tex4 t00 = tex_decode(s, gld(L + (size_t)y0 * wl + x0), t.srgb);
    const Tex4 t10 = tex_decode(s, gld(L + (size_t)y0 * wl + x1), t.srgb);
    const Tex4 t01 = tex_decode(s, gld(L + (size_t)y1 * wl + x0), t.srgb);
    const Tex4 t11 = tex_decode(s, gld(L + (size_t)y1 * wl + x1), t.srgb);
    return tex_lerp(tex_lerp(t00, t10, a), tex_lerp(t01, t11, a), b);
}

__device__ __forceinline__ Tex4 tex_trilinear(const SceneView& s, const TexDesc& t, float lod, float u, float v) {
    const uint32_t last = t.levels - 1u;
    const float d = lod > 0.0f ? lod : 0.0f;
    if (d >= (float)last) return tex_bilinear(s, t, last, u, v);
    const float dl = floorf(d);
    const float frac = d - dl;
    const uint32_t l = (uint32_t)dl;
    const Tex4 lo = tex_bilinear(s, t, l, u, v);
    if (frac == 0.0f) return lo;
    return tex_lerp(lo, tex_bilinear(s, t, l + 1u, u, v), frac);
}

// Screen-space derivatives of the normalised texture coordinates.
struct TexGrad {
    float dsdx, dtdx, dsdy, dtdy;
};

__device__ __forceinline__ Tex4 tex_sample_grad(const SceneView& s, const TexDesc& t, float u, float v,
                                                const TexGrad& g) {
    const float W = (float)t.width, H = (float)t.height;
    const float ax = g.dsdx * W, bx = g.dtdx * H, ay = g.dsdy * W, by = g.dtdy * H;
    const float rx = sqrtf(ax * ax + bx * bx), ry = sqrtf(ay * ay + by * by);
    const float rmax = fmaxf(rx, ry), rmin = fminf(rx, ry);
    float lod = 0.0f;
    uint32_t n = 1;
    if (!(rmax > 0.0f)) {
        lod = 0.0f;  // constant coordinates (or NaN derivatives): the base level
    } else if (!(rmax <= 3.0e38f)) {
        lod = 1.0e30f;  // unbounded footprint: the last level
    } else {
        const float eta = rmin > 0.0f ? fminf(rmax / rmin, kMaxAnisotropy) : kMaxAnisotropy;
        n = (uint32_t)ceilf(eta);
        lod = vgi_log2f(rmax / eta);
    }
    if (n <= 1u) return tex_trilinear(s, t, lod, u, v);
    const bool along_x = rx > ry;
    const float du = along_x ? g.dsdx : g.dsdy, dv = along_x ? g.dtdx : g.dtdy;
    Tex4 sum{0.0f, 0.0f, 0.0f, 0.0f};
    const float np1 = (float)(n + 1u);
    for (uint32_t i = 1; i <= n; ++i) {
        const float o = (float)i / np1 - 0.5f;
        const Tex4 x = tex_trilinear(s, t, lod, u + du * o, v + dv * o);
        sum.r = sum.r + x.r;
        sum.g = sum.g + x.g;
        sum.b = sum.b + x.b;
        sum.a = sum.a + x.a;
    }
    const float fn = (float)n;
    return Tex4{sum.r / fn, sum.g / fn, sum.b / fn, sum.a / fn};
}

}  // namespace vgi
