/*
 * vgi_common.h — host+device building blocks of the ReSTIR GI hot path.
 *
 * Everything here is compiled both for gfx950 (HIP kernels) and for the
 * host (BVH builder, scene/G-buffer helpers).  Floating-point expressions
 * are written in one fixed evaluation order and the library is compiled
 * with -ffp-contract=off, so a value computed here is bit-identical on the
 * host and on the device, and the single-precision build of the oracle
 * (which restates the same reference formulas in the same order) agrees
 * with it bit for bit.
 *
 * Reference sources restated here (violet, HLSL):
 *   rand.hlsl:9-46,82-112       jenkins hash, LCG stream, XXHash32 variant
 *   enc.inc.hlsl:11-63          octahedral normal encodings
 *   gbuffer.hlsl:15-114         G-buffer layer decode
 *   restir/reservoir.hlsl:52-124 reservoir / hit encodings
 *   frame_bindings.hlsl:6-79    FrameParams and reconstruction helpers
 *   brdf.hlsl, sampling.hlsl    BRDF terms and hemisphere sampling
 *   restir/hash_grid.inc.hlsl   hash-grid keys and probing
 */
#pragma once
#include <type_traits>
#include <stdint.h>
#include "vgi_mathf.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VGI_HD __host__ __device__ __forceinline__
#else
#define VGI_HD inline
#endif

namespace vgi {

// ---------------------------------------------------------------------------
// Constants (constants.hlsl:3-8; the HLSL literals are float)
// ---------------------------------------------------------------------------
constexpr float kPi = 3.14159265359f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kOneOverPi = 1.0f / kPi;
constexpr float kF32SignificantlyLarge = 1e19f;
constexpr float kHistoryLenUnit = 1.0f / 255.0f;  // svgf.inc.hlsl:4

// ---------------------------------------------------------------------------
// Bit casts and scalar helpers
// ---------------------------------------------------------------------------
// Loads and stores of device images.  On the device they go through an
// explicit global-memory pointer: images often arrive as pointers read from
// memory (the fused multi-view kernels' job tables), where the compiler
// cannot infer the address space and would emit flat accesses.
// Structs (cells, queries, materials) move as 32-bit words: C++ copies of a
// class object cannot bind an address-space-qualified source.
template <class T>
VGI_HD T gld(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_scalar<T>::value) {
        return *(const __attribute__((address_space(1))) T*)p;
    } else {
        static_assert(sizeof(T) % 4 == 0, "word-wise load");
        T r;
        const __attribute__((address_space(1))) uint32_t* s = (const __attribute__((address_space(1))) uint32_t*)p;
        uint32_t* d = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
        for (size_t k = 0; k < sizeof(T) / 4; ++k) d[k] = s[k];
        return r;
    }
#else
    return *p;
#endif
}
template <class T>
VGI_HD void gst(T* p, T v) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_scalar<T>::value) {
        *(__attribute__((address_space(1))) T*)p = v;
    } else {
        static_assert(sizeof(T) % 4 == 0, "word-wise store");
        __attribute__((address_space(1))) uint32_t* d = (__attribute__((address_space(1))) uint32_t*)p;
        const uint32_t* s = reinterpret_cast<const uint32_t*>(&v);
#pragma unroll
        for (size_t k = 0; k < sizeof(T) / 4; ++k) d[k] = s[k];
    }
#else
    *p = v;
#endif
}

VGI_HD uint32_t asuint(float f) {
    union { float f; uint32_t u; } c;
    c.f = f;
    return c.u;
}
VGI_HD float asfloat(uint32_t u) {
    union { float f; uint32_t u; } c;
    c.u = u;
    return c.f;
}
VGI_HD float fmin_(float a, float b) { return fminf(a, b); }
VGI_HD float fmax_(float a, float b) { return fmaxf(a, b); }
VGI_HD float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
VGI_HD float lerp(float a, float b, float t) { return a + t * (b - a); }
VGI_HD float rcp(float x) { return 1.0f / x; }
VGI_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
VGI_HD uint32_t minu(uint32_t a, uint32_t b) { return a < b ? a : b; }

// HLSL uint(x)/int(x) conversions truncate toward zero; inputs are kept in
// range by the callers (negative or NaN inputs saturate to 0 here so that the
// host build has no undefined behaviour).
VGI_HD uint32_t f2u(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967040.0f) return 0xFFFFFF00u;
    return (uint32_t)x;
}

// ---------------------------------------------------------------------------
// Small vectors.  Component-wise operators; reductions in x, y, z order.
// ---------------------------------------------------------------------------
struct float2_ { float x, y; };
struct float3_ { float x, y, z; };
struct float4_ { float x, y, z, w; };
typedef float2_ f2;
typedef float3_ f3;
typedef float4_ f4;

VGI_HD f2 mk2(float x, float y) { f2 r; r.x = x; r.y = y; return r; }
VGI_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
VGI_HD f3 mk3s(float s) { return mk3(s, s, s); }
VGI_HD f4 mk4(float x, float y, float z, float w) { f4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }

VGI_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
VGI_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
VGI_HD f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
VGI_HD f3 operator/(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
VGI_HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
VGI_HD f3 operator*(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
VGI_HD f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
VGI_HD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
VGI_HD f3 abs3(f3 a) { return mk3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
VGI_HD f3 max3s(f3 a, float s) { return mk3(fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)); }
VGI_HD f3 min3s(f3 a, float s) { return mk3(fminf(a.x, s), fminf(a.y, s), fminf(a.z, s)); }
VGI_HD f3 max3v(f3 a, f3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
VGI_HD f3 min3v(f3 a, f3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
VGI_HD f3 sqrt3(f3 a) { return mk3(sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)); }
VGI_HD f3 lerp3(f3 a, f3 b, float t) { return mk3(lerp(a.x, b.x, t), lerp(a.y, b.y, t), lerp(a.z, b.z, t)); }
VGI_HD float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
VGI_HD float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
VGI_HD float length3(f3 a) { return sqrtf(dot3(a, a)); }
VGI_HD float rsqrt(float x) { return 1.0f / sqrtf(x); }
VGI_HD f3 normalize3(f3 a) { return a * rsqrt(dot3(a, a)); }
VGI_HD f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
VGI_HD float component_max(f3 v) { return fmaxf(fmaxf(v.x, v.y), v.z); }
// util.hlsl:22-25
VGI_HD float luminance(f3 c) { return (c.x * 0.2126f + c.y * 0.7152f) + c.z * 0.0722f; }

// ---------------------------------------------------------------------------
// RNG (rand.hlsl)
// ---------------------------------------------------------------------------
VGI_HD uint32_t jenkins_hash(uint32_t x) {
    x += x << 10;
    x ^= x >> 6;
    x += x << 3;
    x ^= x >> 11;
    x += x << 15;
    return x;
}
VGI_HD uint32_t lcg_init_with_seed(uint32_t px, uint32_t py, uint32_t seed) {
    return jenkins_hash(px ^ (py << 16) ^ seed);
}
VGI_HD float lcg_uint_to_float(uint32_t x) { return asfloat(0x3f800000u | (x >> 9)) - 1.0f; }
VGI_HD uint32_t lcg_xorshift(uint32_t& x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}
VGI_HD float lcg_rand(uint32_t& s) { return lcg_uint_to_float(lcg_xorshift(s)); }

// XXHash32 variant of rand.hlsl:82-112 (not the standard xxh32 finaliser).
struct XXHash32 {
    uint32_t state;
    VGI_HD static XXHash32 init(uint32_t p) {
        XXHash32 r;
        r.state = p + 374761393u;
        return r;
    }
    VGI_HD XXHash32 add(uint32_t p) const {
        uint32_t h = state + p * 3266489917u;
        h = 668265263u * ((h << 17) | (h >> 15));
        XXHash32 r;
        r.state = h;
        return r;
    }
    VGI_HD uint32_t eval() const {
        uint32_t h = state;
        h = 668265263u * ((h << 17) | (h >> 15));
        h = 2246822519u * (h ^ (h >> 15));
        h = 3266489917u * (h ^ (h >> 13));
        return h ^ (h >> 16);
    }
};

// ---------------------------------------------------------------------------
// binary16 conversions (round to nearest even), bit-identical on both sides
// ---------------------------------------------------------------------------
VGI_HD uint32_t f32tof16(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    // v_cvt_f16_f32 rounds to nearest even like the host code below; NaNs are
    // canonicalised to the host code's quiet NaN (a select, not a branch: a
    // branch per conversion splits the caller into blocks and costs it
    // registers where several conversions follow each other)
    const uint32_t h = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)f);
    const uint32_t qnan = ((asuint(f) >> 16) & 0x8000u) | 0x7E00u;
    return f != f ? qnan : h;
#endif
    uint32_t x = asuint(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t absx = x & 0x7FFFFFFFu;
    if (absx >= 0x7F800000u) {  // inf / nan
        return sign | (absx > 0x7F800000u ? 0x7E00u : 0x7C00u);
    }
    if (absx >= 0x477FF000u) return sign | 0x7C00u;  // rounds to >= 65520 -> inf
    if (absx < 0x38800000u) {                          // result is subnormal or zero
        if (absx < 0x33000000u) return sign;           // < 2^-25 rounds to 0
        uint32_t e = absx >> 23;
        uint32_t m = (absx & 0x7FFFFFu) | 0x800000u;
        uint32_t shift = 126u - e;  // e in [102,112]; shift in [14,24]
        uint32_t q = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) q += 1u;
        return sign | q;
    }
    uint32_t e = (absx >> 23) - 112u;
    uint32_t m = absx & 0x7FFFFFu;
    uint32_t q = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q += 1u;
    return sign | q;
}
VGI_HD float f16tof32(uint32_t h) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h);  // exact
#endif
    h &= 0xFFFFu;
    uint32_t sign = (h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1Fu;
    uint32_t m = h & 0x3FFu;
    if (e == 0) {
        // subnormal: m * 2^-24 is exact in binary32
        float v = (float)m * 5.9604644775390625e-8f;
        return sign ? -v : v;
    }
    if (e == 31) return asfloat(sign | 0x7F800000u | (m << 13));
    return asfloat(sign | ((e + 112u) << 23) | (m << 13));
}

// ---------------------------------------------------------------------------
// Octahedral normal encoding (enc.inc.hlsl:11-63)
// ---------------------------------------------------------------------------
VGI_HD f2 oct_wrap(f2 v) {
    return mk2((1.0f - fabsf(v.y)) * (v.x >= 0.0f ? 1.0f : -1.0f),
               (1.0f - fabsf(v.x)) * (v.y >= 0.0f ? 1.0f : -1.0f));
}
VGI_HD f2 normal_encode_oct(f3 n) {
    float s = (fabsf(n.x) + fabsf(n.y)) + fabsf(n.z);
    n = n / s;
    f2 xy = mk2(n.x, n.y);
    if (!(n.z >= 0.0f)) xy = oct_wrap(xy);
    return mk2(xy.x * 0.5f + 0.5f, xy.y * 0.5f + 0.5f);
}
VGI_HD f3 normal_decode_oct(f2 f) {
    f = mk2(f.x * 2.0f - 1.0f, f.y * 2.0f - 1.0f);
    f3 n = mk3(f.x, f.y, (1.0f - fabsf(f.x)) - fabsf(f.y));
    float t = saturate(-n.z);
    n.x += (n.x >= 0.0f) ? -t : t;
    n.y += (n.y >= 0.0f) ? -t : t;
    return normalize3(n);
}
VGI_HD uint32_t normal_encode_oct_u32(f3 n) {
    f2 e = normal_encode_oct(n);
    return f2u(e.x * 65535.0f) | (f2u(e.y * 65535.0f) << 16);
}
VGI_HD f3 normal_decode_oct_u32(uint32_t enc) {
    return normal_decode_oct(mk2((float)(enc & 0xFFFFu) / 65535.0f, (float)((enc >> 16) & 0xFFFFu) / 65535.0f));
}
VGI_HD uint32_t normal_encode_oct_u16(f3 n) {
    f2 e = normal_encode_oct(n);
    return f2u(e.x * 255.0f) | (f2u(e.y * 255.0f) << 8);
}
VGI_HD f3 normal_decode_oct_u16(uint32_t enc) {
    return normal_decode_oct(mk2((float)(enc & 0xFFu) / 255.0f, (float)((enc >> 8) & 0xFFu) / 255.0f));
}

// ---------------------------------------------------------------------------
// G-buffer (gbuffer.hlsl).  Layers are stored as four planes of w*h uint32.
// ---------------------------------------------------------------------------
struct GBufferNormal {
    f3 normal;
    float perceptual_roughness;
};
VGI_HD GBufferNormal decode_gbuffer_normal(uint32_t enc) {
    GBufferNormal g;
    g.normal = normal_decode_oct(mk2((float)(enc & 0xFFFu) / 4095.0f, (float)((enc >> 12) & 0xFFFu) / 4095.0f));
    g.perceptual_roughness = (float)(enc >> 24) / 255.0f;
    return g;
}
VGI_HD f3 decode_gbuffer_color(uint32_t enc) {
    return mk3((float)(enc & 0xFFu) / 255.0f, (float)((enc >> 8) & 0xFFu) / 255.0f,
               (float)((enc >> 16) & 0xFFu) / 255.0f);
}
VGI_HD float decode_gbuffer_metallic(uint32_t enc) { return (float)((enc >> 24) & 0xFFu) / 255.0f; }
VGI_HD f2 decode_gbuffer_fwidth(uint32_t enc) { return mk2(f16tof32(enc & 0xFFFFu), f16tof32(enc >> 16)); }

VGI_HD uint32_t encode_gbuffer_color(f3 color, float metallic) {
    return f2u(color.x * 255.0f) | (f2u(color.y * 255.0f) << 8) | (f2u(color.z * 255.0f) << 16) |
           (f2u(metallic * 255.0f) << 24);
}
VGI_HD uint32_t encode_gbuffer_normal(f3 normal, float perceptual_roughness) {
    f2 e = normal_encode_oct(normal);
    return f2u(e.x * 4095.0f) | (f2u(e.y * 4095.0f) << 12) | (f2u(perceptual_roughness * 255.0f) << 24);
}
VGI_HD uint32_t encode_gbuffer_fwidth(float fwidth_z, float fwidth_n) {
    return f32tof16(fwidth_z) | (f32tof16(fwidth_n) << 16);
}

// ---------------------------------------------------------------------------
// Reservoir encodings (restir/reservoir.hlsl:52-117)
// ---------------------------------------------------------------------------
struct ReservoirSimple {
    uint32_t M;
    float W;
};
VGI_HD uint32_t reservoir_encode_32b(ReservoirSimple r) {
    uint32_t u = minu(r.M, 0xFFu);
    u |= (asuint(r.W) << 1) & 0xFFFFFF00u;
    return u;
}
// Keeps the reference's bit layout: W's low mantissa bits receive M bits 6..7.
VGI_HD ReservoirSimple reservoir_decode_32b(uint32_t u) {
    ReservoirSimple r;
    r.M = u & 0xFFu;
    r.W = asfloat((u >> 1) & 0x7FFFFFE0u);
    return r;
}
struct HitPosNormal {
    f3 pos;
    f3 normal;
};
VGI_HD void hit_encode(f3 pos, f3 normal, uint32_t& ex, uint32_t& ey) {
    ex = f32tof16(pos.x) | (f32tof16(pos.y) << 16);
    ey = f32tof16(pos.z) | (normal_encode_oct_u16(normal) << 16);
}
VGI_HD HitPosNormal hit_decode(uint32_t ex, uint32_t ey) {
    HitPosNormal h;
    h.pos = mk3(f16tof32(ex & 0xFFFFu), f16tof32(ex >> 16), f16tof32(ey & 0xFFFFu));
    h.normal = normal_decode_oct_u16(ey >> 16);
    return h;
}

// ---------------------------------------------------------------------------
// FrameParams (frame_bindings.hlsl:6-25), 384 bytes, column-major matrices.
// ---------------------------------------------------------------------------
struct FrameParams {
    float view_proj[16];
    float inv_view_proj[16];
    float view_pos[4];
    float view_ray_top_left[4];
    float view_ray_right_shift[4];
    float view_ray_down_shift[4];
    float prev_view_proj[16];
    float prev_inv_view_proj[16];
    float prev_view_pos[4];
    float jitter[4];
    float sun_dir[4];
    float sun_inten[4];
};
static_assert(sizeof(FrameParams) == 384, "FrameParams layout");

// mul(M, v) for a column-major 4x4: r_i = ((M0i*x + M1i*y) + M2i*z) + M3i*w
VGI_HD f4 mul44(const float* m, f4 v) {
    f4 r;
    r.x = ((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * v.w;
    r.y = ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * v.w;
    r.z = ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * v.w;
    r.w = ((m[3] * v.x + m[7] * v.y) + m[11] * v.z) + m[15] * v.w;
    return r;
}
VGI_HD f3 xyz(const float* v4) { return mk3(v4[0], v4[1], v4[2]); }

// a / b rounded to nearest, given y = 1 / b rounded to nearest (Markstein's
// division: q = RN(a y); the fused a - b q is exact; RN(q + (a - b q) y) is
// RN(a / b) when nothing overflows or underflows).  Three instructions
// instead of a full IEEE division when y is shared by many quotients; the
// pixel-centre quotients (x + 0.5) / w below are checked against IEEE
// division for every x < w <= 16384.
VGI_HD float div_by_recip(float a, float b, float y) {
    const float q = a * y;
    const float r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}

// cs_depth_to_position (frame_bindings.hlsl:57-61); `m` = inv_view_proj or
// prev_inv_view_proj; depth_bias reproduces the +2^-24 of the sample passes.
VGI_HD f3 depth_to_position(const float* m, uint32_t px, uint32_t py, uint32_t w, uint32_t h, float depth) {
    float su = ((float)px + 0.5f) / (float)w;
    float sv = ((float)py + 0.5f) / (float)h;
    f4 p = mul44(m, mk4(su * 2.0f - 1.0f, sv * 2.0f - 1.0f, depth, 1.0f));
    return mk3(p.x / p.w, p.y / p.w, p.z / p.w);
}
// The same with rw = 1 / w and rh = 1 / h (rounded to nearest) given, for
// loops over many pixels of one image: bit-identical results.
VGI_HD f3 depth_to_position_r(const float* m, uint32_t px, uint32_t py, uint32_t w, uint32_t h, float rw, float rh,
                              float depth) {
    float su = div_by_recip((float)px + 0.5f, (float)w, rw);
    float sv = div_by_recip((float)py + 0.5f, (float)h, rh);
    f4 p = mul44(m, mk4(su * 2.0f - 1.0f, sv * 2.0f - 1.0f, depth, 1.0f));
    return mk3(p.x / p.w, p.y / p.w, p.z / p.w);
}

// ---------------------------------------------------------------------------
// BRDF (brdf.hlsl) and sampling (sampling.hlsl)
// ---------------------------------------------------------------------------
VGI_HD float pow5(float x) {
    float x2 = x * x;
    float x4 = x2 * x2;
    return x4 * x;
}
VGI_HD f3 F_Schlick(float u, f3 f0) {
    float f = pow5(1.0f - u);
    return mk3(f + f0.x * (1.0f - f), f + f0.y * (1.0f - f), f + f0.z * (1.0f - f));
}
VGI_HD float D_GGX(float NoH, float roughness) {
    float a = NoH * roughness;
    float k = roughness / ((1.0f - NoH * NoH) + a * a);
    return (k * k) * kOneOverPi;
}
VGI_HD float vis_smith_G2_height_correlated_GGX(float NoL, float NoV, float roughness) {
    float r2 = roughness * roughness;
    float u_L = NoV * sqrtf((-r2 * NoL + NoL) * NoL + r2);
    float u_V = NoL * sqrtf((-r2 * NoV + NoV) * NoV + r2);
    return 0.5f / (u_L + u_V);
}
VGI_HD float smith_G2_over_G1_height_correlated_GGX(float NoL, float NoV, float roughness) {
    float r2 = roughness * roughness;
    float k_V = sqrtf((-r2 * NoV + NoV) * NoV + r2) / NoV;
    float k_L = sqrtf((-r2 * NoL + NoL) * NoL + r2) / NoL;
    return (1.0f + k_V) / (k_V + k_L);
}
VGI_HD float smith_lambda_GGX_plus_one_mul_cos(float c, float roughness) {
    float r2 = roughness * roughness;
    return (c + sqrtf((-r2 * c + c) * c + r2)) / 2.0f;
}
VGI_HD f3 get_diffuse_rho(f3 base_color, float metallic) { return base_color * (1.0f - metallic); }
VGI_HD f3 get_specular_f0(f3 base_color, float metallic) {
    return mk3(lerp(0.04f, base_color.x, metallic), lerp(0.04f, base_color.y, metallic),
               lerp(0.04f, base_color.z, metallic));
}
// brdf.hlsl:154-177
VGI_HD f3 eval_GGX_Lambertian(f3 v, f3 l, f3 n, float perceptual_roughness, f3 diffuse_rho, f3 specular_f0) {
    f3 h = normalize3(v + l);
    float NoL = saturate(dot3(n, l));
    float NoV = saturate(dot3(n, v));
    float NoH = saturate(dot3(n, h));
    float LoH = saturate(dot3(l, h));
    float roughness = perceptual_roughness * perceptual_roughness;
    float D = D_GGX(NoH, roughness);
    f3 F = F_Schlick(LoH, specular_f0);
    float V = vis_smith_G2_height_correlated_GGX(NoL, NoV, roughness);
    f3 Fr = F * fminf(D * V, kF32SignificantlyLarge);
    f3 Fd = diffuse_rho * kOneOverPi;
    return Fd + Fr;
}
// brdf.hlsl:180-200 (Hirvonen 2019 fit); mul(m, v) with row-major float2x2/3x3.
VGI_HD f3 ggx_brdf_integral_approx(float n_dot_v, float roughness, f3 specular_f0) {
    float a = roughness * roughness;
    float a2 = a * a;
    float a3 = a * a2;
    float c = n_dot_v;
    float c2 = c * c;
    float c3 = c * c2;
    float m0r0 = 0.99044f * 1.0f + -1.28514f * c;
    float m0r1 = 1.29678f * 1.0f + -0.755907f * c;
    float k0 = 1.0f * m0r0 + a * m0r1;
    float m1r0 = (1.0f * 1.0f + 2.92338f * c) + 59.4188f * c3;
    float m1r1 = (20.3225f * 1.0f + -27.0302f * c) + 222.592f * c3;
    float m1r2 = (121.563f * 1.0f + 626.13f * c) + 316.627f * c3;
    k0 /= (1.0f * m1r0 + a * m1r1) + a3 * m1r2;
    float m2r0 = 0.0365463f * 1.0f + 3.32707f * c;
    float m2r1 = 9.0632f * 1.0f + -9.04756f * c;
    float k1 = 1.0f * m2r0 + a * m2r1;
    float m3r0 = (1.0f * 1.0f + 3.59685f * c2) + -1.36772f * c3;
    float m3r1 = (9.04401f * 1.0f + -16.3174f * c2) + 9.22949f * c3;
    float m3r2 = (5.56589f * 1.0f + 19.7886f * c2) + -20.2123f * c3;
    k1 /= (1.0f * m3r0 + a * m3r1) + a3 * m3r2;
    return mk3(k0 + k1 * specular_f0.x, k0 + k1 * specular_f0.y, k0 + k1 * specular_f0.z);
}

// sampling.hlsl:10-26
VGI_HD f4 get_rotation_to_z_from(f3 v) {
    if (v.z < -0.99999f) return mk4(1.0f, 0.0f, 0.0f, 0.0f);
    f4 q = mk4(v.y, -v.x, 0.0f, 1.0f + v.z);
    float inv = rsqrt(((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w);
    return mk4(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
}
VGI_HD f4 invert_rotation(f4 q) { return mk4(-q.x, -q.y, -q.z, q.w); }
VGI_HD f3 rotate_point(f4 q, f3 v) {
    f3 axis = mk3(q.x, q.y, q.z);
    f3 a = axis * (2.0f * dot3(axis, v));
    f3 b = v * (q.w * q.w - dot3(axis, axis));
    f3 c = cross3(axis, v) * (2.0f * q.w);
    return (a + b) + c;
}
VGI_HD f3 sample_hemisphere_uniform(float ux, float uy) {
    float cos_theta = 1.0f - ux;
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    float s, c;
    vgi_sincosf(kTwoPi * uy, &s, &c);
    return mk3(c * sin_theta, s * sin_theta, cos_theta);
}
VGI_HD f3 sample_hemisphere_cosine(float ux, float uy) {
    float sin_theta = sqrtf(ux);
    float cos_theta = sqrtf(1.0f - ux);
    float s, c;
    vgi_sincosf(kTwoPi * uy, &s, &c);
    return mk3(sin_theta * c, sin_theta * s, cos_theta);
}
// raytrace.inc.hlsl:213-218
VGI_HD f3 sample_hemisphere_uniform_with_normal(float ux, float uy, f3 normal) {
    f3 l = sample_hemisphere_uniform(ux, uy);
    f4 q = invert_rotation(get_rotation_to_z_from(normal));
    return rotate_point(q, l);
}

// ---------------------------------------------------------------------------
// Hash grid keys (restir/hash_grid.inc.hlsl:8-105)
// ---------------------------------------------------------------------------
constexpr uint32_t kHashGridNumBuckets = 65536;
constexpr uint32_t kHashGridBucketMask = 0xFFFF;
constexpr uint32_t kHashGridBucketSize = 16;
constexpr uint32_t kHashGridBucketShift = 4;
constexpr uint32_t kHashGridNumCells = kHashGridNumBuckets * kHashGridBucketSize;
constexpr uint32_t kHashGridMaxQueries = 128 * 1024;
constexpr uint32_t kHashGridMaxQueriesShift = 17;
// Appends the ordered cache update samples the frame's queries down to:
// half the buffer, since the candidates of one frame can exceed the last
// frame's by 40% (the frame after a diffuse validation frame).
constexpr uint32_t kHashGridQueryTarget = kHashGridMaxQueries / 2u;

// Threshold word that keeps about kHashGridQueryTarget of `expected` query
// candidates: a candidate is kept if its 32-bit hash is >= the word, which
// happens with probability (2^32 - word) / 2^32.  0 keeps every candidate.
VGI_HD uint32_t query_threshold_for(uint64_t expected) {
    if (expected <= kHashGridQueryTarget) return 0u;
    uint64_t keep = ((uint64_t)kHashGridQueryTarget << 32) / expected;  // < 2^32
    if (keep == 0) keep = 1;
    return (uint32_t)((1ull << 32) - keep);
}

// uint(max(0.5 * log2(d2), 0)) evaluated exactly: for d2 >= 1 with binary
// exponent e, 0.5*log2(d2) lies in [e/2, (e+1)/2), so the truncation is e>>1.
VGI_HD uint32_t hash_grid_cell_lod(f3 pos, f3 cam) {
    f3 d = pos - cam;
    float d2 = dot3(d, d);
    if (!(d2 >= 1.0f)) return 0u;
    uint32_t e = (asuint(d2) >> 23) - 127u;  // d2 >= 1 is normal or inf
    uint32_t lod = e >> 1;
    return lod > 31u ? 31u : lod;
}
struct VertexDescriptor {
    int32_t px, py, pz;
    uint32_t packed_x, packed_y;
    uint32_t lod;
};
VGI_HD VertexDescriptor vertex_descriptor_create(f3 pos, f3 view_pos) {
    VertexDescriptor v;
    v.lod = hash_grid_cell_lod(pos, view_pos);
    float cell_size = (1.0f / 16.0f) * (float)(1u << v.lod);
    float inv = rcp(cell_size);
    v.px = (int32_t)floorf(pos.x * inv);
    v.py = (int32_t)floorf(pos.y * inv);
    v.pz = (int32_t)floorf(pos.z * inv);
    uint32_t ux = (uint32_t)v.px, uy = (uint32_t)v.py, uz = (uint32_t)v.pz;
    v.packed_x = ux ^ (uy << 22);
    v.packed_y = (uy >> 10) ^ (uz << 22);
    return v;
}
VGI_HD uint32_t vertex_descriptor_hash(const VertexDescriptor& v) {
    return XXHash32::init(v.packed_x).add(v.packed_y).add(v.lod).eval() & kHashGridBucketMask;
}
VGI_HD uint32_t vertex_descriptor_checksum(const VertexDescriptor& v) {
    return XXHash32::init((uint32_t)v.py).add((uint32_t)v.pz).add((uint32_t)v.px).add(v.lod).eval() | 1u;
}

}  // namespace vgi
