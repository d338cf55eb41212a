/*
 * vgi_storage.h — per-pixel storage policies.
 *
 * A storage policy decides how the floating-point per-pixel fields of the
 * hot path (radiance RGB, moments/variance pairs, ray length, AO) are laid
 * out in HBM.  Integer encodings (G-buffer layers, packed reservoirs, hit
 * pos/normal, history length as an 8-bit count) are identical in every
 * mode.  All images are row-major w*h arrays of their element type
 * ("array of structs"); a wave of 64 lanes covering an 8x8 tile therefore
 * reads eight contiguous 8-pixel segments per field.
 *
 *   StorageF32: radiance = 3 x f32 (12 B), pair = 2 x f32 (8 B), scalar f32
 *   StorageF16: radiance = 3 x f16 (6 B),  pair = 2 x f16 (4 B), scalar f16
 *
 * Arithmetic is always f32.
 */
#pragma once
#include "vgi_common.h"

namespace vgi {

struct StorageF32 {
    static constexpr int kMode = 0;
    typedef float scalar_t;
    static constexpr int kRadianceBytes = 12;
    static constexpr int kPairBytes = 8;
    static constexpr int kScalarBytes = 4;
    VGI_HD static f3 ld3(const void* p, size_t i) {
        const float* f = (const float*)p + 3 * i;
        return mk3(gld(f), gld(f + 1), gld(f + 2));
    }
    VGI_HD static void st3(void* p, size_t i, f3 v) {
        float* f = (float*)p + 3 * i;
        gst(f, v.x);
        gst(f + 1, v.y);
        gst(f + 2, v.z);
    }
    VGI_HD static f2 ld2(const void* p, size_t i) {
        const float* f = (const float*)p + 2 * i;
        return mk2(gld(f), gld(f + 1));
    }
    VGI_HD static void st2(void* p, size_t i, f2 v) {
        float* f = (float*)p + 2 * i;
        gst(f, v.x);
        gst(f + 1, v.y);
    }
    VGI_HD static float ld1(const void* p, size_t i) { return gld((const float*)p + i); }
    VGI_HD static void st1(void* p, size_t i, float v) { gst((float*)p + i, v); }
    // the value ld3 returns after st3 of v
    VGI_HD static f3 round3(f3 v) { return v; }
};

struct StorageF16 {
    static constexpr int kMode = 1;
    typedef uint16_t scalar_t;
    static constexpr int kRadianceBytes = 6;
    static constexpr int kPairBytes = 4;
    static constexpr int kScalarBytes = 2;
    VGI_HD static f3 ld3(const void* p, size_t i) {
        const uint16_t* h = (const uint16_t*)p + 3 * i;
        return mk3(f16tof32(gld(h)), f16tof32(gld(h + 1)), f16tof32(gld(h + 2)));
    }
    VGI_HD static void st3(void* p, size_t i, f3 v) {
        uint16_t* h = (uint16_t*)p + 3 * i;
        gst(h, (uint16_t)f32tof16(v.x));
        gst(h + 1, (uint16_t)f32tof16(v.y));
        gst(h + 2, (uint16_t)f32tof16(v.z));
    }
    VGI_HD static f2 ld2(const void* p, size_t i) {
        uint32_t u = gld((const uint32_t*)p + i);
        return mk2(f16tof32(u & 0xFFFFu), f16tof32(u >> 16));
    }
    VGI_HD static void st2(void* p, size_t i, f2 v) {
        gst((uint32_t*)p + i, f32tof16(v.x) | (f32tof16(v.y) << 16));
    }
    VGI_HD static float ld1(const void* p, size_t i) { return f16tof32(gld((const uint16_t*)p + i)); }
    VGI_HD static void st1(void* p, size_t i, float v) { gst((uint16_t*)p + i, (uint16_t)f32tof16(v)); }
    // the value ld3 returns after st3 of v
    VGI_HD static f3 round3(f3 v) {
        return mk3(f16tof32(f32tof16(v.x)), f16tof32(f32tof16(v.y)), f16tof32(f32tof16(v.z)));
    }
};

// Bilinear fetch with clamp-to-edge addressing (sampler_linear_clamp), in
// exact f32: texel t = uv*size - 0.5, weights from t - floor(t).
template <class S>
VGI_HD f3 sample_bilinear3(const void* tex, uint32_t w, uint32_t h, float u, float v) {
    float tx = u * (float)w - 0.5f;
    float ty = v * (float)h - 0.5f;
    float x0f = floorf(tx), y0f = floorf(ty);
    float fx = tx - x0f, fy = ty - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int xa = clampi(x0, 0, (int)w - 1), xb = clampi(x0 + 1, 0, (int)w - 1);
    int ya = clampi(y0, 0, (int)h - 1), yb = clampi(y0 + 1, 0, (int)h - 1);
    f3 c00 = S::ld3(tex, (size_t)ya * w + xa);
    f3 c10 = S::ld3(tex, (size_t)ya * w + xb);
    f3 c01 = S::ld3(tex, (size_t)yb * w + xa);
    f3 c11 = S::ld3(tex, (size_t)yb * w + xb);
    f3 top = c00 * (1.0f - fx) + c10 * fx;
    f3 bot = c01 * (1.0f - fx) + c11 * fx;
    return top * (1.0f - fy) + bot * fy;
}

// History length: R8_UNORM in the reference (denoising.rs:122-125).  Stored
// as the 8-bit count round(len * 255) (round half to even), read as k/255.
VGI_HD uint8_t history_len_store(float len) {
    float s = saturate(len) * 255.0f;
    return (uint8_t)rintf(s);
}
VGI_HD float history_len_load(uint8_t k) { return (float)k / 255.0f; }

}  // namespace vgi
