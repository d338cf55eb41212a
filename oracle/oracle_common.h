/*
 * oracle_common.h — ORACLE (test infrastructure only; never linked into the
 * product).  Scalar/vector helpers and the integer encodings of violet's
 * HLSL, restated independently from the product sources.
 *
 * The oracle is compiled twice from the same C source:
 *   REAL = double -> the FP64 oracle (libm transcendentals);
 *   REAL = float  -> the FP32 twin, which evaluates every expression in the
 *                    order written in the reference HLSL, with
 *                    -ffp-contract=off, so that integer decisions taken from
 *                    float values (reservoir selections, neighbour choices,
 *                    hash keys) can be compared bit for bit with the GPU.
 * Only the single-precision sine/cosine and log2 (vgi_mathf.h, Cephes-method
 * routines the GPU also uses) are shared with the product; they are pinned
 * against libm double in tests/test_oracle_kat.py.
 */
#ifndef ORACLE_COMMON_H
#define ORACLE_COMMON_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
typedef ORACLE_REAL real;

#define ORACLE_IS_F32 (sizeof(real) == 4)

#if defined(ORACLE_F32)
#include "../violet_amd/csrc/vgi_mathf.h"
#define R_SQRT sqrtf
#define R_EXP expf
#define R_EXP2 exp2f
#define R_FLOOR floorf
#define R_FABS fabsf
#define R_FMIN fminf
#define R_FMAX fmaxf
#define R_RINT rintf
#define R_CEIL ceilf
#define R_LOG2 vgi_log2f
static inline void r_sincos(real x, real* s, real* c) { vgi_sincosf(x, s, c); }
#else
#define R_SQRT sqrt
#define R_EXP exp
#define R_EXP2 exp2
#define R_FLOOR floor
#define R_FABS fabs
#define R_FMIN fmin
#define R_FMAX fmax
#define R_RINT rint
#define R_CEIL ceil
#define R_LOG2 log2
static inline void r_sincos(real x, real* s, real* c) {
    *s = sin(x);
    *c = cos(x);
}
#endif

#define RC(x) ((real)(x))

/* constants.hlsl */
#define O_PI RC(3.14159265359)
#define O_TWO_PI (RC(2.0) * O_PI)
#define O_ONE_OVER_PI (RC(1.0) / O_PI)
#define O_HIST_UNIT (RC(1.0) / RC(255.0))

typedef struct { real x, y; } v2;
typedef struct { real x, y, z; } v3;
typedef struct { real x, y, z, w; } v4;

static inline v2 V2(real x, real y) { v2 r = {x, y}; return r; }
static inline v3 V3(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 V3s(real s) { v3 r = {s, s, s}; return r; }
static inline v4 V4(real x, real y, real z, real w) { v4 r = {x, y, z, w}; return r; }
static inline v3 add3(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 div3(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scale3(v3 a, real s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 divs3(v3 a, real s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline v3 neg3(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline v3 abs3(v3 a) { return V3(R_FABS(a.x), R_FABS(a.y), R_FABS(a.z)); }
static inline real dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline real len3(v3 a) { return R_SQRT(dot3(a, a)); }
static inline v3 norm3(v3 a) { return scale3(a, RC(1.0) / R_SQRT(dot3(a, a))); }
static inline v3 cross3(v3 a, v3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real sat(real x) { return R_FMIN(R_FMAX(x, RC(0.0)), RC(1.0)); }
static inline real lerp1(real a, real b, real t) { return a + t * (b - a); }
static inline v3 lerp3(v3 a, v3 b, real t) { return V3(lerp1(a.x, b.x, t), lerp1(a.y, b.y, t), lerp1(a.z, b.z, t)); }
static inline real lum(v3 c) { return (c.x * RC(0.2126) + c.y * RC(0.7152)) + c.z * RC(0.0722); }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline real cmax3(v3 v) { return R_FMAX(R_FMAX(v.x, v.y), v.z); }

/* HLSL uint(x) of a non-negative in-range value (truncation). */
static inline uint32_t r2u(real x) {
    if (!(x > 0)) return 0u;
    if (x >= RC(4294967040.0)) return 0xFFFFFF00u;
    return (uint32_t)x;
}

static inline uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ---- rand.hlsl:9-46 ---- */
static inline uint32_t o_jenkins(uint32_t x) {
    x = x + (x << 10);
    x = x ^ (x >> 6);
    x = x + (x << 3);
    x = x ^ (x >> 11);
    x = x + (x << 15);
    return x;
}
static inline uint32_t o_lcg_init(uint32_t px, uint32_t py, uint32_t seed) { return o_jenkins((px ^ (py << 16)) ^ seed); }
static inline uint32_t o_xorshift(uint32_t* s) {
    uint32_t x = *s;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    *s = x;
    return x;
}
/* asfloat(0x3f800000 | x>>9) - 1 == (x>>9) * 2^-23 exactly */
static inline real o_u2f(uint32_t x) { return (real)((double)(x >> 9) * (1.0 / 8388608.0)); }
static inline real o_rand(uint32_t* s) { return o_u2f(o_xorshift(s)); }

/* ---- rand.hlsl:82-112, XXHash32 variant ---- */
static inline uint32_t o_rotl17(uint32_t h) { return (h << 17) | (h >> 15); }
static inline uint32_t o_xxh_init(uint32_t p) { return p + 374761393u; }
static inline uint32_t o_xxh_add(uint32_t h, uint32_t p) { h = h + p * 3266489917u; return 668265263u * o_rotl17(h); }
static inline uint32_t o_xxh_eval(uint32_t h) {
    h = 668265263u * o_rotl17(h);
    h = 2246822519u * (h ^ (h >> 15));
    h = 3266489917u * (h ^ (h >> 13));
    return h ^ (h >> 16);
}

/* ---- binary16 <-> real (round half to even), via frexp/ldexp ---- */
static inline double o_half_to_double(uint32_t h) {
    h &= 0xFFFFu;
    int s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    double v;
    if (e == 0) v = ldexp((double)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = ldexp((double)(m + 1024), e - 25);
    return s ? -v : v;
}
static inline uint32_t o_double_to_half(double x) {
    uint32_t s = signbit(x) ? 0x8000u : 0u;
    double a = fabs(x);
    if (isnan(x)) return s | 0x7E00u;
    if (a >= 65520.0) return s | 0x7C00u;
    if (a < ldexp(1.0, -14)) {
        double q = nearbyint(a * 16777216.0); /* units of 2^-24; default mode is RNE */
        return s | (uint32_t)q;
    }
    int e;
    double f = frexp(a, &e); /* a = f * 2^e, f in [0.5,1) */
    double q = nearbyint(f * 2048.0); /* 11 significant bits */
    if (q >= 2048.0) { q = 1024.0; e += 1; }
    uint32_t he = (uint32_t)(e - 1 + 15);
    if (he >= 31) return s | 0x7C00u;
    return s | (he << 10) | ((uint32_t)q - 1024u);
}

/* ---- enc.inc.hlsl:11-63 ---- */
static inline v2 o_oct_enc(v3 n) {
    real s = (R_FABS(n.x) + R_FABS(n.y)) + R_FABS(n.z);
    n = divs3(n, s);
    real x = n.x, y = n.y;
    if (!(n.z >= RC(0.0))) {
        real wx = (RC(1.0) - R_FABS(n.y)) * (n.x >= RC(0.0) ? RC(1.0) : RC(-1.0));
        real wy = (RC(1.0) - R_FABS(n.x)) * (n.y >= RC(0.0) ? RC(1.0) : RC(-1.0));
        x = wx;
        y = wy;
    }
    return V2(x * RC(0.5) + RC(0.5), y * RC(0.5) + RC(0.5));
}
static inline v3 o_oct_dec(v2 f) {
    real fx = f.x * RC(2.0) - RC(1.0), fy = f.y * RC(2.0) - RC(1.0);
    v3 n = V3(fx, fy, (RC(1.0) - R_FABS(fx)) - R_FABS(fy));
    real t = sat(-n.z);
    n.x = n.x + (n.x >= RC(0.0) ? -t : t);
    n.y = n.y + (n.y >= RC(0.0) ? -t : t);
    return norm3(n);
}
static inline uint32_t o_oct_enc_u16(v3 n) {
    v2 e = o_oct_enc(n);
    return r2u(e.x * RC(255.0)) | (r2u(e.y * RC(255.0)) << 8);
}
static inline v3 o_oct_dec_u16(uint32_t e) {
    return o_oct_dec(V2((real)(e & 0xFFu) / RC(255.0), (real)((e >> 8) & 0xFFu) / RC(255.0)));
}
static inline uint32_t o_oct_enc_u32(v3 n) {
    v2 e = o_oct_enc(n);
    return r2u(e.x * RC(65535.0)) | (r2u(e.y * RC(65535.0)) << 16);
}
static inline v3 o_oct_dec_u32(uint32_t e) {
    return o_oct_dec(V2((real)(e & 0xFFFFu) / RC(65535.0), (real)((e >> 16) & 0xFFFFu) / RC(65535.0)));
}

/* ---- gbuffer.hlsl:15-114 (layer planes L0..L3) ---- */
static inline v3 o_gb_normal(uint32_t l1) {
    return o_oct_dec(V2((real)(l1 & 0xFFFu) / RC(4095.0), (real)((l1 >> 12) & 0xFFFu) / RC(4095.0)));
}
static inline real o_gb_rough(uint32_t l1) { return (real)(l1 >> 24) / RC(255.0); }
static inline v3 o_gb_color(uint32_t l0) {
    return V3((real)(l0 & 0xFFu) / RC(255.0), (real)((l0 >> 8) & 0xFFu) / RC(255.0), (real)((l0 >> 16) & 0xFFu) / RC(255.0));
}
static inline real o_gb_metal(uint32_t l0) { return (real)((l0 >> 24) & 0xFFu) / RC(255.0); }
static inline real o_gb_fwz(uint32_t l2) { return (real)o_half_to_double(l2 & 0xFFFFu); }
static inline real o_gb_fwn(uint32_t l2) { return (real)o_half_to_double(l2 >> 16); }

/* ---- restir/reservoir.hlsl:52-117 ---- */
static inline uint32_t o_rsv_enc(uint32_t M, float W) {
    uint32_t u = M < 0xFFu ? M : 0xFFu;
    return u | ((f_bits(W) << 1) & 0xFFFFFF00u);
}
static inline void o_rsv_dec(uint32_t u, uint32_t* M, float* W) {
    *M = u & 0xFFu;
    *W = bits_f((u >> 1) & 0x7FFFFFE0u);
}
static inline void o_hit_enc(v3 pos, v3 n, uint32_t* ex, uint32_t* ey) {
    *ex = o_double_to_half((double)(float)pos.x) | (o_double_to_half((double)(float)pos.y) << 16);
    *ey = o_double_to_half((double)(float)pos.z) | (o_oct_enc_u16(n) << 16);
}
static inline void o_hit_dec(uint32_t ex, uint32_t ey, v3* pos, v3* n) {
    *pos = V3((real)o_half_to_double(ex & 0xFFFFu), (real)o_half_to_double(ex >> 16), (real)o_half_to_double(ey & 0xFFFFu));
    *n = o_oct_dec_u16(ey >> 16);
}

/* ---- frame_bindings.hlsl:6-79: 96 floats, column-major matrices ---- */
typedef struct {
    real view_proj[16], inv_view_proj[16];
    real view_pos[4], ray_tl[4], ray_r[4], ray_d[4];
    real prev_view_proj[16], prev_inv_view_proj[16];
    real prev_view_pos[4], jitter[4], sun_dir[4], sun_inten[4];
} ofp;
static inline void o_load_fp(const float* f, ofp* o) {
    real* d = (real*)o;
    for (int i = 0; i < 96; ++i) d[i] = (real)f[i];
}
static inline v4 o_mul44(const real* m, v4 v) {
    v4 r;
    r.x = ((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * v.w;
    r.y = ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * v.w;
    r.z = ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * v.w;
    r.w = ((m[3] * v.x + m[7] * v.y) + m[11] * v.z) + m[15] * v.w;
    return r;
}
static inline v3 o_depth_to_pos(const real* inv, int px, int py, int w, int h, real depth) {
    real su = ((real)px + RC(0.5)) / (real)w, sv = ((real)py + RC(0.5)) / (real)h;
    v4 p = o_mul44(inv, V4(su * RC(2.0) - RC(1.0), sv * RC(2.0) - RC(1.0), depth, RC(1.0)));
    return V3(p.x / p.w, p.y / p.w, p.z / p.w);
}
static inline v3 o_xyz(const real* v) { return V3(v[0], v[1], v[2]); }

/* history length R8_UNORM: count = round_half_even(sat(len)*255) */
static inline uint8_t o_len_store(real len) { return (uint8_t)R_RINT(sat(len) * RC(255.0)); }
static inline real o_len_load(uint8_t k) { return (real)k / RC(255.0); }

#endif /* ORACLE_COMMON_H */
