/*
 * vgi_mathf.h — deterministic single-precision sine/cosine.
 *
 * The ReSTIR passes pick neighbour pixels and ray directions from
 * sin/cos of random angles (ind_diff_spatial_resample.hlsl:153-154,
 * ind_spec_spatial_resample.hlsl:224-225, sampling.hlsl:176,195,
 * ind_spec_sample_gen.hlsl:101-102).  A neighbour choice or a ray that
 * differs by one ulp changes which reservoir is merged, so the GPU kernels
 * and the single-precision CPU build of the oracle must evaluate these
 * functions bit-identically.  Vendor libm implementations differ between
 * host and device, so both sides use this one routine.
 *
 * Method: Cody–Waite reduction by pi/4 with a three-part constant and the
 * minimax polynomials of S. L. Moshier's Cephes library (single/sinf.c),
 * evaluated with separate multiplies and adds (no fused multiply-add), so
 * every operation is an IEEE-754 correctly rounded binary32 operation on
 * both host (SSE) and device (gfx950 VALU, compiled with
 * -ffp-contract=off).  Accuracy: < 2 ulp for |x| < 8192 (tested against
 * libm double).  This file is plain C99 so the C oracle can include it.
 *
 * vgi_log2f serves the texture level of detail (log2 of the texel
 * footprint), whose fraction weights two mip levels: Cephes' log2f method
 * (split x = m * 2^e with m in [sqrt(1/2), sqrt(2)), a degree-9 minimax
 * polynomial for log(1 + (m - 1)), and log2(e) applied in two parts), with
 * the same separate multiplies and adds.
 */
#ifndef VGI_MATHF_H
#define VGI_MATHF_H

#if defined(__HIPCC__)
#define VGI_FN __host__ __device__ static inline
#else
#define VGI_FN static inline
#endif

VGI_FN void vgi_sincosf(float x, float *s_out, float *c_out)
{
    const float FOUR_OVER_PI = 1.27323954473516f;
    const float DP1 = 0.78515625f;
    const float DP2 = 2.4187564849853515625e-4f;
    const float DP3 = 3.77489497744594108e-8f;
    int sgn_s = 1;
    float ax = x;
    if (x < 0.0f) {
        sgn_s = -1;
        ax = -x;
    }
    int j = (int)(FOUR_OVER_PI * ax);
    float y = (float)j;
    if (j & 1) {
        j += 1;
        y += 1.0f;
    }
    j &= 7;
    int sgn_c = 1;
    if (j > 3) {
        sgn_s = -sgn_s;
        sgn_c = -sgn_c;
        j -= 4;
    }
    if (j > 1) sgn_c = -sgn_c;
    /* extended-precision modular arithmetic */
    float r = ax - y * DP1;
    r = r - y * DP2;
    r = r - y * DP3;
    float z = r * r;
    /* polynomial for sin on [-pi/4, pi/4] */
    float ps = -1.9515295891e-4f;
    ps = ps * z + 8.3321608736e-3f;
    ps = ps * z + -1.6666654611e-1f;
    ps = ps * z;
    ps = ps * r + r;
    /* polynomial for cos on [-pi/4, pi/4] */
    float pc = 2.443315711809948e-5f;
    pc = pc * z + -1.388731625493765e-3f;
    pc = pc * z + 4.166664568298827e-2f;
    pc = pc * z;
    pc = pc * z;
    float hz = 0.5f * z;
    pc = pc - hz;
    pc = pc + 1.0f;
    float sv, cv;
    if (j == 1 || j == 2) {
        sv = pc;
        cv = ps;
    } else {
        sv = ps;
        cv = pc;
    }
    *s_out = (sgn_s < 0) ? -sv : sv;
    *c_out = (sgn_c < 0) ? -cv : cv;
}

VGI_FN float vgi_sinf(float x)
{
    float s, c;
    vgi_sincosf(x, &s, &c);
    return s;
}

VGI_FN float vgi_cosf(float x)
{
    float s, c;
    vgi_sincosf(x, &s, &c);
    return c;
}

/* log2 of a positive finite float (0, negative and non-finite arguments are
 * the caller's to exclude: the result is then unspecified). */
VGI_FN float vgi_log2f(float x)
{
    const float SQRTHF = 0.707106781186547524f;
    const float LOG2EA = 0.44269504088896340736f; /* log2(e) - 1 */
    unsigned int b;
    __builtin_memcpy(&b, &x, 4);
    int e = (int)((b >> 23) & 0xffu);
    if (e == 0) { /* subnormal: scale into the normal range first */
        x = x * 16777216.0f;
        __builtin_memcpy(&b, &x, 4);
        e = (int)((b >> 23) & 0xffu) - 24;
    }
    e -= 126;
    b = (b & 0x807fffffu) | 0x3f000000u; /* mantissa in [0.5, 1) */
    float m;
    __builtin_memcpy(&m, &b, 4);
    if (m < SQRTHF) {
        e -= 1;
        m = m + m - 1.0f;
    } else {
        m = m - 1.0f;
    }
    float z = m * m;
    float y = 7.0376836292e-2f;
    y = y * m + -1.1514610310e-1f;
    y = y * m + 1.1676998740e-1f;
    y = y * m + -1.2420140846e-1f;
    y = y * m + 1.4249322787e-1f;
    y = y * m + -1.6668057665e-1f;
    y = y * m + 2.0000714765e-1f;
    y = y * m + -2.4999993993e-1f;
    y = y * m + 3.3333331174e-1f;
    y = y * m;
    y = y * z;
    y = y + -0.5f * z;
    float r = y * LOG2EA;
    r = r + m * LOG2EA;
    r = r + y;
    r = r + m;
    r = r + (float)e;
    return r;
}

#endif /* VGI_MATHF_H */
