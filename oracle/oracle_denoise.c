/*
 * oracle_denoise.c — ORACLE (test infrastructure only).  CPU restatement of
 * violet's ReLAX/SVGF denoiser passes, one function per HLSL compute pass,
 * looping over pixels in the order of the reference's dispatch.
 *
 *   oracle_denoise_temporal     denoise/svgf_temporal_filter.hlsl:96-362
 *   oracle_denoise_disocclusion denoise/relax_disocclusion_fix.hlsl:27-122
 *   oracle_denoise_fast_clamp   denoise/relax_fast_history_clamp.hlsl:12-75
 *   oracle_denoise_fallback     denoise/svgf_spatial_fallback_filter.hlsl:146-246
 *   oracle_denoise_atrous       denoise/svgf_atrous_filter.hlsl:24-88
 *   edge_weight                 denoise/svgf.inc.hlsl:57-83
 *
 * Images: row-major w*h; radiance = 3 reals, pairs = 2 reals, depth/ray
 * length = 1 real, history length = u8 count, G-buffer = 4 planes of u32.
 * Sky pixels (depth == 0 / history length 0): the reference leaves the
 * colour outputs unwritten; this restatement writes zeros there (the value
 * the GPU library writes).
 */
#include "oracle_common.h"

static inline v3 ld3(const real* p, size_t i) { return V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st3(real* p, size_t i, v3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
static inline v2 ld2(const real* p, size_t i) { return V2(p[2 * i], p[2 * i + 1]); }
static inline void st2(real* p, size_t i, v2 v) { p[2 * i] = v.x; p[2 * i + 1] = v.y; }

/* pow(x, sigma_n).  The FP64 oracle calls libm pow; the FP32 twin squares
 * log2(sigma_n) times (sigma_n is 128 or 32). */
static inline real o_pow_sigma(real x, int log2_sigma) {
#if defined(ORACLE_F32)
    for (int i = 0; i < log2_sigma; ++i) x = x * x;
    return x;
#else
    return pow(x, (double)(1 << log2_sigma));
#endif
}

typedef struct {
    real depth_center, fwidth_z;
    v3 normal_center;
    real lc[2], lstd[2];
    int use_lumi;
    real sigma_z;
    int log2_sigma_n;
} edge_stop;

static inline void edge_weight(const edge_stop* e, real tap_dist, real depth, v3 normal, real lx, real ly, real* wx,
                               real* wy) {
    const real sigma_l = RC(4.0), eps = RC(1e-8);
    real dist_z = R_FABS(depth - e->depth_center);
    real wz = -dist_z / ((e->sigma_z * e->fwidth_z) * tap_dist + eps);
    real n_dot = sat(dot3(e->normal_center, normal));
    real wn = o_pow_sigma(n_dot, e->log2_sigma_n);
    real lx_log = RC(0.0), ly_log = RC(0.0);
    if (e->use_lumi) {
        lx_log = -R_FABS(lx - e->lc[0]) * (RC(1.0) / (sigma_l * e->lstd[0] + eps));
        ly_log = -R_FABS(ly - e->lc[1]) * (RC(1.0) / (sigma_l * e->lstd[1] + eps));
    }
    *wx = wn * R_EXP(lx_log + wz);
    *wy = wn * R_EXP(ly_log + wz);
}

/* sampler_linear_clamp: texel = uv*size - 0.5, clamp-to-edge, exact weights */
static v3 bilinear3(const real* tex, int w, int h, real u, real v) {
    real tx = u * (real)w - RC(0.5), ty = v * (real)h - RC(0.5);
    real x0f = R_FLOOR(tx), y0f = R_FLOOR(ty);
    real fx = tx - x0f, fy = ty - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    v3 c00 = ld3(tex, (size_t)ya * w + xa), c10 = ld3(tex, (size_t)ya * w + xb);
    v3 c01 = ld3(tex, (size_t)yb * w + xa), c11 = ld3(tex, (size_t)yb * w + xb);
    v3 top = add3(scale3(c00, RC(1.0) - fx), scale3(c10, fx));
    v3 bot = add3(scale3(c01, RC(1.0) - fx), scale3(c11, fx));
    return add3(scale3(top, RC(1.0) - fy), scale3(bot, fy));
}

static inline real dsr_weight(v3 c, v3 mean, v3 dev) {
    real dist = lum(div3(abs3(sub3(c, mean)), dev));
    return R_EXP2(RC(-10.0) * dist);
}

void oracle_denoise_temporal(const float* fpf, int w, int h, const real* depth, const uint32_t* gb,
                             const real* diff_src, const real* spec_src, const real* ray_len, const real* prev_depth,
                             const uint32_t* prev_gb, const real* diff_hist, const real* spec_hist,
                             const real* mom_hist, const uint8_t* len_hist, const real* diff_fast_hist,
                             const real* spec_fast_hist, int has_history, int fast_history, real* o_diff,
                             real* o_spec, real* o_mom, uint8_t* o_len, real* o_diff_fast, real* o_spec_fast) {
    ofp fp;
    o_load_fp(fpf, &fp);
    const size_t plane = (size_t)w * h;
    const real bsx = (real)w, bsy = (real)h;
#pragma omp parallel for schedule(static)
    for (int py = 0; py < h; ++py) {
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real d = depth[i];
            if (d == RC(0.0)) {
                o_len[i] = 0;
                st3(o_diff, i, V3s(0));
                st3(o_spec, i, V3s(0));
                st2(o_mom, i, V2(0, 0));
                if (fast_history) {
                    st3(o_diff_fast, i, V3s(0));
                    st3(o_spec_fast, i, V3s(0));
                }
                continue;
            }
            real sux = ((real)px + RC(0.5)) / bsx, suy = ((real)py + RC(0.5)) / bsy;
            v3 nrm = o_gb_normal(gb[plane + i]);
            real rough = o_gb_rough(gb[plane + i]);
            real fwz = o_gb_fwz(gb[2 * plane + i]), fwn = o_gb_fwn(gb[2 * plane + i]);

            v3 m1 = V3s(0), m2 = V3s(0);
            for (int x = -1; x <= 1; x++)
                for (int y = -1; y <= 1; y++) {
                    int cx = clampi(px + x, 0, w - 1), cy = clampi(py + y, 0, h - 1);
                    v3 s = ld3(spec_src, (size_t)cy * w + cx);
                    m1 = add3(m1, s);
                    m2 = add3(m2, mul3(s, s));
                }
            m1 = scale3(m1, RC(1.0) / RC(9.0));
            m2 = scale3(m2, RC(1.0) / RC(9.0));

            v3 vsh = V3s(0), vsfh = V3s(0);
            v4 ph = o_mul44(fp.inv_view_proj, V4(sux * RC(2.0) - RC(1.0), suy * RC(2.0) - RC(1.0), d, RC(1.0)));
            v3 pos = V3(ph.x / ph.w, ph.y / ph.w, ph.z / ph.w);
            {
                v3 vdir = norm3(sub3(pos, o_xyz(fp.view_pos)));
                v3 vpos = add3(pos, scale3(vdir, ray_len[i]));
                v4 hr = o_mul44(fp.prev_view_proj, V4(vpos.x, vpos.y, vpos.z, RC(1.0)));
                real nx = hr.x / hr.w - fp.jitter[2], ny = hr.y / hr.w - fp.jitter[3];
                real vu = (nx + fp.jitter[0]) * RC(0.5) + RC(0.5), vv = (ny + fp.jitter[1]) * RC(0.5) + RC(0.5);
                if (vu == sat(vu) && vv == sat(vv) && has_history) {
                    vsh = bilinear3(spec_hist, w, h, vu, vv);
                    if (fast_history) vsfh = bilinear3(spec_fast_hist, w, h, vu, vv);
                }
            }
            real mvx, mvy, reproj_depth;
            {
                v4 hp = o_mul44(fp.view_proj, V4(pos.x, pos.y, pos.z, RC(1.0)));
                real nx = hp.x / hp.w - fp.jitter[0], ny = hp.y / hp.w - fp.jitter[1];
                v4 hr = o_mul44(fp.prev_view_proj, V4(pos.x, pos.y, pos.z, RC(1.0)));
                real rx = hr.x / hr.w - fp.jitter[2], ry = hr.y / hr.w - fp.jitter[3];
                mvx = (nx - rx) * RC(0.5);
                mvy = (ny - ry) * RC(0.5);
                reproj_depth = hr.z / hr.w;
            }
            real hu = sux - mvx, hv = suy - mvy;
            v3 dh = V3s(0), sh = V3s(0), dfh = V3s(0), sfh = V3s(0);
            v2 mh = V2(0, 0);
            real hl = RC(0.0);
            int valid = (hu == sat(hu)) && (hv == sat(hv)) && has_history;
            if (valid) {
                real pcx = hu * bsx - RC(0.5), pcy = hv * bsy - RC(0.5);
                real frx = pcx - R_FLOOR(pcx), fry = pcy - R_FLOOR(pcy);
                int bx = (int)pcx, by = (int)pcy;
                real dtol = fwz * RC(4.0) + RC(1e-4);
                real ntol = fwn * RC(16.0) + RC(1e-2);
                real ntol2 = ntol * ntol;
                real b = fry, r = frx, t = RC(1.0) - b, l = RC(1.0) - r;
                real wts[2][2] = {{t * l, t * r}, {b * l, b * r}};
                v3 dsum = V3s(0), ssum = V3s(0), dfs = V3s(0), sfs = V3s(0);
                v2 msum = V2(0, 0);
                real lsum = 0, wsum = 0;
                for (int x = 0; x < 2; ++x)
                    for (int y = 0; y < 2; ++y) {
                        int cx = bx + x, cy = by + y;
                        if (cx < 0 || cy < 0 || cx > w - 1 || cy > h - 1) continue;
                        size_t j = (size_t)cy * w + cx;
                        v3 nd = sub3(o_gb_normal(prev_gb[plane + j]), nrm);
                        if (R_FABS(prev_depth[j] - reproj_depth) < dtol && dot3(nd, nd) < ntol2) {
                            real wt = wts[y][x];
                            wsum += wt;
                            dsum = add3(dsum, scale3(ld3(diff_hist, j), wt));
                            ssum = add3(ssum, scale3(ld3(spec_hist, j), wt));
                            v2 mm = ld2(mom_hist, j);
                            msum = V2(msum.x + wt * mm.x, msum.y + wt * mm.y);
                            lsum += wt * o_len_load(len_hist[j]);
                            if (fast_history) {
                                dfs = add3(dfs, scale3(ld3(diff_fast_hist, j), wt));
                                sfs = add3(sfs, scale3(ld3(spec_fast_hist, j), wt));
                            }
                        }
                    }
                valid = wsum > RC(0.00390625);
                if (valid) {
                    real inv = RC(1.0) / wsum;
                    dh = scale3(dsum, inv);
                    sh = scale3(ssum, inv);
                    mh = V2(msum.x * inv, msum.y * inv);
                    hl = lsum * inv;
                    dfh = scale3(dfs, inv);
                    sfh = scale3(sfs, inv);
                }
            }
            if (!valid) {
                dh = sh = dfh = sfh = V3s(0);
                mh = V2(0, 0);
                hl = 0;
            }
            hl += O_HIST_UNIT;
            v3 ds = ld3(diff_src, i), ss = ld3(spec_src, i);
            real lx = lum(ds), ly = lum(ss);
            const real BF_DIFF = RC(1.0) / RC(20.0), BF_FAST = RC(1.0) / RC(3.0);
            real blur = R_SQRT(sat(rough * RC(6.0)));
            real BF_SPEC = lerp1(RC(7.0) / RC(8.0), BF_DIFF, blur);
            real BF_SPEC_FAST = lerp1(RC(1.0), BF_FAST, blur);
            real la = O_HIST_UNIT / hl;
            real ax = R_FMAX(BF_DIFF, la), ay = R_FMAX(BF_SPEC, la);
            v3 dfilt = lerp3(dh, ds, ax);
            v2 mom = V2(lerp1(mh.x, lx * lx, ax), lerp1(mh.y, ly * ly, ay));
            real afx = R_FMAX(BF_FAST, la), afy = R_FMAX(BF_SPEC_FAST, la);
            v3 dfilt_fast = lerp3(dfh, ds, afx);
            v3 dev = abs3(sub3(m2, mul3(m1, m1)));
            dev = V3(R_FMAX(R_SQRT(dev.x), RC(1e-19)), R_FMAX(R_SQRT(dev.y), RC(1e-19)), R_FMAX(R_SQRT(dev.z), RC(1e-19)));
            real wv = dsr_weight(vsh, m1, dev), wh = dsr_weight(sh, m1, dev);
            v3 dual = add3(scale3(vsh, wv), scale3(sh, wh));
            dual = divs3(dual, R_FMAX(wv + wh, RC(1e-37)));
            v3 sfilt = lerp3(dual, ss, ay);
            st3(o_diff, i, dfilt);
            st3(o_spec, i, sfilt);
            st2(o_mom, i, mom);
            o_len[i] = o_len_store(hl);
            if (fast_history) {
                real wvf = dsr_weight(vsfh, m1, dev), whf = dsr_weight(sfh, m1, dev);
                v3 dualf = add3(scale3(vsfh, wvf), scale3(sfh, whf));
                dualf = divs3(dualf, R_FMAX(wvf + whf, RC(1e-37)));
                st3(o_diff_fast, i, dfilt_fast);
                st3(o_spec_fast, i, lerp3(dualf, ss, afy));
            }
        }
    }
}

/* history length of the 2x2 quad aligned at even coordinates (QuadReadAcross) */
static inline real quad_max_len(const uint8_t* len, int w, int h, int px, int py) {
    int qx = px & ~1, qy = py & ~1;
    real m = RC(0.0);
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            int x = qx + dx, y = qy + dy;
            real v = (x < w && y < h) ? o_len_load(len[(size_t)y * w + x]) : RC(0.0);
            m = R_FMAX(m, v);
        }
    return m;
}

void oracle_denoise_disocclusion(int w, int h, const real* depth, const uint32_t* gb, const uint8_t* len_tex,
                                 const real* diff, const real* spec, const real* mom, real* o_diff, real* o_spec,
                                 real* o_mom) {
    const size_t plane = (size_t)w * h;
#pragma omp parallel for schedule(static)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real len = o_len_load(len_tex[i]);
            if (len == RC(0.0)) {
                st3(o_diff, i, V3s(0));
                st3(o_spec, i, V3s(0));
                st2(o_mom, i, V2(0, 0));
                continue;
            }
            len = quad_max_len(len_tex, w, h, px, py);
            if (len < RC(4.0) * O_HIST_UNIT) {
                edge_stop e;
                memset(&e, 0, sizeof(e));
                e.depth_center = depth[i];
                e.fwidth_z = o_gb_fwz(gb[2 * plane + i]);
                e.normal_center = o_gb_normal(gb[plane + i]);
                e.sigma_z = RC(5.0);
                e.log2_sigma_n = 5;
                real wsx = 0, wsy = 0;
                v3 dsum = V3s(0), ssum = V3s(0);
                v2 msum = V2(0, 0);
                int stride = (int)((RC(14.0) * O_HIST_UNIT) / len);
                for (int y = -2; y <= 2; y++)
                    for (int x = -2; x <= 2; x++) {
                        int ox = x * stride, oy = y * stride;
                        int cx = px + ox, cy = py + oy;
                        if (cx < 0 || cy < 0 || cx > w - 1 || cy > h - 1) continue;
                        size_t j = (size_t)cy * w + cx;
                        real fox = (real)ox, foy = (real)oy, wx, wy;
                        edge_weight(&e, R_SQRT(fox * fox + foy * foy), depth[j], o_gb_normal(gb[plane + j]), 0, 0, &wx,
                                    &wy);
                        v2 mm = ld2(mom, j);
                        wsx += wx;
                        wsy += wy;
                        dsum = add3(dsum, scale3(ld3(diff, j), wx));
                        ssum = add3(ssum, scale3(ld3(spec, j), wy));
                        msum = V2(msum.x + mm.x * wx, msum.y + mm.y * wy);
                    }
                st3(o_diff, i, divs3(dsum, wsx));
                st3(o_spec, i, divs3(ssum, wsy));
                st2(o_mom, i, V2(msum.x / wsx, msum.y / wsy));
            } else {
                st3(o_diff, i, ld3(diff, i));
                st3(o_spec, i, ld3(spec, i));
                st2(o_mom, i, ld2(mom, i));
            }
        }
}

void oracle_denoise_fallback(int w, int h, const real* depth, const uint32_t* gb, const uint8_t* len_tex,
                             const real* diff, const real* spec, const real* mom, real* o_diff, real* o_spec,
                             real* o_var) {
    const size_t plane = (size_t)w * h;
#pragma omp parallel for schedule(static)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            real len = o_len_load(len_tex[i]);
            if (len == RC(0.0)) {
                st3(o_diff, i, V3s(0));
                st3(o_spec, i, V3s(0));
                st2(o_var, i, V2(0, 0));
                continue;
            }
            len = quad_max_len(len_tex, w, h, px, py);
            v3 dv, sv;
            v2 sm;
            if (len < RC(4.0) * O_HIST_UNIT) {
                edge_stop e;
                memset(&e, 0, sizeof(e));
                e.depth_center = depth[i];
                e.fwidth_z = o_gb_fwz(gb[2 * plane + i]) * RC(3.0);
                e.normal_center = o_gb_normal(gb[plane + i]);
                e.sigma_z = RC(1.0);
                e.log2_sigma_n = 7;
                real wsx = 0, wsy = 0;
                v3 dsum = V3s(0), ssum = V3s(0);
                v2 msum = V2(0, 0);
                for (int y = -3; y <= 3; y++)
                    for (int x = -3; x <= 3; x++) {
                        int cx = px + x, cy = py + y;
                        if (cx < 0 || cy < 0 || cx > w - 1 || cy > h - 1) continue;
                        size_t j = (size_t)cy * w + cx;
                        real fx = (real)x, fy = (real)y, wx, wy;
                        edge_weight(&e, R_SQRT(fx * fx + fy * fy), depth[j], o_gb_normal(gb[plane + j]), 0, 0, &wx, &wy);
                        v2 mm = ld2(mom, j);
                        wsx += wx;
                        wsy += wx;
                        dsum = add3(dsum, scale3(ld3(diff, j), wx));
                        ssum = add3(ssum, scale3(ld3(spec, j), wx));
                        msum = V2(msum.x + mm.x * wx, msum.y + mm.y * wx);
                    }
                dv = divs3(dsum, wsx);
                sv = divs3(ssum, wsy);
                sm = V2(msum.x / wsx, msum.y / wsy);
            } else {
                dv = ld3(diff, i);
                sv = ld3(spec, i);
                sm = ld2(mom, i);
            }
            real l1x = lum(dv), l1y = lum(sv);
            v2 var = V2(R_FMAX(sm.x - l1x * l1x, RC(0.0)), R_FMAX(sm.y - l1y * l1y, RC(0.0)));
            real boost = R_FMAX(RC(1.0), (RC(4.0) * O_HIST_UNIT) / len);
            st3(o_diff, i, dv);
            st3(o_spec, i, sv);
            st2(o_var, i, V2(var.x * boost, var.y * boost));
        }
}

static inline v3 ycocg(v3 c) {
    return V3((c.x * RC(0.25) + c.y * RC(0.5)) + c.z * RC(0.25), c.x * RC(0.5) - c.z * RC(0.5),
              (c.x * RC(-0.25) + c.y * RC(0.5)) - c.z * RC(0.25));
}
static inline v3 ycocg_inv(v3 c) {
    real t = c.x - c.z;
    return V3(t + c.y, c.x + c.z, t - c.y);
}
static inline v3 clamp3v(v3 v, v3 lo, v3 hi) {
    return V3(R_FMIN(R_FMAX(v.x, lo.x), hi.x), R_FMIN(R_FMAX(v.y, lo.y), hi.y), R_FMIN(R_FMAX(v.z, lo.z), hi.z));
}

/* in place on rw_diff / rw_spec */
void oracle_denoise_fast_clamp(int w, int h, const real* diff_fast, const real* spec_fast, real* rw_diff,
                               real* rw_spec) {
    size_t n = (size_t)w * h;
    real* od = (real*)malloc(sizeof(real) * 3 * n);
    real* os = (real*)malloc(sizeof(real) * 3 * n);
#pragma omp parallel for schedule(static)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            v3 dm1 = V3s(0), dm2 = V3s(0), sm1 = V3s(0), sm2 = V3s(0);
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    size_t j = (size_t)clampi(py + y, 0, h - 1) * w + clampi(px + x, 0, w - 1);
                    v3 d = ycocg(ld3(diff_fast, j)), s = ycocg(ld3(spec_fast, j));
                    dm1 = add3(dm1, d);
                    dm2 = add3(dm2, mul3(d, d));
                    sm1 = add3(sm1, s);
                    sm2 = add3(sm2, mul3(s, s));
                }
            dm1 = divs3(dm1, RC(9.0));
            dm2 = divs3(dm2, RC(9.0));
            sm1 = divs3(sm1, RC(9.0));
            sm2 = divs3(sm2, RC(9.0));
            v3 dd = sub3(dm2, mul3(dm1, dm1)), sd = sub3(sm2, mul3(sm1, sm1));
            dd = scale3(V3(R_SQRT(dd.x), R_SQRT(dd.y), R_SQRT(dd.z)), RC(0.75));
            sd = scale3(V3(R_SQRT(sd.x), R_SQRT(sd.y), R_SQRT(sd.z)), RC(0.75));
            v3 d = ycocg(ld3(rw_diff, i)), s = ycocg(ld3(rw_spec, i));
            d = clamp3v(d, sub3(dm1, dd), add3(dm1, dd));
            s = clamp3v(s, sub3(sm1, sd), add3(sm1, sd));
            st3(od, i, ycocg_inv(d));
            st3(os, i, ycocg_inv(s));
        }
    memcpy(rw_diff, od, sizeof(real) * 3 * n);
    memcpy(rw_spec, os, sizeof(real) * 3 * n);
    free(od);
    free(os);
}

void oracle_denoise_atrous(int w, int h, int step, const real* depth, const uint32_t* gb, const real* diff,
                           const real* spec, const real* var, real* o_diff, real* o_spec, real* o_var) {
    const size_t plane = (size_t)w * h;
    const real kw[3] = {RC(1.0), RC(2.0) / RC(3.0), RC(1.0) / RC(6.0)};
#pragma omp parallel for schedule(static)
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            v3 dc = ld3(diff, i), sc = ld3(spec, i);
            v2 vc = ld2(var, i);
            edge_stop e;
            memset(&e, 0, sizeof(e));
            e.depth_center = depth[i];
            e.fwidth_z = o_gb_fwz(gb[2 * plane + i]);
            e.normal_center = o_gb_normal(gb[plane + i]);
            e.sigma_z = RC(1.0);
            e.log2_sigma_n = 7;
            e.use_lumi = 1;
            e.lc[0] = lum(dc);
            e.lc[1] = lum(sc);
            e.lstd[0] = R_SQRT(vc.x);
            e.lstd[1] = R_SQRT(vc.y);
            v3 dsum = dc, ssum = sc;
            v2 vsum = vc;
            real wsx = RC(1.0), wsy = RC(1.0);
            for (int y = -2; y <= 2; y++) {
                real kwy = kw[y < 0 ? -y : y];
                for (int x = -2; x <= 2; x++) {
                    int ox = x * step, oy = y * step;
                    int cx = px + ox, cy = py + oy;
                    int valid = !(cx < 0 || cy < 0 || cx > w - 1 || cy > h - 1);
                    int not_center = (x != 0) && (y != 0); /* svgf_atrous_filter.hlsl:54 */
                    if (!(valid && not_center)) continue;
                    size_t j = (size_t)cy * w + cx;
                    real kxy = kwy * kw[x < 0 ? -x : x];
                    v3 dv = ld3(diff, j), sv = ld3(spec, j);
                    real fox = (real)ox, foy = (real)oy, ex, ey;
                    edge_weight(&e, R_SQRT(fox * fox + foy * foy), depth[j], o_gb_normal(gb[plane + j]), lum(dv), lum(sv),
                                &ex, &ey);
                    real wx = kxy * ex, wy = kxy * ey;
                    v2 vv = ld2(var, j);
                    dsum = add3(dsum, scale3(dv, wx));
                    ssum = add3(ssum, scale3(sv, wy));
                    vsum = V2(vsum.x + (wx * wx) * vv.x, vsum.y + (wy * wy) * vv.y);
                    wsx += wx;
                    wsy += wy;
                }
            }
            st3(o_diff, i, divs3(dsum, wsx));
            st3(o_spec, i, divs3(ssum, wsy));
            st2(o_var, i, V2(vsum.x / (wsx * wsx), vsum.y / (wsy * wsy)));
        }
}
