/*
 * oracle_trace.c — ORACLE (test infrastructure only).  Closest-hit and
 * any-hit ray queries restated from what violet asks of TraceRay
 * (restir/raytrace.inc.hlsl:31-56,67-82: opaque geometry, two-sided,
 * t in [TMin, TMax], shadow rays accept the first hit), with the
 * Moller-Trumbore test in REAL precision and ties broken to the lowest
 * triangle id.  Two independent drivers:
 *   oracle_trace_brute  — every ray against every triangle (small scenes);
 *   oracle_trace_bvh    — an object-median binary BVH built here in double
 *                         (different structure from the product's SAH BVH;
 *                         boxes padded so pruning is conservative).
 * Rays: 8 floats {o, tmin, d, tmax}; hits: {t, u, v} as REAL + id.
 */
#include <float.h>

#include "oracle_common.h"

static inline int o_tri_test(v3 o, v3 d, v3 v0, v3 e1, v3 e2, real tmin, real* t, real* u, real* v) {
    v3 pvec = cross3(d, e2);
    real det = dot3(e1, pvec);
    if (det == RC(0.0)) return 0;
    real inv = RC(1.0) / det;
    v3 tvec = sub3(o, v0);
    *u = dot3(tvec, pvec) * inv;
    if (!(*u >= RC(0.0) && *u <= RC(1.0))) return 0;
    v3 qvec = cross3(tvec, e1);
    *v = dot3(d, qvec) * inv;
    if (!(*v >= RC(0.0) && *u + *v <= RC(1.0))) return 0;
    *t = dot3(e2, qvec) * inv;
    return *t >= tmin;
}

static inline void tri_verts(const float* P, const uint32_t* I, uint32_t t, v3* v0, v3* e1, v3* e2) {
    const float* a = P + 3 * I[3 * t];
    const float* b = P + 3 * I[3 * t + 1];
    const float* c = P + 3 * I[3 * t + 2];
    /* edges formed in single precision from the stored f32 positions */
    *v0 = V3(a[0], a[1], a[2]);
    *e1 = V3((real)(b[0] - a[0]), (real)(b[1] - a[1]), (real)(b[2] - a[2]));
    *e2 = V3((real)(c[0] - a[0]), (real)(c[1] - a[1]), (real)(c[2] - a[2]));
#if !defined(ORACLE_F32)
    *e1 = V3((real)b[0] - a[0], (real)b[1] - a[1], (real)b[2] - a[2]);
    *e2 = V3((real)c[0] - a[0], (real)c[1] - a[1], (real)c[2] - a[2]);
#endif
}

typedef struct {
    real t, u, v;
    uint32_t id;
} ohit;

static inline void consider(ohit* h, v3 o, v3 d, real tmin, const float* P, const uint32_t* I, uint32_t t) {
    v3 v0, e1, e2;
    tri_verts(P, I, t, &v0, &e1, &e2);
    real tt, u, v;
    if (o_tri_test(o, d, v0, e1, e2, tmin, &tt, &u, &v)) {
        if (tt < h->t || (tt == h->t && t < h->id)) {
            h->t = tt;
            h->u = u;
            h->v = v;
            h->id = t;
        }
    }
}

void oracle_trace_brute(const float* P, const uint32_t* I, uint32_t nt, const float* rays, uint32_t n, int any_hit,
                        real* thv, uint32_t* ids) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < (int64_t)n; ++r) {
        const float* ry = rays + 8 * r;
        v3 o = V3(ry[0], ry[1], ry[2]), d = V3(ry[4], ry[5], ry[6]);
        ohit h = {(real)ry[7], 0, 0, 0xFFFFFFFFu};
        for (uint32_t t = 0; t < nt; ++t) {
            consider(&h, o, d, (real)ry[3], P, I, t);
            if (any_hit && h.id != 0xFFFFFFFFu) break;
        }
        thv[3 * r] = h.t;
        thv[3 * r + 1] = h.u;
        thv[3 * r + 2] = h.v;
        ids[r] = h.id;
    }
}

/* ---- independent median-split BVH (double boxes) ---- */
typedef struct {
    double lo[3], hi[3];
    int32_t left, right; /* children, or -1 */
    uint32_t first, count;
} onode;

typedef struct {
    onode* nodes;
    uint32_t n, cap;
    uint32_t* tri;
    const float* P;
    const uint32_t* I;
    double* cent;
} obvh;

static void tri_box(const obvh* b, uint32_t t, double* lo, double* hi) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = DBL_MAX;
        hi[a] = -DBL_MAX;
    }
    for (int k = 0; k < 3; ++k) {
        const float* p = b->P + 3 * b->I[3 * t + k];
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a]) lo[a] = p[a];
            if (p[a] > hi[a]) hi[a] = p[a];
        }
    }
}

static int g_axis;
static const double* g_cent;
static int cmp_cent(const void* x, const void* y) {
    double a = g_cent[3 * *(const uint32_t*)x + g_axis], c = g_cent[3 * *(const uint32_t*)y + g_axis];
    return a < c ? -1 : (a > c ? 1 : 0);
}

static int32_t obvh_build(obvh* b, uint32_t first, uint32_t count) {
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->nodes = (onode*)realloc(b->nodes, sizeof(onode) * b->cap);
    }
    int32_t id = (int32_t)b->n++;
    onode nd;
    for (int a = 0; a < 3; ++a) {
        nd.lo[a] = DBL_MAX;
        nd.hi[a] = -DBL_MAX;
    }
    double clo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, chi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint32_t i = first; i < first + count; ++i) {
        double lo[3], hi[3];
        tri_box(b, b->tri[i], lo, hi);
        for (int a = 0; a < 3; ++a) {
            if (lo[a] < nd.lo[a]) nd.lo[a] = lo[a];
            if (hi[a] > nd.hi[a]) nd.hi[a] = hi[a];
            double c = b->cent[3 * b->tri[i] + a];
            if (c < clo[a]) clo[a] = c;
            if (c > chi[a]) chi[a] = c;
        }
    }
    double s = 1.0;
    for (int a = 0; a < 3; ++a) {
        if (fabs(nd.lo[a]) > s) s = fabs(nd.lo[a]);
        if (fabs(nd.hi[a]) > s) s = fabs(nd.hi[a]);
    }
    for (int a = 0; a < 3; ++a) {
        nd.lo[a] -= s * 1e-5;
        nd.hi[a] += s * 1e-5;
    }
    nd.first = first;
    nd.count = count;
    nd.left = nd.right = -1;
    if (count > 8) {
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
        g_axis = axis;
        g_cent = b->cent;
        qsort(b->tri + first, count, sizeof(uint32_t), cmp_cent);
        uint32_t half = count / 2;
        int32_t l = obvh_build(b, first, half);
        int32_t r = obvh_build(b, first + half, count - half);
        nd.left = l;
        nd.right = r;
        nd.count = 0;
    }
    b->nodes[id] = nd;
    return id;
}

static int box_hit(const onode* nd, const double* o, const double* idir, double tmin, double tmax) {
    double tn = tmin, tf = tmax;
    for (int a = 0; a < 3; ++a) {
        double t0 = (nd->lo[a] - o[a]) * idir[a], t1 = (nd->hi[a] - o[a]) * idir[a];
        if (isnan(t0) || isnan(t1)) continue;
        double lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
        if (lo > tn) tn = lo;
        if (hi < tf) tf = hi;
    }
    return tn <= tf * (1.0 + 1e-9) + 1e-12;
}

void oracle_trace_bvh(const float* P, const uint32_t* I, uint32_t nt, const float* rays, uint32_t n, int any_hit,
                      real* thv, uint32_t* ids) {
    obvh b;
    memset(&b, 0, sizeof(b));
    b.P = P;
    b.I = I;
    b.tri = (uint32_t*)malloc(sizeof(uint32_t) * nt);
    b.cent = (double*)malloc(sizeof(double) * 3 * nt);
    for (uint32_t t = 0; t < nt; ++t) {
        b.tri[t] = t;
        double lo[3], hi[3];
        tri_box(&b, t, lo, hi);
        for (int a = 0; a < 3; ++a) b.cent[3 * t + a] = 0.5 * (lo[a] + hi[a]);
    }
    obvh_build(&b, 0, nt);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < (int64_t)n; ++r) {
        const float* ry = rays + 8 * r;
        v3 o = V3(ry[0], ry[1], ry[2]), d = V3(ry[4], ry[5], ry[6]);
        double od[3] = {ry[0], ry[1], ry[2]};
        double idir[3] = {1.0 / ry[4], 1.0 / ry[5], 1.0 / ry[6]};
        ohit h = {(real)ry[7], 0, 0, 0xFFFFFFFFu};
        int32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const onode* nd = &b.nodes[stack[--sp]];
            if (!box_hit(nd, od, idir, ry[3], (double)h.t)) continue;
            if (nd->left < 0) {
                for (uint32_t i = nd->first; i < nd->first + nd->count; ++i) consider(&h, o, d, (real)ry[3], P, I, b.tri[i]);
                if (any_hit && h.id != 0xFFFFFFFFu) break;
            } else {
                stack[sp++] = nd->left;
                stack[sp++] = nd->right;
            }
        }
        thv[3 * r] = h.t;
        thv[3 * r + 1] = h.u;
        thv[3 * r + 2] = h.v;
        ids[r] = h.id;
    }
    free(b.nodes);
    free(b.tri);
    free(b.cent);
}
