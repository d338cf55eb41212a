/*
 * vgi_traverse.h — device-side BVH traversal (closest hit and any hit).
 *
 * One lane per ray over the 4-wide BVH of 64-byte quantised nodes.  The
 * traversal stack lives in LDS, interleaved across the 64 lanes of the wave
 * (entry k of lane l at [k*64 + l]) so that a push or pop by all lanes
 * touches 64 consecutive dwords (no bank conflicts).  At a node the hit
 * children are ordered by entry distance; the nearest is visited next and
 * the others are pushed farthest first.  Boxes are tested with the slab
 * method on the node's 8-bit grid: per axis the ray's distance to grid line
 * q is q * (step / d) + (org - o) / d.  One byte permute turns two planes'
 * bytes into binary16 subnormals q * 2^-24 (exact), and one mixed-precision
 * fused multiply-add per plane (v_fma_mix_f32, a binary16 operand widened
 * exactly) multiplies them by the node's stored step * 2^24: the same
 * q * step * |1/d| product as a byte-to-float conversion followed by an FMA,
 * in 1.5 instructions per plane instead of 2 (conservative within the
 * boxes' padding, see slab_q).
 */
#pragma once
#include "vgi_bvh.h"

namespace vgi {

// Explicit address spaces for the traversal's memory traffic: the scene
// arrays are global memory and the stack is LDS, also when the scene arrives
// through a pointer loaded from memory (the fused multi-view kernels read it
// from a job table, where the compiler cannot infer the address space and
// would otherwise emit flat accesses that also count against lgkmcnt).
typedef __attribute__((address_space(3))) int32_t* LdsStack;
typedef float Vf4 __attribute__((ext_vector_type(4)));
// 16-byte load of scene data (element k of a float4 array) from global memory.
__device__ __forceinline__ Vf4 ld_global16(const void* p, int k) {
    return ((const __attribute__((address_space(1))) Vf4*)p)[k];
}

// The node and triangle arrays are read through buffer descriptors (wave-
// uniform base and size in scalar registers, a 32-bit byte offset per
// lane): the addresses of a node's seven 16-byte loads then cost one 32-bit
// add each instead of 64-bit address arithmetic, and fewer registers.
// Descriptor word 3 of the gfx9 family for untyped dword access; loads
// beyond `bytes` return zeros.  vgi_scene_create keeps both arrays below
// 2^31 bytes.
typedef __amdgpu_buffer_rsrc_t BufRsrc;
__device__ __forceinline__ BufRsrc scene_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ Vf4 ld_buf16(BufRsrc r, uint32_t byte_off) {
    return __builtin_bit_cast(Vf4, __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0));
}

struct RayInv {
    f3 o, d, idir;
};

// A direction component of magnitude below 1e-20 gives a finite inverse of
// +-1e20 (slab_nf evaluates P * idir - o * idir, where an infinite inverse
// would turn every plane of that axis into inf - inf = NaN and stop the
// axis from culling anything).  The slab interval of such an axis is then
// (-huge, +huge) for an origin between the planes and empty otherwise, as
// for an exact zero.
__device__ __forceinline__ float safe_dir(float v) { return fabsf(v) < 1e-20f ? copysignf(1e-20f, v) : v; }

__device__ __forceinline__ RayInv make_ray(f3 o, f3 d) {
    RayInv r;
    r.o = o;
    r.d = d;
    r.idir = mk3(1.0f / safe_dir(d.x), 1.0f / safe_dir(d.y), 1.0f / safe_dir(d.z));
    return r;
}

// Bytes 2j and 2j + 1 of a plane word (children 2j and 2j + 1) as the low
// and high halves of a binary16 pair: byte q becomes the subnormal
// q * 2^-24, exactly (one v_perm_b32 placing a zero byte above each).
// Kernels keep the default binary16 denormal mode (inputs and outputs not
// flushed), which the mixed FMA below relies on.
// The permute selectors: held in scalar registers (an opaque scalar move,
// so the compiler does not rematerialise the literals into two vector
// registers kept live across the traversal).
struct PairSel {
    uint32_t s01, s23;
};
__device__ __forceinline__ PairSel pair_selectors() {
    PairSel p;
    asm volatile("s_mov_b32 %0, 0x0C010C00" : "=s"(p.s01));
    asm volatile("s_mov_b32 %0, 0x0C030C02" : "=s"(p.s23));
    return p;
}
template <int j>
__device__ __forceinline__ uint32_t qpair(uint32_t w, const PairSel& sel) {
    return __builtin_amdgcn_perm(0u, w, j ? sel.s23 : sel.s01);
}
// Child k's grid line from the pairs of children 0-1 (p01) and 2-3 (p23),
// as the binary16 q * 2^-24 widened to float (folded into v_fma_mix_f32).
template <int k>
__device__ __forceinline__ float qline(uint32_t p01, uint32_t p23) {
    const uint32_t p = k < 2 ? p01 : p23;
    return (float)__builtin_bit_cast(_Float16, (uint16_t)((k & 1) ? p >> 16 : p & 0xFFFFu));
}

// A ray's near and far plane pairs of one node (children 0-1 and 2-3 per
// axis and side).
struct NodePlanes {
    uint32_t nx01, nx23, fx01, fx23, ny01, ny23, fy01, fy23, nz01, nz23, fz01, fz23;
};

// Entry distance of child k of a quantised node, or +inf if missed.  Per
// axis, the near and far planes are those of the lo or hi word (near = lo
// where the direction component is >= 0, else hi) and the distance to grid
// line q is (q * 2^-24) * a + b with a = (step * 2^24) * idir and b =
// (org - o) * idir; both scalings are by powers of two, so the product is
// q * step * idir as exactly as for step * idir.  Against the exact
// distance (org + q * step - o) * idir this errs by about an ulp of
// |org - o| in world units plus the final rounding, inside the 2^-16
// relative padding of every child box, so the test stays conservative.  An
// unused child (lo = 255, hi = 0) has its near plane beyond its far one on
// every axis.
template <int k>
__device__ __forceinline__ float slab_q(f3 a, f3 b, const NodePlanes& p, float tmin, float tmax) {
    const float tnx = __builtin_fmaf(qline<k>(p.nx01, p.nx23), a.x, b.x);
    const float tfx = __builtin_fmaf(qline<k>(p.fx01, p.fx23), a.x, b.x);
    const float tny = __builtin_fmaf(qline<k>(p.ny01, p.ny23), a.y, b.y);
    const float tfy = __builtin_fmaf(qline<k>(p.fy01, p.fy23), a.y, b.y);
    const float tnz = __builtin_fmaf(qline<k>(p.nz01, p.nz23), a.z, b.z);
    const float tfz = __builtin_fmaf(qline<k>(p.fz01, p.fz23), a.z, b.z);
    const float tn = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
    const float tf = fminf(fminf(fminf(tfx, tfy), tfz), tmax);
    return tn <= tf ? tn : INFINITY;
}

__device__ __forceinline__ void count_traversal(const SceneView& s, bool any_hit, uint32_t nodes, uint32_t tris) {
    unsigned long long* c = s.counters + 4u * ((blockIdx.x + 7u * blockIdx.y) % (unsigned)kTraceCounterSlots);
    atomicAdd(c + (any_hit ? 1 : 0), 1ull);
    atomicAdd(c + 2, (unsigned long long)nodes);
    atomicAdd(c + 3, (unsigned long long)tris);
}

// Compare-exchange of (entry distance, link) pairs without branches.
__device__ __forceinline__ void sort_pair(float& ta, int32_t& la, float& tb, int32_t& lb) {
    const bool sw = tb < ta;
    const float t0 = sw ? tb : ta, t1 = sw ? ta : tb;
    const int32_t l0 = sw ? lb : la, l1 = sw ? la : lb;
    ta = t0;
    tb = t1;
    la = l0;
    lb = l1;
}

// Traversal finished.  Never a leaf link: leaf links are ~(first << 4 | count)
// with first + count <= number of triangles < 2^27.
constexpr int32_t kTraversalDone = INT32_MIN;

// Closest hit over [tmin, tmax]; ties in t resolve to the lowest triangle id.
// `stack` points at this lane's slot 0 (stride 64 between entries).
// kCount: the instrumented build (traversal statistics) that adds this ray's
// node and triangle counts to s.counters.
//
// Loop structure (while-while with postponed leaves): lanes step through
// interior nodes, each setting aside the first leaf it reaches and going on
// down the tree, until every lane of the wave holds a leaf; then all lanes
// intersect their leaves together.  Interior-node and triangle code thus run
// with most lanes active instead of alternating per iteration.  Boxes are
// tested against the near and far plane of each axis, chosen once per ray by
// the sign of the direction.  The hits do not depend on the visiting order
// (closest t, then lowest id), only the work does.
template <bool kAnyHit, bool kCount = false>
__device__ __forceinline__ Hit traverse(const SceneView& s, f3 o, f3 d, float tmin, float tmax, LdsStack stack) {
    const RayInv r = make_ray(o, d);
    // which plane word of each axis is the near one (lo where the direction
    // component is >= 0, else hi)
    const bool px = r.idir.x >= 0.0f, py = r.idir.y >= 0.0f, pz = r.idir.z >= 0.0f;
    const PairSel sel = pair_selectors();
    const BufRsrc nodes = scene_rsrc(s.nodes, s.num_nodes * (uint32_t)sizeof(BvhNode));
    const BufRsrc tris = scene_rsrc(s.tris, s.num_tris * (uint32_t)sizeof(BvhTri));
    Hit h;
    h.t = tmax;
    h.u = h.v = 0.0f;
    h.tri = 0xFFFFFFFFu;
    int sp = 0;
    int32_t node = s.root;  // next interior node, a leaf link, or kTraversalDone
    int32_t leaf = 0;       // postponed leaf link (< 0), 0 for none
    // Each node and leaf of a tree is reached at most once (a node has at
    // most 4 children); the budget only guards against a corrupted BVH
    // turning into an endless loop.
    uint32_t budget = 5u * s.num_nodes + 8u;
    uint32_t n_nodes = 0, n_tris = 0;
    int32_t spill[kStackMaxEntries - kLdsStackEntries];
    const int depth = (int)s.stack_depth;  // <= kStackMaxEntries for every scene vgi_scene_create builds
    auto push = [&](int32_t x) {
        if (sp < (int)kLdsStackEntries)
            stack[sp * 64] = x;
        else
            spill[sp - (int)kLdsStackEntries] = x;
        ++sp;
    };
    auto pop = [&]() -> int32_t {
        if (sp == 0) return kTraversalDone;
        --sp;
        return sp < (int)kLdsStackEntries ? stack[sp * 64] : spill[sp - (int)kLdsStackEntries];
    };
    while (node != kTraversalDone || leaf != 0) {
        while (node >= 0) {
            if (budget-- == 0u) {
                node = kTraversalDone;
                leaf = 0;
                break;
            }
            if (kCount) ++n_nodes;
            const uint32_t nb = (uint32_t)node * (uint32_t)sizeof(BvhNode);
            const Vf4 g0 = ld_buf16(nodes, nb);        // org.xyz, step.x * 2^24
            const Vf4 g1 = ld_buf16(nodes, nb + 16u);  // step.yz * 2^24, q lo.x, q hi.x
            const Vf4 g2 = ld_buf16(nodes, nb + 32u);  // q lo.y, hi.y, lo.z, hi.z
            const Vf4 lk = ld_buf16(nodes, nb + 48u);
            // The links are first needed after the box tests, and left alone
            // the compiler sinks their load to that point: a second round
            // trip per node.  Using them here issues all four loads together.
            asm volatile("" ::"v"(lk));
            const f3 a = mk3(g0.w * r.idir.x, g1.x * r.idir.y, g1.y * r.idir.z);
            const f3 b = mk3((g0.x - r.o.x) * r.idir.x, (g0.y - r.o.y) * r.idir.y, (g0.z - r.o.z) * r.idir.z);
            const uint32_t qlx = __float_as_uint(g1.z), qhx = __float_as_uint(g1.w);
            const uint32_t qly = __float_as_uint(g2.x), qhy = __float_as_uint(g2.y);
            const uint32_t qlz = __float_as_uint(g2.z), qhz = __float_as_uint(g2.w);
            const uint32_t qnx = px ? qlx : qhx, qfx = px ? qhx : qlx;
            const uint32_t qny = py ? qly : qhy, qfy = py ? qhy : qly;
            const uint32_t qnz = pz ? qlz : qhz, qfz = pz ? qhz : qlz;
            const NodePlanes pl{qpair<0>(qnx, sel), qpair<1>(qnx, sel), qpair<0>(qfx, sel), qpair<1>(qfx, sel),
                                qpair<0>(qny, sel), qpair<1>(qny, sel), qpair<0>(qfy, sel), qpair<1>(qfy, sel),
                                qpair<0>(qnz, sel), qpair<1>(qnz, sel), qpair<0>(qfz, sel), qpair<1>(qfz, sel)};
            float t0 = slab_q<0>(a, b, pl, tmin, h.t);
            float t1 = slab_q<1>(a, b, pl, tmin, h.t);
            float t2 = slab_q<2>(a, b, pl, tmin, h.t);
            float t3 = slab_q<3>(a, b, pl, tmin, h.t);
            int32_t l0 = __float_as_int(lk.x), l1 = __float_as_int(lk.y);
            int32_t l2 = __float_as_int(lk.z), l3 = __float_as_int(lk.w);
            if (kAnyHit) {
                // an occlusion test may visit the hit children in any order:
                // descend into the first one, push the others unsorted
                bool have = false;
                int32_t next = kTraversalDone;
                auto take = [&](float t, int32_t l) {
                    if (t == INFINITY) return;
                    if (!have) {
                        next = l;
                        have = true;
                    } else if (sp < depth) {
                        push(l);
                    }
                };
                take(t0, l0);
                take(t1, l1);
                take(t2, l2);
                take(t3, l3);
                node = have ? next : pop();
                if (node < 0 && node != kTraversalDone && leaf == 0) {
                    leaf = node;
                    node = pop();
                }
                if (__all(leaf != 0)) break;
                continue;
            }
            // ascending entry distance (misses are +inf): a 5-exchange network
            sort_pair(t0, l0, t1, l1);
            sort_pair(t2, l2, t3, l3);
            sort_pair(t0, l0, t2, l2);
            sort_pair(t1, l1, t3, l3);
            sort_pair(t1, l1, t2, l2);
            if (t0 != INFINITY) {
                if (sp + 3 <= (int)kLdsStackEntries) {
                    // all three fit in the LDS part: write every slot and
                    // advance past the hit ones (no branch per child; a
                    // slot written for a missed child lies above the top)
                    stack[sp * 64] = l3;
                    sp += t3 != INFINITY;
                    stack[sp * 64] = l2;
                    sp += t2 != INFINITY;
                    stack[sp * 64] = l1;
                    sp += t1 != INFINITY;
                } else {
                    // the depth test always passes for a BVH from
                    // vgi_scene_create: the stack never outgrows 3 far
                    // children per level
                    if (t3 != INFINITY && sp < depth) push(l3);
                    if (t2 != INFINITY && sp < depth) push(l2);
                    if (t1 != INFINITY && sp < depth) push(l1);
                }
                node = l0;
            } else {
                node = pop();
            }
            if (node < 0 && node != kTraversalDone && leaf == 0) {
                leaf = node;
                node = pop();
            }
            if (__all(leaf != 0)) break;
        }
        while (leaf != 0) {
            if (budget-- == 0u) {
                node = kTraversalDone;
                leaf = 0;
                break;
            }
            const uint32_t enc = (uint32_t)(~leaf);
            const uint32_t first = enc >> 4, cnt = enc & 15u;
            for (uint32_t k = 0; k < cnt; ++k) {
                if (kCount) ++n_tris;
                const uint32_t tb = (first + k) * (uint32_t)sizeof(BvhTri);
                Vf4 a = ld_buf16(tris, tb), e1 = ld_buf16(tris, tb + 16u), e2 = ld_buf16(tris, tb + 32u);
                float t, u, v;
                if (intersect_tri(o, d, mk3(a.x, a.y, a.z), mk3(e1.x, e1.y, e1.z), mk3(e2.x, e2.y, e2.z), tmin, t, u,
                                  v)) {
                    uint32_t id = __float_as_uint(a.w);
                    if (t < h.t || (t == h.t && id < h.tri)) {
                        h.t = t;
                        h.u = u;
                        h.v = v;
                        h.tri = id;
                        if (kAnyHit) {
                            if (kCount) count_traversal(s, true, n_nodes, n_tris);
                            return h;
                        }
                    }
                }
            }
            leaf = 0;
            if (node < 0 && node != kTraversalDone) {
                leaf = node;
                node = pop();
            }
        }
    }
    if (kCount) count_traversal(s, kAnyHit, n_nodes, n_tris);
    return h;
}

template <bool kCount = false>
__device__ __forceinline__ Hit trace_closest(const SceneView& s, f3 o, f3 d, float tmin, float tmax, LdsStack stack) {
    return traverse<false, kCount>(s, o, d, tmin, tmax, stack);
}
// Shadow ray: true if something is hit on [tmin, tmax] (trace_shadow returns !this).
template <bool kCount = false>
__device__ __forceinline__ bool trace_any(const SceneView& s, f3 o, f3 d, float tmin, float tmax, LdsStack stack) {
    return traverse<true, kCount>(s, o, d, tmin, tmax, stack).tri != 0xFFFFFFFFu;
}

// Per-wave LDS traversal stack: SceneView::stack_depth entries x 64 lanes in
// dynamic LDS (launch with traversal_lds_bytes(scene) bytes: at most
// kLdsStackEntries entries per lane; deeper entries go to private memory).
// A ray at a 4-wide node of level d holds at most 3(d+1) stacked children, so
// stack_depth entries always suffice.
#define VGI_TRAVERSAL_STACK(name)                        \
    extern __shared__ int32_t name##_mem[];              \
    LdsStack name = (LdsStack)(name##_mem) + (threadIdx.x & 63u)

}  // namespace vgi
