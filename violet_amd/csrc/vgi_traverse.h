/*
 * vgi_traverse.h — device-side BVH traversal (closest hit and any hit).
 *
 * One lane per ray over the 4-wide BVH.  The traversal stack lives in LDS,
 * interleaved across the 64 lanes of the wave (entry k of lane l at
 * [k*64 + l]) so that a push or pop by all lanes touches 64 consecutive
 * dwords (no bank conflicts).  At a node the hit children are ordered by
 * entry distance; the nearest is visited next and the others are pushed
 * farthest first.  Boxes are tested with the slab method, one fused
 * multiply-add per plane (conservative within the boxes' padding, see
 * slab_nf); NaNs from 0*inf drop out of the max/min, which only widens the
 * interval.
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

// Entry distance of a box given each axis' near and far plane for this ray
// (near = lo where the direction component is >= 0, else hi), or +inf if
// missed.  A plane's distance is one fused multiply-add, P * idir + oid with
// oid = -o * idir: against the exact (P - o) * idir it errs by at most an
// ulp of o * idir plus the final rounding, i.e. a world distance of about
// ulp(|o|), far inside the 2^-16 relative padding of every child box, so
// the test stays conservative.  A NaN from 0 * inf (origin on a plane of a
// box the ray runs parallel to) drops out of the max/min, which only widens
// the interval.
__device__ __forceinline__ float slab_nf(const RayInv& r, f3 oid, float nx, float fx, float ny, float fy, float nz,
                                         float fz, float tmin, float tmax) {
    const float tnx = __builtin_fmaf(nx, r.idir.x, oid.x), tfx = __builtin_fmaf(fx, r.idir.x, oid.x);
    const float tny = __builtin_fmaf(ny, r.idir.y, oid.y), tfy = __builtin_fmaf(fy, r.idir.y, oid.y);
    const float tnz = __builtin_fmaf(nz, r.idir.z, oid.z), tfz = __builtin_fmaf(fz, r.idir.z, oid.z);
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
    // byte offset of each axis' near plane in a node (float4 slot lo = 2 axis,
    // hi = 2 axis + 1; the far plane is the other one, offset ^ 16)
    const uint32_t nx = r.idir.x >= 0.0f ? 0u : 16u, ny = r.idir.y >= 0.0f ? 32u : 48u, nz = r.idir.z >= 0.0f ? 64u : 80u;
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
            const Vf4 nrx = ld_buf16(nodes, nb + nx), frx = ld_buf16(nodes, nb + (nx ^ 16u));
            const Vf4 nry = ld_buf16(nodes, nb + ny), fry = ld_buf16(nodes, nb + (ny ^ 16u));
            const Vf4 nrz = ld_buf16(nodes, nb + nz), frz = ld_buf16(nodes, nb + (nz ^ 16u));
            const Vf4 lk = ld_buf16(nodes, nb + 96u);
            // -o * idir per node: kept out of loop-invariant hoisting (three
            // more live registers across the traversal cost more than three
            // multiplies per node)
            f3 oo = r.o;
            asm volatile("" : "+v"(oo.x), "+v"(oo.y), "+v"(oo.z));
            const f3 oid = mk3(-(oo.x * r.idir.x), -(oo.y * r.idir.y), -(oo.z * r.idir.z));
            float t0 = slab_nf(r, oid, nrx.x, frx.x, nry.x, fry.x, nrz.x, frz.x, tmin, h.t);
            float t1 = slab_nf(r, oid, nrx.y, frx.y, nry.y, fry.y, nrz.y, frz.y, tmin, h.t);
            float t2 = slab_nf(r, oid, nrx.z, frx.z, nry.z, fry.z, nrz.z, frz.z, tmin, h.t);
            float t3 = slab_nf(r, oid, nrx.w, frx.w, nry.w, fry.w, nrz.w, frz.w, tmin, h.t);
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
