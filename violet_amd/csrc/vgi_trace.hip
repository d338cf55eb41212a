/*
 * vgi_trace.hip — batch ray queries against the scene BVH
 * (TraceRay without shading; used by hosts and by the hit-id parity tests).
 */
#include <utility>
#include <vector>

#include "vgi_internal.h"
#include "vgi_traverse.h"

namespace vgi {
const SceneView& scene_view(const vgi_scene* s);

template <bool kCount>
__global__ void __launch_bounds__(64, kTraceWaves) k_trace_rays(SceneView s, const float* rays, uint32_t n, uint32_t any_hit,
                                                   uint32_t* hits) {
    VGI_TRAVERSAL_STACK(stack);
    uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + 8 * (size_t)i;
    f3 o = mk3(r[0], r[1], r[2]);
    f3 d = mk3(r[4], r[5], r[6]);
    Hit h = any_hit ? traverse<true, kCount>(s, o, d, r[3], r[7], stack) : traverse<false, kCount>(s, o, d, r[3], r[7], stack);
    uint32_t* out = hits + 4 * (size_t)i;
    out[0] = __float_as_uint(h.t);
    out[1] = __float_as_uint(h.u);
    out[2] = __float_as_uint(h.v);
    out[3] = h.tri;
}

}  // namespace vgi

using namespace vgi;

extern "C" vgi_status vgi_trace_rays(const vgi_scene* scene, const float* rays, uint32_t num_rays, uint32_t any_hit,
                                     uint32_t* hits, void* stream) {
    VGI_REQUIRE(scene && rays && hits, "vgi_trace_rays: null argument");
    if (num_rays == 0) return VGI_OK;
    hipStream_t st = (hipStream_t)stream;
    const char* name = any_hit ? "trace_rays_any" : "trace_rays_closest";
    ProfScope prof(name, st, 0.0);
    SceneView sv = scene_view(scene);
    sv.counters = trace_counters();
    const dim3 g((num_rays + 63) / 64);
    if (sv.counters)
        hipLaunchKernelGGL(k_trace_rays<true>, g, dim3(64), traversal_lds_bytes(sv), st, sv, rays, num_rays, any_hit, hits);
    else
        hipLaunchKernelGGL(k_trace_rays<false>, g, dim3(64), traversal_lds_bytes(sv), st, sv, rays, num_rays, any_hit, hits);
    trace_stats_collect(name, st);
    return check_launch("k_trace_rays");
}

// ---------------------------------------------------------------------------
// Dependent-fetch probe: every lane chases its own chain of 128-byte records
// (one random cycle through the buffer), so each step is one round trip of
// eight 16-byte loads whose address depends on the previous step, the access
// pattern of a BVH traversal's node fetches.  Launched like the trace kernels.
// ---------------------------------------------------------------------------
namespace vgi {
namespace {

__global__ void __launch_bounds__(64, kTraceWaves) k_chase(const uint4* __restrict__ nodes, uint32_t n, uint32_t steps,
                                                           uint32_t group_shift, uint32_t* __restrict__ sink) {
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    // lanes of one group (2^group_shift consecutive lanes) share a chain
    uint32_t idx = (uint32_t)(((uint64_t)(lane >> group_shift) * 2654435761ull) % n);
    uint32_t acc = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        const uint4* p = nodes + (size_t)idx * 8u;
        const uint4 v0 = p[0], v1 = p[1], v2 = p[2], v3 = p[3], v4 = p[4], v5 = p[5], v6 = p[6], v7 = p[7];
        acc += (v0.y ^ v1.x ^ v2.z ^ v3.w) + (v4.x ^ v5.y ^ v6.z ^ v7.w);
        idx = v0.x;
    }
    sink[lane] = acc + idx;
}

}  // namespace
}  // namespace vgi

extern "C" vgi_status vgi_probe_dependent_fetch(size_t bytes, uint32_t steps, uint32_t lanes_per_chain,
                                                double* steps_per_s, void* stream) {
    VGI_REQUIRE(steps_per_s && steps > 0 && steps <= 4096, "vgi_probe_dependent_fetch: bad arguments");
    VGI_REQUIRE(lanes_per_chain >= 1 && lanes_per_chain <= 64 && (lanes_per_chain & (lanes_per_chain - 1)) == 0,
                "vgi_probe_dependent_fetch: lanes_per_chain %u is not a power of two in 1..64", lanes_per_chain);
    const uint32_t group_shift = (uint32_t)__builtin_ctz(lanes_per_chain);
    const size_t n = bytes / 128u;
    VGI_REQUIRE(n >= 2 && n <= (1u << 26), "vgi_probe_dependent_fetch: %zu bytes out of range", bytes);
    hipStream_t st = (hipStream_t)stream;
    // one cycle through all records (Sattolo's shuffle), link in word 0
    std::vector<uint32_t> host(n * 32u);
    std::vector<uint32_t> perm(n);
    for (size_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    uint64_t rs = 0x9E3779B97F4A7C15ull;
    for (size_t i = n - 1; i > 0; --i) {
        rs ^= rs << 13, rs ^= rs >> 7, rs ^= rs << 17;
        const size_t j = (size_t)(rs % i);
        std::swap(perm[i], perm[j]);
    }
    for (size_t i = 0; i < n; ++i) {
        uint32_t* rec = &host[i * 32u];
        for (int k = 0; k < 32; ++k) rec[k] = (uint32_t)(i * 32u + k) * 0x9E3779B1u;
        rec[0] = perm[i];
    }
    int dev = 0, cus = 0;
    VGI_TRY_HIP(hipGetDevice(&dev));
    VGI_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t groups = (uint32_t)cus * 4u * (uint32_t)kTraceWaves;  // 64-lane groups filling every SIMD
    DevBuf nodes, sink;
    VGI_CHECK(nodes.alloc(host.size() * 4u));
    VGI_CHECK(sink.alloc((size_t)groups * 64u * 4u));
    VGI_TRY_HIP(hipMemcpyAsync(nodes.p, host.data(), host.size() * 4u, hipMemcpyHostToDevice, st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    VGI_TRY_HIP(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        return check_hip(he, "hipEventCreate");
    }
    float ms = 0.0f;
    // an untimed launch warms the caches with the buffer
    hipLaunchKernelGGL(k_chase, dim3(groups), dim3(64), 0, st, (const uint4*)nodes.p, (uint32_t)n, steps,
                       group_shift, sink.as<uint32_t>());
    if ((he = hipEventRecord(e0, st)) == hipSuccess) {
        hipLaunchKernelGGL(k_chase, dim3(groups), dim3(64), 0, st, (const uint4*)nodes.p, (uint32_t)n, steps,
                           group_shift, sink.as<uint32_t>());
        if ((he = hipEventRecord(e1, st)) == hipSuccess && (he = hipEventSynchronize(e1)) == hipSuccess)
            he = hipEventElapsedTime(&ms, e0, e1);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (he != hipSuccess) return check_hip(he, "vgi_probe_dependent_fetch");
    VGI_CHECK(check_launch("k_chase"));
    *steps_per_s = (double)groups * 64.0 * (double)steps / ((double)ms * 1e-3);
    return VGI_OK;
}
