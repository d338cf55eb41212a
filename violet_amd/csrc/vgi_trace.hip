/*
 * vgi_trace.hip — batch ray queries against the scene BVH
 * (TraceRay without shading; used by hosts and by the hit-id parity tests).
 */
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
