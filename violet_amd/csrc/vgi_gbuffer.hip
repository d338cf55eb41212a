/*
 * vgi_gbuffer.hip — primary-visibility G-buffer producer.
 *
 * Stands in for violet's raster G-buffer pass (gbuffer_pass.rs:45-137,
 * mesh_gbuffer.hlsl:62-126): one ray through each pixel centre, the hit's
 * reverse-Z depth (view_proj applied to the interpolated hit position), and
 * the four G-buffer layers encoded as gbuffer.hlsl:54-77.  fwidth_z and
 * fwidth_n are |ddx| + |ddy| with fine differences inside each aligned 2x2
 * quad, taken with lane shuffles: the quad's four pixels are lanes 4q..4q+3
 * (x partner = lane ^ 1, y partner = lane ^ 2), exactly the lanes a raster
 * pixel shader's derivatives come from.  A partner without geometry (or
 * outside the image) contributes 0.  Material textures are sampled with the
 * texture-coordinate derivatives of the same quad, evaluated on the hit
 * triangle's plane (vgi_radiance.h shade_hit_t).
 */
#include "vgi_batch_kernel.h"
#include "vgi_internal.h"
#include "vgi_radiance.h"

namespace vgi {
const SceneView& scene_view(const vgi_scene* s);

// Direction of the ray through pixel (px, py)'s centre.
__device__ __forceinline__ f3 primary_dir(const FrameParams& fp, f3 o, uint32_t px, uint32_t py, uint32_t w,
                                          uint32_t h) {
    float ndx = ((float)px + 0.5f) / (float)w * 2.0f - 1.0f;
    float ndy = ((float)py + 0.5f) / (float)h * 2.0f - 1.0f;
    f4 e = mul44(fp.inv_view_proj, mk4(ndx, ndy, 1.0f, 1.0f));
    f3 end = mk3(e.x / e.w, e.y / e.w, e.z / e.w);
    return normalize3(end - o);
}

template <bool kCount>
__device__ __forceinline__ void k_gbuffer_body(const FrameParams& fp, const SceneView& s, const vgi_gbuffer_io& io, uint32_t bx_, uint32_t by_) {
    VGI_TRAVERSAL_STACK(stack);
    uint32_t px, py;
    quad_tile_pixel_at<8>(bx_, by_, px, py);
    const uint32_t w = io.width, h = io.height;
    const bool inside = px < w && py < h;
    bool geo = false;
    float depth = 0.0f;
    f3 n = mk3s(0.0f);
    SurfaceHit sh;
    if (inside) {
        const f3 o = xyz(fp.view_pos);
        const f3 d = primary_dir(fp, o, px, py, w, h);
        Hit hit = trace_closest<kCount>(s, o, d, 0.05f, 100000.0f, stack);
        PrimaryFootprint pf;
        pf.o = o;
        pf.dir_x = primary_dir(fp, o, px ^ 1u, py, w, h);
        pf.dir_y = primary_dir(fp, o, px, py ^ 1u, w, h);
        pf.x_odd = (px & 1u) != 0;
        pf.y_odd = (py & 1u) != 0;
        sh = shade_hit_t<true>(s, hit, d, &pf);
        if (!sh.missed) {
            f4 c = mul44(fp.view_proj, mk4(sh.position.x, sh.position.y, sh.position.z, 1.0f));
            depth = c.z / c.w;
            geo = depth > 0.0f;
            n = sh.normal;
        }
    }
    // quad partners (all 64 lanes take part in the shuffles)
    float dx = __shfl_xor(depth, 1), dy = __shfl_xor(depth, 2);
    int gx = __shfl_xor((int)geo, 1), gy = __shfl_xor((int)geo, 2);
    f3 nx = mk3(__shfl_xor(n.x, 1), __shfl_xor(n.y, 1), __shfl_xor(n.z, 1));
    f3 ny = mk3(__shfl_xor(n.x, 2), __shfl_xor(n.y, 2), __shfl_xor(n.z, 2));
    if (!inside) return;
    const size_t i = (size_t)py * w + px;
    const size_t plane = (size_t)w * h;
    if (!geo) {
        G(io.rw_depth)[i] = 0.0f;
        G(io.rw_gbuffer)[i] = 0u;
        G(io.rw_gbuffer)[plane + i] = 0u;
        G(io.rw_gbuffer)[2 * plane + i] = 0u;
        G(io.rw_gbuffer)[3 * plane + i] = 0u;
        return;
    }
    float fz = (gx ? fabsf(dx - depth) : 0.0f) + (gy ? fabsf(dy - depth) : 0.0f);
    f3 ax = gx ? abs3(nx - n) : mk3s(0.0f);
    f3 ay = gy ? abs3(ny - n) : mk3s(0.0f);
    float fn = length3(ax + ay);
    G(io.rw_depth)[i] = depth;
    G(io.rw_gbuffer)[i] = encode_gbuffer_color(sh.base_color, sh.metallic);
    G(io.rw_gbuffer)[plane + i] = encode_gbuffer_normal(n, sh.roughness);
    G(io.rw_gbuffer)[2 * plane + i] = encode_gbuffer_fwidth(fz, fn);
    G(io.rw_gbuffer)[3 * plane + i] = 1u;
}

// 6 waves per SIMD (80 VGPRs) rather than the sample passes' 8: the pass
// alone is 1.5% slower, but it runs beside the hash-grid cache update, and
// with fewer resident G-buffer waves the update's kernels get the room to
// overlap it (the frame is 0.4% faster).
constexpr int kGBufferWaves = 6;
template <bool kCount>
__global__ void __launch_bounds__(64, kGBufferWaves) k_gbuffer(FrameParams fp, SceneView s, vgi_gbuffer_io io) {
    k_gbuffer_body<kCount>(fp, s, io, blockIdx.x, blockIdx.y);
}

// Fused multi-view launch (vgi_batch.h): every view traces the shared scene.
template <class>
struct JobGBuffer {
    FrameParams fp;
    SceneView s;
    vgi_gbuffer_io io;
    __device__ static void run(const JobGBuffer& j, uint32_t bx, uint32_t by) {
        k_gbuffer_body<false>(j.fp, j.s, j.io, bx, by);
    }
};
template <class S>
struct BatchMinWaves<JobGBuffer<S>> {
    static constexpr int value = kTraceWaves;
};

static vgi_status fill_gbuffer(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobGBuffer>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        VGI_REQUIRE(r.scene && valid_dims(r.io.gbuffer.width, r.io.gbuffer.height), "batched G-buffer: bad job");
        auto& j = t->jobs[v];
        memcpy(&j.fp, &r.fp, sizeof(j.fp));
        j.s = scene_view(r.scene);
        j.s.counters = nullptr;
        j.io = r.io.gbuffer;
        batch_add_tiles(t->hd, v, (j.io.width + 7) / 8, (j.io.height + 7) / 8);
    }
    return VGI_OK;
}

const BatchOps* batch_ops_gbuffer(int kind) {
    static const BatchOps g{sizeof(BatchTable<JobGBuffer<StorageF32>>), fill_gbuffer, batch_launch<JobGBuffer>, 64, true};
    return kind == kPassGBuffer ? &g : nullptr;
}

}  // namespace vgi

using namespace vgi;

extern "C" vgi_status vgi_pass_gbuffer_primary(const vgi_frame_params* fp, const vgi_scene* scene,
                                               const vgi_gbuffer_io* io, void* stream) {
    VGI_REQUIRE(fp && scene && io, "vgi_pass_gbuffer_primary: null argument");
    VGI_REQUIRE(valid_dims(io->width, io->height), "vgi_pass_gbuffer_primary: bad size");
    VGI_REQUIRE(io->rw_depth && io->rw_gbuffer, "vgi_pass_gbuffer_primary: missing binding");
    FrameParams f;
    memcpy(&f, fp, sizeof(f));
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("gbuffer_primary", st, (double)io->width * io->height * 20.0);
    SceneView sv = scene_view(scene);
    sv.counters = trace_counters();
    const dim3 g = tile8_grid(io->width, io->height);
    if (sv.counters)
        hipLaunchKernelGGL(k_gbuffer<true>, g, dim3(64), traversal_lds_bytes(sv), st, f, sv, *io);
    else
        hipLaunchKernelGGL(k_gbuffer<false>, g, dim3(64), traversal_lds_bytes(sv), st, f, sv, *io);
    trace_stats_collect("gbuffer_primary", st);
    return check_launch("k_gbuffer");
}
