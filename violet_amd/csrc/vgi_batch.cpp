/*
 * vgi_batch.cpp — batched frames over many views (no reference counterpart:
 * violet renders one view per frame, restir_render_loop.rs:262-268).
 *
 * Each view is an ordinary vgi_renderer (its own reservoirs, denoiser
 * history and radiance cache).  A batched frame plans every view with the
 * renderer's own sequencer into per-slot records, writes one table per slot
 * into a pinned host arena, uploads the arena with a single copy and runs
 * every slot as one persistent launch over all views (vgi_batch.h).  The
 * launch shapes depend only on the batch configuration, the storage mode
 * and the scene's traversal-stack size, so the whole frame can be captured
 * once as a HIP graph and replayed; the tables it reads are the only
 * per-frame input.  Two arenas alternate so the host can fill one while the
 * device still reads the other.
 */
#include <new>
#include <vector>

#include "vgi_batch.h"
#include "vgi_bvh.h"

using namespace vgi;

namespace vgi {
const SceneView& scene_view(const vgi_scene* s);

static const BatchOps* batch_ops(int kind) {
    const BatchOps* o = nullptr;
    if (!o) o = batch_ops_util(kind);
    if (!o) o = batch_ops_gbuffer(kind);
    if (!o) o = batch_ops_hashgrid(kind);
    if (!o) o = batch_ops_sample(kind);
    if (!o) o = batch_ops_resample(kind);
    if (!o) o = batch_ops_denoise(kind);
    if (!o) o = batch_ops_compose(kind);
    return o;
}

static const char* kind_name(int kind) {
    static const char* n[kPassKindCount] = {
        "batch_clear",           "batch_gbuffer_primary",      "batch_hash_grid_decay",
        "batch_hash_grid_query_select", "batch_hash_grid_query_compact", "batch_hash_grid_raygen",
        "batch_ind_diff_sample_gen", "batch_ind_diff_sample_validate", "batch_raytraced_ao_gen",
        "batch_raytraced_ao_filter", "batch_ind_diff_temporal_resample", "batch_ind_diff_spatial_resample",
        "batch_ind_spec_sample_gen", "batch_ind_spec_sample_validate", "batch_ind_spec_temporal_resample",
        "batch_ind_spec_spatial_resample", "batch_denoise_temporal", "batch_denoise_disocclusion_fix",
        "batch_denoise_fast_history_clamp", "batch_denoise_spatial_fallback", "batch_denoise_atrous",
        "batch_raytraced_shadow", "batch_final_lighting", "batch_temporal_aa",
        "batch_hash_grid_query_select_ordered", "batch_hash_grid_query_compact_ordered",
        "batch_hash_grid_raygen_ordered"};
    return kind >= 0 && kind < kPassKindCount ? n[kind] : "batch_unknown";
}
}  // namespace vgi

struct vgi_batch {
    vgi_batch_config cfg{};
    size_t offsets[kSlotCount] = {};
    size_t arena_bytes = 0;
    void* host[2] = {nullptr, nullptr};
    void* dev[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    int flip = 0;
    std::vector<RecordSink> recs;
    hipStream_t cap = nullptr;  // capture/replay stream (graph mode)
    hipEvent_t ev_in = nullptr;
    // second lane (concurrent_lanes): independent slot chains beside the main stream
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_cross = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint32_t key_storage[2] = {0, 0};
    uint32_t key_shape[2][2] = {{0, 0}, {0, 0}};  // launch shapes (vgi_tuning) the graph was captured with
    size_t key_lds[2] = {0, 0};
    // One-group-per-tile slots launch as many groups as the largest tile
    // count seen so far for the slot (a graph is re-captured when it grows).
    uint32_t grid_bound[kSlotCount];
    uint32_t key_grid[2][kSlotCount];

    ~vgi_batch() {
        for (int a = 0; a < 2; ++a) {
            if (done[a]) {
                (void)hipEventSynchronize(done[a]);
                (void)hipEventDestroy(done[a]);
            }
            if (exec[a]) (void)hipGraphExecDestroy(exec[a]);
            if (host[a]) (void)hipHostFree(host[a]);
            if (dev[a]) (void)hipFree(dev[a]);
        }
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_cross) (void)hipEventDestroy(ev_cross);
        if (side) (void)hipStreamDestroy(side);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

extern "C" vgi_status vgi_batch_default_config(vgi_batch_config* cfg) {
    VGI_REQUIRE(cfg, "vgi_batch_default_config: null");
    memset(cfg, 0, sizeof(*cfg));
    cfg->max_views = kMaxBatchViews;
    cfg->pixel_groups = 0;  // one group per tile
    cfg->trace_groups = 0;
    cfg->atrous_groups = 0;
    cfg->use_graph = 0;
    cfg->concurrent_lanes = 1;
    return VGI_OK;
}

extern "C" vgi_status vgi_batch_create(const vgi_batch_config* cfg, vgi_batch** out) {
    VGI_REQUIRE(out, "vgi_batch_create: null out");
    *out = nullptr;
    vgi_batch_config c;
    if (cfg) {
        c = *cfg;
    } else {
        vgi_status s = vgi_batch_default_config(&c);
        if (s != VGI_OK) return s;
    }
    VGI_REQUIRE(c.max_views >= 1 && c.max_views <= (uint32_t)kMaxBatchViews, "vgi_batch_create: max_views %u not in 1..%d",
                c.max_views, kMaxBatchViews);
    VGI_REQUIRE(c.pixel_groups <= (1u << 20) && c.trace_groups <= (1u << 20) && c.atrous_groups <= (1u << 20),
                "vgi_batch_create: persistent group counts must be at most 2^20");
    vgi_batch* b = new (std::nothrow) vgi_batch();
    if (!b) return set_error(VGI_E_NOMEM, "vgi_batch_create: out of host memory");
    b->cfg = c;
    size_t off = 0;
    for (int s = 0; s < kSlotCount; ++s) {
        const BatchOps* o = batch_ops(slot_kind(s));
        if (!o) {
            delete b;
            return set_error(VGI_E_INVALID, "internal: no batched launch for pass kind %d", slot_kind(s));
        }
        b->offsets[s] = off;
        off += (o->table_bytes + 255) & ~(size_t)255;
    }
    b->arena_bytes = off;
    for (int s = 0; s < kSlotCount; ++s) b->grid_bound[s] = 1;
    memset(b->key_grid, 0, sizeof(b->key_grid));
    hipError_t e = hipSuccess;
    for (int a = 0; a < 2 && e == hipSuccess; ++a) {
        e = hipHostMalloc(&b->host[a], off, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&b->dev[a], off);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done[a], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
    if (e == hipSuccess && c.use_graph) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e == hipSuccess && c.concurrent_lanes) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    // lane events order two streams of this device only (no system-scope release)
    const unsigned lane_ev = hipEventDisableTiming | hipEventDisableSystemFence;
    if (e == hipSuccess && c.concurrent_lanes) e = hipEventCreateWithFlags(&b->ev_fork, lane_ev);
    if (e == hipSuccess && c.concurrent_lanes) e = hipEventCreateWithFlags(&b->ev_join, lane_ev);
    if (e == hipSuccess && c.concurrent_lanes) e = hipEventCreateWithFlags(&b->ev_cross, lane_ev);
    if (e != hipSuccess) {
        delete b;
        return check_hip(e, "vgi_batch_create");
    }
    b->recs.resize(c.max_views);
    *out = b;
    return VGI_OK;
}

extern "C" vgi_status vgi_batch_destroy(vgi_batch* b) {
    delete b;
    return VGI_OK;
}

// Persistent group count of a slot's launch, or 0 for one group per tile.
static uint32_t persistent_groups(const vgi_batch* b, const BatchOps* o) {
    if (o->threads == 256) return b->cfg.atrous_groups;
    return o->trace ? b->cfg.trace_groups : b->cfg.pixel_groups;
}
static uint32_t groups_for(const vgi_batch* b, int slot, const BatchOps* o) {
    const uint32_t g = persistent_groups(b, o);
    return g ? g : b->grid_bound[slot];
}

// Upload of the arena and every slot's launch on `main` and the side lane
// (all slots in graph mode, so the captured topology never changes; only
// occupied ones otherwise), joined back into `main` at the end.
static vgi_status issue_slots(vgi_batch* b, int a, uint32_t storage, size_t lds, Lanes& lanes, bool all_slots,
                              bool profile) {
    VGI_TRY_HIP(hipMemcpyAsync(b->dev[a], b->host[a], b->arena_bytes, hipMemcpyHostToDevice, lanes.main));
    for (int s = 0; s < kSlotCount; ++s) {
        const int kind = slot_kind(s);
        const BatchOps* o = batch_ops(kind);
        const BatchHeader* hd = (const BatchHeader*)((const char*)b->host[a] + b->offsets[s]);
        if (!all_slots && hd->njobs == 0) continue;
        hipStream_t st;
        vgi_status r = lanes.enter(s, st);
        if (r != VGI_OK) return r;
        if (profile) {
            ProfScope prof(kind_name(kind), st, 0.0);
            r = o->launch((const char*)b->dev[a] + b->offsets[s], storage, groups_for(b, s, o), o->trace ? lds : 0, st);
        } else {
            r = o->launch((const char*)b->dev[a] + b->offsets[s], storage, groups_for(b, s, o), o->trace ? lds : 0, st);
        }
        if (r == VGI_OK) r = lanes.leave(s, st);
        if (r != VGI_OK) return r;
    }
    return VGI_OK;
}

static vgi_status issue_frame(vgi_batch* b, int a, uint32_t storage, size_t lds, hipStream_t st, bool all_slots,
                              bool profile) {
    Lanes lanes;
    lanes.main = st;
    lanes.side = b->side;
    lanes.ev_fork = b->ev_fork;
    lanes.ev_join = b->ev_join;
    lanes.ev_cross = b->ev_cross;
    vgi_status r = issue_slots(b, a, storage, lds, lanes, all_slots, profile);
    vgi_status j = lanes.finish();
    return r != VGI_OK ? r : j;
}

extern "C" vgi_status vgi_batch_add_passes(vgi_batch* b, const vgi_scene* scene, uint32_t num_views,
                                           vgi_renderer* const* views, const vgi_frame_params* fps,
                                           const uint32_t* frame_indices, vgi_renderer_output* outs, void* stream) {
    VGI_REQUIRE(b && scene && views && fps && frame_indices && outs, "vgi_batch_add_passes: null argument");
    VGI_REQUIRE(num_views >= 1 && num_views <= b->cfg.max_views, "vgi_batch_add_passes: %u views (max %u)", num_views,
                b->cfg.max_views);
    hipStream_t st = (hipStream_t)stream;
    uint32_t storage = 0;
    for (uint32_t v = 0; v < num_views; ++v) {
        VGI_REQUIRE(views[v], "vgi_batch_add_passes: null view %u", v);
        for (uint32_t u = 0; u < v; ++u)
            VGI_REQUIRE(views[u] != views[v], "vgi_batch_add_passes: view %u appears twice", v);
        const uint32_t sv = vgi_renderer_storage(views[v]);
        if (v == 0) storage = sv;
        VGI_REQUIRE(sv == storage, "vgi_batch_add_passes: views mix storage modes");
    }
    // 1. plan every view (advances each view's host-side state)
    for (uint32_t v = 0; v < num_views; ++v) {
        b->recs[v].clear();
        vgi_status s = vgi_renderer_plan(views[v], scene, &fps[v], frame_indices[v], &outs[v], b->recs[v]);
        if (s != VGI_OK) return s;
    }
    // 2. fill the tables of the free arena
    const int a = b->flip;
    if (b->used[a]) VGI_TRY_HIP(hipEventSynchronize(b->done[a]));
    const PassRec* list[kMaxBatchViews];
    for (int s = 0; s < kSlotCount; ++s) {
        int n = 0;
        for (uint32_t v = 0; v < num_views; ++v)
            if (b->recs[v].has[s]) list[n++] = &b->recs[v].recs[s];
        const BatchOps* o = batch_ops(slot_kind(s));
        vgi_status r = o->fill(list, n, (char*)b->host[a] + b->offsets[s]);
        if (r != VGI_OK) return r;
        const BatchHeader* hd = (const BatchHeader*)((const char*)b->host[a] + b->offsets[s]);
        const uint32_t tiles = hd->tile_begin[hd->njobs];
        if (persistent_groups(b, o) == 0 && tiles > b->grid_bound[s]) b->grid_bound[s] = tiles;
    }
    // 3. upload + fused launches
    const size_t lds = traversal_lds_bytes(scene_view(scene));
    if (!b->cfg.use_graph) {
        vgi_status r = issue_frame(b, a, storage, lds, st, false, true);
        if (r != VGI_OK) return r;
        VGI_TRY_HIP(hipEventRecord(b->done[a], st));
    } else {
        const uint32_t shape[2] = {tuning_atrous_block(), tuning_trace_tile_width()};
        if (!b->exec[a] || b->key_storage[a] != storage || b->key_lds[a] != lds ||
            memcmp(b->key_grid[a], b->grid_bound, sizeof(b->grid_bound)) != 0 ||
            memcmp(b->key_shape[a], shape, sizeof(shape)) != 0) {
            if (b->exec[a]) {
                VGI_TRY_HIP(hipGraphExecDestroy(b->exec[a]));
                b->exec[a] = nullptr;
            }
            hipGraph_t g = nullptr;
            VGI_TRY_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
            vgi_status r = issue_frame(b, a, storage, lds, b->cap, true, false);
            hipError_t e = hipStreamEndCapture(b->cap, &g);
            if (r != VGI_OK) {
                if (g) (void)hipGraphDestroy(g);
                return r;
            }
            VGI_TRY_HIP(e);
            e = hipGraphInstantiate(&b->exec[a], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            VGI_TRY_HIP(e);
            b->key_storage[a] = storage;
            b->key_lds[a] = lds;
            memcpy(b->key_grid[a], b->grid_bound, sizeof(b->grid_bound));
            memcpy(b->key_shape[a], shape, sizeof(shape));
        }
        VGI_TRY_HIP(hipEventRecord(b->ev_in, st));
        VGI_TRY_HIP(hipStreamWaitEvent(b->cap, b->ev_in, 0));
        VGI_TRY_HIP(hipGraphLaunch(b->exec[a], b->cap));
        VGI_TRY_HIP(hipEventRecord(b->done[a], b->cap));
        VGI_TRY_HIP(hipStreamWaitEvent(st, b->done[a], 0));
    }
    b->used[a] = true;
    b->flip = 1 - a;
    return VGI_OK;
}
