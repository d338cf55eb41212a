/*
 * vgi_batch.h — tables of fused multi-view launches.
 *
 * A batched frame runs each sequence slot of many views as ONE persistent
 * launch: a fixed grid of workgroups walks the concatenated tile lists of
 * all views of the slot.  The per-view arguments of a slot live in a
 * BatchTable in device memory (uploaded once per frame); a view that does
 * not issue the slot this frame has no tiles.  Because the launch shapes do
 * not depend on the views, a batched frame can be captured once as a HIP
 * graph and replayed with new tables.
 */
#pragma once
#include "vgi_passes.h"

namespace vgi {

constexpr int kMaxBatchViews = 64;

struct BatchHeader {
    uint32_t njobs;
    uint32_t tiles_x[kMaxBatchViews];             // tile-grid width of each job
    uint32_t tile_begin[kMaxBatchViews + 1];      // prefix sum of tile counts
};

template <class Job>
struct BatchTable {
    BatchHeader hd;
    Job jobs[kMaxBatchViews];
};

// Appends a job of `tiles_x` x `tiles_y` tiles to the header.
inline void batch_add_tiles(BatchHeader& hd, uint32_t v, uint32_t tiles_x, uint32_t tiles_y) {
    hd.tiles_x[v] = tiles_x ? tiles_x : 1u;
    hd.tile_begin[v + 1] = hd.tile_begin[v] + tiles_x * tiles_y;
    hd.njobs = v + 1;
}
inline void batch_add_pixel_tiles(BatchHeader& hd, uint32_t v, uint32_t w, uint32_t h) {
    const dim3 g = tile_grid(w, h);
    batch_add_tiles(hd, v, g.x, g.y);
}

// What a pass kind contributes to batching: the size of its table, the host
// fill of a table from the views' records, and the fused launch.
struct BatchOps {
    size_t table_bytes;
    vgi_status (*fill)(const PassRec* const* recs, int n, void* host_table);
    vgi_status (*launch)(const void* dev_table, uint32_t storage, uint32_t grid, size_t lds, hipStream_t st);
    int threads;  // workgroup size of the fused kernel
    bool trace;   // traces rays (LDS traversal stack, trace grid size)
};

// Per-file registries (nullptr for kinds a file does not own).
const BatchOps* batch_ops_denoise(int kind);
const BatchOps* batch_ops_resample(int kind);
const BatchOps* batch_ops_sample(int kind);
const BatchOps* batch_ops_hashgrid(int kind);
const BatchOps* batch_ops_compose(int kind);
const BatchOps* batch_ops_gbuffer(int kind);
const BatchOps* batch_ops_util(int kind);

}  // namespace vgi
