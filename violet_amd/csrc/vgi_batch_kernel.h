/*
 * vgi_batch_kernel.h — the persistent multi-view kernel (device side of
 * vgi_batch.h).  Workgroup b handles tiles b, b + gridDim.x, ... of the
 * concatenated tile list; the owning job is found by binary search of the
 * prefix sums (wave-uniform, so the job's arguments are scalar loads).
 */
#pragma once
#include "vgi_batch.h"
#include "vgi_runtime.h"
#include "vgi_storage.h"

namespace vgi {

// Waves per SIMD a job's fused kernel is compiled for (a register budget;
// 1 = no constraint).  Specialised next to the job where it matters.
template <class Job>
struct BatchMinWaves {
    static constexpr int value = 1;
};

// kBarrier: the job body shares LDS across its workgroup, so the next tile
// must wait until every thread is done with this one.  kXcdOrder: walk the
// tiles in XCD-aware order (xcd_remap), for bodies whose consecutive tiles
// share cache lines.
template <class Job, int kThreads, bool kBarrier, bool kXcdOrder>
__global__ void __launch_bounds__(kThreads, BatchMinWaves<Job>::value) k_batch(const BatchTable<Job>* __restrict__ t) {
    const uint32_t n = t->hd.njobs;
    const uint32_t total = t->hd.tile_begin[n];
    const uint32_t first = kXcdOrder ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    for (uint32_t b = first; b < total; b += gridDim.x) {
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (t->hd.tile_begin[mid] <= b)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t local = b - t->hd.tile_begin[lo];
        const uint32_t tx = t->hd.tiles_x[lo];
        Job::run(t->jobs[lo], local % tx, local / tx);
        if (kBarrier) __syncthreads();
    }
}

template <template <class> class J, int kThreads = 64, bool kBarrier = false, bool kXcdOrder = false>
vgi_status batch_launch(const void* dev, uint32_t storage, uint32_t grid, size_t lds, hipStream_t st) {
    if (storage == VGI_STORAGE_F32)
        hipLaunchKernelGGL((k_batch<J<StorageF32>, kThreads, kBarrier, kXcdOrder>), dim3(grid), dim3(kThreads), lds,
                           st, (const BatchTable<J<StorageF32>>*)dev);
    else
        hipLaunchKernelGGL((k_batch<J<StorageF16>, kThreads, kBarrier, kXcdOrder>), dim3(grid), dim3(kThreads), lds,
                           st, (const BatchTable<J<StorageF16>>*)dev);
    return check_launch("k_batch");
}

template <template <class> class J>
BatchTable<J<StorageF32>>* batch_table(void* host) {
    BatchTable<J<StorageF32>>* t = (BatchTable<J<StorageF32>>*)host;
    t->hd.njobs = 0;
    t->hd.tile_begin[0] = 0;
    return t;
}

}  // namespace vgi
