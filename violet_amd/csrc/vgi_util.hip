/*
 * vgi_util.hip — buffer fills (util/clear_buffer.hlsl:10-17 and the
 * zero-fill used for absent histories).  Kernels rather than
 * hipMemsetAsync so that a frame can be captured into a hipGraph.
 */
#include <algorithm>
#include <chrono>
#include <new>
#include <thread>

#include "vgi_batch_kernel.h"
#include "vgi_internal.h"

namespace vgi {

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

vgi_status fill_u32(uint32_t* p, size_t count, uint32_t value, hipStream_t st) {
    if (count == 0) return VGI_OK;
    size_t blocks = (count + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, p, count, value);
    return check_launch("k_fill_u32");
}

vgi_status fill_zero(void* p, size_t bytes, hipStream_t st) {
    if (bytes == 0) return VGI_OK;
    if ((bytes & 3) == 0 && ((uintptr_t)p & 3) == 0) return fill_u32((uint32_t*)p, bytes / 4, 0u, st);
    size_t blocks = (bytes + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)blocks), dim3(256), 0, st, (uint8_t*)p, bytes, (uint8_t)0);
    return check_launch("k_fill_u8");
}

// Fused multi-view zero fill (vgi_batch.h): a tile is 1024 words; the byte
// tail of a length that is not a multiple of 4 goes to the first tile.
constexpr uint32_t kClearTileWords = 1024;
template <class>
struct JobClear {
    uint32_t* words;
    size_t nwords;
    uint8_t* tail;
    uint32_t tail_bytes;
    uint32_t value;
    __device__ static void run(const JobClear& j, uint32_t bx, uint32_t) {
        const size_t base = (size_t)bx * kClearTileWords;
#pragma unroll
        for (uint32_t k = 0; k < kClearTileWords / 64; ++k) {
            const size_t i = base + k * 64u + (threadIdx.x & 63u);
            if (i < j.nwords) j.words[i] = j.value;
        }
        if (bx == 0 && threadIdx.x < j.tail_bytes) j.tail[threadIdx.x] = 0u;
    }
};

static vgi_status fill_clear(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobClear>(host);
    for (int v = 0; v < n; ++v) {
        const PassRec& r = *recs[v];
        VGI_REQUIRE(((uintptr_t)r.clear_ptr & 3u) == 0 || r.clear_bytes == 0, "batched clear: unaligned buffer");
        auto& j = t->jobs[v];
        j.words = (uint32_t*)r.clear_ptr;
        j.nwords = r.clear_bytes / 4;
        j.tail = (uint8_t*)r.clear_ptr + 4 * j.nwords;
        j.tail_bytes = (uint32_t)(r.clear_bytes & 3u);
        j.value = r.clear_value;
        VGI_REQUIRE(r.clear_value == 0 || j.tail_bytes == 0, "batched fill: buffer not whole words");
        const size_t tiles = std::max<size_t>(1, (j.nwords + kClearTileWords - 1) / kClearTileWords);
        VGI_REQUIRE(tiles < (1u << 30), "batched clear: buffer too large");
        batch_add_tiles(t->hd, v, (uint32_t)tiles, 1);
    }
    return VGI_OK;
}

const BatchOps* batch_ops_util(int kind) {
    static const BatchOps clear{sizeof(BatchTable<JobClear<StorageF32>>), fill_clear, batch_launch<JobClear>, 64, false};
    return kind == kPassClear ? &clear : nullptr;
}

}  // namespace vgi

using namespace vgi;

struct vgi_fence {
    hipEvent_t ev = nullptr;
    bool recorded = false;
};

extern "C" vgi_status vgi_fence_create(vgi_fence** out) {
    VGI_REQUIRE(out, "vgi_fence_create: null out");
    *out = nullptr;
    vgi_fence* f = new (std::nothrow) vgi_fence();
    if (!f) return set_error(VGI_E_NOMEM, "vgi_fence_create: out of host memory");
    const hipError_t e = hipEventCreateWithFlags(&f->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete f;
        return check_hip(e, "vgi_fence_create");
    }
    *out = f;
    return VGI_OK;
}

extern "C" vgi_status vgi_fence_destroy(vgi_fence* f) {
    if (!f) return VGI_OK;
    if (f->ev) (void)hipEventDestroy(f->ev);
    delete f;
    return VGI_OK;
}

extern "C" vgi_status vgi_fence_record(vgi_fence* f, void* stream) {
    VGI_REQUIRE(f, "vgi_fence_record: null fence");
    VGI_TRY_HIP(hipEventRecord(f->ev, (hipStream_t)stream));
    f->recorded = true;
    return VGI_OK;
}

// Polls the event: back-to-back for the first 100 us, then sleeping with a
// doubling interval up to 1 ms, so short frames are seen at once and a stuck
// one costs little host time.
extern "C" vgi_status vgi_fence_wait(vgi_fence* f, uint32_t timeout_ms, double* waited_ms) {
    VGI_REQUIRE(f, "vgi_fence_wait: null fence");
    using clk = std::chrono::steady_clock;
    const clk::time_point t0 = clk::now();
    const clk::time_point deadline = t0 + std::chrono::milliseconds(timeout_ms);
    auto elapsed_ms = [&] { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); };
    vgi_status st = VGI_OK;
    std::chrono::microseconds nap(0);
    while (f->recorded) {
        const hipError_t e = hipEventQuery(f->ev);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) {
            st = check_hip(e, "vgi_fence_wait");
            break;
        }
        const clk::time_point now = clk::now();
        if (now >= deadline) {
            st = set_error(VGI_E_TIMEOUT, "vgi_fence_wait: the marked work did not finish within %u ms", timeout_ms);
            break;
        }
        if (now - t0 > std::chrono::microseconds(100)) {
            nap = nap.count() ? std::min(nap * 2, std::chrono::microseconds(1000)) : std::chrono::microseconds(20);
            std::this_thread::sleep_for(std::min<clk::duration>(nap, deadline - now));
        }
    }
    if (waited_ms) *waited_ms = elapsed_ms();
    return st;
}

extern "C" vgi_status vgi_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    VGI_REQUIRE(dst && src, "vgi_memcpy_d2d: null pointer");
    if (bytes == 0) return VGI_OK;
    VGI_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return VGI_OK;
}
