/*
 * vgi_runtime.h — host-side helpers shared by the C-ABI translation units:
 * error reporting (thread-local message + status codes), HIP call checking,
 * and the tile mapping used by every per-pixel kernel.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "../../include/vgi.h"
#include "vgi_common.h"

#ifndef VGI_PIX_TILE_W
#define VGI_PIX_TILE_W 32u
#endif

namespace vgi {

vgi_status set_error(vgi_status code, const char* fmt, ...);
vgi_status check_hip(hipError_t e, const char* what);

#define VGI_TRY_HIP(expr)                                        \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return ::vgi::check_hip(_e, #expr); \
    } while (0)

#define VGI_REQUIRE(cond, ...)                                              \
    do {                                                                    \
        if (!(cond)) return ::vgi::set_error(VGI_E_INVALID, __VA_ARGS__); \
    } while (0)

// Propagates a non-OK status (the callee has already set the message).
#define VGI_CHECK(expr)                          \
    do {                                         \
        vgi_status _s = (expr);                  \
        if (_s != VGI_OK) return _s;             \
    } while (0)

// Launch-after check: reports a launch configuration error synchronously.
vgi_status check_launch(const char* kernel);

constexpr uint32_t kMaxDim = 16384;

// Per-pass GPU timer (active only while vgi_profiler_enable(1)); records a
// hipEvent before construction's caller launches and one at destruction.
class ProfScope {
   public:
    ProfScope(const char* name, hipStream_t s, double bytes);
    ~ProfScope();

   private:
    int id_ = -1;
    hipStream_t s_ = nullptr;
    hipEvent_t a_ = nullptr;
    double bytes_ = 0;
};

// Traversal statistics: kTraceCounterSlots slots of {closest rays, any-hit
// rays, interior nodes, triangles}, spread over slots to limit contention.
constexpr int kTraceCounterSlots = 64;
// Traversal statistics: the device counter buffer while they are enabled
// (else null), and the per-pass collection after an instrumented launch
// (waits for `s`, adds the counts to the pass's profiler entry, clears them).
unsigned long long* trace_counters();
void trace_stats_collect(const char* name, hipStream_t s);

inline bool valid_dims(uint32_t w, uint32_t h) { return w > 0 && h > 0 && w <= kMaxDim && h <= kMaxDim; }

// Pixel tile of the per-pixel gather kernels: one wave64 covers kPixTileW x
// kPixTileH pixels as 2x2 quads (see tile_pixel_at).  Wide tiles make each
// tap of a wave read whole rows of 128 bytes from every 4-byte image
// instead of eight 32-byte pieces (measured at 1920x1080: the disocclusion
// fix 1.9x, the temporal filter 1.5x faster than with 8 x 8 tiles); the
// ray-tracing passes keep 8 x 8 tiles (square ones trace faster).
constexpr uint32_t kPixTileW = VGI_PIX_TILE_W;
constexpr uint32_t kPixTileH = 64u / kPixTileW;
static_assert(kPixTileW == 8u || kPixTileW == 16u || kPixTileW == 32u, "pixel tile width");
inline dim3 tile_grid(uint32_t w, uint32_t h) {
    return dim3((w + kPixTileW - 1) / kPixTileW, (h + kPixTileH - 1) / kPixTileH, 1);
}
inline dim3 tile8_grid(uint32_t w, uint32_t h) { return dim3((w + 7) / 8, (h + 7) / 8, 1); }

// Current launch shapes (vgi_tuning_set).
uint32_t tuning_atrous_block();
uint32_t tuning_trace_tile_width();
// Tile grid of the sample passes for a tile `tw` pixels wide and 64 / tw high.
inline dim3 trace_tile_grid(uint32_t w, uint32_t h, uint32_t tw) {
    const uint32_t th = 64u / tw;
    return dim3((w + tw - 1) / tw, (h + th - 1) / th, 1);
}

}  // namespace vgi

#if defined(__HIPCC__)
namespace vgi {
// Thread -> pixel mapping for an 8x8 tile processed by one wave64: lanes are
// grouped into 2x2 quads (lane>>2 = quad index in a 4x4 grid of quads,
// lane&3 = position in the quad), as in the quad remap of
// relax_disocclusion_fix.hlsl:36-40.  Quad members are lanes 4q..4q+3, so a
// quad reduction is two xor-shuffles.
template <uint32_t TW>
__device__ __forceinline__ void quad_tile_pixel_at(uint32_t bx, uint32_t by, uint32_t& px, uint32_t& py) {
    constexpr uint32_t qw = TW / 2u, th = 64u / TW;  // quads per tile row, tile height
    uint32_t lane = threadIdx.x & 63u;
    uint32_t q = lane >> 2;
    uint32_t qi = lane & 3u;
    px = bx * TW + ((q % qw) << 1) + (qi & 1u);
    py = by * th + ((q / qw) << 1) + (qi >> 1);
}
__device__ __forceinline__ void tile_pixel_at(uint32_t bx, uint32_t by, uint32_t& px, uint32_t& py) {
    quad_tile_pixel_at<kPixTileW>(bx, by, px, py);
}
__device__ __forceinline__ void tile_pixel(uint32_t& px, uint32_t& py) { tile_pixel_at(blockIdx.x, blockIdx.y, px, py); }

// XCD-aware workgroup order.  MI355X dispatches consecutive workgroup ids
// round-robin over its 8 XCDs, each with its own L2.  xcd_remap maps the
// dispatch id p in [0, n) to a logical id such that every XCD gets one
// contiguous range of logical ids (the first n % 8 XCDs one more).  The
// a-trous pass orders its tiles so that the residue classes of one region
// are consecutive, and this keeps them in one L2 (measured: the plain tile
// passes, whose neighbours are contiguous already, get slower with it).
constexpr uint32_t kNumXcd = 8;
__device__ __forceinline__ uint32_t xcd_remap(uint32_t p, uint32_t n) {
    const uint32_t q = n / kNumXcd, r = n % kNumXcd;
    const uint32_t xcd = p % kNumXcd, k = p / kNumXcd;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

// An image pointer as a global-memory pointer (see gld in vgi_common.h):
// G(io.image)[i] reads and writes global memory even when the pointer was
// loaded from a job table rather than passed as a kernel argument.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* G(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// Device-scope relaxed atomics on global memory.  Through a global pointer
// the compiler's atomic optimizer folds a wave's same-address adds into one
// atomic; through a flat pointer it issues one per lane (a 64-fold
// contention on hot counters such as the cache's query counter).
__device__ __forceinline__ uint32_t g_atomic_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t g_atomic_max(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_max(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t g_atomic_cas(uint32_t* p, uint32_t expected, uint32_t desired) {
    __hip_atomic_compare_exchange_strong(G(p), &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected;
}
}  // namespace vgi
#endif
