/*
 * vgi_internal.h — internal host-side declarations shared by the library's
 * translation units (device buffers, storage sizes, kernel launchers).
 */
#pragma once
#include "vgi_runtime.h"

namespace vgi {

// Owned device allocation (hipMalloc'd, freed on destruction).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    vgi_status alloc(size_t n) {
        release();
        if (n == 0) return VGI_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            return check_hip(e, "hipMalloc");
        }
        bytes = n;
        return VGI_OK;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

// One pixel as the diffuse spatial pass reads it, written by the diffuse
// temporal pass in the renderer's own frames: depth bits, G-buffer layer 1,
// the reservoir word, the two hit words and the hit radiance as f32 (the
// value the spatial pass would read back from the storage mode's radiance
// image).  32 bytes: a spatial tap is one aligned pair of 16-byte loads from
// one cache line instead of loads from five images, and its geometry test
// and payload arrive together.
struct alignas(32) DiffTap {
    uint32_t w[8];
};

// The diffuse temporal and spatial passes with the tap records (taps NULL:
// exactly the C entry points vgi_pass_ind_diff_{temporal,spatial}_resample).
vgi_status launch_diff_temporal(const vgi_frame_params* fp, const vgi_ind_diff_temporal_io* io,
                                const vgi_pc_ind_diff_temporal* pc, const vgi_resample_params* params, DiffTap* taps,
                                const uint32_t* gbuffer_color, hipStream_t st);
vgi_status launch_diff_spatial(const vgi_frame_params* fp, const vgi_ind_diff_spatial_io* io,
                               const vgi_pc_ind_diff_spatial* pc, const vgi_resample_params* params,
                               const DiffTap* taps, hipStream_t st);

inline size_t radiance_bytes(uint32_t storage) { return storage == VGI_STORAGE_F16 ? 6 : 12; }
inline size_t pair_bytes(uint32_t storage) { return storage == VGI_STORAGE_F16 ? 4 : 8; }
inline size_t scalar_bytes(uint32_t storage) { return storage == VGI_STORAGE_F16 ? 2 : 4; }

// Zero-fill with a kernel (no hipMemsetAsync, so the call can sit inside a
// captured hipGraph).
vgi_status fill_zero(void* p, size_t bytes, hipStream_t st);
// Fill a u32 buffer with a value (util/clear_buffer.hlsl:10-17).
vgi_status fill_u32(uint32_t* p, size_t count, uint32_t value, hipStream_t st);

}  // namespace vgi
