/*
 * vgi_util.hip — buffer fills (util/clear_buffer.hlsl:10-17 and the
 * zero-fill used for absent histories).  Kernels rather than
 * hipMemsetAsync so that a frame can be captured into a hipGraph.
 */
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

}  // namespace vgi

using namespace vgi;

extern "C" vgi_status vgi_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    VGI_REQUIRE(dst && src, "vgi_memcpy_d2d: null pointer");
    if (bytes == 0) return VGI_OK;
    VGI_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return VGI_OK;
}
