// Plain device-to-device copy kernel (16-byte vectors, grid-stride). Its
// HBM counters calibrate the traffic measurements of the other kernels:
// it reads and writes exactly `bytes` each.
#include <hip/hip_runtime.h>

#include "api_internal.h"

namespace rrhip {
namespace {

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src,
                                                   uint4* __restrict__ dst, uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

}  // namespace
}  // namespace rrhip

extern "C" int rrhip_device_copy(void* stream, void* dst, const void* src, uint64_t bytes) {
  if (bytes == 0) return RRHIP_OK;
  if (!dst || !src || (bytes % 16) != 0 || ((uintptr_t)dst % 16) != 0 || ((uintptr_t)src % 16) != 0)
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_device_copy: null, unaligned or non-multiple-of-16 buffers");
  const uint64_t n16 = bytes / 16;
  const uint64_t want = (n16 + 255) / 256;
  const unsigned blocks = (unsigned)(want < 65536 ? want : 65536);
  hipLaunchKernelGGL(rrhip::copy_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     (const uint4*)src, (uint4*)dst, n16);
  RRHIP_HIP_TRY(hipGetLastError());
  return RRHIP_OK;
}
