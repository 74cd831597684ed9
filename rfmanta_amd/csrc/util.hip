// Device utilities: a streaming copy (16-byte vector loads and stores, grid-stride) used
// by callers that move ray batches between buffers, and as the known-size reference
// when HBM read/write counters are calibrated.

#include <hip/hip_runtime.h>

#include "mh_device.h"

namespace mh {
namespace {

__global__ __launch_bounds__(256) void k_copy16(const int4* __restrict__ src,
                                               int4* __restrict__ dst, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy1(const uint8_t* __restrict__ src,
                                              uint8_t* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

int launch_copy(void* dst, const void* src, size_t bytes, int cu_count, void* stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int64_t n16 = (int64_t)(bytes / 16);
  if (n16 > 0) {
    const int64_t want = (n16 + 255) / 256;
    const int64_t cap = (int64_t)cu_count * 8;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)(want < cap ? want : cap)), dim3(256), 0, s,
                       reinterpret_cast<const int4*>(src), reinterpret_cast<int4*>(dst), n16);
  }
  const int64_t tail = (int64_t)(bytes - (size_t)n16 * 16);
  if (tail > 0)
    hipLaunchKernelGGL(k_copy1, dim3(1), dim3(256), 0, s,
                       reinterpret_cast<const uint8_t*>(src) + n16 * 16,
                       reinterpret_cast<uint8_t*>(dst) + n16 * 16, tail);
  return (int)hipGetLastError();
}

}  // namespace mh
