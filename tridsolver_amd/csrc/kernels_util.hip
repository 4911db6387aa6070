// kernels_util.hip - measurement helper of libtridhip.so: a plain streaming
// copy whose bytes moved are known exactly. It gives the HBM bandwidth the
// solver's roofline fraction is quoted against, and the reference kernel that
// calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters.
#include "../../include/tridsolver_hip.h"
#include <algorithm>

#include "trid_internal.h"

namespace trid {
namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int COPY_BLOCK = 256;
constexpr int COPY_UNROLL = 4;  // 16 KiB per workgroup, all loads before stores

__global__ __launch_bounds__(COPY_BLOCK) void stream_copy_kernel(
    const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, size_t n16) {
  const size_t base = (size_t)blockIdx.x * (COPY_BLOCK * COPY_UNROLL) + threadIdx.x;
  u32x4 v[COPY_UNROLL];
#pragma unroll
  for (int k = 0; k < COPY_UNROLL; ++k) {
    const size_t i = base + (size_t)k * COPY_BLOCK;
    if (i < n16) v[k] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int k = 0; k < COPY_UNROLL; ++k) {
    const size_t i = base + (size_t)k * COPY_BLOCK;
    if (i < n16) __builtin_nontemporal_store(v[k], dst + i);
  }
}

}  // namespace
}  // namespace trid

EXTERN_C tridStatus_t tridHipStreamCopy(const void *src, void *dst, size_t bytes) {
  if (!src || !dst || bytes % 16 != 0 || (reinterpret_cast<uintptr_t>(src) % 16) ||
      (reinterpret_cast<uintptr_t>(dst) % 16))
    return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  tridStatus_t s = trid::check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (bytes == 0) return TRID_STATUS_SUCCESS;
  const size_t n16 = bytes / 16;
  const size_t per_block = (size_t)trid::COPY_BLOCK * trid::COPY_UNROLL;
  const size_t blocks = (n16 + per_block - 1) / per_block;
  if (blocks > 0x7fffffffu) return TRID_STATUS_INVALID_VALUE;
  hipLaunchKernelGGL(trid::stream_copy_kernel, dim3((unsigned)blocks), dim3(trid::COPY_BLOCK), 0,
                     trid::current_stream(), static_cast<const trid::u32x4 *>(src),
                     static_cast<trid::u32x4 *>(dst), n16);
  return hipGetLastError() == hipSuccess ? TRID_STATUS_SUCCESS
                                         : TRID_STATUS_EXECUTION_FAILED;
}
