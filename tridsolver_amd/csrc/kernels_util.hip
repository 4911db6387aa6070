// kernels_util.hip - measurement helper of libtridhip.so: a plain streaming
// copy whose bytes moved are known exactly. It gives the HBM bandwidth the
// solver's roofline fraction is quoted against, and the reference kernel that
// calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters.
#include "../../include/tridsolver_hip.h"
#include <algorithm>

#include "trid_internal.h"

namespace trid {
namespace {

__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4 *__restrict__ src,
                                                           uint4 *__restrict__ dst,
                                                           size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
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
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t n16 = bytes / 16;
  size_t blocks = std::min<size_t>((n16 + 255) / 256, (size_t)cus * 16);
  hipLaunchKernelGGL(trid::stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     trid::current_stream(), static_cast<const uint4 *>(src),
                     static_cast<uint4 *>(dst), n16);
  return hipGetLastError() == hipSuccess ? TRID_STATUS_SUCCESS
                                         : TRID_STATUS_EXECUTION_FAILED;
}
