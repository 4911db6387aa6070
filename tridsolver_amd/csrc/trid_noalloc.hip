// trid_noalloc.hip - tridDmtsvRaggedBatchNoAlloc: the FP64 ragged batch
// solve for callers that cannot allocate device memory (code under hipGraph
// capture, tight serving loops, harnesses that forbid hipMalloc). No
// counterpart in the reference.
//
// The ragged solve plans in a workspace of ragged_workspace_bytes(nsys)
// bytes (a small header plus one 32-bit slot per system), which
// tridDmtsvRaggedBatch takes from the per-stream workspace and grows with
// hipMalloc. Here it is the file-scope device array `scratch`, sized for
// NOALLOC_MAX_SYSTEMS systems; each GPU the library runs on holds its own copy
// until the process ends. Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): the device
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and kept in g_scratch.
//
// The planning kernels clear what must start at zero (counters, work queue,
// error flag) at the start of every call, so a call does not depend on what
// the previous one left. Every call on one GPU uses that GPU's single copy:
// calls that overlap on one GPU (concurrent streams or host threads) race.

#include <mutex>

#include "../../include/tridsolver_hip.h"
#include "trid_internal.h"
#include "trid_tune.h"

namespace trid {
namespace {

constexpr int64_t NOALLOC_MAX_SYSTEMS = TRID_HIP_NOALLOC_MAX_SYSTEMS;
constexpr size_t NOALLOC_SCRATCH_BYTES = RAGGED_PLAN_HEADER_BYTES + 4 * NOALLOC_MAX_SYSTEMS;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[NOALLOC_DEVICES] = {};

// Device address of this device's `scratch` (looked up on the first call on
// the device, with it current).
tridStatus_t scratch_of(int dev, void **out) {
  if (dev < 0 || dev >= NOALLOC_DEVICES) return TRID_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      record_hip_error(e);
      return TRID_STATUS_EXECUTION_FAILED;
    }
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return TRID_STATUS_SUCCESS;
}

}  // namespace
}  // namespace trid

using namespace trid;

EXTERN_C tridStatus_t tridDmtsvRaggedBatchNoAlloc(const TridParams *ctx, const double *a,
                                                  const double *b, const double *c, double *d,
                                                  const int64_t *offsets, int64_t nsys,
                                                  int64_t max_len) {
  static_assert(NOALLOC_SCRATCH_BYTES >= RAGGED_PLAN_HEADER_BYTES + 4 * NOALLOC_MAX_SYSTEMS,
                "scratch holds the plan of the largest batch");
  if (!a || !b || !c || !d || !offsets || nsys < 0 || max_len < 0)
    return TRID_STATUS_INVALID_VALUE;
  if (nsys > NOALLOC_MAX_SYSTEMS) return TRID_STATUS_INVALID_VALUE;  // beyond the scratch
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (nsys == 0) return finish(ctx, st, hipSuccess);
  // the tuned shape if one is cached for this device and length bound, else
  // the default: never a tuning run (it allocates)
  int M = 0, per_cu = 1, launches = 1;
  ragged_tuned_config<double>(dev, max_len, &M, &per_cu, &launches);
  const int supported = ragged_max_len(1, M);
  if (supported == 0 || max_len > supported) return TRID_STATUS_INVALID_VALUE;
  if (ragged_workspace_bytes(nsys) > NOALLOC_SCRATCH_BYTES) return TRID_STATUS_INVALID_VALUE;
  void *ws = nullptr;
  if ((s = scratch_of(dev, &ws)) != TRID_STATUS_SUCCESS) return s;
  HybridCfg rec;
  rec.M = M; rec.P = 0; rec.W = launches; rec.G = per_cu;
  record_config(5, &rec);
  s = finish(ctx, st, launch_ragged<double, false>(M, a, b, c, d, nullptr, offsets, nsys, per_cu,
                                                   launches, ws, st, dev));
  if (s == TRID_STATUS_SUCCESS && (ctx == nullptr || ctx->sync != 0)) {
    int flag = 0;
    if (ragged_error_flag(ws, &flag) != hipSuccess) return TRID_STATUS_EXECUTION_FAILED;
    if (flag) return TRID_STATUS_INVALID_VALUE;  // a system longer than supported
  }
  return s;
}
