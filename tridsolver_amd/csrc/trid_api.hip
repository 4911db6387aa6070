// trid_api.hip - the C ABI of libtridhip.so: argument checks, device checks,
// stream/workspace handling and the choice of kernel for each call.
//
// Entry points and their reference counterparts:
//   tridDmtsvStridedBatch      scalar/include/tridsolver.h:38-42
//   tridSmtsvStridedBatch      scalar/include/tridsolver.h:44-48
//   tridDmtsvStridedBatchInc   scalar/include/tridsolver.h:57-62
//   tridSmtsvStridedBatchInc   scalar/include/tridsolver.h:64-68
// The reference's CUDA body (scalar/src/cuda/trid_cuda.cu:150-454) dispatches
// on solvedim and ctx->opts, transposes x through cuBLAS geam into static
// buffers, and always synchronises; here x is solved in place by a contiguous
// kernel, opts are hints, and ctx->sync is honoured.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/tridsolver.h"
#include "../../include/tridsolver_hip.h"
#include "trid_internal.h"
#include "trid_tune.h"

namespace trid {
namespace {

thread_local hipStream_t tl_stream = nullptr;
thread_local int tl_last_algo = 0, tl_last_m = 0, tl_last_p = 0, tl_last_w = 0,
                 tl_last_g = 0, tl_last_v = 0;
thread_local HybridCfg tl_forced;
thread_local hipError_t tl_hip_error = hipSuccess;

std::mutex g_mu;
struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};
// keyed by (device, stream, role): role 0 is the solve's and the ragged
// plan's workspace, role 1 the partitioned solve's saved boundary rows
std::map<std::tuple<int, hipStream_t, int>, Workspace> g_ws;
// Buffers replaced by a larger one. A hipGraph captured earlier may still
// point at them, so they are only freed by tridHipReleaseWorkspaces().
std::map<int, std::vector<void *>> g_retired;
std::map<int, tridStatus_t> g_arch;

}  // namespace

hipStream_t current_stream() { return tl_stream; }

tridStatus_t check_device(int *dev_out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TRID_STATUS_NOT_INITIALIZED;
  if (dev_out) *dev_out = dev;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_arch.find(dev);
  if (it != g_arch.end()) return it->second;
  hipDeviceProp_t prop;
  tridStatus_t st = TRID_STATUS_SUCCESS;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    st = TRID_STATUS_NOT_INITIALIZED;
  } else if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    st = TRID_STATUS_ARCH_MISMATCH;  // code objects are built for gfx950 only
  }
  g_arch[dev] = st;
  return st;
}

// Returns a workspace of at least `bytes` for (device, stream). Grows it with
// hipMalloc outside stream capture; inside capture a too-small workspace is
// an error (TRID_STATUS_INTERNAL_ERROR) because allocation is not capturable.
tridStatus_t get_workspace(int dev, hipStream_t st, size_t bytes, void **out, int role) {
  std::lock_guard<std::mutex> lk(g_mu);
  Workspace &ws = g_ws[std::make_tuple(dev, st, role)];
  if (ws.bytes >= bytes) {
    *out = ws.ptr;
    return TRID_STATUS_SUCCESS;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess) return TRID_STATUS_EXECUTION_FAILED;
  if (cap != hipStreamCaptureStatusNone) return TRID_STATUS_INTERNAL_ERROR;
  if (ws.ptr) {
    g_retired[dev].push_back(ws.ptr);
    ws.ptr = nullptr;
    ws.bytes = 0;
  }
  // Round up to 64 MiB to avoid regrowing for slightly larger calls.
  const size_t gran = size_t(64) << 20;
  const size_t want = (bytes + gran - 1) / gran * gran;
  if (hipMalloc(&ws.ptr, want) != hipSuccess) {
    (void)hipGetLastError();
    ws.ptr = nullptr;
    return TRID_STATUS_ALLOC_FAILED;
  }
  ws.bytes = want;
  *out = ws.ptr;
  return TRID_STATUS_SUCCESS;
}

// Resolves the C API's (ndim, solvedim, dims, pads) into a Geom, following
// the reference's per-solvedim strides (scalar/src/cpu/trid_cpu.cpp:363-454).
// *empty is set when there is nothing to solve (a zero extent).
tridStatus_t make_geom(int ndim, int solvedim, const int *dims, const int *pads,
                       Geom *g, bool *empty) {
  if (!dims || !pads || ndim < 1 || ndim > MAXDIM || solvedim < 0 || solvedim >= ndim)
    return TRID_STATUS_INVALID_VALUE;
  int64_t dm[3] = {1, 1, 1}, pd[3] = {1, 1, 1};
  *empty = false;
  for (int i = 0; i < ndim; ++i) {
    if (dims[i] < 0 || pads[i] < dims[i] || pads[i] < 1) return TRID_STATUS_INVALID_VALUE;
    if (dims[i] == 0) *empty = true;
    dm[i] = dims[i];
    pd[i] = pads[i];
  }
  // launch grids have at most one workgroup per system and HIP grids are
  // 32-bit: at most 2^31 - 1 systems per call
  int64_t systems = 1;
  for (int k = 0; k < 3; ++k)
    if (k != solvedim) systems *= dm[k];
  if (systems > INT32_MAX) return TRID_STATUS_INVALID_VALUE;
  g->N = dm[solvedim];
  if (solvedim == 0) {
    g->n0 = dm[1]; g->n1 = dm[2]; g->st0 = pd[0]; g->st1 = pd[0] * pd[1]; g->stride = 1;
  } else if (solvedim == 1) {
    g->n0 = dm[0]; g->n1 = dm[2]; g->st0 = 1; g->st1 = pd[0] * pd[1]; g->stride = pd[0];
  } else {
    g->n0 = dm[0]; g->n1 = dm[1]; g->st0 = 1; g->st1 = pd[0]; g->stride = pd[0] * pd[1];
  }
  return TRID_STATUS_SUCCESS;
}

void record_hip_error(hipError_t e) { tl_hip_error = e; }

tridStatus_t finish(const TridParams *ctx, hipStream_t st, hipError_t launched) {
  if (launched != hipSuccess) {
    record_hip_error(launched);
    return TRID_STATUS_EXECUTION_FAILED;
  }
  const bool sync = ctx == nullptr || ctx->sync != 0;
  if (sync) {
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      record_hip_error(e);
      return TRID_STATUS_EXECUTION_FAILED;
    }
  }
  return TRID_STATUS_SUCCESS;
}

void record_config(int algo, const HybridCfg *cfg) {
  tl_last_algo = algo;
  tl_last_m = cfg ? cfg->M : 0;
  tl_last_p = cfg ? cfg->P : 0;
  tl_last_w = cfg ? cfg->W : 0;
  tl_last_g = cfg ? cfg->G : 0;
  tl_last_v = cfg ? cfg->V : 0;
}

const HybridCfg &forced_config() { return tl_forced; }

// Chooses and launches the kernel for one batched solve.
template <typename REAL, bool INC>
tridStatus_t solve(const TridParams *ctx, const REAL *a, const REAL *b, const REAL *c,
                   REAL *d, REAL *u, int ndim, int solvedim, const int *dims,
                   const int *pads) {
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (!a || !b || !c || !d || (INC && !u)) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = tl_stream;
  if (empty) return finish(ctx, st, hipSuccess);

  const int hint = ctx ? ctx->opts[solvedim] : 0;
  HybridCfg cfg;
  bool use_hybrid = false, use_long = false;
  if (tl_forced.M != 0) {
    cfg = tl_forced;
    if (!hybrid_cfg_valid<REAL>(g, cfg)) return TRID_STATUS_INVALID_VALUE;
    use_hybrid = true;
  } else if (hint == 1 || hint == 2) {
    use_hybrid = false;
  } else {
    // 12: longer than one workgroup holds, the hybrid with an outside reduced solve
    const int algo = tuned_solve_algo<REAL, INC>(dev, g, &cfg);
    use_hybrid = algo == 3;
    use_long = algo == 12;
  }

  if (use_hybrid) {
    record_config(3, &cfg);
    return finish(ctx, st, for_each_grid_chunk(g, cfg.W, cfg.W * cfg.P,
                                               [&](const Geom &sub, int64_t o) {
      return launch_hybrid<REAL, INC>(sub, cfg, a + o, b + o, c + o, d + o, advance(u, o), st);
    }));
  }
  if (use_long) {
    record_config(12, nullptr);
    const size_t need = long_workspace_bytes<REAL>(g);
    void *lws = nullptr;
    if ((s = get_workspace(dev, st, need, &lws)) != TRID_STATUS_SUCCESS) return s;
    return finish(ctx, st, launch_long<REAL, INC>(dev, g, a, b, c, d, u, lws, need, st));
  }
  record_config(1, nullptr);
  const size_t one = sizeof(REAL) * (size_t)g.nsys() * (size_t)g.N;
  void *ws = nullptr;
  if ((s = get_workspace(dev, st, INC ? 2 * one : one, &ws)) != TRID_STATUS_SUCCESS) return s;
  REAL *cws = static_cast<REAL *>(ws);
  REAL *dws = INC ? cws + (size_t)g.nsys() * (size_t)g.N : nullptr;
  return finish(ctx, st, launch_lane<REAL, INC>(g, a, b, c, d, u, cws, dws, st));
}

// Workspace bytes the per-lane path would take for this call on this thread,
// 0 when the call runs a hybrid kernel (the decision solve() makes). A
// caller that captures such a call reserves this much for the capturing
// stream first.
template <typename REAL, bool INC>
size_t lane_workspace_need(int dev, const TridParams *ctx, int ndim, int solvedim,
                           const int *dims, const int *pads) {
  Geom g;
  bool empty = false;
  if (make_geom(ndim, solvedim, dims, pads, &g, &empty) != TRID_STATUS_SUCCESS || empty)
    return 0;
  if (tl_forced.M != 0) return 0;
  const int hint = ctx ? ctx->opts[solvedim] : 0;
  HybridCfg cfg;
  if (hint != 1 && hint != 2) {
    const int algo = tuned_solve_algo<REAL, INC>(dev, g, &cfg);
    if (algo == 3) return 0;
    if (algo == 12) return long_workspace_bytes<REAL>(g);
  }
  const size_t one = sizeof(REAL) * (size_t)g.nsys() * (size_t)g.N;
  return INC ? 2 * one : one;
}

// Frees the workspaces of a stream that is being destroyed (after its work
// has completed) and forgets them.
void drop_stream_workspaces(int dev, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto it = g_ws.begin(); it != g_ws.end();) {
    if (std::get<0>(it->first) == dev && std::get<1>(it->first) == st) {
      (void)hipFree(it->second.ptr);
      it = g_ws.erase(it);
    } else {
      ++it;
    }
  }
}

#define TRID_SOLVE_INST(REAL, INC)                                                          \
  template tridStatus_t solve<REAL, INC>(const TridParams *, const REAL *, const REAL *,   \
                                         const REAL *, REAL *, REAL *, int, int,           \
                                         const int *, const int *);
TRID_SOLVE_INST(float, false)
TRID_SOLVE_INST(float, true)
TRID_SOLVE_INST(double, false)
TRID_SOLVE_INST(double, true)
template size_t lane_workspace_need<float, false>(int, const TridParams *, int, int, const int *,
                                                  const int *);
template size_t lane_workspace_need<float, true>(int, const TridParams *, int, int, const int *,
                                                 const int *);
template size_t lane_workspace_need<double, false>(int, const TridParams *, int, int,
                                                   const int *, const int *);
template size_t lane_workspace_need<double, true>(int, const TridParams *, int, int, const int *,
                                                  const int *);

}  // namespace trid

using namespace trid;

// ---------------------------------------------------------------------------
// Reference API
// ---------------------------------------------------------------------------
EXTERN_C tridStatus_t tridDmtsvStridedBatch(const TridParams *ctx, const double *a,
                                            const double *b, const double *c, double *d,
                                            int ndim, int solvedim, const int *dims,
                                            const int *pads) {
  return solve<double, false>(ctx, a, b, c, d, nullptr, ndim, solvedim, dims, pads);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatch(const TridParams *ctx, const float *a,
                                            const float *b, const float *c, float *d,
                                            int ndim, int solvedim, const int *dims,
                                            const int *pads) {
  return solve<float, false>(ctx, a, b, c, d, nullptr, ndim, solvedim, dims, pads);
}

EXTERN_C tridStatus_t tridDmtsvStridedBatchInc(const TridParams *ctx, const double *a,
                                               const double *b, const double *c,
                                               double *d, double *u, int ndim,
                                               int solvedim, const int *dims,
                                               const int *pads) {
  return solve<double, true>(ctx, a, b, c, d, u, ndim, solvedim, dims, pads);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchInc(const TridParams *ctx, const float *a,
                                               const float *b, const float *c, float *d,
                                               float *u, int ndim, int solvedim,
                                               const int *dims, const int *pads) {
  return solve<float, true>(ctx, a, b, c, d, u, ndim, solvedim, dims, pads);
}

// ---------------------------------------------------------------------------
// Extensions
// ---------------------------------------------------------------------------
EXTERN_C const char *tridHipLastErrorString(void) {
  return tl_hip_error == hipSuccess ? "no error" : hipGetErrorString(tl_hip_error);
}

EXTERN_C tridStatus_t tridHipSetStream(void *stream) {
  tl_stream = static_cast<hipStream_t>(stream);
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipGetStream(void **stream) {
  if (!stream) return TRID_STATUS_INVALID_VALUE;
  *stream = tl_stream;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipReserveWorkspace(size_t bytes) {
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  void *p = nullptr;
  return get_workspace(dev, tl_stream, bytes, &p);
}

EXTERN_C tridStatus_t tridHipReleaseWorkspaces(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TRID_STATUS_NOT_INITIALIZED;
  if (hipDeviceSynchronize() != hipSuccess) return TRID_STATUS_EXECUTION_FAILED;
  std::lock_guard<std::mutex> lk(g_mu);
  for (void *p : g_retired[dev]) (void)hipFree(p);
  g_retired[dev].clear();
  for (auto it = g_ws.begin(); it != g_ws.end();) {
    if (std::get<0>(it->first) == dev) {
      (void)hipFree(it->second.ptr);
      it = g_ws.erase(it);
    } else {
      ++it;
    }
  }
  return TRID_STATUS_SUCCESS;
}

EXTERN_C size_t tridHipWorkspaceBytes(int fp64, int inc, int ndim, int solvedim,
                                      const int *dims) {
  Geom g;
  bool empty = false;
  if (make_geom(ndim, solvedim, dims, dims, &g, &empty) != TRID_STATUS_SUCCESS || empty)
    return 0;
  const size_t one = (fp64 ? 8 : 4) * (size_t)g.nsys() * (size_t)g.N;
  const size_t lane = inc ? 2 * one : one;
  // systems longer than one workgroup holds take the long-system path by
  // default, with its own workspace
  HybridCfg cfg;
  const bool fits = fp64 ? hybrid_default_cfg<double>(g, &cfg) : hybrid_default_cfg<float>(g, &cfg);
  if (fits) return lane;
  const size_t lng = fp64 ? long_workspace_bytes<double>(g) : long_workspace_bytes<float>(g);
  return std::max(lane, lng);
}

EXTERN_C tridStatus_t tridHipLastConfig(int *algorithm, int *M, int *P, int *W, int *G) {
  if (algorithm) {  // hybrid variants: 1-3 report 6-8, 4 reports 13
    *algorithm = tl_last_algo == 3 && tl_last_v > 0 ? (tl_last_v == 4 ? 13 : 5 + tl_last_v)
                                                    : tl_last_algo;
  }
  if (M) *M = tl_last_m;
  if (P) *P = tl_last_p;
  if (W) *W = tl_last_w;
  if (G) *G = tl_last_g;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipForceHybridConfig(int M, int P, int W, int G, int variant) {
  if (M == 0) {
    tl_forced = HybridCfg();
    return TRID_STATUS_SUCCESS;
  }
  if (M < 0 || P < 1 || W < 1 || G < 1 || variant < 0 || variant > 4)
    return TRID_STATUS_INVALID_VALUE;
  tl_forced.M = M;
  tl_forced.P = P;
  tl_forced.W = W;
  tl_forced.G = G;
  tl_forced.V = variant;
  return TRID_STATUS_SUCCESS;
}
