// trid_internal.h - host-side declarations shared by the C ABI layer and the
// kernel translation units of libtridhip.so. Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <atomic>
#include <cstddef>
#include <cstdint>

#include "../../include/trid_common.h"
#include "../../include/trid_params.h"

namespace trid {

// Shape of one batched solve, after the (ndim, solvedim, dims, pads) of the
// C API have been resolved (the reference does the same per solvedim in
// scalar/src/cpu/trid_cpu.cpp:369-454):
//   system s = i0 + n0 * i1 starts at element i0 * st0 + i1 * st1 and its
//   element n is at that base + n * stride.
// For solvedim 1 and 2, i0 is the x index (st0 = 1), so neighbouring lanes
// read neighbouring addresses. For solvedim 0, i0 is y (st0 = pads[0]) and the
// system itself is contiguous (stride = 1).
struct Geom {
  int64_t N;       // system length (dims[solvedim])
  int64_t stride;  // element stride inside a system
  int64_t n0, n1;  // systems = n0 * n1
  int64_t st0, st1;
  __host__ __device__ int64_t nsys() const { return n0 * n1; }
};

// Tuning of the Thomas-PCR hybrid for one call: M rows per lane, P lanes per
// system, W systems per workgroup (block = W * P threads), and G, the number
// of groups the workgroups are dealt into so that consecutive system groups
// run on the same XCD at the same time (workgroups are dispatched to the 8
// XCDs round-robin; G = 1 keeps the plain order).
// V selects the kernel variant: 0 the plain hybrid (all loads before the
// sweep); 1 and 2 the streamed strided variants (rows loaded a few ahead of
// the sweep; kernels_stream.hip); 3 the strided variant with `a` staged in
// LDS (kernels_wide.hip). For the mixed-precision solve, 3 is the M = 16
// kernel with the FP64 a and c operands in LDS (kernels_mixed.hip); for the
// solve of FP32-stored systems, the wide strided kernel (kernels_mixed32.hip).
struct HybridCfg {
  int M = 0;
  int P = 0;
  int W = 0;
  int G = 8;
  int V = 0;
};

enum class Algo : int { Auto = 0, Lane = 1, Hybrid = 3 };

// Elements from the first element of a batch to its last, pads included
// (the size of an array in the batch's layout), and the grid of the
// grid-stride elementwise kernels over its nsys * N elements (256 threads).
inline size_t batch_span(const Geom &g) {
  return (size_t)((g.n0 - 1) * g.st0 + (g.n1 - 1) * g.st1 + (g.N - 1) * g.stride + 1);
}
inline unsigned batch_elem_blocks(const Geom &g) {
  const int64_t total = g.nsys() * g.N;
  const int64_t blocks = (total + 255) / 256;
  return (unsigned)(blocks < (int64_t(1) << 20) ? blocks : (int64_t(1) << 20));
}

// ---- launchers (kernels_solve.hip) --------------------------------------
// Per-lane Thomas, any length. cws: N * nsys reals of workspace; dws: same
// size, only needed for Inc (the plain solve keeps d' in d itself).
template <typename REAL, bool INC>
hipError_t launch_lane(const Geom &g, const REAL *a, const REAL *b, const REAL *c,
                       REAL *d, REAL *u, REAL *cws, REAL *dws, hipStream_t st);

// Thomas-PCR hybrid; contiguous systems (stride == 1) use the LDS-staged
// loader, strided systems load per lane with x-adjacent lanes.
template <typename REAL, bool INC>
hipError_t launch_hybrid(const Geom &g, const HybridCfg &cfg, const REAL *a,
                         const REAL *b, const REAL *c, REAL *d, REAL *u,
                         hipStream_t st);

// ---- shared host helpers (trid_api.hip) ---------------------------------
// The batched solve behind tridXmtsvStridedBatch[Inc] (shape choice, launch,
// ctx->sync).
template <typename REAL, bool INC>
tridStatus_t solve(const TridParams *ctx, const REAL *a, const REAL *b, const REAL *c,
                   REAL *d, REAL *u, int ndim, int solvedim, const int *dims,
                   const int *pads);
hipStream_t current_stream();
tridStatus_t check_device(int *dev);
// role 0: the solve's / ragged plan's workspace; 1: partitioned-solve saves;
// 2: the transposed coefficients of the long-system adjoint, the promoted
// coefficients of the long-system FP32-stored solve
tridStatus_t get_workspace(int dev, hipStream_t st, size_t bytes, void **out, int role = 0);
// Role-0 workspace bytes the per-lane path would take for this call (0 when
// it runs a hybrid kernel), and the release of a destroyed stream's buffers.
template <typename REAL, bool INC>
size_t lane_workspace_need(int dev, const TridParams *ctx, int ndim, int solvedim,
                           const int *dims, const int *pads);
void drop_stream_workspaces(int dev, hipStream_t st);

// ---- systems longer than one workgroup holds (kernels_long.hip) -----------
// The hybrid with its reduced system solved by a second batched solve;
// workspace: long_workspace_bytes(g) bytes.
template <typename REAL>
size_t long_workspace_bytes(const Geom &g);
template <typename REAL, bool INC>
hipError_t launch_long(int dev, const Geom &g, const REAL *a, const REAL *b, const REAL *c,
                       REAL *d, REAL *u, void *ws, size_t ws_bytes, hipStream_t st);

// ---- streamed strided variant (kernels_stream.hip) ------------------------
template <typename REAL>
bool stream_cfg_valid(const Geom &g, const HybridCfg &c);
template <typename REAL, bool INC>
hipError_t launch_stream(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                         const REAL *c, REAL *d, REAL *u, hipStream_t st);

// ---- wide strided variant (kernels_wide.hip) ------------------------------
template <typename REAL>
bool wide_cfg_valid(const Geom &g, const HybridCfg &c);
template <typename REAL, bool INC>
hipError_t launch_wide(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                       const REAL *c, REAL *d, REAL *u, hipStream_t st);

// ---- mixed precision (kernels_mixed.hip), FP64 operands ------------------
bool mixed_cfg_valid(const Geom &g, const HybridCfg &c);
bool mixed_default_cfg(const Geom &g, HybridCfg *cfg);
hipError_t launch_mixed(const Geom &g, const HybridCfg &cfg, int refine, const double *a,
                        const double *b, const double *c, double *d, hipStream_t st);

// ---- FP64 solution of systems stored in FP32 (kernels_mixed32.hip) -------
// refine: the refinements the shape is to run (more than one needs a
// general-loop kernel of the shape)
bool mixed32_cfg_valid(const Geom &g, const HybridCfg &c, int refine);
bool mixed32_default_cfg(const Geom &g, int refine, HybridCfg *cfg);
hipError_t launch_mixed32(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                          const float *b, const float *c, const float *d, double *x,
                          hipStream_t st);
// Systems no such kernel holds: a, b, c promoted to FP64 arrays A, B, C of
// batch_span(g) elements in the batch's layout, and d to x, which the FP64
// batched solve of (A, B, C) then overwrites with the solution.
hipError_t launch_mixed32_promote(const Geom &g, const float *a, const float *b, const float *c,
                                  const float *d, double *A, double *B, double *C, double *x,
                                  hipStream_t st);

// ---- adjoint (kernels_adjoint.hip) ----------------------------------------
template <typename REAL>
bool adjoint_cfg_valid(const Geom &g, const HybridCfg &c);
template <typename REAL>
bool adjoint_default_cfg(const Geom &g, HybridCfg *cfg);
template <typename REAL>
hipError_t launch_adjoint(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                          const REAL *c, const REAL *x, const REAL *w, REAL *lam, REAL *ga,
                          REAL *gb, REAL *gc, hipStream_t st);
// Systems longer than the adjoint kernel holds: sub[n] = c[n-1], sup[n] =
// a[n+1] and lam = w, in the inputs' layout (arrays of batch_span(g)
// elements); after the batched solve of (sub, b, sup) into lam, the
// requested gradients from lam and x.
template <typename REAL>
hipError_t launch_adjoint_transpose(const Geom &g, const REAL *a, const REAL *c, const REAL *w,
                                    REAL *sub, REAL *sup, REAL *lam, hipStream_t st);
template <typename REAL>
hipError_t launch_adjoint_grads(const Geom &g, const REAL *x, const REAL *lam, REAL *ga,
                                REAL *gb, REAL *gc, hipStream_t st);

// ---- ragged batch (kernels_ragged.hip) ------------------------------------
// A ragged batch plans in ragged_workspace_bytes(S) bytes: this header (plan
// counters and tables) followed by one 32-bit slot per system.
constexpr size_t RAGGED_PLAN_HEADER_BYTES = 512;
int ragged_mixed_max_len(int M);
int ragged_mixed32_max_len();
hipError_t launch_ragged_mixed32(const float *a, const float *b, const float *c, const float *d,
                                 double *x, const int64_t *off, int64_t S, int per_cu,
                                 int launches, int refine, void *ws, hipStream_t st, int dev);
hipError_t launch_ragged_mixed(int M, const double *a, const double *b, const double *c,
                               double *d, const int64_t *off, int64_t S, int per_cu,
                               int launches, int refine, void *ws, hipStream_t st, int dev);
size_t ragged_workspace_bytes(int64_t S);
int ragged_max_len(int fp64, int M);
template <typename REAL, bool INC>
hipError_t launch_ragged(int M, const REAL *a, const REAL *b, const REAL *c, REAL *d, REAL *u,
                         const int64_t *off, int64_t S, int per_cu, int launches, void *ws,
                         hipStream_t st, int dev);
hipError_t ragged_error_flag(void *ws, int *flag);

// The dynamic-LDS limit granted to one kernel, per device (the attribute is
// set on the current device; a process may drive several).
constexpr int LDS_GRANT_DEVICES = 64;
struct LdsGrant {
  std::atomic<size_t> bytes[LDS_GRANT_DEVICES] = {};
};

// Raises a kernel's dynamic-LDS limit on the current device to `lds` bytes
// when it exceeds the default 64 KiB and what was granted there before (not
// capturable: the first call with a given size must happen outside hipGraph
// capture).
inline hipError_t ensure_lds(const void *kern, size_t lds, LdsGrant &granted) {
  if (lds <= 65536) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= LDS_GRANT_DEVICES) return hipErrorInvalidDevice;
  if (lds <= granted.bytes[dev].load()) return hipSuccess;
  e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) granted.bytes[dev].store(lds);
  return e;
}

// A HIP grid holds fewer than 2^32 work-items (gridDim.x * blockDim.x). The
// batched kernels launch ceil(n0 / W) * n1 workgroups of T threads, which can
// pass that for 2^31 short systems. This calls fn(sub, off) once per
// sub-batch whose grid fits, where `sub` covers a range of i1 (or, when one
// i1 slice alone is too large, a range of i0 in whole workgroups) and `off`
// is the element offset of its first system: the caller launches on `sub`
// with every array pointer advanced by `off`. One call when the grid fits.
template <class F>
hipError_t for_each_grid_chunk(const Geom &g, int W, int T, F &&fn) {
  const int64_t max_blocks = (int64_t)UINT32_MAX / T;
  const int64_t nb0 = (g.n0 + W - 1) / W;
  if (nb0 * g.n1 <= max_blocks) return fn(g, int64_t(0));
  if (nb0 <= max_blocks) {
    const int64_t per = max_blocks / nb0;
    for (int64_t i1 = 0; i1 < g.n1; i1 += per) {
      Geom s = g;
      s.n1 = g.n1 - i1 < per ? g.n1 - i1 : per;
      const hipError_t e = fn(s, i1 * g.st1);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  const int64_t per0 = max_blocks * W;
  for (int64_t i1 = 0; i1 < g.n1; ++i1) {
    for (int64_t i0 = 0; i0 < g.n0; i0 += per0) {
      Geom s = g;
      s.n0 = g.n0 - i0 < per0 ? g.n0 - i0 : per0;
      s.n1 = 1;
      const hipError_t e = fn(s, i0 * g.st0 + i1 * g.st1);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

// p + off, or nullptr for an absent (null) array.
template <typename T>
inline T *advance(T *p, int64_t off) {
  return p ? p + off : nullptr;
}

// ---- shared entry-point helpers (trid_api.hip) ----------------------------
// Records the HIP error behind a failed call (tridHipLastErrorString()).
void record_hip_error(hipError_t e);
tridStatus_t make_geom(int ndim, int solvedim, const int *dims, const int *pads, Geom *g,
                       bool *empty);
tridStatus_t finish(const TridParams *ctx, hipStream_t st, hipError_t launched);
void record_config(int algo, const HybridCfg *cfg);
const HybridCfg &forced_config();

// Picks a valid hybrid configuration for g, or returns false if the systems
// are too long for one workgroup.
template <typename REAL>
bool hybrid_default_cfg(const Geom &g, HybridCfg *cfg);
// The default among plain-kernel shapes only (V = 0).
template <typename REAL>
bool hybrid_plain_cfg(const Geom &g, HybridCfg *cfg);
template <typename REAL>
bool hybrid_cfg_valid(const Geom &g, const HybridCfg &cfg);

}  // namespace trid
