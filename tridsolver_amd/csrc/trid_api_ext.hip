// trid_api_ext.hip - C ABI entry points of the extensions built on the
// batched solve: FP32 sweeps with FP64 refinement, the adjoint, and ragged
// batches. None has a counterpart in the reference (see include/tridsolver_hip.h).

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "../../include/tridsolver_hip.h"
#include "trid_internal.h"
#include "trid_tune.h"

namespace trid {
namespace {

thread_local int tl_ragged_m = 0, tl_ragged_per_cu = 1, tl_ragged_launches = 1;

template <typename REAL, bool INC>
tridStatus_t ragged(const TridParams *ctx, const REAL *a, const REAL *b, const REAL *c, REAL *d,
                    REAL *u, const int64_t *offsets, int64_t nsys, int64_t max_len) {
  if (!a || !b || !c || !d || !offsets || (INC && !u) || nsys < 0 || nsys > INT32_MAX ||
      max_len < 0)
    return TRID_STATUS_INVALID_VALUE;  // the plan's system list holds 32-bit indices
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (nsys == 0) return finish(ctx, st, hipSuccess);
  int M = tl_ragged_m, per_cu = tl_ragged_per_cu, launches = tl_ragged_launches;
  if (M == 0) ragged_tuned_config<REAL>(dev, max_len, &M, &per_cu, &launches);
  const int supported = ragged_max_len(sizeof(REAL) == 8, M);
  if (supported == 0 || max_len > supported) return TRID_STATUS_INVALID_VALUE;
  // the three-launch split is built for 16 rows per lane
  if (launches == 3 && M != 16) return TRID_STATUS_INVALID_VALUE;
  void *ws = nullptr;
  if ((s = get_workspace(dev, st, ragged_workspace_bytes(nsys), &ws)) != TRID_STATUS_SUCCESS)
    return s;
  HybridCfg rec;
  rec.M = M; rec.P = 0; rec.W = launches; rec.G = per_cu;
  record_config(5, &rec);
  s = finish(ctx, st, launch_ragged<REAL, INC>(M, a, b, c, d, u, offsets, nsys, per_cu, launches,
                                              ws, st, dev));
  if (s == TRID_STATUS_SUCCESS && (ctx == nullptr || ctx->sync != 0)) {
    int flag = 0;
    if (ragged_error_flag(ws, &flag) != hipSuccess) return TRID_STATUS_EXECUTION_FAILED;
    if (flag) return TRID_STATUS_INVALID_VALUE;  // a system longer than supported
  }
  return s;
}

tridStatus_t ragged_mixed(const TridParams *ctx, const double *a, const double *b,
                          const double *c, double *d, const int64_t *offsets, int64_t nsys,
                          int64_t max_len, int refine) {
  if (!a || !b || !c || !d || !offsets || nsys < 0 || nsys > INT32_MAX || max_len < 0 ||
      refine < 0 || refine > 8)
    return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (nsys == 0) return finish(ctx, st, hipSuccess);
  // rows per lane: the forced value, else 16 (a and c in LDS, up to 8192
  // rows; M = 4 and 8 keep every operand in registers, up to 2048 / 4096)
  const bool forced = tl_ragged_m != 0;
  const int M = forced ? tl_ragged_m : 16;
  const int supported = ragged_mixed_max_len(M);
  if (supported == 0 || max_len > supported) return TRID_STATUS_INVALID_VALUE;
  void *ws = nullptr;
  if ((s = get_workspace(dev, st, ragged_workspace_bytes(nsys), &ws)) != TRID_STATUS_SUCCESS)
    return s;
  HybridCfg rec;
  // M = 4 and 8 always run in one launch; the mixed solve has no
  // three-launch split. Unforced, batches with systems over 4096 rows take
  // two launches (those systems in 512-thread workgroups, then the rest in
  // 256-thread ones, two per CU): 0.56 against 0.58 ms on 40,000 systems of
  // 8-8192 rows.
  const int launches = M != 16 ? 1 : forced ? tl_ragged_launches : (max_len > 4096 ? 2 : 1);
  if (launches == 3) return TRID_STATUS_INVALID_VALUE;
  rec.M = M; rec.P = 0; rec.W = launches; rec.G = tl_ragged_per_cu;
  record_config(9, &rec);
  s = finish(ctx, st, launch_ragged_mixed(M, a, b, c, d, offsets, nsys, tl_ragged_per_cu,
                                          launches, refine, ws, st, dev));
  if (s == TRID_STATUS_SUCCESS && (ctx == nullptr || ctx->sync != 0)) {
    int flag = 0;
    if (ragged_error_flag(ws, &flag) != hipSuccess) return TRID_STATUS_EXECUTION_FAILED;
    if (flag) return TRID_STATUS_INVALID_VALUE;  // a system longer than supported
  }
  return s;
}

template <typename REAL>
tridStatus_t adjoint(const TridParams *ctx, const REAL *a, const REAL *b, const REAL *c,
                     const REAL *x, const REAL *w, REAL *lam, REAL *ga, REAL *gb, REAL *gc,
                     int ndim, int solvedim, const int *dims, const int *pads) {
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  // x is read only for the coefficient gradients
  if (!a || !b || !c || !w || !lam || (!x && (ga || gb || gc))) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (empty) return finish(ctx, st, hipSuccess);
  HybridCfg cfg;
  const HybridCfg &forced = forced_config();
  if (forced.M != 0) {
    cfg = forced;
    if (!adjoint_cfg_valid<REAL>(g, cfg)) return TRID_STATUS_INVALID_VALUE;
  } else if (!adjoint_tuned_config<REAL>(dev, g, ga || gb || gc, &cfg)) {
    // longer than the adjoint kernel holds: the transpose's off-diagonals
    // into the stream's role-2 workspace, A^T lam = w by the batched solve
    // (which takes any length), then the gradients
    const size_t span = batch_span(g);
    void *ws = nullptr;
    if ((s = get_workspace(dev, st, 2 * span * sizeof(REAL), &ws, 2)) != TRID_STATUS_SUCCESS)
      return s;
    REAL *sub = static_cast<REAL *>(ws), *sup = sub + span;
    hipError_t e = launch_adjoint_transpose<REAL>(g, a, c, w, sub, sup, lam, st);
    if (e != hipSuccess) return finish(ctx, st, e);
    TridParams inner{};  // the solve's defaults, no synchronisation
    if ((s = solve<REAL, false>(&inner, sub, b, sup, lam, nullptr, ndim, solvedim, dims, pads)) !=
        TRID_STATUS_SUCCESS)
      return s;
    record_config(14, nullptr);
    return finish(ctx, st, (ga || gb || gc) ? launch_adjoint_grads<REAL>(g, x, lam, ga, gb, gc, st)
                                            : hipSuccess);
  }
  record_config(4, &cfg);
  return finish(ctx, st, for_each_grid_chunk(g, cfg.W, cfg.W * cfg.P,
                                             [&](const Geom &sub, int64_t o) {
    return launch_adjoint<REAL>(sub, cfg, a + o, b + o, c + o, advance(x, o), w + o, lam + o,
                                advance(ga, o), advance(gb, o), advance(gc, o), st);
  }));
}

tridStatus_t ragged_mixed32(const TridParams *ctx, const float *a, const float *b,
                            const float *c, const float *d, double *x, const int64_t *offsets,
                            int64_t nsys, int64_t max_len, int refine) {
  if (!a || !b || !c || !d || !x || !offsets || nsys < 0 || nsys > INT32_MAX || max_len < 0 ||
      refine < 0 || refine > 8)
    return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (nsys == 0) return finish(ctx, st, hipSuccess);
  if (max_len > ragged_mixed32_max_len()) return TRID_STATUS_INVALID_VALUE;
  void *ws = nullptr;
  if ((s = get_workspace(dev, st, ragged_workspace_bytes(nsys), &ws)) != TRID_STATUS_SUCCESS)
    return s;
  // launches: forced (1 or 2), else two where systems may exceed 4096 rows
  // (as the FP64-stored mixed solve)
  const int launches = tl_ragged_m ? tl_ragged_launches : (max_len > 4096 ? 2 : 1);
  if (launches != 1 && launches != 2) return TRID_STATUS_INVALID_VALUE;
  HybridCfg rec;
  rec.M = 16; rec.P = 0; rec.W = launches; rec.G = tl_ragged_per_cu;
  record_config(11, &rec);
  s = finish(ctx, st, launch_ragged_mixed32(a, b, c, d, x, offsets, nsys, tl_ragged_per_cu,
                                            launches, refine, ws, st, dev));
  if (s == TRID_STATUS_SUCCESS && (ctx == nullptr || ctx->sync != 0)) {
    int flag = 0;
    if (ragged_error_flag(ws, &flag) != hipSuccess) return TRID_STATUS_EXECUTION_FAILED;
    if (flag) return TRID_STATUS_INVALID_VALUE;  // a system longer than supported
  }
  return s;
}

}  // namespace
}  // namespace trid

using namespace trid;

EXTERN_C tridStatus_t tridDmtsvStridedBatchMixed(const TridParams *ctx, const double *a,
                                                 const double *b, const double *c, double *d,
                                                 int ndim, int solvedim, const int *dims,
                                                 const int *pads, int refinements) {
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (!a || !b || !c || !d || refinements < 0 || refinements > 8) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (empty) return finish(ctx, st, hipSuccess);
  HybridCfg cfg;
  const HybridCfg &forced = forced_config();
  if (forced.M != 0) {
    cfg = forced;
    if (!mixed_cfg_valid(g, cfg)) return TRID_STATUS_INVALID_VALUE;
  } else if (!mixed_tuned_config(dev, g, &cfg)) {
    // systems no mixed kernel holds: the FP64 batched solve (any length, at
    // least the accuracy the refinements reach)
    TridParams inner{};  // the solve's defaults, no synchronisation
    if ((s = solve<double, false>(&inner, a, b, c, d, nullptr, ndim, solvedim, dims, pads)) !=
        TRID_STATUS_SUCCESS)
      return s;
    record_config(15, nullptr);
    return finish(ctx, st, hipSuccess);
  }
  record_config(2, &cfg);
  return finish(ctx, st, for_each_grid_chunk(g, cfg.W, cfg.W * cfg.P,
                                             [&](const Geom &sub, int64_t o) {
    return launch_mixed(sub, cfg, refinements, a + o, b + o, c + o, d + o, st);
  }));
}

EXTERN_C tridStatus_t tridDmtsvStridedBatchMixedF32(const TridParams *ctx, const float *a,
                                                    const float *b, const float *c,
                                                    const float *d, double *x, int ndim,
                                                    int solvedim, const int *dims,
                                                    const int *pads, int refinements) {
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (!a || !b || !c || !d || !x || refinements < 0 || refinements > 8)
    return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (empty) return finish(ctx, st, hipSuccess);
  HybridCfg cfg;
  const HybridCfg &forced = forced_config();
  if (forced.M != 0) {
    cfg = forced;
    if (!mixed32_cfg_valid(g, cfg, refinements)) return TRID_STATUS_INVALID_VALUE;
  } else if (!mixed32_tuned_config(dev, g, refinements, &cfg)) {
    // systems no such kernel holds: a, b, c promoted into the stream's role-2
    // workspace and d into x, then the FP64 batched solve of the promoted
    // system (any length) in x
    const size_t span = batch_span(g);
    void *ws = nullptr;
    if ((s = get_workspace(dev, st, 3 * span * sizeof(double), &ws, 2)) != TRID_STATUS_SUCCESS)
      return s;
    double *A = static_cast<double *>(ws), *B = A + span, *C = B + span;
    const hipError_t e = launch_mixed32_promote(g, a, b, c, d, A, B, C, x, st);
    if (e != hipSuccess) return finish(ctx, st, e);
    TridParams inner{};  // the solve's defaults, no synchronisation
    if ((s = solve<double, false>(&inner, A, B, C, x, nullptr, ndim, solvedim, dims, pads)) !=
        TRID_STATUS_SUCCESS)
      return s;
    record_config(16, nullptr);
    return finish(ctx, st, hipSuccess);
  }
  record_config(10, &cfg);
  return finish(ctx, st, for_each_grid_chunk(g, cfg.W, cfg.W * cfg.P,
                                             [&](const Geom &sub, int64_t o) {
    return launch_mixed32(sub, cfg, refinements, a + o, b + o, c + o, d + o, x + o, st);
  }));
}

EXTERN_C tridStatus_t tridDmtsvStridedBatchAdjoint(const TridParams *ctx, const double *a,
                                                   const double *b, const double *c,
                                                   const double *x, const double *w,
                                                   double *lam, double *grad_a,
                                                   double *grad_b, double *grad_c, int ndim,
                                                   int solvedim, const int *dims,
                                                   const int *pads) {
  return adjoint<double>(ctx, a, b, c, x, w, lam, grad_a, grad_b, grad_c, ndim, solvedim,
                         dims, pads);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchAdjoint(const TridParams *ctx, const float *a,
                                                   const float *b, const float *c,
                                                   const float *x, const float *w, float *lam,
                                                   float *grad_a, float *grad_b, float *grad_c,
                                                   int ndim, int solvedim, const int *dims,
                                                   const int *pads) {
  return adjoint<float>(ctx, a, b, c, x, w, lam, grad_a, grad_b, grad_c, ndim, solvedim, dims,
                        pads);
}

EXTERN_C tridStatus_t tridDmtsvRaggedBatch(const TridParams *ctx, const double *a,
                                           const double *b, const double *c, double *d,
                                           const int64_t *offsets, int64_t nsys,
                                           int64_t max_len) {
  return ragged<double, false>(ctx, a, b, c, d, nullptr, offsets, nsys, max_len);
}

EXTERN_C tridStatus_t tridSmtsvRaggedBatch(const TridParams *ctx, const float *a,
                                           const float *b, const float *c, float *d,
                                           const int64_t *offsets, int64_t nsys,
                                           int64_t max_len) {
  return ragged<float, false>(ctx, a, b, c, d, nullptr, offsets, nsys, max_len);
}

EXTERN_C tridStatus_t tridDmtsvRaggedBatchInc(const TridParams *ctx, const double *a,
                                              const double *b, const double *c, double *d,
                                              double *u, const int64_t *offsets, int64_t nsys,
                                              int64_t max_len) {
  return ragged<double, true>(ctx, a, b, c, d, u, offsets, nsys, max_len);
}

EXTERN_C tridStatus_t tridSmtsvRaggedBatchInc(const TridParams *ctx, const float *a,
                                              const float *b, const float *c, float *d,
                                              float *u, const int64_t *offsets, int64_t nsys,
                                              int64_t max_len) {
  return ragged<float, true>(ctx, a, b, c, d, u, offsets, nsys, max_len);
}

EXTERN_C tridStatus_t tridHipForceRaggedConfig(int M, int per_cu, int launches) {
  if (M == 0) {
    tl_ragged_m = 0;
    tl_ragged_per_cu = 1;
    tl_ragged_launches = 1;
    return TRID_STATUS_SUCCESS;
  }
  if ((M != 4 && M != 8 && M != 16) || per_cu < 1 || per_cu > 8 || launches < 1 ||
      launches > 3)
    return TRID_STATUS_INVALID_VALUE;
  tl_ragged_m = M;
  tl_ragged_per_cu = per_cu;
  tl_ragged_launches = launches;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C size_t tridHipRaggedWorkspaceBytes(int64_t nsys) { return ragged_workspace_bytes(nsys); }

EXTERN_C tridStatus_t tridDmtsvRaggedBatchMixedF32(const TridParams *ctx, const float *a,
                                                  const float *b, const float *c,
                                                  const float *d, double *x,
                                                  const int64_t *offsets, int64_t nsys,
                                                  int64_t max_len, int refinements) {
  return ragged_mixed32(ctx, a, b, c, d, x, offsets, nsys, max_len, refinements);
}

EXTERN_C tridStatus_t tridDmtsvRaggedBatchMixed(const TridParams *ctx, const double *a,
                                                const double *b, const double *c, double *d,
                                                const int64_t *offsets, int64_t nsys,
                                                int64_t max_len, int refinements) {
  return ragged_mixed(ctx, a, b, c, d, offsets, nsys, max_len, refinements);
}
