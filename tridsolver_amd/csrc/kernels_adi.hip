// kernels_adi.hip - the ADI heat-diffusion caller of the batched solve.
//
// The reference app (apps/adi/src/adi_cuda.cu:282-326) runs, per timestep,
//   preproc -> tridXmtsvStridedBatch(solvedim 0) -> (solvedim 1)
//           -> tridXmtsvStridedBatchInc(solvedim 2, u += x),
// with a device synchronisation inside every solve. Here the step is four
// stream-ordered launches with no host synchronisation, so a whole step can be
// captured once into a hipGraph and replayed (tridAdiRun*).
//
// preproc (reference apps/adi/src/preproc.cu:40-90, CPU apps/adi/src/preproc.hpp:172-214):
//   boundary points:  d = 0, a = 0, b = 1, c = 0
//   interior points:  d = lambda * (sum of the 6 neighbours - 6 u),
//                     a = c = -0.5 lambda, b = 1 + lambda.
// The reference writes the same a, b, c into nine arrays (x, y and z); the
// kernel writes each distinct output pointer once, so a caller may pass one
// coefficient set for all three directions (4 arrays written instead of 10).
// Each thread owns one (i, j) column and walks a chunk of k with the three
// planes k-1, k, k+1 of u in registers.

#include <algorithm>

#include "../../include/tridsolver.h"
#include "../../include/tridsolver_adi.h"
#include "../../include/tridsolver_hip.h"
#include "trid_internal.h"

namespace trid {
namespace {

constexpr int PRE_BX = 64, PRE_BY = 4, PRE_KCHUNK = 16;

template <typename REAL>
struct Coef {
  REAL *a, *b, *c;
};

template <typename REAL>
__global__ __launch_bounds__(PRE_BX *PRE_BY) void adi_preproc_kernel(
    REAL lambda, const REAL *__restrict__ u, REAL *__restrict__ du, Coef<REAL> c0,
    Coef<REAL> c1, Coef<REAL> c2, int nx, int ny, int nz) {
  const int i = blockIdx.x * PRE_BX + threadIdx.x;
  const int j = blockIdx.y * PRE_BY + threadIdx.y;
  const int k0 = blockIdx.z * PRE_KCHUNK;
  if (i >= nx || j >= ny || k0 >= nz) return;
  const int k1 = min(nz, k0 + PRE_KCHUNK);
  const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
  const bool edge_ij = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
  int64_t ind = (int64_t)k0 * sz + (int64_t)j * sy + i;
  REAL um = 0, uc = u[ind], up = 0;
  if (!edge_ij && k0 > 0) um = u[ind - sz];
  const REAL ca = REAL(-0.5) * lambda, cb = REAL(1) + lambda;
  for (int k = k0; k < k1; ++k, ind += sz) {
    const bool inner = !edge_ij && k > 0 && k < nz - 1;
    if (!edge_ij && k < nz - 1) up = u[ind + sz];
    REAL d, a, b, c;
    if (inner) {
      // lambda * (u[-1] + u[+1] + u[-nx] + u[+nx] + u[-nx*ny] + u[+nx*ny] - 6 u)
      d = lambda * (u[ind - sx] + u[ind + sx] + u[ind - sy] + u[ind + sy] + um + up -
                    REAL(6) * uc);
      a = ca;
      b = cb;
      c = ca;
    } else {
      d = 0; a = 0; b = 1; c = 0;
    }
    du[ind] = d;
    c0.a[ind] = a; c0.b[ind] = b; c0.c[ind] = c;
    if (c1.a) { c1.a[ind] = a; c1.b[ind] = b; c1.c[ind] = c; }
    if (c2.a) { c2.a[ind] = a; c2.b[ind] = b; c2.c[ind] = c; }
    um = uc;
    uc = up;
  }
}

template <typename REAL>
tridStatus_t preproc(REAL lambda, const REAL *u, REAL *du, REAL *ax, REAL *bx, REAL *cx,
                     REAL *ay, REAL *by, REAL *cy, REAL *az, REAL *bz, REAL *cz, int nx,
                     int ny, int nz) {
  if (!u || !du || !ax || !bx || !cx || nx < 1 || ny < 1 || nz < 1)
    return TRID_STATUS_INVALID_VALUE;
  // A direction whose three pointers are NULL, or equal to an earlier set,
  // is not written again.
  auto same = [](REAL *p, REAL *q, REAL *r, REAL *P, REAL *Q, REAL *R) {
    return p == P && q == Q && r == R;
  };
  Coef<REAL> c0{ax, bx, cx}, c1{nullptr, nullptr, nullptr}, c2{nullptr, nullptr, nullptr};
  const bool y_null = !ay && !by && !cy, z_null = !az && !bz && !cz;
  if ((!y_null && (!ay || !by || !cy)) || (!z_null && (!az || !bz || !cz)))
    return TRID_STATUS_INVALID_VALUE;
  if (!y_null && !same(ay, by, cy, ax, bx, cx)) c1 = {ay, by, cy};
  if (!z_null && !same(az, bz, cz, ax, bx, cx) && !same(az, bz, cz, ay, by, cy))
    c2 = {az, bz, cz};
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  dim3 grid((nx + PRE_BX - 1) / PRE_BX, (ny + PRE_BY - 1) / PRE_BY,
            (nz + PRE_KCHUNK - 1) / PRE_KCHUNK);
  hipLaunchKernelGGL(adi_preproc_kernel<REAL>, grid, dim3(PRE_BX, PRE_BY), 0,
                     current_stream(), lambda, u, du, c0, c1, c2, nx, ny, nz);
  return hipGetLastError() == hipSuccess ? TRID_STATUS_SUCCESS
                                         : TRID_STATUS_EXECUTION_FAILED;
}

// One step, stream-ordered, no host synchronisation (ctx->sync forced to 0
// for the solves; the caller decides when to wait).
template <typename REAL>
tridStatus_t step(const TridParams *ctx, REAL lambda, REAL *u, REAL *du, REAL *a,
                  REAL *b, REAL *c, int nx, int ny, int nz);

template <>
tridStatus_t step<double>(const TridParams *ctx, double lambda, double *u, double *du,
                          double *a, double *b, double *c, int nx, int ny, int nz) {
  TridParams p = ctx ? *ctx : TridParams{nullptr, {0, 0, 0}, 0};
  p.sync = 0;
  const int dims[3] = {nx, ny, nz};
  tridStatus_t s = preproc<double>(lambda, u, du, a, b, c, nullptr, nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nx, ny, nz);
  if (s == TRID_STATUS_SUCCESS) s = tridDmtsvStridedBatch(&p, a, b, c, du, 3, 0, dims, dims);
  if (s == TRID_STATUS_SUCCESS) s = tridDmtsvStridedBatch(&p, a, b, c, du, 3, 1, dims, dims);
  if (s == TRID_STATUS_SUCCESS)
    s = tridDmtsvStridedBatchInc(&p, a, b, c, du, u, 3, 2, dims, dims);
  return s;
}

template <>
tridStatus_t step<float>(const TridParams *ctx, float lambda, float *u, float *du,
                         float *a, float *b, float *c, int nx, int ny, int nz) {
  TridParams p = ctx ? *ctx : TridParams{nullptr, {0, 0, 0}, 0};
  p.sync = 0;
  const int dims[3] = {nx, ny, nz};
  tridStatus_t s = preproc<float>(lambda, u, du, a, b, c, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nx, ny, nz);
  if (s == TRID_STATUS_SUCCESS) s = tridSmtsvStridedBatch(&p, a, b, c, du, 3, 0, dims, dims);
  if (s == TRID_STATUS_SUCCESS) s = tridSmtsvStridedBatch(&p, a, b, c, du, 3, 1, dims, dims);
  if (s == TRID_STATUS_SUCCESS)
    s = tridSmtsvStridedBatchInc(&p, a, b, c, du, u, 3, 2, dims, dims);
  return s;
}

template <typename REAL>
tridStatus_t run(const TridParams *ctx, REAL lambda, REAL *u, REAL *du, REAL *a, REAL *b,
                 REAL *c, int nx, int ny, int nz, int iters, int use_graph) {
  if (iters < 0) return TRID_STATUS_INVALID_VALUE;
  hipStream_t st = current_stream();
  tridStatus_t s = TRID_STATUS_SUCCESS;
  const bool sync = ctx == nullptr || ctx->sync != 0;
  if (!use_graph || iters == 0) {
    for (int it = 0; it < iters && s == TRID_STATUS_SUCCESS; ++it)
      s = step<REAL>(ctx, lambda, u, du, a, b, c, nx, ny, nz);
  } else {
    // Capture one step on a private stream (the null stream cannot be
    // captured), replay it `iters` times on the caller's stream.
    hipStream_t cap = nullptr;
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess)
      return TRID_STATUS_EXECUTION_FAILED;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *prev = nullptr;
    tridHipGetStream(&prev);
    if (s == TRID_STATUS_SUCCESS) {
      if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        s = TRID_STATUS_EXECUTION_FAILED;
      } else {
        tridHipSetStream(cap);
        tridStatus_t s_step = step<REAL>(ctx, lambda, u, du, a, b, c, nx, ny, nz);
        tridHipSetStream(prev);
        hipError_t e = hipStreamEndCapture(cap, &graph);
        if (s_step != TRID_STATUS_SUCCESS) s = s_step;
        else if (e != hipSuccess) s = TRID_STATUS_EXECUTION_FAILED;
      }
    }
    if (s == TRID_STATUS_SUCCESS &&
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      s = TRID_STATUS_EXECUTION_FAILED;
    for (int it = 0; it < iters && s == TRID_STATUS_SUCCESS; ++it)
      if (hipGraphLaunch(exec, st) != hipSuccess) s = TRID_STATUS_EXECUTION_FAILED;
    // Wait for the replays before the executable graph is destroyed.
    if (exec) {
      if (hipStreamSynchronize(st) != hipSuccess && s == TRID_STATUS_SUCCESS)
        s = TRID_STATUS_EXECUTION_FAILED;
      (void)hipGraphExecDestroy(exec);
    }
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cap);
  }
  if (s == TRID_STATUS_SUCCESS && sync && hipStreamSynchronize(st) != hipSuccess)
    s = TRID_STATUS_EXECUTION_FAILED;
  return s;
}

}  // namespace
}  // namespace trid

using namespace trid;

EXTERN_C tridStatus_t tridAdiPreprocD(double lambda, const double *u, double *du,
                                      double *ax, double *bx, double *cx, double *ay,
                                      double *by, double *cy, double *az, double *bz,
                                      double *cz, int nx, int ny, int nz) {
  return preproc<double>(lambda, u, du, ax, bx, cx, ay, by, cy, az, bz, cz, nx, ny, nz);
}

EXTERN_C tridStatus_t tridAdiPreprocS(float lambda, const float *u, float *du, float *ax,
                                      float *bx, float *cx, float *ay, float *by,
                                      float *cy, float *az, float *bz, float *cz, int nx,
                                      int ny, int nz) {
  return preproc<float>(lambda, u, du, ax, bx, cx, ay, by, cy, az, bz, cz, nx, ny, nz);
}

EXTERN_C tridStatus_t tridAdiStepD(const TridParams *ctx, double lambda, double *u,
                                   double *du, double *a, double *b, double *c, int nx,
                                   int ny, int nz) {
  tridStatus_t s = step<double>(ctx, lambda, u, du, a, b, c, nx, ny, nz);
  if (s == TRID_STATUS_SUCCESS && (ctx == nullptr || ctx->sync != 0) &&
      hipStreamSynchronize(current_stream()) != hipSuccess)
    s = TRID_STATUS_EXECUTION_FAILED;
  return s;
}

EXTERN_C tridStatus_t tridAdiStepS(const TridParams *ctx, float lambda, float *u, float *du,
                                   float *a, float *b, float *c, int nx, int ny, int nz) {
  tridStatus_t s = step<float>(ctx, lambda, u, du, a, b, c, nx, ny, nz);
  if (s == TRID_STATUS_SUCCESS && (ctx == nullptr || ctx->sync != 0) &&
      hipStreamSynchronize(current_stream()) != hipSuccess)
    s = TRID_STATUS_EXECUTION_FAILED;
  return s;
}

EXTERN_C tridStatus_t tridAdiRunD(const TridParams *ctx, double lambda, double *u,
                                  double *du, double *a, double *b, double *c, int nx,
                                  int ny, int nz, int iters, int use_graph) {
  return run<double>(ctx, lambda, u, du, a, b, c, nx, ny, nz, iters, use_graph);
}

EXTERN_C tridStatus_t tridAdiRunS(const TridParams *ctx, float lambda, float *u, float *du,
                                  float *a, float *b, float *c, int nx, int ny, int nz,
                                  int iters, int use_graph) {
  return run<float>(ctx, lambda, u, du, a, b, c, nx, ny, nz, iters, use_graph);
}
