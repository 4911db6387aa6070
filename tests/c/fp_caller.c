/*
 * fp_caller.c - a caller written in the reference ADI app's style: the
 * precision is chosen at compile time by FPPREC, arrays are declared with the
 * FP type of trid_common.h, literals carry the F suffix, allocations are
 * rounded up to CUDA_ALIGN_BYTE, and the solver symbol follows FPPREC
 * (apps/adi/src/adi_cuda.cu:47,110-170,294-323; adi_cpu.cpp:235-266).
 *
 * Build: cc -DFPPREC=0 (float) or -DFPPREC=1 (double) -I include ...
 *
 * It fills a 24 x 20 x 18 grid with the ADI coefficients (a = c = -lambda/2,
 * b = 1 + lambda inside; identity rows on the boundary) and a right-hand side
 * in U(-1, 1), then runs the app's x, y, z-Inc sweep pattern with ctx
 * {NULL, {0, 0, 0}, 1} and checks each sweep on the host by its residual
 * max|d - A x| / max|d| (1e-13 at double, 1e-5 at float). Exit status 0 when
 * every check passes.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "tridsolver.h"

#define CAT_(x, s) x##s
#define CAT(x, s) CAT_(x, s)
#define LIT(x) CAT(x, F) /* 0.5f at FPPREC 0, 0.5 at FPPREC 1 */

#if FPPREC == 0
#  define SOLVE tridSmtsvStridedBatch
#  define SOLVE_INC tridSmtsvStridedBatchInc
#  define TOL 1e-5
#else
#  define SOLVE tridDmtsvStridedBatch
#  define SOLVE_INC tridDmtsvStridedBatchInc
#  define TOL 1e-13
#endif

#define NX 24
#define NY 20
#define NZ 18

static uint64_t lcg_state = 777;
static FP uni(void) { /* U(-1, 1) */
  lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
  return (FP)((double)(lcg_state >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}

static double residual(const FP *a, const FP *b, const FP *c, const FP *d, const FP *x,
                       int solvedim) {
  const int dims[3] = {NX, NY, NZ};
  const size_t st[3] = {1, NX, (size_t)NX * NY};
  const int o0 = solvedim == 0 ? 1 : 0, o1 = solvedim == 2 ? 1 : 2;
  const int N = dims[solvedim];
  double rmax = 0, dmax = 0;
  for (int i1 = 0; i1 < dims[o1]; ++i1)
    for (int i0 = 0; i0 < dims[o0]; ++i0) {
      const size_t base = i0 * st[o0] + i1 * st[o1];
      for (int n = 0; n < N; ++n) {
        const size_t k = base + n * st[solvedim];
        double r = (double)d[k] - (double)b[k] * x[k];
        if (n > 0) r -= (double)a[k] * x[k - st[solvedim]];
        if (n < N - 1) r -= (double)c[k] * x[k + st[solvedim]];
        if (fabs(r) > rmax) rmax = fabs(r);
        if (fabs((double)d[k]) > dmax) dmax = fabs((double)d[k]);
      }
    }
  return rmax / dmax;
}

int main(void) {
  const size_t n = (size_t)NX * NY * NZ;
  /* allocation rounded up to the reference's alignment unit */
  const size_t bytes = (n * sizeof(FP) + CUDA_ALIGN_BYTE - 1) / CUDA_ALIGN_BYTE * CUDA_ALIGN_BYTE;
  const FP lambda = LIT(1.0);
  FP *h_a = (FP *)calloc(bytes, 1), *h_b = (FP *)calloc(bytes, 1), *h_c = (FP *)calloc(bytes, 1);
  FP *h_d = (FP *)calloc(bytes, 1), *h_u = (FP *)calloc(bytes, 1), *h_x = (FP *)calloc(bytes, 1);
  for (int k = 0; k < NZ; ++k)
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        const size_t ind = i + NX * (j + (size_t)NY * k);
        const int edge = i == 0 || i == NX - 1 || j == 0 || j == NY - 1 || k == 0 || k == NZ - 1;
        h_a[ind] = edge ? LIT(0.0) : LIT(-0.5) * lambda;
        h_b[ind] = edge ? LIT(1.0) : LIT(1.0) + lambda;
        h_c[ind] = edge ? LIT(0.0) : LIT(-0.5) * lambda;
        h_d[ind] = uni();
        h_u[ind] = uni();
      }
  FP *d_a, *d_b, *d_c, *d_d, *d_u;
  if (hipMalloc((void **)&d_a, bytes) != hipSuccess || hipMalloc((void **)&d_b, bytes) != hipSuccess ||
      hipMalloc((void **)&d_c, bytes) != hipSuccess || hipMalloc((void **)&d_d, bytes) != hipSuccess ||
      hipMalloc((void **)&d_u, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  hipMemcpy(d_a, h_a, bytes, hipMemcpyHostToDevice);
  hipMemcpy(d_b, h_b, bytes, hipMemcpyHostToDevice);
  hipMemcpy(d_c, h_c, bytes, hipMemcpyHostToDevice);
  TridParams params;
  memset(&params, 0, sizeof params);
  params.sync = 1;
  const int dims[3] = {NX, NY, NZ}, pads[3] = {NX, NY, NZ};
  int failed = 0;
  for (int sd = 0; sd < 3; ++sd) {
    hipMemcpy(d_d, h_d, bytes, hipMemcpyHostToDevice);
    hipMemcpy(d_u, h_u, bytes, hipMemcpyHostToDevice);
    tridStatus_t st;
    if (sd < 2) {
      st = SOLVE(&params, d_a, d_b, d_c, d_d, 3, sd, dims, pads);
      hipMemcpy(h_x, d_d, bytes, hipMemcpyDeviceToHost);
    } else {
      st = SOLVE_INC(&params, d_a, d_b, d_c, d_d, d_u, 3, sd, dims, pads);
      hipMemcpy(h_x, d_u, bytes, hipMemcpyDeviceToHost);
      for (size_t i = 0; i < n; ++i) h_x[i] -= h_u[i];
    }
    const double r = residual(h_a, h_b, h_c, h_d, h_x, sd);
    printf("FPPREC %d sizeof(FP) %d solvedim %d: status %d residual %.2e\n", FPPREC,
           (int)sizeof(FP), sd, (int)st, r);
    if (st != TRID_STATUS_SUCCESS || !(r <= TOL)) failed = 1;
  }
  hipFree(d_a); hipFree(d_b); hipFree(d_c); hipFree(d_d); hipFree(d_u);
  free(h_a); free(h_b); free(h_c); free(h_d); free(h_u); free(h_x);
  printf(failed ? "fp_caller: FAILED\n" : "fp_caller: OK\n");
  return failed;
}
