/*
 * drop_in.c - a C caller of libtridhip.so written the way the reference's
 * own C/CUDA callers use the library (apps/adi/src/adi_cuda.cu:294-323,
 * scalar/test/test_cuda.cu:21-34): device arrays from the HIP runtime, this
 * library's include/tridsolver.h (which declares the reference's four solver
 * symbols and types), ctx with opts and sync. The same file also compiles as
 * C++ against the reference's own scalar/include headers with FPPREC=1.
 *
 * It solves a diagonally dominant 40 x 36 x 33 FP64 batch along x, y and z
 * (plain, plain, Inc: the ADI sweep pattern), an FP32 solve along y, and the
 * same x solve inside a halo-padded allocation, and checks each result on
 * the host by its residual max|d - A x| / max|d| (1e-13 FP64, 1e-5 FP32).
 * Exit status 0 when every check passes.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "tridsolver.h"

#define NX 40
#define NY 36
#define NZ 33

static uint64_t lcg_state = 12345;
static double uni(void) { /* U(-1, 1) */
  lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(lcg_state >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static int check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return 1;
  }
  return 0;
}

/* max|d - A x| / max|d| over every system along solvedim (3-D, pads = dims + pad,
 * first element at `off`). */
static double residual(const double *a, const double *b, const double *c, const double *d,
                       const double *x, const int *dims, const int *pads, size_t off,
                       int solvedim) {
  const size_t st[3] = {1, (size_t)pads[0], (size_t)pads[0] * pads[1]};
  const int o0 = solvedim == 0 ? 1 : 0, o1 = solvedim == 2 ? 1 : 2;
  const int N = dims[solvedim];
  double rmax = 0, dmax = 0;
  for (int i1 = 0; i1 < dims[o1]; ++i1)
    for (int i0 = 0; i0 < dims[o0]; ++i0) {
      const size_t base = off + i0 * st[o0] + i1 * st[o1];
      for (int n = 0; n < N; ++n) {
        const size_t k = base + n * st[solvedim];
        double r = d[k] - b[k] * x[k];
        if (n > 0) r -= a[k] * x[k - st[solvedim]];
        if (n < N - 1) r -= c[k] * x[k + st[solvedim]];
        if (fabs(r) > rmax) rmax = fabs(r);
        if (fabs(d[k]) > dmax) dmax = fabs(d[k]);
      }
    }
  return rmax / dmax;
}

static int run(int pad) {
  const int dims[3] = {NX, NY, NZ};
  const int pads[3] = {NX + 2 * pad, NY + 2 * pad, NZ + 2 * pad};
  const size_t n = (size_t)pads[0] * pads[1] * pads[2];
  const size_t off = pad ? 1 + (size_t)pads[0] + (size_t)pads[0] * pads[1] : 0;
  double *h[6];
  for (int i = 0; i < 6; ++i) h[i] = (double *)calloc(n, sizeof(double));
  double *ha = h[0], *hb = h[1], *hc = h[2], *hd = h[3], *hu = h[4], *hx = h[5];
  for (size_t i = 0; i < n; ++i) {
    ha[i] = uni();
    hc[i] = uni();
    hb[i] = 2.0 + fabs(ha[i]) + fabs(hc[i]) + 0.5 * (uni() + 1.0);
    hd[i] = uni();
    hu[i] = uni();
  }
  double *da, *db, *dc, *dd, *du;
  const size_t bytes = n * sizeof(double);
  if (check_hip(hipMalloc((void **)&da, bytes), "hipMalloc") ||
      check_hip(hipMalloc((void **)&db, bytes), "hipMalloc") ||
      check_hip(hipMalloc((void **)&dc, bytes), "hipMalloc") ||
      check_hip(hipMalloc((void **)&dd, bytes), "hipMalloc") ||
      check_hip(hipMalloc((void **)&du, bytes), "hipMalloc"))
    return 1;
  hipMemcpy(da, ha, bytes, hipMemcpyHostToDevice);
  hipMemcpy(db, hb, bytes, hipMemcpyHostToDevice);
  hipMemcpy(dc, hc, bytes, hipMemcpyHostToDevice);
  TridParams p;
  memset(&p, 0, sizeof p);
  p.sync = 1;
  int failed = 0;
  for (int sd = 0; sd < 3; ++sd) {
    hipMemcpy(dd, hd, bytes, hipMemcpyHostToDevice);
    hipMemcpy(du, hu, bytes, hipMemcpyHostToDevice);
    tridStatus_t st;
    if (sd < 2) {
      st = tridDmtsvStridedBatch(&p, da + off, db + off, dc + off, dd + off, 3, sd, dims, pads);
      hipMemcpy(hx, dd, bytes, hipMemcpyDeviceToHost);
    } else {  /* u += x, d read only */
      st = tridDmtsvStridedBatchInc(&p, da + off, db + off, dc + off, dd + off, du + off, 3, sd,
                                    dims, pads);
      hipMemcpy(hx, du, bytes, hipMemcpyDeviceToHost);
      for (size_t i = 0; i < n; ++i) hx[i] -= hu[i];
    }
    if (st != TRID_STATUS_SUCCESS) {
      fprintf(stderr, "solve sd=%d pad=%d: status %d\n", sd, pad, (int)st);
      return 1;
    }
    const double r = residual(ha, hb, hc, hd, hx, dims, pads, off, sd);
    printf("FP64 %s solvedim %d pad %d: residual %.2e\n", sd < 2 ? "plain" : "Inc  ", sd, pad, r);
    if (!(r <= 1e-13)) failed = 1;
  }
  /* FP32 along y: the same data rounded to float */
  {
    float *f[5];
    for (int i = 0; i < 5; ++i) f[i] = (float *)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
      f[0][i] = (float)ha[i]; f[1][i] = (float)hb[i]; f[2][i] = (float)hc[i]; f[3][i] = (float)hd[i];
    }
    float *fa = (float *)da, *fb = (float *)db, *fc = (float *)dc, *fd = (float *)dd;
    hipMemcpy(fa, f[0], n * sizeof(float), hipMemcpyHostToDevice);
    hipMemcpy(fb, f[1], n * sizeof(float), hipMemcpyHostToDevice);
    hipMemcpy(fc, f[2], n * sizeof(float), hipMemcpyHostToDevice);
    hipMemcpy(fd, f[3], n * sizeof(float), hipMemcpyHostToDevice);
    tridStatus_t st = tridSmtsvStridedBatch(&p, fa + off, fb + off, fc + off, fd + off, 3, 1, dims,
                                            pads);
    hipMemcpy(f[4], fd, n * sizeof(float), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < n; ++i) {
      ha[i] = f[0][i]; hb[i] = f[1][i]; hc[i] = f[2][i]; hd[i] = f[3][i]; hx[i] = f[4][i];
    }
    const double r = residual(ha, hb, hc, hd, hx, dims, pads, off, 1);
    printf("FP32 plain solvedim 1 pad %d: status %d residual %.2e\n", pad, (int)st, r);
    if (st != TRID_STATUS_SUCCESS || !(r <= 1e-5)) failed = 1;
    for (int i = 0; i < 5; ++i) free(f[i]);
  }
  /* bad arguments come back as a status, not an exit */
  {
    const int bad[3] = {NX, NY, 0 - 1};
    tridStatus_t st = tridDmtsvStridedBatch(&p, da, db, dc, dd, 3, 0, bad, pads);
    printf("invalid dims: status %d\n", (int)st);
    if (st != TRID_STATUS_INVALID_VALUE) failed = 1;
  }
  hipFree(da); hipFree(db); hipFree(dc); hipFree(dd); hipFree(du);
  for (int i = 0; i < 6; ++i) free(h[i]);
  return failed;
}

int main(void) {
  int failed = run(0);
  failed |= run(1);
  printf(failed ? "drop_in: FAILED\n" : "drop_in: OK\n");
  return failed;
}
