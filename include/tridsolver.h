/*
 * tridsolver.h - batched scalar tridiagonal solves along one axis of a 1-3 D
 * array, on an AMD Instinct MI355X (gfx950).
 *
 * Drop-in for the reference's public API: the same four C symbols with the
 * same argument types and meaning (reference: scalar/include/tridsolver.h:38-68).
 *
 * Data contract (identical to the reference):
 *   - a, b, c, d (and u) are device pointers to arrays of extent pads[0..ndim-1]
 *     (x fastest); element (i,j,k) is at i + pads[0]*(j + pads[1]*k).
 *     Pointers may point into a larger (halo-padded) allocation.
 *   - One system per index of the non-solve dimensions; its length is
 *     dims[solvedim] and its element stride is 1, pads[0] or pads[0]*pads[1].
 *   - Row n of a system:  a[n] x[n-1] + b[n] x[n] + c[n] x[n+1] = d[n];
 *     a[0] and c[len-1] are never used.
 *   - 1 <= ndim <= MAXDIM, 0 <= solvedim < ndim, 1 <= dims[i] <= pads[i].
 *
 * Differences from the reference (all documented, none changes results):
 *   - invalid arguments return TRID_STATUS_INVALID_VALUE instead of being
 *     undefined behaviour; HIP failures return TRID_STATUS_EXECUTION_FAILED
 *     instead of exit(-1);
 *   - there is no N_MAX=1024 length limit; a call solves at most 2^31 - 1
 *     systems (the product of the non-solve dimensions);
 *   - ctx may be NULL; ctx->sync == 0 makes the call asynchronous on the
 *     library stream (see tridsolver_hip.h);
 *   - Inc never writes d on any algorithm (the reference's x options 2/3
 *     drop INC and option 4 adds into d: trid_cuda.cu:94-117).
 */
#ifndef TRIDSOLVER_H__
#define TRIDSOLVER_H__

#include "trid_params.h"
#include "trid_common.h"

/* Solve every system along `solvedim`; the solution overwrites d.
 * Replaces reference scalar/include/tridsolver.h:38-42 (CPU body
 * scalar/src/cpu/trid_cpu.cpp:507-514, CUDA body scalar/src/cuda/trid_cuda.cu:428-435). */
EXTERN_C
tridStatus_t tridDmtsvStridedBatch(const TridParams *ctx, const double *a,
                                   const double *b, const double *c, double *d,
                                   int ndim, int solvedim, const int *dims,
                                   const int *pads);

/* Single precision of the above.
 * Replaces reference scalar/include/tridsolver.h:44-48 (trid_cpu.cpp:459-466,
 * trid_cuda.cu:419-426). */
EXTERN_C
tridStatus_t tridSmtsvStridedBatch(const TridParams *ctx, const float *a,
                                   const float *b, const float *c, float *d,
                                   int ndim, int solvedim, const int *dims,
                                   const int *pads);

/* Solve every system along `solvedim` and add the solution into u; d is read
 * only. Replaces reference scalar/include/tridsolver.h:57-62 (trid_cpu.cpp:516-524,
 * trid_cuda.cu:446-454). */
EXTERN_C
tridStatus_t tridDmtsvStridedBatchInc(const TridParams *ctx, const double *a,
                                      const double *b, const double *c,
                                      double *d, double *u, int ndim,
                                      int solvedim, const int *dims,
                                      const int *pads);

/* Single precision of the above.
 * Replaces reference scalar/include/tridsolver.h:64-68 (trid_cpu.cpp:468-475,
 * trid_cuda.cu:437-444). */
EXTERN_C
tridStatus_t tridSmtsvStridedBatchInc(const TridParams *ctx, const float *a,
                                      const float *b, const float *c, float *d,
                                      float *u, int ndim, int solvedim,
                                      const int *dims, const int *pads);

#endif
