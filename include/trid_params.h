/*
 * trid_params.h - the per-call context.
 *
 * Same members, order and size as the reference's TridParams
 * (reference: scalar/include/trid_params.h:31-35): 24 bytes on LP64.
 *
 * Meaning in this library:
 *   mpi_params  ignored (single-GPU library); kept for layout compatibility.
 *   opts[d]     algorithm hint for solvedim d:
 *                 0  automatic (autotuned choice, the default)
 *                 1  one Thomas recurrence per lane (any length; uses a
 *                    workspace for the forward-sweep coefficients)
 *                 2  same as 1 (the reference's shared-memory x variant)
 *                 3  Thomas-PCR hybrid (system held in registers, reduced
 *                    system solved by PCR in LDS)
 *                 4  same as 3 (the reference's x hybrid option)
 *               Unknown values fall back to 0. The reference maps the same
 *               numbers to its CUDA kernels (scalar/src/cuda/trid_cuda.cu:87-126,
 *               :250-285).
 *   sync        nonzero: the call returns after the results are complete
 *               (the reference's behaviour, trid_cuda.cu:409-410).
 *               zero: the work is only enqueued on the library's stream
 *               (stream-ordered; required under hipGraph capture).
 * A NULL context means {NULL, {0,0,0}, 1}.
 */
#ifndef TRID_PARAMS_H__
#define TRID_PARAMS_H__

#include "trid_common.h"

struct TridParams {
  void *mpi_params;
  int opts[MAXDIM];
  int sync;
};

#ifndef __cplusplus
typedef struct TridParams TridParams;
#endif

#endif
