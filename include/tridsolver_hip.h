/*
 * tridsolver_hip.h - MI355X extensions of the tridsolver C ABI.
 *
 * None of these has a counterpart in the reference's single-node API
 * (scalar/include/tridsolver.h); each exists for a caller the reference cannot
 * serve. Plain C types only: a hipStream_t is passed as void*.
 */
#ifndef TRIDSOLVER_HIP_H__
#define TRIDSOLVER_HIP_H__

#include <stddef.h>
#include <stdint.h>

#include "trid_common.h"
#include "trid_params.h"

/* ---- execution control ---------------------------------------------------
 * The reference runs on the default stream and synchronises the device after
 * every call (scalar/src/cuda/trid_cuda.cu:409-410). Here every call of the
 * calling thread is enqueued on the thread's library stream (default: the
 * null stream); ctx->sync decides whether the call waits for it. */

/* Sets the stream used by this thread's subsequent calls (NULL = null stream). */
EXTERN_C tridStatus_t tridHipSetStream(void *stream);
EXTERN_C tridStatus_t tridHipGetStream(void **stream);

/* The per-lane Thomas path keeps its forward coefficients in a workspace that
 * the library grows on demand (one per device and stream). Inside hipGraph
 * capture it cannot allocate: reserve enough beforehand with this call
 * (bytes for the current device and this thread's stream). Calls that share
 * one stream share its workspace, which is safe because they are ordered. */
EXTERN_C tridStatus_t tridHipReserveWorkspace(size_t bytes);
/* Frees every workspace of the current device (synchronises the device). */
EXTERN_C tridStatus_t tridHipReleaseWorkspaces(void);
/* Workspace bytes the per-lane path needs for one call. */
EXTERN_C size_t tridHipWorkspaceBytes(int fp64, int inc, int ndim, int solvedim,
                                      const int *dims);

/* ---- introspection and tuning --------------------------------------------
 * algorithm ids: 1 = per-lane Thomas, 3 = Thomas-PCR hybrid. */

/* What the last solve of this thread ran: algorithm, and for the hybrid its
 * rows per lane M, lanes per system P, systems per workgroup W and XCD
 * grouping G (consecutive system groups dealt to one XCD; 1 = plain order). */
EXTERN_C tridStatus_t tridHipLastConfig(int *algorithm, int *M, int *P, int *W, int *G);

/* Forces the hybrid shape for this thread's subsequent solves (M = 0 clears).
 * A shape that does not fit a call (P*M < length, too many threads or LDS)
 * makes that call return TRID_STATUS_INVALID_VALUE. */
EXTERN_C tridStatus_t tridHipForceHybridConfig(int M, int P, int W, int G);

/* ---- measurement ---------------------------------------------------------
 * No counterpart in the reference. Streams `bytes` (a multiple of 16, both
 * pointers 16-byte aligned) from src to dst on this thread's stream,
 * asynchronously. Used to measure the HBM bandwidth the solver's roofline is
 * quoted against and to calibrate hardware byte counters. */
EXTERN_C tridStatus_t tridHipStreamCopy(const void *src, void *dst, size_t bytes);

#endif
