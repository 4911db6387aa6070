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

/* Text of the HIP error behind this thread's last TRID_STATUS_EXECUTION_FAILED
 * ("no error" if none); a static string. */
EXTERN_C const char *tridHipLastErrorString(void);

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
/* Workspace bytes one call may need: the per-lane path's, or, for systems
 * longer than one workgroup holds (above 16384 contiguous / 2048 strided
 * FP64 rows), the larger of that and the long-system path's. */
EXTERN_C size_t tridHipWorkspaceBytes(int fp64, int inc, int ndim, int solvedim,
                                      const int *dims);

/* ---- introspection and tuning --------------------------------------------
 */

/* What the last solve of this thread ran: algorithm (1 per-lane Thomas,
 * 2 mixed precision, 3 hybrid, 4 adjoint, 5 ragged, 6 streamed strided
 * hybrid, 7 streamed strided hybrid with LDS storage, 8 wide strided hybrid,
 * 9 ragged mixed precision, 10 FP64 solution of FP32-stored systems, 11 the
 * same for a ragged batch, 12 systems longer than one workgroup holds: the
 * hybrid with its reduced system solved by a second batched solve, 13
 * contiguous systems of 4097-16384 rows, one per workgroup, with a in LDS,
 * 14 the adjoint of systems longer than its kernel holds: the transposed
 * system solved by the batched solve, 15 the mixed-precision call on
 * systems no mixed kernel holds: the FP64 batched solve, 16 the same for
 * FP32-stored systems: promoted to FP64, then the FP64 batched solve), and
 * for the hybrid its
 * rows per lane M, lanes per system P, systems per workgroup W and XCD
 * grouping G (consecutive system groups dealt to one XCD; 1 = plain order).
 * For the ragged solves: M, W = solve launches, G = workgroups per CU. */
EXTERN_C tridStatus_t tridHipLastConfig(int *algorithm, int *M, int *P, int *W, int *G);

/* Forces the hybrid shape for this thread's subsequent solves (M = 0 clears).
 * variant 0: plain hybrid; 1: streamed strided variant (strided systems,
 * FP64 M = 8, FP32 M = 16, up to 1024 threads); 2: as 1 with one sweep
 * value per row kept in LDS (arrays spanning < 4 GiB); 3: strided variant
 * with a staged in LDS by global->LDS loads (x-neighbouring strided systems,
 * W * sizeof = 128 or 256 bytes, M = 16/32, W * P a multiple of 64, arrays
 * spanning < 4 GiB); 4: contiguous systems, one per workgroup of P threads,
 * with a moved into LDS in chunk order by global->LDS loads (M = 32, W = 1;
 * P = 256 for 4097-8192 rows, P = 512 for 8193-16384). A shape that does not
 * fit a call (P*M < length, too many threads or LDS) makes that call return
 * TRID_STATUS_INVALID_VALUE. */
EXTERN_C tridStatus_t tridHipForceHybridConfig(int M, int P, int W, int G, int variant);

/* ---- autotuner -------------------------------------------------------------
 * No counterpart in the reference, whose kernel shapes are compile-time tables
 * (scalar/src/cuda/trid_thomaspcr.hpp:521-601). The shape of each call -- the
 * per-lane Thomas kernel, or the hybrid's rows per lane M (the Thomas / PCR
 * split), systems per workgroup W and XCD grouping G -- comes from a
 * per-device cache keyed by the exact geometry; uncached geometries use
 * measured defaults. Tuning times every valid shape on scratch arrays of the
 * same geometry (the caller's data is never touched) and caches the fastest.
 * Tuning allocates device memory and synchronises: call it outside hipGraph
 * capture. */

/* mode 0: never tune implicitly (default; environment TRIDSOLVER_AUTOTUNE=1
 * selects mode 1 at start-up); mode 1: tune every new geometry on its first
 * call outside stream capture. */
EXTERN_C tridStatus_t tridHipSetAutotuneMode(int mode);
/* Tunes the batched solve (fp64: 1 double / 0 float; inc: the Inc variant)
 * for one geometry; *best_ms (may be NULL) receives the winner's time. */
EXTERN_C tridStatus_t tridHipAutotune(int fp64, int inc, int ndim, int solvedim,
                                      const int *dims, const int *pads, float *best_ms);
/* Tunes tridDmtsvStridedBatchMixed (2 refinements) for one geometry. */
EXTERN_C tridStatus_t tridHipAutotuneMixed(int ndim, int solvedim, const int *dims,
                                           const int *pads, float *best_ms);
/* Tunes tridDmtsvStridedBatchMixedF32 for one geometry: the shape for one
 * refinement and the shape for more (timed at 2). *best_ms is the first's
 * time. */
EXTERN_C tridStatus_t tridHipAutotuneMixedF32(int ndim, int solvedim, const int *dims,
                                              const int *pads, float *best_ms);
/* Tunes trid{D,S}mtsvStridedBatchAdjoint for one geometry (fp64: 1 double /
 * 0 float; grads: 1 with the coefficient gradients, 0 for lambda only).
 * Systems longer than the adjoint kernel holds have nothing to tune (the
 * transposed batched solve runs; its solve is tuned as tridHipAutotune
 * tunes it): the call records that and sets *best_ms to 0. */
EXTERN_C tridStatus_t tridHipAutotuneAdjoint(int fp64, int grads, int ndim, int solvedim,
                                             const int *dims, const int *pads, float *best_ms);
/* Tunes the ragged kernel (rows per lane, workgroups per CU) on the caller's
 * device offsets with scratch values; cached per (precision, max_len). */
EXTERN_C tridStatus_t tridHipAutotuneRagged(int fp64, const int64_t *offsets, int64_t nsys,
                                            int64_t max_len);
/* Saves / loads the current device's cache as text; clears all entries. */
EXTERN_C tridStatus_t tridHipAutotuneSave(const char *path);
EXTERN_C tridStatus_t tridHipAutotuneLoad(const char *path);
EXTERN_C tridStatus_t tridHipAutotuneClear(void);

/* ---- FP32 sweeps with FP64 refinement ------------------------------------
 * No counterpart in the reference. Same arguments and layout as
 * tridDmtsvStridedBatch (scalar/include/tridsolver.h:38-42) plus the number
 * of refinement steps (0..8; 2 reaches FP64 accuracy on diagonally dominant
 * systems). The matrix is factored and the sweeps run in FP32; the residual
 * d - A x and the accumulation x += dx are FP64. One pass over HBM: a, b, c,
 * d are read once and d is overwritten with the FP64 solution. The mixed
 * kernels hold systems of up to 8192 rows along a contiguous dimension (16
 * rows per lane, 512 lanes) and 4096 along a strided one (8 rows per lane);
 * longer systems are solved by the FP64 batched solve (FP64 throughout, at
 * least as accurate). */
EXTERN_C tridStatus_t tridDmtsvStridedBatchMixed(const TridParams *ctx, const double *a,
                                                 const double *b, const double *c, double *d,
                                                 int ndim, int solvedim, const int *dims,
                                                 const int *pads, int refinements);

/* FP64 solution of a batch stored in FP32. No counterpart in the reference.
 * a, b, c, d are FP32 arrays in the layout of tridSmtsvStridedBatch
 * (scalar/include/tridsolver.h:44-48) and x an FP64 array of the same layout
 * (same dims, pads and element offsets); d is not written. The system solved
 * is the one whose coefficients and right-hand side are the FP32 values,
 * each exact in FP64: the matrix is factored and the sweeps run in FP32, the
 * residual d - A x and x += dx in FP64, `refinements` (0..8) times; 1 reaches
 * about 1e-13 of the FP64 solution on diagonally dominant systems. One pass
 * over HBM: 16 B read and 8 B written per point, against 40 B for the FP64
 * solve of the same system stored in FP64. The kernels hold systems of up to
 * 8192 rows; longer systems are promoted to FP64
 * in the stream's workspace (24 B per point) and solved by the FP64 batched
 * solve: the FP64 solution of the same promoted system, at about 88 B of
 * traffic per point. Under stream capture that workspace must already exist:
 * run the same call once on the capturing stream beforehand. */
EXTERN_C tridStatus_t tridDmtsvStridedBatchMixedF32(const TridParams *ctx, const float *a,
                                                    const float *b, const float *c,
                                                    const float *d, double *x, int ndim,
                                                    int solvedim, const int *dims,
                                                    const int *pads, int refinements);

/* ---- adjoint ---------------------------------------------------------------
 * No counterpart in the reference. For the solve x = A^-1 d along `solvedim`
 * (same layout arguments as tridDmtsvStridedBatch) and w = dL/dx, solves
 * A^T lam = w and writes lam (= dL/dd) and, when the pointers are not NULL,
 * dL/da = -lam * x[n-1], dL/db = -lam * x[n], dL/dc = -lam * x[n+1]; the
 * entries a[0] and c[N-1] of every system get 0 (they are unused by the
 * solve). x must be the forward solution; it is read only for those three
 * gradients and may be NULL when all three pointers are NULL (the solve for
 * lam alone then moves the forward solve's bytes). Any length: systems
 * longer than the adjoint kernel holds (contiguous, 4096 FP64 / 8192 FP32
 * rows; strided, 8192) take three steps -- the transpose's off-diagonals
 * written to the stream's workspace, A^T lam = w by the batched solve, then
 * the gradients -- about 128 B per point with all three gradients at FP64.
 * Under stream capture that workspace must already exist: run the same call
 * once on the capturing stream beforehand. */
EXTERN_C tridStatus_t tridDmtsvStridedBatchAdjoint(const TridParams *ctx, const double *a,
                                                   const double *b, const double *c,
                                                   const double *x, const double *w,
                                                   double *lam, double *grad_a,
                                                   double *grad_b, double *grad_c, int ndim,
                                                   int solvedim, const int *dims,
                                                   const int *pads);
EXTERN_C tridStatus_t tridSmtsvStridedBatchAdjoint(const TridParams *ctx, const float *a,
                                                   const float *b, const float *c,
                                                   const float *x, const float *w, float *lam,
                                                   float *grad_a, float *grad_b, float *grad_c,
                                                   int ndim, int solvedim, const int *dims,
                                                   const int *pads);

/* ---- ragged batches --------------------------------------------------------
 * No counterpart in the reference (its API needs one length for all systems).
 * nsys independent systems stored back to back: system s occupies elements
 * offsets[s] .. offsets[s+1]-1 of a, b, c, d (offsets: nsys+1 int64 values in
 * device memory, non-decreasing; nsys <= 2^31 - 1). Within a system the data
 * contract of tridsolver.h holds (a[first] and c[last] unused). max_len is an
 * upper bound of every length, known to the caller (at most 8192); a longer
 * system found on the device is skipped and, for a synchronous call, reported
 * as TRID_STATUS_INVALID_VALUE (an asynchronous call cannot report it). Four
 * small planning kernels, then one persistent launch that takes every system
 * from a work queue (tridHipForceRaggedConfig can split it in two); capturable.
 * Needs tridHipRaggedWorkspaceBytes(nsys) of the stream's workspace (reserve
 * it before a hipGraph capture). */
EXTERN_C tridStatus_t tridDmtsvRaggedBatch(const TridParams *ctx, const double *a,
                                           const double *b, const double *c, double *d,
                                           const int64_t *offsets, int64_t nsys,
                                           int64_t max_len);
EXTERN_C tridStatus_t tridSmtsvRaggedBatch(const TridParams *ctx, const float *a,
                                           const float *b, const float *c, float *d,
                                           const int64_t *offsets, int64_t nsys,
                                           int64_t max_len);
/* Same, adding the solutions into u (d is read only). */
EXTERN_C tridStatus_t tridDmtsvRaggedBatchInc(const TridParams *ctx, const double *a,
                                              const double *b, const double *c, double *d,
                                              double *u, const int64_t *offsets, int64_t nsys,
                                              int64_t max_len);
EXTERN_C tridStatus_t tridSmtsvRaggedBatchInc(const TridParams *ctx, const float *a,
                                              const float *b, const float *c, float *d,
                                              float *u, const int64_t *offsets, int64_t nsys,
                                              int64_t max_len);
/* FP64 ragged batch by FP32 sweeps plus `refinements` (0..8) FP64
 * iterative-refinement passes, in one pass over HBM (40 B per point, as the
 * FP64 ragged solve). Lengths up to 8192 (16 rows per lane, a and c kept in
 * LDS); otherwise as tridDmtsvRaggedBatch. Forced rows per lane
 * (tridHipForceRaggedConfig) 4 or 8 keep every operand in registers and take
 * lengths up to 2048 / 4096. */
EXTERN_C tridStatus_t tridDmtsvRaggedBatchMixed(const TridParams *ctx, const double *a,
                                                const double *b, const double *c, double *d,
                                                const int64_t *offsets, int64_t nsys,
                                                int64_t max_len, int refinements);
/* FP64 solution of a ragged batch stored in FP32 (as
 * tridDmtsvStridedBatchMixedF32 for strided batches): FP32 a, b, c, d with
 * the layout of tridSmtsvRaggedBatch, FP64 x with the same offsets; d is not
 * written. FP32 factorisation and sweeps, FP64 residual and accumulation,
 * `refinements` (0..8) times, in one persistent launch after the planning
 * kernels: 16 B read and 8 B written per point. Lengths up to 8192;
 * otherwise as tridDmtsvRaggedBatch. */
EXTERN_C tridStatus_t tridDmtsvRaggedBatchMixedF32(const TridParams *ctx, const float *a,
                                                  const float *b, const float *c,
                                                  const float *d, double *x,
                                                  const int64_t *offsets, int64_t nsys,
                                                  int64_t max_len, int refinements);
EXTERN_C size_t tridHipRaggedWorkspaceBytes(int64_t nsys);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tridDmtsvRaggedBatch with its plan in a fixed array the
 * library holds per GPU (file-scope device memory of about 4 MiB, kept until
 * the process ends) instead of the stream's workspace. It calls no hipMalloc,
 * hipFree or allocating routine and never runs the autotuner (it takes a
 * cached shape or the default). nsys may be at most
 * TRID_HIP_NOALLOC_MAX_SYSTEMS; larger batches return
 * TRID_STATUS_INVALID_VALUE before anything is launched. Every call on one
 * GPU plans in that GPU's single array: calls that overlap on one GPU (other
 * streams or host threads) race and must be serialised by the caller.
 * Results equal tridDmtsvRaggedBatch's. */
#define TRID_HIP_NOALLOC_MAX_SYSTEMS 1048576
EXTERN_C tridStatus_t tridDmtsvRaggedBatchNoAlloc(const TridParams *ctx, const double *a,
                                                  const double *b, const double *c, double *d,
                                                  const int64_t *offsets, int64_t nsys,
                                                  int64_t max_len);
/* Forces the ragged kernel's rows per lane (4, 8 or 16), workgroups per CU
 * and solve launches (1; 2: systems needing more than half a workgroup's
 * lanes, then the rest in half-size workgroups, two per CU; 3, at M = 16
 * only: one launch of 256-thread workgroups taking the systems over 4096 rows
 * at 32 rows per lane, then the rest) for this thread's calls; M = 0 restores
 * the tuned choice. The mixed ragged solve takes the same settings (its M = 4
 * and 8 always run in one launch; 3 is refused). */
EXTERN_C tridStatus_t tridHipForceRaggedConfig(int M, int per_cu, int launches);

/* ---- measurement ---------------------------------------------------------
 * No counterpart in the reference. Streams `bytes` (a multiple of 16, both
 * pointers 16-byte aligned) from src to dst on this thread's stream,
 * asynchronously. Used to measure the HBM bandwidth the solver's roofline is
 * quoted against and to calibrate hardware byte counters. */
EXTERN_C tridStatus_t tridHipStreamCopy(const void *src, void *dst, size_t bytes);
/* The same copy in a chosen shape: `threads` per workgroup (64..1024, whole
 * waves), `vectors` 16-byte vectors per thread (1, 2 or 4, all loaded before
 * the stores), non-temporal (1) or plain (0) accesses. A sweep of these gives
 * the device's best copy bandwidth. */
EXTERN_C tridStatus_t tridHipStreamCopyShape(const void *src, void *dst, size_t bytes,
                                             int threads, int vectors, int nontemporal);
/* The loads and stores of the strided FP64 solve with seg_bytes-wide row
 * segments (128 or 256: 16 or 32 x-neighbouring systems per workgroup) and M
 * rows per lane (16 or 32) along solvedim 1 or 2 of a dense dims[0..2] grid:
 * `a` by global->LDS loads, b, c, d into registers, d = a + b + c + d written
 * back; G as the hybrid's XCD grouping. Reads 32 B and writes 8 B per point,
 * like the solve, with none of its arithmetic: the bandwidth that access
 * pattern allows. Grids below 4 GiB per array; workgroups of at most 1024
 * threads (512 at M = 32). solvedim 0 (seg_bytes 0, M = 4, 8 or 16): the
 * contiguous solve's pattern, P = ceil(dims[0] / M) lanes of M rows per
 * system (lane p: rows p + P*r), 256 / P systems per workgroup, every array
 * read once and d written. Asynchronous on this thread's stream. */
EXTERN_C tridStatus_t tridHipAccessProbe(const double *a, const double *b, const double *c,
                                         double *d, int solvedim, const int *dims, int seg_bytes,
                                         int M, int G);

/* ---- systems split across shards ------------------------------------------
 * No counterpart in the reference's single-GPU API; the role of its
 * distributed solver (scalar/src/cuda/trid_cuda_mpi.cu:454-519), for callers
 * that hold each system in G contiguous pieces (one per GPU). Shard `shard` of
 * `nshards` holds rows [o_k, o_k + dims[solvedim]) of every system in its own
 * arrays (same ndim, solvedim, other dims; its own pads). Every shard needs at
 * least 2 rows. a[0] of shard 0 and c[N-1] of the last shard are ignored as in
 * the plain solve; on the other shards they couple to the neighbour shard.
 *
 *   1. tridXmtsvStridedBatchPartBoundary: writes 6 reals per system to `bnd`,
 *      laid out [6][nsys] (systems in the order of the non-solve dimensions,
 *      first fastest), reading a, b, c, d (none is written);
 *   2. the caller gathers the nshards blocks shard-major into `bnd_all`
 *      (nshards * 6 * nsys reals; an allgather in a distributed run);
 *   3. tridXmtsvStridedBatch[Inc]PartSolve: the shard's rows of the solution,
 *      in d (or added to u; d then stays untouched), by the batched solve.
 * Steps 1 and 3 follow ctx like the plain solve (sync, opts hints). */
EXTERN_C tridStatus_t tridDmtsvStridedBatchPartBoundary(const TridParams *ctx, const double *a,
                                                        const double *b, const double *c,
                                                        const double *d, int ndim,
                                                        int solvedim, const int *dims,
                                                        const int *pads, int shard,
                                                        int nshards, double *bnd);
EXTERN_C tridStatus_t tridSmtsvStridedBatchPartBoundary(const TridParams *ctx, const float *a,
                                                        const float *b, const float *c,
                                                        const float *d, int ndim,
                                                        int solvedim, const int *dims,
                                                        const int *pads, int shard,
                                                        int nshards, float *bnd);
EXTERN_C tridStatus_t tridDmtsvStridedBatchPartSolve(const TridParams *ctx, const double *a,
                                                     const double *b, const double *c,
                                                     double *d, int ndim, int solvedim,
                                                     const int *dims, const int *pads,
                                                     int shard, int nshards,
                                                     const double *bnd_all);
EXTERN_C tridStatus_t tridSmtsvStridedBatchPartSolve(const TridParams *ctx, const float *a,
                                                     const float *b, const float *c, float *d,
                                                     int ndim, int solvedim, const int *dims,
                                                     const int *pads, int shard, int nshards,
                                                     const float *bnd_all);
EXTERN_C tridStatus_t tridDmtsvStridedBatchIncPartSolve(const TridParams *ctx, const double *a,
                                                        const double *b, const double *c,
                                                        double *d, double *u, int ndim,
                                                        int solvedim, const int *dims,
                                                        const int *pads, int shard,
                                                        int nshards, const double *bnd_all);
EXTERN_C tridStatus_t tridSmtsvStridedBatchIncPartSolve(const TridParams *ctx, const float *a,
                                                        const float *b, const float *c,
                                                        float *d, float *u, int ndim,
                                                        int solvedim, const int *dims,
                                                        const int *pads, int shard,
                                                        int nshards, const float *bnd_all);

#endif
