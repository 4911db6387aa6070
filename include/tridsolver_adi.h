/*
 * tridsolver_adi.h - the ADI heat-diffusion timestep that calls the batched
 * solve along x, y and z (the reference app apps/adi), on one MI355X.
 *
 * Arrays are dense nx*ny*nz device arrays, x fastest. One timestep is
 *   preproc(u) -> du, a, b, c;  solve x, solve y on du;  u += solve z of du.
 */
#ifndef TRIDSOLVER_ADI_H__
#define TRIDSOLVER_ADI_H__

#include "trid_common.h"
#include "trid_params.h"

/* Right-hand side and coefficients of one step, enqueued on this thread's
 * library stream. Same arguments as the reference's preproc kernel
 * (apps/adi/src/preproc.cu:40); ay..cz (and az..cz) may be NULL or equal to
 * an earlier set, and are then not written again. */
EXTERN_C tridStatus_t tridAdiPreprocD(double lambda, const double *u, double *du,
                                      double *ax, double *bx, double *cx, double *ay,
                                      double *by, double *cy, double *az, double *bz,
                                      double *cz, int nx, int ny, int nz);
EXTERN_C tridStatus_t tridAdiPreprocS(float lambda, const float *u, float *du, float *ax,
                                      float *bx, float *cx, float *ay, float *by,
                                      float *cy, float *az, float *bz, float *cz, int nx,
                                      int ny, int nz);

/* One timestep (reference loop body apps/adi/src/adi_cuda.cu:283-325) with one
 * shared coefficient set a, b, c. ctx as for the solves (opts are hints for
 * the three sweeps); ctx->sync != 0 or ctx == NULL waits for completion. */
EXTERN_C tridStatus_t tridAdiStepD(const TridParams *ctx, double lambda, double *u,
                                   double *du, double *a, double *b, double *c, int nx,
                                   int ny, int nz);
EXTERN_C tridStatus_t tridAdiStepS(const TridParams *ctx, float lambda, float *u, float *du,
                                   float *a, float *b, float *c, int nx, int ny, int nz);

/* `iters` timesteps (reference loop apps/adi/src/adi_cuda.cu:282-326). With
 * use_graph != 0 one step is captured into a hipGraph on a private stream and
 * replayed `iters` times on this thread's stream; the call then waits for the
 * replays before it returns (the executable graph is destroyed). */
EXTERN_C tridStatus_t tridAdiRunD(const TridParams *ctx, double lambda, double *u,
                                  double *du, double *a, double *b, double *c, int nx,
                                  int ny, int nz, int iters, int use_graph);
EXTERN_C tridStatus_t tridAdiRunS(const TridParams *ctx, float lambda, float *u, float *du,
                                  float *a, float *b, float *c, int nx, int ny, int nz,
                                  int iters, int use_graph);

/* Matrix-free timestep (no counterpart in the reference; its app writes nine
 * coefficient arrays, apps/adi/src/preproc.cu:77-85, from values fixed by
 * position, :59-73). The ADI coefficients depend only on the grid position,
 * so each sweep computes a, b, c from (i, j, k), and the x sweep computes its
 * right-hand side from u: 56 B per point per step at FP64 instead of 168,
 * same results bit for bit as tridAdiStep. du is workspace. */
EXTERN_C tridStatus_t tridAdiStepFusedD(double lambda, double *u, double *du, int nx, int ny,
                                        int nz);
EXTERN_C tridStatus_t tridAdiStepFusedS(float lambda, float *u, float *du, int nx, int ny,
                                        int nz);
EXTERN_C tridStatus_t tridAdiRunFusedD(const TridParams *ctx, double lambda, double *u,
                                       double *du, int nx, int ny, int nz, int iters,
                                       int use_graph);
EXTERN_C tridStatus_t tridAdiRunFusedS(const TridParams *ctx, float lambda, float *u, float *du,
                                       int nx, int ny, int nz, int iters, int use_graph);

#endif
