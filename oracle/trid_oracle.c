/*
 * trid_oracle.c - CPU oracle for the tridsolver_amd parity tests.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, to check and time against. The
 * product library (tridsolver_amd/lib/libtridhip.so) never links or calls it.
 *
 * It restates, in plain C, the reference's CPU algorithm for the scalar batched
 * tridiagonal solve and its ADI caller. Every function names the reference
 * lines it follows. Build with -ffp-contract=off so that each expression is
 * rounded exactly as written (see oracle/Makefile).
 *
 * Pinned by (tests/test_oracle.py):
 *   orc_thomas_*      fixtures written by tests/golden/gen_fixtures.py (scipy
 *                     SuperLU solves, the reference test generator's method),
 *                     bit-exact agreement with the reference CPU library built
 *                     from source (oracle/_ref), dense-LU brute force for
 *                     N <= 64, closed forms (N=1, N=2 Cramer, a=c=0), and the
 *                     manufactured-solution residual bound.
 *   orc_batch_*       same fixtures in padded and unpadded layouts; bit-exact
 *                     against oracle/_ref on every solvedim and on Inc.
 *   orc_adi_*         oracle/_ref driven through the ADI call pattern
 *                     (preproc -> x -> y -> z Inc), and the closed-form
 *                     boundary/interior coefficients of preproc.
 *   orc_adjoint_d     the identity <w, A^-1 d> = <A^-T w, d>, central finite
 *                     differences, and a dense transpose solve.
 *   orc_ragged_d      per-system dense LU and orc_thomas_d on each slice.
 *   orc_ragged_s      per-system dense LU (FP32 tolerance), bit-exact
 *                     against orc_batch_s on each slice, and the Inc form.
 *   orc_*_ld          long-double solve; used only to report errors. Pinned
 *                     by the dense-LU brute force on small systems.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Scalar Thomas, the reference's canonical arithmetic order.               */
/* Follows scalar/src/cpu/trid_cpu.cpp:312-353 (trid_scalar) and             */
/* apps/adi/src/adi_orig.c:103-145 (trid_cpu).                               */
/*   forward:  bb = 1/b0; cc = bb*c0; dd = bb*d0;                            */
/*             bb = b - a*cc; dd = d - a*dd; bb = 1/bb; cc = bb*c; dd = bb*dd */
/*   backward: dd = d2 - c2*dd                                               */
/* `fused` selects the forward form of trid_x_transpose on AVX-512           */
/* (trid_cpu.cpp:139-141): bb = fnmadd(a,cc,b), dd = fnmadd(a,dd,d).         */
/* c2/d2 are caller-provided scratch of length N.                            */
/* ------------------------------------------------------------------------ */
#define DEFINE_THOMAS(NAME, REAL, ONE)                                          \
  static void NAME(const REAL *a, const REAL *b, const REAL *c, REAL *d,        \
                   REAL *u, int64_t N, int64_t stride, int inc, int fused,      \
                   REAL *c2, REAL *d2) {                                        \
    int64_t i, ind = 0;                                                         \
    REAL aa, bb, cc, dd;                                                        \
    bb = ONE / b[0];                                                            \
    cc = bb * c[0];                                                             \
    dd = bb * d[0];                                                             \
    c2[0] = cc;                                                                 \
    d2[0] = dd;                                                                 \
    for (i = 1; i < N; i++) {                                                   \
      ind = ind + stride;                                                       \
      aa = a[ind];                                                              \
      if (fused) {                                                              \
        bb = FMA_##REAL(-aa, cc, b[ind]);                                       \
        dd = FMA_##REAL(-aa, dd, d[ind]);                                       \
      } else {                                                                  \
        bb = b[ind] - aa * cc;                                                  \
        dd = d[ind] - aa * dd;                                                  \
      }                                                                         \
      bb = ONE / bb;                                                            \
      cc = bb * c[ind];                                                         \
      dd = bb * dd;                                                             \
      c2[i] = cc;                                                               \
      d2[i] = dd;                                                               \
    }                                                                           \
    if (inc)                                                                    \
      u[ind] += dd;                                                             \
    else                                                                        \
      d[ind] = dd;                                                              \
    for (i = N - 2; i >= 0; i--) {                                              \
      ind = ind - stride;                                                       \
      dd = d2[i] - c2[i] * dd;                                                  \
      if (inc)                                                                  \
        u[ind] += dd;                                                           \
      else                                                                      \
        d[ind] = dd;                                                            \
    }                                                                           \
  }

#define FMA_double(x, y, z) fma((x), (y), (z))
#define FMA_float(x, y, z) fmaf((x), (y), (z))
typedef long double ldouble;
#define FMA_ldouble(x, y, z) fmal((x), (y), (z))

DEFINE_THOMAS(thomas_d, double, 1.0)
DEFINE_THOMAS(thomas_s, float, 1.0f)

/* Single system, public: restates trid_scalar{D,S} (trid_cpu.cpp:526-531,477-482)
 * plus the Inc form (trid_cpu.cpp:341-351). */
void orc_thomas_d(const double *a, const double *b, const double *c, double *d,
                  double *u, int64_t N, int64_t stride, int inc) {
  double *s = (double *)malloc(sizeof(double) * 2 * (size_t)N);
  thomas_d(a, b, c, d, u, N, stride, inc, 0, s, s + N);
  free(s);
}

void orc_thomas_s(const float *a, const float *b, const float *c, float *d,
                  float *u, int64_t N, int64_t stride, int inc) {
  float *s = (float *)malloc(sizeof(float) * 2 * (size_t)N);
  thomas_s(a, b, c, d, u, N, stride, inc, 0, s, s + N);
  free(s);
}

/* ------------------------------------------------------------------------ */
/* Batch dispatch. Follows tridMultiDimBatchSolve, trid_cpu.cpp:358-455:     */
/*   dims/pads padded to 3 with 1 (:363-368);                                */
/*   solvedim 0: systems (j,k), base k*p0*p1 + j*p0, stride 1 (:369-403);    */
/*   solvedim 1: systems (i,k), base k*p0*p1 + i,   stride p0 (:404-427);    */
/*   solvedim 2: systems (i,j), base j*p0 + i,      stride p0*p1 (:428-454). */
/* simd_vec > 0 reproduces which x rows the reference routes through         */
/* trid_x_transpose (pads[0] % simd_vec == 0 and j < ROUND_DOWN(dims[1]),    */
/* :374-382): those rows use the fused forward form. simd_vec = 0 gives the  */
/* plain trid_scalar arithmetic everywhere.                                  */
/* nthreads <= 0: OpenMP default.                                            */
/* ------------------------------------------------------------------------ */
#define DEFINE_BATCH(NAME, THOMAS, REAL)                                        \
  int NAME(const REAL *a, const REAL *b, const REAL *c, REAL *d, REAL *u,       \
           int ndim, int solvedim, const int *dims_p, const int *pads_p,        \
           int inc, int simd_vec, int nthreads) {                               \
    int64_t dims[3] = {1, 1, 1}, pads[3] = {1, 1, 1};                           \
    if (ndim < 1 || ndim > 3 || solvedim < 0 || solvedim >= ndim) return 3;     \
    for (int i = 0; i < ndim; i++) {                                            \
      dims[i] = dims_p[i];                                                      \
      pads[i] = pads_p[i];                                                      \
    }                                                                           \
    int64_t n0, n1, st0, st1, stride, N = dims[solvedim];                       \
    if (solvedim == 0) {                                                        \
      n0 = dims[1]; n1 = dims[2]; st0 = pads[0]; st1 = pads[0] * pads[1];       \
      stride = 1;                                                               \
    } else if (solvedim == 1) {                                                 \
      n0 = dims[0]; n1 = dims[2]; st0 = 1; st1 = pads[0] * pads[1];             \
      stride = pads[0];                                                         \
    } else {                                                                    \
      n0 = dims[0]; n1 = dims[1]; st0 = 1; st1 = pads[0];                       \
      stride = pads[0] * pads[1];                                               \
    }                                                                           \
    int fused_rows = 0;                                                         \
    if (solvedim == 0 && simd_vec > 0 && pads[0] % simd_vec == 0)               \
      fused_rows = (int)((dims[1] / simd_vec) * simd_vec);                      \
    int64_t nsys = n0 * n1;                                                     \
    int nt = nthreads;                                                          \
    if (nt <= 0) nt = omp_max_threads_();                                       \
    _Pragma("omp parallel num_threads(nt)")                                     \
    {                                                                           \
      REAL *s = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)N);                   \
      _Pragma("omp for schedule(static)")                                       \
      for (int64_t q = 0; q < nsys; q++) {                                      \
        int64_t i0 = q % n0, i1 = q / n0;                                       \
        int64_t base = i0 * st0 + i1 * st1;                                     \
        int fused = (solvedim == 0 && i0 < fused_rows);                         \
        THOMAS(a + base, b + base, c + base, d + base, u ? u + base : 0, N,     \
               stride, inc, fused, s, s + N);                                   \
      }                                                                         \
      free(s);                                                                  \
    }                                                                           \
    return 0;                                                                   \
  }

static int omp_max_threads_(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

DEFINE_BATCH(orc_batch_d, thomas_d, double)
DEFINE_BATCH(orc_batch_s, thomas_s, float)

/* ------------------------------------------------------------------------ */
/* Long-double error reference: same Thomas recurrence evaluated in          */
/* long double from double (or float) inputs; results rounded to double.     */
/* Used only to report max error of both the oracle and the GPU results.     */
/* ------------------------------------------------------------------------ */
static void thomas_ld(const double *a, const double *b, const double *c,
                      const double *d, double *x, int64_t N, int64_t stride,
                      ldouble *c2, ldouble *d2) {
  ldouble bb = 1.0L / (ldouble)b[0];
  ldouble cc = bb * (ldouble)c[0];
  ldouble dd = bb * (ldouble)d[0];
  c2[0] = cc;
  d2[0] = dd;
  int64_t ind = 0;
  for (int64_t i = 1; i < N; i++) {
    ind += stride;
    ldouble aa = a[ind];
    bb = 1.0L / ((ldouble)b[ind] - aa * cc);
    cc = bb * (ldouble)c[ind];
    dd = bb * ((ldouble)d[ind] - aa * dd);
    c2[i] = cc;
    d2[i] = dd;
  }
  x[ind] = (double)dd;
  for (int64_t i = N - 2; i >= 0; i--) {
    ind -= stride;
    dd = d2[i] - c2[i] * dd;
    x[ind] = (double)dd;
  }
}

int orc_batch_ld(const double *a, const double *b, const double *c,
                 const double *d, double *x, int ndim, int solvedim,
                 const int *dims_p, const int *pads_p, int nthreads) {
  int64_t dims[3] = {1, 1, 1}, pads[3] = {1, 1, 1};
  if (ndim < 1 || ndim > 3 || solvedim < 0 || solvedim >= ndim) return 3;
  for (int i = 0; i < ndim; i++) {
    dims[i] = dims_p[i];
    pads[i] = pads_p[i];
  }
  int64_t n0, n1, st0, st1, stride, N = dims[solvedim];
  if (solvedim == 0) {
    n0 = dims[1]; n1 = dims[2]; st0 = pads[0]; st1 = pads[0] * pads[1]; stride = 1;
  } else if (solvedim == 1) {
    n0 = dims[0]; n1 = dims[2]; st0 = 1; st1 = pads[0] * pads[1]; stride = pads[0];
  } else {
    n0 = dims[0]; n1 = dims[1]; st0 = 1; st1 = pads[0]; stride = pads[0] * pads[1];
  }
  int64_t nsys = n0 * n1;
  int nt = nthreads > 0 ? nthreads : omp_max_threads_();
#pragma omp parallel num_threads(nt)
  {
    ldouble *s = (ldouble *)malloc(sizeof(ldouble) * 2 * (size_t)N);
#pragma omp for schedule(static)
    for (int64_t q = 0; q < nsys; q++) {
      int64_t base = (q % n0) * st0 + (q / n0) * st1;
      thomas_ld(a + base, b + base, c + base, d + base, x + base, N, stride, s,
                s + N);
    }
    free(s);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Ragged batch: systems concatenated back to back; system s occupies        */
/* [offsets[s], offsets[s+1]). Each system is solved by the canonical        */
/* trid_scalar arithmetic (trid_cpu.cpp:312-353) with stride 1. The          */
/* reference has no ragged API; this is the per-system semantics the ragged  */
/* extension must reproduce. inc: u += x instead of d = x.                   */
/* ------------------------------------------------------------------------ */
void orc_ragged_d(const double *a, const double *b, const double *c, double *d,
                  double *u, const int64_t *offsets, int64_t nsys, int inc,
                  int nthreads) {
  int64_t maxlen = 1;
  for (int64_t s = 0; s < nsys; s++) {
    int64_t L = offsets[s + 1] - offsets[s];
    if (L > maxlen) maxlen = L;
  }
  int nt = nthreads > 0 ? nthreads : omp_max_threads_();
#pragma omp parallel num_threads(nt)
  {
    double *s2 = (double *)malloc(sizeof(double) * 2 * (size_t)maxlen);
#pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsys; s++) {
      int64_t o = offsets[s], L = offsets[s + 1] - o;
      if (L > 0)
        thomas_d(a + o, b + o, c + o, d + o, u ? u + o : 0, L, 1, inc, 0, s2,
                 s2 + maxlen);
    }
    free(s2);
  }
}

void orc_ragged_s(const float *a, const float *b, const float *c, float *d,
                  float *u, const int64_t *offsets, int64_t nsys, int inc,
                  int nthreads) {
  int64_t maxlen = 1;
  for (int64_t s = 0; s < nsys; s++) {
    int64_t L = offsets[s + 1] - offsets[s];
    if (L > maxlen) maxlen = L;
  }
  int nt = nthreads > 0 ? nthreads : omp_max_threads_();
#pragma omp parallel num_threads(nt)
  {
    float *s2 = (float *)malloc(sizeof(float) * 2 * (size_t)maxlen);
#pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsys; s++) {
      int64_t o = offsets[s], L = offsets[s + 1] - o;
      if (L > 0)
        thomas_s(a + o, b + o, c + o, d + o, u ? u + o : 0, L, 1, inc, 0, s2,
                 s2 + maxlen);
    }
    free(s2);
  }
}

/* Long-double error reference of a ragged FP64 batch (x = A^-1 d per      */
/* system, rounded to double); empty systems are skipped.                   */
void orc_ragged_ld(const double *a, const double *b, const double *c,
                   const double *d, double *x, const int64_t *offsets,
                   int64_t nsys, int nthreads) {
  int64_t maxlen = 1;
  for (int64_t s = 0; s < nsys; s++) {
    int64_t L = offsets[s + 1] - offsets[s];
    if (L > maxlen) maxlen = L;
  }
  int nt = nthreads > 0 ? nthreads : omp_max_threads_();
#pragma omp parallel num_threads(nt)
  {
    ldouble *s2 = (ldouble *)malloc(sizeof(ldouble) * 2 * (size_t)maxlen);
#pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsys; s++) {
      int64_t o = offsets[s], L = offsets[s + 1] - o;
      if (L > 0)
        thomas_ld(a + o, b + o, c + o, d + o, x + o, L, 1, s2, s2 + maxlen);
    }
    free(s2);
  }
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the batched solve. For one system A x = d with A = tridiag     */
/* (a, b, c) (a[0], c[N-1] unused), and w = dL/dx, the adjoint solves        */
/*   A^T lam = w,  A^T = tridiag(sub = c[n-1], diag = b[n], sup = a[n+1])    */
/* and forms dL/dd = lam, dL/da[n] = -lam[n] x[n-1] (n >= 1),                */
/* dL/db[n] = -lam[n] x[n], dL/dc[n] = -lam[n] x[n+1] (n <= N-2); the        */
/* entries a[0], c[N-1] get 0. The transpose solve uses the canonical        */
/* Thomas arithmetic (trid_cpu.cpp:312-353) on the transposed coefficients.  */
/* The reference has no adjoint; this is the mathematics the extension must  */
/* reproduce. Any of ga/gb/gc may be NULL.                                   */
/* ------------------------------------------------------------------------ */
int orc_adjoint_d(const double *a, const double *b, const double *c,
                  const double *x, const double *w, double *lam, double *ga,
                  double *gb, double *gc, int ndim, int solvedim,
                  const int *dims_p, const int *pads_p, int nthreads) {
  int64_t dims[3] = {1, 1, 1}, pads[3] = {1, 1, 1};
  if (ndim < 1 || ndim > 3 || solvedim < 0 || solvedim >= ndim) return 3;
  for (int i = 0; i < ndim; i++) {
    dims[i] = dims_p[i];
    pads[i] = pads_p[i];
  }
  int64_t n0, n1, st0, st1, stride, N = dims[solvedim];
  if (solvedim == 0) {
    n0 = dims[1]; n1 = dims[2]; st0 = pads[0]; st1 = pads[0] * pads[1]; stride = 1;
  } else if (solvedim == 1) {
    n0 = dims[0]; n1 = dims[2]; st0 = 1; st1 = pads[0] * pads[1]; stride = pads[0];
  } else {
    n0 = dims[0]; n1 = dims[1]; st0 = 1; st1 = pads[0]; stride = pads[0] * pads[1];
  }
  int64_t nsys = n0 * n1;
  int nt = nthreads > 0 ? nthreads : omp_max_threads_();
#pragma omp parallel num_threads(nt)
  {
    double *buf = (double *)malloc(sizeof(double) * 6 * (size_t)N);
    double *ta = buf, *tb = buf + N, *tc = buf + 2 * N, *tl = buf + 3 * N;
    double *s2 = buf + 4 * N;
#pragma omp for schedule(static)
    for (int64_t q = 0; q < nsys; q++) {
      int64_t base = (q % n0) * st0 + (q / n0) * st1;
      for (int64_t n = 0; n < N; n++) {
        int64_t o = base + n * stride;
        ta[n] = n > 0 ? c[o - stride] : 0.0;
        tb[n] = b[o];
        tc[n] = n < N - 1 ? a[o + stride] : 0.0;
        tl[n] = w[o];
      }
      thomas_d(ta, tb, tc, tl, 0, N, 1, 0, 0, s2, s2 + N);
      for (int64_t n = 0; n < N; n++) {
        int64_t o = base + n * stride;
        double l = tl[n];
        lam[o] = l;
        if (ga) ga[o] = n > 0 ? -l * x[o - stride] : 0.0;
        if (gb) gb[o] = -l * x[o];
        if (gc) gc[o] = n < N - 1 ? -l * x[o + stride] : 0.0;
      }
    }
    free(buf);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* ADI heat-diffusion timestep, restating apps/adi/src/adi_orig.c:147-242    */
/* (and the GPU preproc, apps/adi/src/preproc.cu:40-90):                     */
/*   preproc: on the boundary d=0, a=0, b=1, c=0; inside                     */
/*            d = lambda*(u[-1]+u[+1]+u[-nx]+u[+nx]+u[-nx*ny]+u[+nx*ny]      */
/*                        - 6*u), a = c = -0.5*lambda, b = 1 + lambda;       */
/*            the same a,b,c are used by all three directions;               */
/*   solve x (stride 1), y (stride nx), z (stride nx*ny), each on du;        */
/*   u += du.                                                                */
/* The reference writes identical a,b,c into 9 arrays; the restatement keeps */
/* one copy because the values are identical.                               */
/* ------------------------------------------------------------------------ */
#define DEFINE_PREPROC(NAME, REAL, C6, CH, C1)                                   \
  void NAME(REAL lambda, const REAL *u, REAL *du, REAL *a, REAL *b, REAL *c,    \
            int nx, int ny, int nz) {                                           \
    _Pragma("omp parallel for collapse(2)")                                     \
    for (int k = 0; k < nz; k++)                                                \
      for (int j = 0; j < ny; j++)                                              \
        for (int i = 0; i < nx; i++) {                                          \
          int64_t ind = (int64_t)k * nx * ny + (int64_t)j * nx + i;            \
          REAL aa, bb, cc, dd;                                                  \
          if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 ||      \
              k == nz - 1) {                                                    \
            dd = 0; aa = 0; bb = 1; cc = 0;                                     \
          } else {                                                              \
            int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;                    \
            dd = lambda * (u[ind - sx] + u[ind + sx] + u[ind - sy] +           \
                           u[ind + sy] + u[ind - sz] + u[ind + sz] -           \
                           C6 * u[ind]);                                        \
            aa = -CH * lambda;                                                  \
            bb = C1 + lambda;                                                   \
            cc = -CH * lambda;                                                  \
          }                                                                     \
          du[ind] = dd; a[ind] = aa; b[ind] = bb; c[ind] = cc;                  \
        }                                                                       \
  }

DEFINE_PREPROC(orc_adi_preproc_d, double, 6.0f, 0.5f, 1.0f)
DEFINE_PREPROC(orc_adi_preproc_s, float, 6.0f, 0.5f, 1.0f)

#define DEFINE_ADI(NAME, PREPROC, BATCH, REAL)                                  \
  void NAME(REAL lambda, REAL *u, int nx, int ny, int nz, int iters,           \
            int nthreads) {                                                     \
    size_t n = (size_t)nx * ny * nz;                                            \
    REAL *du = (REAL *)malloc(sizeof(REAL) * n);                                \
    REAL *a = (REAL *)malloc(sizeof(REAL) * n);                                 \
    REAL *b = (REAL *)malloc(sizeof(REAL) * n);                                 \
    REAL *c = (REAL *)malloc(sizeof(REAL) * n);                                 \
    int dims[3] = {nx, ny, nz};                                                 \
    for (int it = 0; it < iters; it++) {                                        \
      PREPROC(lambda, u, du, a, b, c, nx, ny, nz);                              \
      BATCH(a, b, c, du, 0, 3, 0, dims, dims, 0, 0, nthreads);                  \
      BATCH(a, b, c, du, 0, 3, 1, dims, dims, 0, 0, nthreads);                  \
      BATCH(a, b, c, du, u, 3, 2, dims, dims, 1, 0, nthreads);                  \
    }                                                                           \
    free(du); free(a); free(b); free(c);                                        \
  }

DEFINE_ADI(orc_adi_d, orc_adi_preproc_d, orc_batch_d, double)
DEFINE_ADI(orc_adi_s, orc_adi_preproc_s, orc_batch_s, float)

/* ------------------------------------------------------------------------ */
/* Residual in long double: r = max_n |d_n - (a_n x_{n-1} + b_n x_n +        */
/* c_n x_{n+1})| over all systems, with a_0 and c_{N-1} ignored.             */
/* Returns max|r| and writes max|d| to *dmax.                                */
/* ------------------------------------------------------------------------ */
double orc_residual_d(const double *a, const double *b, const double *c,
                      const double *d, const double *x, int ndim, int solvedim,
                      const int *dims_p, const int *pads_p, double *dmax) {
  int64_t dims[3] = {1, 1, 1}, pads[3] = {1, 1, 1};
  for (int i = 0; i < ndim; i++) {
    dims[i] = dims_p[i];
    pads[i] = pads_p[i];
  }
  int64_t n0, n1, st0, st1, stride, N = dims[solvedim];
  if (solvedim == 0) {
    n0 = dims[1]; n1 = dims[2]; st0 = pads[0]; st1 = pads[0] * pads[1]; stride = 1;
  } else if (solvedim == 1) {
    n0 = dims[0]; n1 = dims[2]; st0 = 1; st1 = pads[0] * pads[1]; stride = pads[0];
  } else {
    n0 = dims[0]; n1 = dims[1]; st0 = 1; st1 = pads[0]; stride = pads[0] * pads[1];
  }
  double rmax = 0, dm = 0;
  int64_t nsys = n0 * n1;
#pragma omp parallel for reduction(max : rmax, dm)
  for (int64_t q = 0; q < nsys; q++) {
    int64_t base = (q % n0) * st0 + (q / n0) * st1;
    for (int64_t n = 0; n < N; n++) {
      int64_t o = base + n * stride;
      ldouble r = (ldouble)b[o] * x[o];
      if (n > 0) r += (ldouble)a[o] * x[o - stride];
      if (n < N - 1) r += (ldouble)c[o] * x[o + stride];
      r = (ldouble)d[o] - r;
      double ar = (double)fabsl(r);
      if (ar > rmax) rmax = ar;
      if (fabs(d[o]) > dm) dm = fabs(d[o]);
    }
  }
  if (dmax) *dmax = dm;
  return rmax;
}
