// kernels_part.hip - batched solve of systems split into contiguous shards
// along the solve dimension, as G GPUs of a distributed run hold them (the
// role of the reference's distributed solver, scalar/src/cuda/trid_cuda_mpi.cu
// :454-519 and scalar/src/cpu/trid_mpi_cpu.hpp:44-136, whose per-rank step is
// a modified Thomas pass, an exchange of boundary rows and a reduced solve).
//
// Per shard k of n >= 2 rows, F = its first unknown, E = its last:
//  boundary  interior rows 1..n-2 are written as x_i = dd_i - aa_i F - cc_i E
//            by one top-down sweep that stores nothing: its recurrence gives
//            row n-2 (x_i = D_i - A_i F - C_i x_{i+1}), and substituting each
//            new row into row 1's running form x_1 = P + Q F + R x_{i+1}
//            (P += R D_i, Q -= R A_i, R = -R C_i) gives row 1 at the end, with
//            x_{n-1} = E; rows 0 and n-1 then read,
//            normalised by their diagonal,
//                alpha E_{k-1} + F + gamma E       = delta
//                alpha' F      + E + gamma' F_{k+1} = delta'
//            -- 6 reals per system (alpha = 0 on shard 0, gamma' = 0 on the
//            last shard), written [6][nsys];
//  exchange  the caller gathers every shard's 6 x nsys block, shard-major
//            (an allgather over RCCL in a distributed run);
//  solve     the 2G unknowns (F_0, E_0, F_1, ...) form a tridiagonal system
//            with unit diagonal; each system's lane gets E_{k-1} and F_{k+1}
//            by twisted elimination (top-down and bottom-up sweeps meeting
//            at the wanted rows, one pass each for both; O(G), no storage),
//            subtracts a_0 E_{k-1} from d_0 and c_{n-1} F_{k+1} from
//            d_{n-1}, and the shard's own rows are solved by the batched
//            solve (any shape it picks). For Inc,
//            the two rows of d are saved before and restored after, so d
//            stays untouched as in the plain Inc solve.
// Traffic per shard point: boundary reads a, b, c, d once (32 B FP64), the
// solve moves its usual 40 (48 Inc); the exchanged data is 6 reals per
// system and shard.

#include <algorithm>

#include "../../include/tridsolver_hip.h"
#include "trid_internal.h"

namespace trid {
namespace {

constexpr int PART_BLOCK = 256;
constexpr int PART_K = 8;  // rows fetched per batch in the sweeps

template <typename REAL>
__global__ __launch_bounds__(PART_BLOCK) void part_boundary_kernel(
    Geom g, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, const REAL *__restrict__ d, int first, int last,
    REAL *__restrict__ bnd) {
  const int64_t ns = g.nsys();
  const int64_t s = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
  if (s >= ns) return;
  const int64_t base = (s % g.n0) * g.st0 + (s / g.n0) * g.st1;
  const int64_t n = g.N, str = g.stride;
  a += base; b += base; c += base; d += base;

  REAL dd1, aa1, cc1, ddm, aam, ccm;
  if (n == 2) {
    dd1 = 0; aa1 = 0; cc1 = -1;  // row 1 is E itself
    ddm = 0; aam = -1; ccm = 0;  // row n-2 is F itself
  } else {
    // top-down over rows 1..n-2: x_i = D - A F - C x_{i+1}; row 1 carried
    // along as x_1 = P + Q F + R x_{i+1}
    REAL D = d[str] / b[str], A = a[str] / b[str], C = c[str] / b[str];
    REAL P = D, Q = -A, R = -C;
    int64_t i = 2;
    for (; i + PART_K <= n - 1; i += PART_K) {
      REAL av[PART_K], bv[PART_K], cv[PART_K], dv[PART_K];
#pragma unroll
      for (int k = 0; k < PART_K; ++k) {
        const int64_t o = (i + k) * str;
        av[k] = a[o]; bv[k] = b[o]; cv[k] = c[o]; dv[k] = d[o];
      }
#pragma unroll
      for (int k = 0; k < PART_K; ++k) {
        const REAL inv = REAL(1) / (bv[k] - av[k] * C);
        D = (dv[k] - av[k] * D) * inv;
        A = -av[k] * A * inv;
        C = cv[k] * inv;
        P += R * D;
        Q -= R * A;
        R = -R * C;
      }
    }
    for (; i < n - 1; ++i) {
      const int64_t o = i * str;
      const REAL ai = a[o];
      const REAL inv = REAL(1) / (b[o] - ai * C);
      D = (d[o] - ai * D) * inv;
      A = -ai * A * inv;
      C = c[o] * inv;
      P += R * D;
      Q -= R * A;
      R = -R * C;
    }
    ddm = D; aam = A; ccm = C;
    dd1 = P; aa1 = -Q; cc1 = -R;  // x_1 = dd1 - aa1 F - cc1 E
  }
  const int64_t oe = (n - 1) * str;
  const REAL a0 = first ? REAL(0) : a[0];
  const REAL cn = last ? REAL(0) : c[oe];
  const REAL r0 = REAL(1) / (b[0] - c[0] * aa1);
  const REAL rE = REAL(1) / (b[oe] - a[oe] * ccm);
  bnd[0 * ns + s] = a0 * r0;
  bnd[1 * ns + s] = -c[0] * cc1 * r0;
  bnd[2 * ns + s] = (d[0] - c[0] * dd1) * r0;
  bnd[3 * ns + s] = -a[oe] * aam * rE;
  bnd[4 * ns + s] = cn * rE;
  bnd[5 * ns + s] = (d[oe] - a[oe] * ddm) * rE;
}

// Row r of the reduced system of system s: F-row of shard r/2 (r even) or
// E-row (r odd); unit diagonal.
template <typename REAL>
__device__ __forceinline__ void reduced_row(const REAL *__restrict__ all, int64_t ns, int64_t s,
                                            int r, REAL &sub, REAL &sup, REAL &rhs) {
  const REAL *rec = all + (int64_t)(r >> 1) * 6 * ns + s;
  const int f = (r & 1) * 3;
  sub = rec[(f + 0) * ns];
  sup = rec[(f + 1) * ns];
  rhs = rec[(f + 2) * ns];
}

// Unknown j of the reduced system: top-down elimination of rows 0..j-1,
// bottom-up of rows R-1..j+1, row j solved with both.
template <typename REAL>
__device__ REAL twisted(const REAL *__restrict__ all, int64_t ns, int64_t s, int R, int j) {
  REAL cf = 0, df = 0;
  for (int i = 0; i < j; ++i) {
    REAL sub, sup, rhs;
    reduced_row(all, ns, s, i, sub, sup, rhs);
    const REAL inv = REAL(1) / (REAL(1) - sub * cf);
    df = (rhs - sub * df) * inv;
    cf = sup * inv;
  }
  REAL ar = 0, dr = 0;
  for (int i = R - 1; i > j; --i) {
    REAL sub, sup, rhs;
    reduced_row(all, ns, s, i, sub, sup, rhs);
    const REAL inv = REAL(1) / (REAL(1) - sup * ar);
    dr = (rhs - sup * dr) * inv;
    ar = sub * inv;
  }
  REAL sub, sup, rhs;
  reduced_row(all, ns, s, j, sub, sup, rhs);
  return (rhs - sub * df - sup * dr) / (REAL(1) - sub * cf - sup * ar);
}

// Unknowns j1 < j2 of the reduced system at once: one top-down sweep over
// rows 0..j2-1 (its state before row j1 kept) and one bottom-up sweep over
// rows R-1..j1+1 (its state before row j2 kept), so each row is read once
// (rows j1..j2 twice). Each unknown gets exactly the operations of
// twisted(j), hence the same value.
template <typename REAL>
__device__ void twisted2(const REAL *__restrict__ all, int64_t ns, int64_t s, int R, int j1,
                         int j2, REAL &x1, REAL &x2) {
  REAL cf = 0, df = 0, cf1 = 0, df1 = 0;
  for (int i = 0; i < j2; ++i) {
    if (i == j1) { cf1 = cf; df1 = df; }
    REAL sub, sup, rhs;
    reduced_row(all, ns, s, i, sub, sup, rhs);
    const REAL inv = REAL(1) / (REAL(1) - sub * cf);
    df = (rhs - sub * df) * inv;
    cf = sup * inv;
  }
  REAL ar = 0, dr = 0, ar2 = 0, dr2 = 0;
  for (int i = R - 1; i > j1; --i) {
    if (i == j2) { ar2 = ar; dr2 = dr; }
    REAL sub, sup, rhs;
    reduced_row(all, ns, s, i, sub, sup, rhs);
    const REAL inv = REAL(1) / (REAL(1) - sup * ar);
    dr = (rhs - sup * dr) * inv;
    ar = sub * inv;
  }
  REAL sub, sup, rhs;
  reduced_row(all, ns, s, j1, sub, sup, rhs);
  x1 = (rhs - sub * df1 - sup * dr) / (REAL(1) - sub * cf1 - sup * ar);
  reduced_row(all, ns, s, j2, sub, sup, rhs);
  x2 = (rhs - sub * df - sup * dr2) / (REAL(1) - sub * cf - sup * ar2);
}

template <typename REAL, bool INC>
__global__ __launch_bounds__(PART_BLOCK) void part_adjust_kernel(
    Geom g, const REAL *__restrict__ a, const REAL *__restrict__ c, REAL *__restrict__ d,
    const REAL *__restrict__ all, int G, int k, REAL *__restrict__ save) {
  const int64_t ns = g.nsys();
  const int64_t s = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
  if (s >= ns) return;
  const int64_t base = (s % g.n0) * g.st0 + (s / g.n0) * g.st1;
  const int64_t oe = base + (g.N - 1) * g.stride;
  const REAL d0 = d[base], dE = d[oe];
  if (INC) {
    save[s] = d0;
    save[ns + s] = dE;
  }
  const int R = 2 * G;
  if (k > 0 && k < G - 1) {  // an interior shard needs E_{k-1} and F_{k+1}
    REAL e_prev, f_next;
    twisted2(all, ns, s, R, 2 * k - 1, 2 * k + 2, e_prev, f_next);
    d[base] = d0 - a[base] * e_prev;
    d[oe] = dE - c[oe] * f_next;
    return;
  }
  if (k > 0) d[base] = d0 - a[base] * twisted(all, ns, s, R, 2 * k - 1);
  if (k < G - 1) d[oe] = dE - c[oe] * twisted(all, ns, s, R, 2 * k + 2);
}

template <typename REAL>
__global__ __launch_bounds__(PART_BLOCK) void part_restore_kernel(Geom g, REAL *__restrict__ d,
                                                                  const REAL *__restrict__ save) {
  const int64_t ns = g.nsys();
  const int64_t s = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
  if (s >= ns) return;
  const int64_t base = (s % g.n0) * g.st0 + (s / g.n0) * g.st1;
  d[base] = save[s];
  d[base + (g.N - 1) * g.stride] = save[ns + s];
}

unsigned part_blocks(const Geom &g) {
  return (unsigned)((g.nsys() + PART_BLOCK - 1) / PART_BLOCK);
}

tridStatus_t part_args(int ndim, int solvedim, const int *dims, const int *pads, int shard,
                       int nshards, Geom *g, bool *empty) {
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, g, empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (nshards < 1 || shard < 0 || shard >= nshards) return TRID_STATUS_INVALID_VALUE;
  if (!*empty && g->N < 2) return TRID_STATUS_INVALID_VALUE;  // every shard >= 2 rows
  return TRID_STATUS_SUCCESS;
}

template <typename REAL>
tridStatus_t boundary(const TridParams *ctx, const REAL *a, const REAL *b, const REAL *c,
                      const REAL *d, int ndim, int solvedim, const int *dims, const int *pads,
                      int shard, int nshards, REAL *bnd) {
  Geom g;
  bool empty = false;
  tridStatus_t s = part_args(ndim, solvedim, dims, pads, shard, nshards, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (!a || !b || !c || !d || !bnd) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (empty) return finish(ctx, st, hipSuccess);
  hipLaunchKernelGGL(part_boundary_kernel<REAL>, dim3(part_blocks(g)), dim3(PART_BLOCK), 0, st,
                     g, a, b, c, d, shard == 0 ? 1 : 0, shard == nshards - 1 ? 1 : 0, bnd);
  return finish(ctx, st, hipGetLastError());
}

template <typename REAL, bool INC>
tridStatus_t part_solve(const TridParams *ctx, const REAL *a, const REAL *b, const REAL *c,
                        REAL *d, REAL *u, int ndim, int solvedim, const int *dims,
                        const int *pads, int shard, int nshards, const REAL *all) {
  Geom g;
  bool empty = false;
  tridStatus_t s = part_args(ndim, solvedim, dims, pads, shard, nshards, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (!a || !b || !c || !d || !all || (INC && !u)) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  hipStream_t st = current_stream();
  if (empty) return finish(ctx, st, hipSuccess);
  REAL *save = nullptr;
  if (INC) {
    void *p = nullptr;
    s = get_workspace(dev, st, 2 * sizeof(REAL) * (size_t)g.nsys(), &p, 1);
    if (s != TRID_STATUS_SUCCESS) return s;
    save = static_cast<REAL *>(p);
  }
  hipLaunchKernelGGL((part_adjust_kernel<REAL, INC>), dim3(part_blocks(g)), dim3(PART_BLOCK), 0,
                     st, g, a, c, d, all, nshards, shard, save);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return finish(ctx, st, e);
  // the shard's own rows, stream-ordered after the adjustment
  TridParams local = ctx ? *ctx : TridParams{nullptr, {0, 0, 0}, 1};
  local.sync = 0;
  s = solve<REAL, INC>(&local, a, b, c, d, u, ndim, solvedim, dims, pads);
  if (INC) {  // d back as it was, also when the solve failed
    hipLaunchKernelGGL(part_restore_kernel<REAL>, dim3(part_blocks(g)), dim3(PART_BLOCK), 0,
                       st, g, d, save);
    e = hipGetLastError();
  }
  if (s != TRID_STATUS_SUCCESS) return s;
  return finish(ctx, st, e);
}

}  // namespace
}  // namespace trid

using namespace trid;

EXTERN_C tridStatus_t tridDmtsvStridedBatchPartBoundary(const TridParams *ctx, const double *a,
                                                        const double *b, const double *c,
                                                        const double *d, int ndim,
                                                        int solvedim, const int *dims,
                                                        const int *pads, int shard,
                                                        int nshards, double *bnd) {
  return boundary<double>(ctx, a, b, c, d, ndim, solvedim, dims, pads, shard, nshards, bnd);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchPartBoundary(const TridParams *ctx, const float *a,
                                                        const float *b, const float *c,
                                                        const float *d, int ndim,
                                                        int solvedim, const int *dims,
                                                        const int *pads, int shard,
                                                        int nshards, float *bnd) {
  return boundary<float>(ctx, a, b, c, d, ndim, solvedim, dims, pads, shard, nshards, bnd);
}

EXTERN_C tridStatus_t tridDmtsvStridedBatchPartSolve(const TridParams *ctx, const double *a,
                                                     const double *b, const double *c,
                                                     double *d, int ndim, int solvedim,
                                                     const int *dims, const int *pads,
                                                     int shard, int nshards,
                                                     const double *bnd_all) {
  return part_solve<double, false>(ctx, a, b, c, d, nullptr, ndim, solvedim, dims, pads, shard,
                                   nshards, bnd_all);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchPartSolve(const TridParams *ctx, const float *a,
                                                     const float *b, const float *c, float *d,
                                                     int ndim, int solvedim, const int *dims,
                                                     const int *pads, int shard, int nshards,
                                                     const float *bnd_all) {
  return part_solve<float, false>(ctx, a, b, c, d, nullptr, ndim, solvedim, dims, pads, shard,
                                  nshards, bnd_all);
}

EXTERN_C tridStatus_t tridDmtsvStridedBatchIncPartSolve(const TridParams *ctx, const double *a,
                                                        const double *b, const double *c,
                                                        double *d, double *u, int ndim,
                                                        int solvedim, const int *dims,
                                                        const int *pads, int shard,
                                                        int nshards, const double *bnd_all) {
  return part_solve<double, true>(ctx, a, b, c, d, u, ndim, solvedim, dims, pads, shard,
                                  nshards, bnd_all);
}

EXTERN_C tridStatus_t tridSmtsvStridedBatchIncPartSolve(const TridParams *ctx, const float *a,
                                                        const float *b, const float *c,
                                                        float *d, float *u, int ndim,
                                                        int solvedim, const int *dims,
                                                        const int *pads, int shard,
                                                        int nshards, const float *bnd_all) {
  return part_solve<float, true>(ctx, a, b, c, d, u, ndim, solvedim, dims, pads, shard, nshards,
                                 bnd_all);
}
