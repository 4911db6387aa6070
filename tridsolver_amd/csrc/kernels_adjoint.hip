// kernels_adjoint.hip - adjoint of the batched solve (no counterpart in the
// reference).
//
// For every system A x = d (A = tridiag(a, b, c), a[0] and c[N-1] unused) and
// an incoming gradient w = dL/dx, the adjoint solves
//     A^T lam = w,   A^T = tridiag(sub = c[n-1], diag = b[n], sup = a[n+1]),
// and forms
//     dL/dd = lam,  dL/da[n] = -lam[n] x[n-1] (0 at n = 0),
//     dL/db[n] = -lam[n] x[n],  dL/dc[n] = -lam[n] x[n+1] (0 at n = N-1),
// with x the forward solution. The transpose is solved by the same Thomas-PCR
// hybrid as the forward solve (hybrid_core.hpp): the shifted coefficients are
// fetched directly (row n of the transpose reads c at n-1 and a at n+1, so
// every element is still read once), and the neighbour chunks' boundary x
// values come through LDS. HBM traffic: read a, b, c, w, x; write lam and
// each requested gradient (72 B per point at FP64 with all three gradients).
// When no coefficient gradient is requested (GRAD false) x is not read: the
// solve for lam alone moves the forward solve's 40 B per point.
// Systems longer than adjoint_kernel holds take three steps instead: the
// transpose's off-diagonals written out (adj_transpose_kernel), A^T lam = w
// solved by the batched solve, which takes any length, and the gradients
// (adj_grad_kernel): about 128 B per point with all three gradients.

#include <algorithm>
#include <atomic>

#include "../../include/tridsolver_hip.h"
#include "hybrid_core.hpp"
#include "trid_internal.h"

namespace trid {
namespace {

template <typename REAL, int M>
struct AdjBounds {
  static constexpr int threads = sizeof(REAL) * M >= 64 ? 512 : 1024;
};

template <typename REAL, int M, bool CONTIG, bool GRAD, bool NT>
__global__ __launch_bounds__((AdjBounds<REAL, M>::threads)) void adjoint_kernel(
    Geom g, int P, int W, int G, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, const REAL *__restrict__ x, const REAL *__restrict__ w,
    REAL *__restrict__ lam, REAL *__restrict__ ga, REAL *__restrict__ gb,
    REAL *__restrict__ gc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<CONTIG>(g, P, W, G);
  const size_t tile_bytes = CONTIG ? sizeof(REAL) * (size_t)W * P * (M + 1) : 0;
  REAL *tile = reinterpret_cast<REAL *>(smem_raw) + (size_t)L.w * P * (M + 1);
  hyb::Ring ring{smem_raw + tile_bytes, (int)(3 * sizeof(REAL) * L.T), 0};

  REAL ra[M], rb[M], rc[M], rd[M], rx[M];
  hyb::fetch<M, CONTIG, NT>(L, c, ra, REAL(0), -1, 0, L.N);  // sub of A^T: c[n-1]
  hyb::fetch<M, CONTIG, NT>(L, b, rb, REAL(1), 0, 0, L.N);
  hyb::fetch<M, CONTIG, NT>(L, a, rc, REAL(0), +1, 1, L.N);  // sup of A^T: a[n+1]
  hyb::fetch<M, CONTIG, NT>(L, w, rd, REAL(0), 0, 0, L.N);
  if (GRAD) hyb::fetch<M, CONTIG, NT>(L, x, rx, REAL(0), 0, 0, L.N);
  hyb::to_chunk<M, CONTIG>(L, ra, tile);
  hyb::to_chunk<M, CONTIG>(L, rb, tile);
  hyb::to_chunk<M, CONTIG>(L, rc, tile);
  hyb::to_chunk<M, CONTIG>(L, rd, tile);
  if (GRAD) hyb::to_chunk<M, CONTIG>(L, rx, tile);
  REAL l[M];
  hyb::solve_chunk<REAL, M, CONTIG>(L, ring, ra, rb, rc, rd, l);
  if (!GRAD) {
    hyb::to_fetch<M, CONTIG>(L, l, tile);
    hyb::store<M, CONTIG, NT>(L, lam, l);
    return;
  }

  // x of the neighbouring chunks' boundary rows
  REAL xprev = 0, xnext = 0;
  {
    REAL *q = ring.next<REAL>();
    q[L.t] = rx[0];
    q[L.T + L.t] = rx[M - 1];
    __syncthreads();
    if (L.p > 0) xprev = q[L.T + L.t - L.pstep];
    if (L.p + 1 < L.P) xnext = q[L.t + L.pstep];
  }
  const int64_t row0 = (int64_t)L.p * M;  // global row of chunk slot 0
  // Outputs go through the tile one at a time (contiguous systems).
  if (ga) {
    REAL v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const REAL xl = r > 0 ? rx[r - 1] : xprev;
      v[r] = (row0 + r == 0) ? REAL(0) : -l[r] * xl;
    }
    hyb::to_fetch<M, CONTIG>(L, v, tile);
    hyb::store<M, CONTIG, NT>(L, ga, v);
  }
  if (gb) {
    REAL v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) v[r] = -l[r] * rx[r];
    hyb::to_fetch<M, CONTIG>(L, v, tile);
    hyb::store<M, CONTIG, NT>(L, gb, v);
  }
  if (gc) {
    REAL v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const REAL xr = r < M - 1 ? rx[r + 1] : xnext;
      v[r] = (row0 + r >= L.N - 1) ? REAL(0) : -l[r] * xr;
    }
    hyb::to_fetch<M, CONTIG>(L, v, tile);
    hyb::store<M, CONTIG, NT>(L, gc, v);
  }
  hyb::to_fetch<M, CONTIG>(L, l, tile);
  hyb::store<M, CONTIG, NT>(L, lam, l);
}

// Nontemporal loads and stores except for strided systems whose rows are
// 64 KiB or more apart (256^3 FP64, all gradients: x / y 0.233 / 0.233 ms
// with them against 0.242 / 0.252, z 0.294 against 0.282).
template <typename REAL, int M, bool CONTIG, bool GRAD, bool NT>
hipError_t launch_adjoint_nt(const Geom &g, const HybridCfg &cfg, const REAL *a,
                             const REAL *b, const REAL *c, const REAL *x, const REAL *w,
                             REAL *lam, REAL *ga, REAL *gb, REAL *gc, hipStream_t st) {
  const int T = cfg.W * cfg.P;
  size_t lds = hyb::ring_bytes<REAL>(T, 3);
  if (CONTIG) lds += sizeof(REAL) * (size_t)T * (M + 1);
  auto kern = adjoint_kernel<REAL, M, CONTIG, GRAD, NT>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.W, cfg.G,
                     a, b, c, x, w, lam, ga, gb, gc);
  return hipGetLastError();
}

template <typename REAL, int M, bool CONTIG, bool GRAD>
hipError_t launch_adjoint_m(const Geom &g, const HybridCfg &cfg, const REAL *a,
                            const REAL *b, const REAL *c, const REAL *x, const REAL *w,
                            REAL *lam, REAL *ga, REAL *gb, REAL *gc, hipStream_t st) {
  if (CONTIG || g.stride * (int64_t)sizeof(REAL) < (64 << 10))
    return launch_adjoint_nt<REAL, M, CONTIG, GRAD, true>(g, cfg, a, b, c, x, w, lam, ga, gb,
                                                          gc, st);
  return launch_adjoint_nt<REAL, M, CONTIG, GRAD, false>(g, cfg, a, b, c, x, w, lam, ga, gb,
                                                         gc, st);
}

template <typename REAL>
int adj_max_threads(int M) {
  switch (M) {
    case 1: return AdjBounds<REAL, 1>::threads;
    case 2: return AdjBounds<REAL, 2>::threads;
    case 4: return AdjBounds<REAL, 4>::threads;
    case 8: return AdjBounds<REAL, 8>::threads;
    case 16: return AdjBounds<REAL, 16>::threads;
    default: return 0;
  }
}

// ---- systems longer than the adjoint kernel holds -------------------------
// The transpose's off-diagonals in the layout of the inputs, sub[n] = c[n-1]
// and sup[n] = a[n+1] (0 where the row has none), and lam = w, the right-hand
// side the batched solve then overwrites with the solution of A^T lam = w.
template <typename REAL, bool CONTIG>
__global__ __launch_bounds__(256) void adj_transpose_kernel(
    Geom g, const REAL *__restrict__ a, const REAL *__restrict__ c, const REAL *__restrict__ w,
    REAL *__restrict__ sub, REAL *__restrict__ sup, REAL *__restrict__ lam) {
  const int64_t total = g.nsys() * g.N;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    int64_t k;
    const int64_t i = hyb::batch_elem<CONTIG>(g, e, &k);
    sub[i] = k > 0 ? c[i - g.stride] : REAL(0);
    sup[i] = k + 1 < g.N ? a[i + g.stride] : REAL(0);
    lam[i] = w[i];
  }
}

// dL/da[n] = -lam[n] x[n-1] (0 at n = 0), dL/db[n] = -lam[n] x[n],
// dL/dc[n] = -lam[n] x[n+1] (0 at n = N-1), each where its output is given.
template <typename REAL, bool CONTIG>
__global__ __launch_bounds__(256) void adj_grad_kernel(Geom g, const REAL *__restrict__ x,
                                                       const REAL *__restrict__ lam,
                                                       REAL *__restrict__ ga,
                                                       REAL *__restrict__ gb,
                                                       REAL *__restrict__ gc) {
  const int64_t total = g.nsys() * g.N;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    int64_t k;
    const int64_t i = hyb::batch_elem<CONTIG>(g, e, &k);
    const REAL l = lam[i];
    if (ga) ga[i] = k > 0 ? -l * x[i - g.stride] : REAL(0);
    if (gb) gb[i] = -l * x[i];
    if (gc) gc[i] = k + 1 < g.N ? -l * x[i + g.stride] : REAL(0);
  }
}

}  // namespace

template <typename REAL>
hipError_t launch_adjoint_transpose(const Geom &g, const REAL *a, const REAL *c, const REAL *w,
                                    REAL *sub, REAL *sup, REAL *lam, hipStream_t st) {
  if (g.stride == 1)
    hipLaunchKernelGGL((adj_transpose_kernel<REAL, true>), dim3(batch_elem_blocks(g)), dim3(256),
                       0, st, g, a, c, w, sub, sup, lam);
  else
    hipLaunchKernelGGL((adj_transpose_kernel<REAL, false>), dim3(batch_elem_blocks(g)),
                       dim3(256), 0, st, g, a, c, w, sub, sup, lam);
  return hipGetLastError();
}

template <typename REAL>
hipError_t launch_adjoint_grads(const Geom &g, const REAL *x, const REAL *lam, REAL *ga,
                                REAL *gb, REAL *gc, hipStream_t st) {
  if (g.stride == 1)
    hipLaunchKernelGGL((adj_grad_kernel<REAL, true>), dim3(batch_elem_blocks(g)), dim3(256), 0,
                       st, g, x, lam, ga, gb, gc);
  else
    hipLaunchKernelGGL((adj_grad_kernel<REAL, false>), dim3(batch_elem_blocks(g)), dim3(256), 0,
                       st, g, x, lam, ga, gb, gc);
  return hipGetLastError();
}

template <typename REAL>
bool adjoint_cfg_valid(const Geom &g, const HybridCfg &c) {
  // FP64 M = 16 on contiguous systems spills a few registers; not offered.
  if (sizeof(REAL) == 8 && c.M == 16 && g.stride == 1) return false;
  const int maxT = adj_max_threads<REAL>(c.M);
  if (maxT == 0 || c.P < 1 || c.W < 1 || c.G < 1) return false;
  if ((int64_t)c.P * c.M < g.N) return false;
  const int64_t T = (int64_t)c.W * c.P;
  if (T > maxT) return false;
  size_t lds = hyb::ring_bytes<REAL>((int)T, 3);
  if (g.stride == 1) lds += sizeof(REAL) * (size_t)T * (c.M + 1);
  return lds <= 160 * 1024;
}

template <typename REAL>
bool adjoint_default_cfg(const Geom &g, HybridCfg *cfg) {
  // Same shapes as the forward solve where they fit the adjoint's register
  // budget, else the next valid one.
  HybridCfg c;
  if (hybrid_plain_cfg<REAL>(g, &c)) {
    for (int k = 0; k < 4 && !adjoint_cfg_valid<REAL>(g, c); ++k) {
      if (g.stride == 1 && c.W > 1) c.W /= 2;
      else if (c.M < 16) { c.M *= 2; c.P = (int)((g.N + c.M - 1) / c.M); }
    }
    if (adjoint_cfg_valid<REAL>(g, c)) { *cfg = c; return true; }
  }
  static const int Ms[] = {8, 4, 16, 2, 1};
  for (int W = g.stride == 1 ? 1 : 128 / (int)sizeof(REAL); W >= 1; W /= 2)
    for (int M : Ms) {
      c.M = M;
      c.P = (int)((g.N + M - 1) / M);
      c.W = W;
      c.G = 8;
      if (adjoint_cfg_valid<REAL>(g, c)) { *cfg = c; return true; }
    }
  return false;
}

template <typename REAL>
hipError_t launch_adjoint(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                          const REAL *c, const REAL *x, const REAL *w, REAL *lam, REAL *ga,
                          REAL *gb, REAL *gc, hipStream_t st) {
  const bool contig = g.stride == 1, grad = ga || gb || gc;
#define TRID_ADJ_CASE(MM)                                                                  \
  case MM:                                                                                 \
    if (grad)                                                                              \
      return contig ? launch_adjoint_m<REAL, MM, true, true>(g, cfg, a, b, c, x, w, lam, ga, \
                                                             gb, gc, st)                   \
                    : launch_adjoint_m<REAL, MM, false, true>(g, cfg, a, b, c, x, w, lam, \
                                                              ga, gb, gc, st);             \
    return contig ? launch_adjoint_m<REAL, MM, true, false>(g, cfg, a, b, c, x, w, lam, ga, \
                                                            gb, gc, st)                    \
                  : launch_adjoint_m<REAL, MM, false, false>(g, cfg, a, b, c, x, w, lam, ga, \
                                                             gb, gc, st);
  switch (cfg.M) {
    TRID_ADJ_CASE(1)
    TRID_ADJ_CASE(2)
    TRID_ADJ_CASE(4)
    TRID_ADJ_CASE(8)
    TRID_ADJ_CASE(16)
    default:
      return hipErrorInvalidValue;
  }
#undef TRID_ADJ_CASE
}

template bool adjoint_cfg_valid<float>(const Geom &, const HybridCfg &);
template bool adjoint_cfg_valid<double>(const Geom &, const HybridCfg &);
template bool adjoint_default_cfg<float>(const Geom &, HybridCfg *);
template bool adjoint_default_cfg<double>(const Geom &, HybridCfg *);
template hipError_t launch_adjoint<float>(const Geom &, const HybridCfg &, const float *,
                                          const float *, const float *, const float *,
                                          const float *, float *, float *, float *, float *,
                                          hipStream_t);
template hipError_t launch_adjoint<double>(const Geom &, const HybridCfg &, const double *,
                                           const double *, const double *, const double *,
                                           const double *, double *, double *, double *,
                                           double *, hipStream_t);
#define TRID_ADJ_LONG_INST(REAL)                                                            \
  template hipError_t launch_adjoint_transpose<REAL>(const Geom &, const REAL *, const REAL *, \
                                                     const REAL *, REAL *, REAL *, REAL *,     \
                                                     hipStream_t);                             \
  template hipError_t launch_adjoint_grads<REAL>(const Geom &, const REAL *, const REAL *,    \
                                                 REAL *, REAL *, REAL *, hipStream_t);
TRID_ADJ_LONG_INST(float)
TRID_ADJ_LONG_INST(double)
#undef TRID_ADJ_LONG_INST

}  // namespace trid
