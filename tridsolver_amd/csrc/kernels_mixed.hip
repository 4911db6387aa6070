// kernels_mixed.hip - FP64 batched solve by FP32 sweeps plus FP64 iterative
// refinement, fused into one pass over HBM (no counterpart in the reference).
//
// For each system (same chunking, LDS staging and reduced PCR as the hybrid
// in kernels_solve.hip):
//   1. the FP64 operands a, b, c, d of the lane's chunk are loaded once and
//      kept in registers;
//   2. the matrix is factored in FP32: the spike sweeps' multipliers
//      (1/beta_i, ga_i, ap_i, gp_i per row) and, for every PCR step, the
//      pre-step coefficients (A, C, 1/(1 - A Cm - C Ap)) -- none of these
//      depend on the right-hand side;
//   3. x0 = FP32 solve of (float) d with that factorisation (right-hand-side
//      recurrences only: de_i = (rhs_i - a_i de_{i-1}) / beta_i, backward
//      dp_i = de_i - ga_i dp_{i+1}, PCR on D only, substitution);
//   4. `refine` times: r = d - A x in FP64 (neighbour chunks' boundary x via
//      LDS), dx = FP32 solve of (float) r, x += dx in FP64;
//   5. x is written in FP64.
// The residual and the accumulation are the only FP64 arithmetic; HBM traffic
// is one read of a, b, c, d and one write of d (40 B per point), the same as
// the FP64 solve.

#include <algorithm>
#include <atomic>

#include "../../include/tridsolver_hip.h"
#include "hybrid_core.hpp"
#include "mixed_core.hpp"
#include "trid_internal.h"

namespace trid {
namespace {

constexpr int MIXED_THREADS = 512;

template <int M, bool CONTIG>
__global__ __launch_bounds__(MIXED_THREADS) void mixed_kernel(
    Geom g, int P, int W, int G, int refine, const double *__restrict__ a,
    const double *__restrict__ b, const double *__restrict__ c, double *__restrict__ d) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<CONTIG>(g, P, W, G);
  const size_t tile_bytes = CONTIG ? sizeof(double) * (size_t)W * P * (M + 1) : 0;
  double *tile = reinterpret_cast<double *>(smem_raw) + (size_t)L.w * P * (M + 1);
  hyb::Ring ring{smem_raw + tile_bytes, (int)(3 * sizeof(double) * L.T), 0};
  double A[M], B[M], Cc[M], D[M];
  hyb::fetch_abcd<M, CONTIG>(L, a, b, c, d, A, B, Cc, D);
  hyb::to_chunk<M, CONTIG>(L, A, tile);
  hyb::to_chunk<M, CONTIG>(L, B, tile);
  hyb::to_chunk<M, CONTIG>(L, Cc, tile);
  hyb::to_chunk<M, CONTIG>(L, D, tile);

  double X[M];
  mixed_chunk<M>(L, ring, A, B, Cc, D, refine, X);
  hyb::to_fetch<M, CONTIG>(L, X, tile);
  hyb::store<M, CONTIG>(L, d, X);
}

template <int M, bool CONTIG>
hipError_t launch_mixed_m(const Geom &g, const HybridCfg &cfg, int refine, const double *a,
                          const double *b, const double *c, double *d, hipStream_t st) {
  const int T = cfg.W * cfg.P;
  size_t lds = hyb::ring_bytes<double>(T, 3);
  if (CONTIG) lds += sizeof(double) * (size_t)T * (M + 1);
  auto kern = mixed_kernel<M, CONTIG>;
  static std::atomic<size_t> lds_granted{0};
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.W, cfg.G,
                     refine, a, b, c, d);
  return hipGetLastError();
}

}  // namespace

bool mixed_cfg_valid(const Geom &g, const HybridCfg &c) {
  if (c.M != 4 && c.M != 8) return false;  // M = 16 exceeds the register file
  if (c.P < 1 || c.W < 1 || c.G < 1) return false;
  if ((int64_t)c.P * c.M < g.N || c.P > (1 << MIXED_MAX_STEPS)) return false;
  const int64_t T = (int64_t)c.W * c.P;
  if (T > MIXED_THREADS) return false;
  size_t lds = hyb::ring_bytes<double>((int)T, 3);
  if (g.stride == 1) lds += sizeof(double) * (size_t)T * (c.M + 1);
  return lds <= 160 * 1024;
}

bool mixed_default_cfg(const Geom &g, HybridCfg *cfg) {
  const bool contig = g.stride == 1;
  static const int Ms[] = {4, 8};
  for (int M : Ms) {
    HybridCfg c;
    c.M = M;
    c.P = (int)((g.N + M - 1) / M);
    if (contig) {
      c.W = (int)std::max<int64_t>(1, 256 / c.P);
    } else {
      c.W = 16;
      while (c.W > 1 && (int64_t)c.W * c.P > MIXED_THREADS) c.W /= 2;
    }
    if (mixed_cfg_valid(g, c)) { *cfg = c; return true; }
  }
  return false;
}

hipError_t launch_mixed(const Geom &g, const HybridCfg &cfg, int refine, const double *a,
                        const double *b, const double *c, double *d, hipStream_t st) {
  const bool contig = g.stride == 1;
  switch (cfg.M) {
    case 4: return contig ? launch_mixed_m<4, true>(g, cfg, refine, a, b, c, d, st)
                          : launch_mixed_m<4, false>(g, cfg, refine, a, b, c, d, st);
    case 8: return contig ? launch_mixed_m<8, true>(g, cfg, refine, a, b, c, d, st)
                          : launch_mixed_m<8, false>(g, cfg, refine, a, b, c, d, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace trid
