// kernels_mixed.hip - FP64 batched solve by FP32 sweeps plus FP64 iterative
// refinement, fused into one pass over HBM (no counterpart in the reference).
//
// For each system (same chunking, LDS staging and reduced PCR as the hybrid
// in kernels_solve.hip):
//   1. the FP64 operands a, b, c, d of the lane's chunk are loaded once and
//      kept on chip (registers; a and c in LDS for the M = 16 kernels);
//   2. the matrix is factored in FP32 (mixed_core.hpp): the forward sweep's
//      multipliers (1/beta_i, ga_i and the spike al_i per row), the chunk's
//      first row in reduced form and, for every PCR step, the pre-step
//      coefficients (A, C, 1/(1 - A Cm - C Ap)) -- none of these depend on
//      the right-hand side;
//   3. x0 = FP32 solve of (float) d with that factorisation (right-hand-side
//      recurrences only: de_i = (rhs_i - a_i de_{i-1}) / beta_i, the first
//      row's dp_0, PCR on D only, then x_i = de_i - al_i x_{s-1} -
//      ga_i x_{i+1} from the chunk's end);
//   4. `refine` times: r = d - A x in FP64 (neighbour chunks' boundary x via
//      LDS), dx = FP32 solve of (float) r, x += dx in FP64;
//   5. x is written in FP64.
// The residual and the accumulation are the only FP64 arithmetic; the FP32
// factorisation takes its reciprocals from the hardware's approximate
// reciprocal (mixed_core.hpp, FAST), whose error the refinement removes. HBM
// traffic is one read of a, b, c, d and one write of d (40 B per point), the
// same as the FP64 solve.

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
  hyb::fetch_abcd<M, CONTIG, true>(L, a, b, c, d, A, B, Cc, D);
  hyb::to_chunk<M, CONTIG>(L, A, tile);
  hyb::to_chunk<M, CONTIG>(L, B, tile);
  hyb::to_chunk<M, CONTIG>(L, Cc, tile);
  hyb::to_chunk<M, CONTIG>(L, D, tile);

  double X[M];
  mixed_chunk<M>(L, ring, A, B, Cc, D, refine, X);
  hyb::to_fetch<M, CONTIG>(L, X, tile);
  hyb::store<M, CONTIG, true>(L, d, X);
}

template <int M, bool CONTIG>
hipError_t launch_mixed_m(const Geom &g, const HybridCfg &cfg, int refine, const double *a,
                          const double *b, const double *c, double *d, hipStream_t st) {
  const int T = cfg.W * cfg.P;
  size_t lds = hyb::ring_bytes<double>(T, 3);
  if (CONTIG) lds += sizeof(double) * (size_t)T * (M + 1);
  auto kern = mixed_kernel<M, CONTIG>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.W, cfg.G,
                     refine, a, b, c, d);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// V = 3: M = 16 rows per lane with the FP64 a and c operands in LDS, b and d
// in registers (the four operands plus the FP32 factors of 16 rows exceed
// the register file). Contiguous systems: per-system tiles of P*(M+1)
// doubles (mixed_contig_lds). Strided systems: tiles [row][system] filled by
// global->LDS loads, SEG = W*8 = 64 or 128-byte rows, as in kernels_wide.hip.
// ---------------------------------------------------------------------------
constexpr int MIXED_LDS_M = 16;
constexpr size_t MIXED_LDS_MAX = 160 * 1024;

size_t mixed_lds_bytes(const Geom &g, int P, int W) {
  const size_t T = (size_t)W * P;
  const size_t ring = hyb::ring_bytes<double>((int)T, 2);
  if (g.stride == 1) return 2 * sizeof(double) * T * (MIXED_LDS_M + 1) + ring;
  return 2 * sizeof(double) * (size_t)P * MIXED_LDS_M * W + ring;
}

__global__ __launch_bounds__(MIXED_THREADS) void mixed_tile_kernel(
    Geom g, int P, int W, int G, int refine, const double *__restrict__ a,
    const double *__restrict__ b, const double *__restrict__ c, double *__restrict__ d) {
  constexpr int M = MIXED_LDS_M;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<true>(g, P, W, G);
  double *tA = reinterpret_cast<double *>(smem_raw) + (size_t)L.w * P * (M + 1);
  double *tC = tA + (size_t)L.T * (M + 1);
  hyb::Ring ring{smem_raw + 2 * sizeof(double) * (size_t)L.T * (M + 1),
                 (int)(2 * sizeof(double) * L.T), 0};
  mixed_contig_lds<M, MIXED_MAX_STEPS, true>(L, ring, a, b, c, d, tA, tC, refine, [] {},
                                       [&] { return L.base; });
}

typedef __attribute__((address_space(3))) void lds_void;

template <int SEG>
__global__ __launch_bounds__(MIXED_THREADS) void mixed_wide_kernel(
    Geom g, int P, int G, int refine, const double *__restrict__ a,
    const double *__restrict__ b, const double *__restrict__ c, double *__restrict__ d) {
  constexpr int M = MIXED_LDS_M;
  constexpr int W = SEG / (int)sizeof(double);
  constexpr int K = 256 / SEG;  // rows per wave-wide global->LDS load
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<false>(g, P, W, G);
  const int T = L.T;
  const int64_t R = (int64_t)P * M;  // rows per tile (>= N)
  double *tA = reinterpret_cast<double *>(smem_raw);
  double *tC = tA + (size_t)R * W;
  hyb::Ring ring{smem_raw + 2 * sizeof(double) * (size_t)R * W, (int)(2 * sizeof(double) * T), 0};

  // (0) a rows 1..N-1 and c rows 0..N-2 -> tiles [row][system] by
  // global->LDS loads. One wave instruction writes 256 contiguous LDS bytes,
  // lane l its bytes 4l..4l+3: K rows of SEG bytes, lane l serving row
  // (first row + 4l / SEG), system (4l % SEG) / 8. Systems beyond n0 are
  // masked off; the slots no load writes (a row 0, c row N-1, rows >= N,
  // systems beyond n0) are cleared by their own lanes below.
  {
    const int lane = L.t & 63, wv = L.t >> 6, nw = T >> 6;
    const int sub = (lane * 4) / SEG, byte = (lane * 4) % SEG;
    const int64_t first = L.i0 - L.w;
    const bool sys_ok = first + byte / (int)sizeof(double) < g.n0;
    const int64_t row_bytes = L.stride * (int64_t)sizeof(double);
    const int64_t o = (first + L.i1 * g.st1) * (int64_t)sizeof(double) + byte + sub * row_bytes;
    const char *srcA = reinterpret_cast<const char *>(a) + o;
    const char *srcC = reinterpret_cast<const char *>(c) + o;
    if (sys_ok) {
      for (int64_t row = (int64_t)K * wv; row < L.N; row += (int64_t)K * nw) {
        const int64_t rr = row + sub;
        if (rr >= 1 && rr < L.N)
          __builtin_amdgcn_global_load_lds(srcA + row * row_bytes, (lds_void *)(tA + row * W), 4,
                                           0, 0);
        if (rr < L.N - 1)
          __builtin_amdgcn_global_load_lds(srcC + row * row_bytes, (lds_void *)(tC + row * W), 4,
                                           0, 0);
      }
    }
  }

  // b, d of the lane's rows into registers (identity rows past N), 32-bit
  // byte offsets from the uniform bases (extent < 4 GiB, checked by
  // mixed_cfg_valid).
  const int64_t row0 = (int64_t)L.p * M;
  const uint32_t ob = (uint32_t)((L.base + row0 * L.stride) * (int64_t)sizeof(double));
  const uint32_t os = (uint32_t)(L.stride * (int64_t)sizeof(double));
  auto at = [](const double *p, uint32_t off) -> double {
    return hyb::ld<true>(reinterpret_cast<const double *>(reinterpret_cast<const char *>(p) + off));
  };
  const bool full = L.active && row0 + M <= L.N;
  double B[M], D[M];
  if (full) {
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const uint32_t o = ob + (uint32_t)r * os;
      B[r] = at(b, o);
      D[r] = at(d, o);
    }
  } else {
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const bool ok = L.active && row0 + r < L.N;
      const uint32_t o = ob + (uint32_t)r * os;
      B[r] = ok ? at(b, o) : 1.0;
      D[r] = ok ? at(d, o) : 0.0;
    }
  }
  double *sa = tA + row0 * W + L.w;  // sa[r * W]: this lane's row r
  double *sc = tC + row0 * W + L.w;
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int64_t row = row0 + r;
    if (!L.active || row == 0 || row >= L.N) sa[r * W] = 0.0;
    if (!L.active || row >= L.N - 1) sc[r * W] = 0.0;
  }
  // the tiles are written by other waves' loads: wait for this wave's, then
  // for every wave
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  double X[M];
  mixed_chunk_acc<M, MIXED_MAX_STEPS, true, false, true>(
      L, ring, [&](int r) { return sa[r * W]; }, B, [&](int r) { return sc[r * W]; }, D, refine,
      X);
  if (!L.active) return;
  // store offsets recomputed (the empty asm hides that they equal the load
  // offsets, which would otherwise stay live across the solve)
  uint32_t obs = ob;
  asm volatile("" : "+v"(obs));
  const int rows = full ? M : (int)std::min<int64_t>(M, std::max<int64_t>(0, L.N - row0));
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < rows)
      hyb::st<true>(reinterpret_cast<double *>(reinterpret_cast<char *>(d) + obs + (uint32_t)r * os), X[r]);
  }
}

hipError_t launch_mixed_lds(const Geom &g, const HybridCfg &cfg, int refine, const double *a,
                            const double *b, const double *c, double *d, hipStream_t st) {
  const int T = cfg.W * cfg.P;
  const size_t lds = mixed_lds_bytes(g, cfg.P, cfg.W);
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  auto go = [&](const void *kern, LdsGrant &granted, auto launch) {
    hipError_t e = ensure_lds(kern, lds, granted);
    if (e != hipSuccess) return e;
    launch();
    return hipGetLastError();
  };
  if (g.stride == 1) {
    static LdsGrant granted;
    return go(reinterpret_cast<const void *>(mixed_tile_kernel), granted, [&] {
      hipLaunchKernelGGL(mixed_tile_kernel, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P,
                         cfg.W, cfg.G, refine, a, b, c, d);
    });
  }
  if (cfg.W * (int)sizeof(double) == 128) {
    static LdsGrant granted;
    return go(reinterpret_cast<const void *>(mixed_wide_kernel<128>), granted, [&] {
      hipLaunchKernelGGL(mixed_wide_kernel<128>, dim3((unsigned)blocks), dim3(T), lds, st, g,
                         cfg.P, cfg.G, refine, a, b, c, d);
    });
  }
  static LdsGrant granted;
  return go(reinterpret_cast<const void *>(mixed_wide_kernel<64>), granted, [&] {
    hipLaunchKernelGGL(mixed_wide_kernel<64>, dim3((unsigned)blocks), dim3(T), lds, st, g,
                       cfg.P, cfg.G, refine, a, b, c, d);
  });
}

bool mixed_lds_cfg_valid(const Geom &g, const HybridCfg &c) {
  if (c.M != MIXED_LDS_M || c.P < 1 || c.W < 1 || c.G < 1) return false;
  if ((int64_t)c.P * c.M < g.N || c.P > (1 << MIXED_MAX_STEPS)) return false;
  const int64_t T = (int64_t)c.W * c.P;
  if (T > MIXED_THREADS) return false;
  if (g.stride != 1) {
    const int seg = c.W * (int)sizeof(double);
    if (g.st0 != 1 || (seg != 64 && seg != 128) || T % 64 != 0) return false;
    // 32-bit byte offsets: every address a lane forms lies below 4 GiB
    const int64_t nb0 = (g.n0 + c.W - 1) / c.W * c.W;
    const int64_t last =
        (nb0 - 1) + (g.n1 - 1) * g.st1 + ((int64_t)c.P * c.M - 1) * g.stride;
    if ((last + 1) * (int64_t)sizeof(double) > (int64_t)0xFFFFFFFFll) return false;
  }
  return mixed_lds_bytes(g, c.P, c.W) <= MIXED_LDS_MAX;
}

// Default V = 3 shape: contiguous, 256-thread workgroups (two per CU) up to
// 4096 rows, one system of up to 512 lanes beyond; strided, 128-byte rows
// while 16 systems fit 512 threads and the LDS (N <= 512), else 64-byte rows
// (N <= 1024).
bool mixed_lds_default_cfg(const Geom &g, HybridCfg *cfg) {
  HybridCfg c;
  c.M = MIXED_LDS_M;
  c.V = 3;
  c.P = (int)((g.N + c.M - 1) / c.M);
  if (g.stride == 1) {
    c.W = (int)std::max<int64_t>(1, 256 / c.P);
    if (mixed_lds_cfg_valid(g, c)) { *cfg = c; return true; }
    return false;
  }
  for (int W : {16, 8}) {
    HybridCfg s = c;
    s.W = W;
    const int per_wave = 64 / W;
    s.P = (c.P + per_wave - 1) / per_wave * per_wave;
    if (mixed_lds_cfg_valid(g, s)) { *cfg = s; return true; }
  }
  return false;
}

}  // namespace

bool mixed_cfg_valid(const Geom &g, const HybridCfg &c) {
  if (c.V == 3) return mixed_lds_cfg_valid(g, c);
  if (c.V != 0) return false;
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
  if (mixed_lds_default_cfg(g, cfg)) return true;
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
  if (cfg.V == 3) return launch_mixed_lds(g, cfg, refine, a, b, c, d, st);
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
