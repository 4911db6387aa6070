// kernels_mixed32.hip - FP64 solution of batched systems stored in FP32, by
// FP32 sweeps plus FP64 iterative refinement, in one pass over HBM (no
// counterpart in the reference).
//
// The system solved is the one whose coefficients and right-hand side are
// the FP32 values, each exactly representable in FP64: A x = d with a, b, c
// and d promoted. The factorisation and the sweeps run on the FP32 values as
// they are; the residual d - A x is formed in FP64 from the promoted values
// and x += dx is accumulated in FP64 (mixed_core.hpp). x goes to a separate
// FP64 array of the operands' layout. The FP32 factorisation takes its
// reciprocals from the hardware's approximate reciprocal (mixed_core.hpp,
// FAST). HBM traffic per point: 16 B read (a, b, c, d) and 8 B written (x),
// against the 40 B the FP64 solve of the same system moves when it is
// stored in FP64.
//
// Three kernels:
//   mixed32_kernel       every operand in registers; contiguous systems go
//                        through an LDS tile per system (float operands in,
//                        FP64 x out), strided ones take W x-neighbours.
//   mixed32_lds_kernel   a and c kept in LDS (2*M registers fewer a lane):
//                        contiguous systems in their transpose tile,
//                        strided ones in per-lane slots.
//   mixed32_wide_kernel  strided systems, 32 x-neighbours per workgroup, so
//                        every row is one 128-byte segment (what keeps the
//                        z solve's address translations per byte low), or 16
//                        (64-byte rows); `a` and `c` go straight from HBM
//                        into LDS tiles [row][system] by global->LDS loads
//                        (one dword per lane: one FP32 element), b and d
//                        into registers.
// More than one refinement keeps x in FP64 and d through the loop (the
// general loop of mixed_core.hpp): only the shapes whose general kernel fits
// the register file without spilling offer it (general_ok).
// Systems none of them holds (above 8192 rows) are promoted to FP64 (mixed32_promote_kernel) and solved by the
// FP64 batched solve: the FP64 solution of the same promoted system, at 24 B
// of workspace and about 88 B of traffic per point.

#include <algorithm>
#include <atomic>

#include "../../include/tridsolver_hip.h"
#include "hybrid_core.hpp"
#include "mixed_core.hpp"
#include "trid_internal.h"

namespace trid {
namespace {

constexpr int M32_THREADS = 512;
constexpr size_t M32_LDS_MAX = 160 * 1024;

template <int M, bool CONTIG, bool ONE>
__global__ __launch_bounds__(M32_THREADS) void mixed32_kernel(
    Geom g, int P, int W, int G, int refine, const float *__restrict__ a,
    const float *__restrict__ b, const float *__restrict__ c, const float *__restrict__ d,
    double *__restrict__ x) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<CONTIG>(g, P, W, G);
  // contiguous: one P*(M+1)-double tile per system; the FP32 operands use
  // the first half of it, X (FP64) all of it
  const size_t tile_bytes = CONTIG ? sizeof(double) * (size_t)W * P * (M + 1) : 0;
  double *dtile = reinterpret_cast<double *>(smem_raw) + (size_t)L.w * P * (M + 1);
  float *ftile = reinterpret_cast<float *>(dtile);
  hyb::Ring ring{smem_raw + tile_bytes, (int)(2 * sizeof(double) * L.T), 0};
  float A[M], B[M], Cc[M], D[M];
  hyb::fetch_abcd<M, CONTIG, true>(L, a, b, c, d, A, B, Cc, D);
  hyb::to_chunk<M, CONTIG>(L, A, ftile);
  hyb::to_chunk<M, CONTIG>(L, B, ftile);
  hyb::to_chunk<M, CONTIG>(L, Cc, ftile);
  hyb::to_chunk<M, CONTIG>(L, D, ftile);
  double X[M];
  mixed_chunk_acc<M, MIXED_MAX_STEPS, false, ONE, true>(
      L, ring, [&](int r) { return A[r]; }, B, [&](int r) { return Cc[r]; }, D, refine, X);
  hyb::to_fetch<M, CONTIG>(L, X, dtile);
  hyb::store<M, CONTIG, true>(L, x, X);
}

// a and c kept in LDS, read from there by the solve, which frees 2*M
// registers a lane against mixed32_kernel. Contiguous systems: b and d go to
// chunk order through the system's tile, then a and c are written into it
// (skewed, the two FP32 halves of the P*(M+1)-double tile); x's transpose
// reuses the whole tile once every lane has read a and c. Strided systems:
// the lane's own a and c slots, [row][thread] (no exchange between lanes).
// STEPS: PCR steps, P <= 2^STEPS.
template <int M, int STEPS, bool ONE, bool CONTIG>
__global__ __launch_bounds__(M32_THREADS) void mixed32_lds_kernel(
    Geom g, int P, int W, int G, int refine, const float *__restrict__ a,
    const float *__restrict__ b, const float *__restrict__ c, const float *__restrict__ d,
    double *__restrict__ x) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<CONTIG>(g, P, W, G);
  if constexpr (!CONTIG) {
    const int T = L.T, t = L.t;
    float *tA = reinterpret_cast<float *>(smem_raw), *tC = tA + (size_t)M * T;
    hyb::Ring ring{smem_raw + 2 * sizeof(float) * (size_t)M * T, (int)(2 * sizeof(double) * T),
                   0};
    float A[M], B[M], Cc[M], D[M];
    hyb::fetch_abcd<M, false, true>(L, a, b, c, d, A, B, Cc, D);
#pragma unroll
    for (int r = 0; r < M; ++r) {
      tA[r * T + t] = A[r];
      tC[r * T + t] = Cc[r];
    }
    // the solve reads a and c back from LDS; without this the compiler keeps
    // the stored registers live instead
    asm volatile("" ::: "memory");
    double X[M];
    mixed_chunk_acc<M, STEPS, false, ONE, true>(
        L, ring, [&](int r) { return tA[r * T + t]; }, B, [&](int r) { return tC[r * T + t]; },
        D, refine, X);
    hyb::store<M, false, true>(L, x, X);
    return;
  }
  const int tile = P * (M + 1);
  double *dtile = reinterpret_cast<double *>(smem_raw) + (size_t)L.w * tile;
  float *tA = reinterpret_cast<float *>(dtile), *tC = tA + tile;
  hyb::Ring ring{smem_raw + sizeof(double) * (size_t)L.T * (M + 1),
                 (int)(2 * sizeof(double) * L.T), 0};
  {
    float A[M], B[M], Cc[M], D[M];
    hyb::fetch_abcd<M, true, true>(L, a, b, c, d, A, B, Cc, D);
    hyb::to_chunk<M, true>(L, B, tA);
    hyb::to_chunk<M, true>(L, D, tA);
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const int e = hyb::skew<M>(L.p + L.P * r);
      tA[e] = A[r];
      tC[e] = Cc[r];
    }
    __syncthreads();
    const float *ca = tA + L.p * (M + 1), *cc = tC + L.p * (M + 1);
    double X[M];
    mixed_chunk_acc<M, STEPS, false, ONE, true>(
        L, ring, [&](int r) { return ca[r]; }, B, [&](int r) { return cc[r]; }, D, refine, X);
    __syncthreads();  // a and c are read to the end before x overwrites the tile
    hyb::to_fetch<M, true>(L, X, dtile);
    hyb::store<M, true, true>(L, x, X);
  }
}

typedef __attribute__((address_space(3))) void lds_void;

// Threads per workgroup of the wide kernel with W FP32 x-neighbours (W = 32:
// 128-byte rows, W = 16: 64-byte rows): M = 32 rows per lane at 16 lanes
// per system (two waves per SIMD), M = 16 at 32 lanes. Built: (M, W) =
// (32, 32), (32, 16) and (16, 16).
template <int M, int W>
struct M32Wide {
  static constexpr int threads = (M >= 32 ? 16 : 32) * W;
  static constexpr int steps = M >= 32 ? 4 : 5;  // PCR steps: P <= threads / W
};

template <int M, int W, bool ONE>
__global__ __launch_bounds__((M32Wide<M, W>::threads)) void mixed32_wide_kernel(
    Geom g, int P, int G, int refine, const float *__restrict__ a, const float *__restrict__ b,
    const float *__restrict__ c, const float *__restrict__ d, double *__restrict__ x) {
  constexpr int SEG = W * (int)sizeof(float);
  constexpr int K = 256 / SEG;  // rows per wave-wide global->LDS load
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<false>(g, P, W, G);
  const int T = L.T;
  const int64_t R = (int64_t)P * M;  // rows per tile (>= N)
  float *tA = reinterpret_cast<float *>(smem_raw);
  float *tC = tA + (size_t)R * W;
  hyb::Ring ring{smem_raw + 2 * sizeof(float) * (size_t)R * W, (int)(2 * sizeof(double) * T),
                 0};

  // (0) a rows 1..N-1 and c rows 0..N-2 -> tiles [row][system]. One wave
  // instruction writes 256 contiguous LDS bytes, lane l its bytes 4l..4l+3:
  // K rows of SEG bytes; lane l serves row (first row + 4l / SEG) and system
  // (4l % SEG) / 4. Systems beyond n0 are masked off; the slots no load
  // writes (a row 0, c row N-1, rows >= N, systems beyond n0) are cleared by
  // their own lanes below.
  {
    const int lane = L.t & 63, wv = L.t >> 6, nw = T >> 6;
    const int sub = (lane * 4) / SEG, byte = (lane * 4) % SEG;
    const int64_t first = L.i0 - L.w;
    const bool sys_ok = first + byte / (int)sizeof(float) < g.n0;
    const int64_t row_bytes = L.stride * (int64_t)sizeof(float);
    const int64_t o = (first + L.i1 * g.st1) * (int64_t)sizeof(float) + byte + sub * row_bytes;
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

  // b, d of the lane's rows into registers (identity rows past N). Element
  // offsets are 32-bit (the launch checks that x's extent is below 4 GiB,
  // so the FP32 arrays' is below 2 GiB).
  const int64_t row0 = (int64_t)L.p * M;
  const uint32_t oe = (uint32_t)(L.base + row0 * L.stride);
  const uint32_t se = (uint32_t)L.stride;
  const bool full = L.active && row0 + M <= L.N;
  float B[M], D[M];
  if (full) {
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const uint32_t e = oe + (uint32_t)r * se;
      B[r] = b[e];
      D[r] = d[e];
    }
  } else {
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const bool ok = L.active && row0 + r < L.N;
      const uint32_t e = oe + (uint32_t)r * se;
      B[r] = ok ? b[e] : 1.0f;
      D[r] = ok ? d[e] : 0.0f;
    }
  }
  float *sa = tA + row0 * W + L.w;  // sa[r * W]: this lane's row r
  float *sc = tC + row0 * W + L.w;
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int64_t row = row0 + r;
    if (!L.active || row == 0 || row >= L.N) sa[r * W] = 0.0f;
    if (!L.active || row >= L.N - 1) sc[r * W] = 0.0f;
  }
  // the tiles are written by other waves' loads: wait for this wave's, then
  // for every wave
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  double X[M];
  mixed_chunk_acc<M, M32Wide<M, W>::steps, false, ONE, true>(
      L, ring, [&](int r) { return sa[r * W]; }, B, [&](int r) { return sc[r * W]; }, D, refine,
      X);
  if (!L.active) return;
  // store offsets recomputed (the empty asm hides that they equal the load
  // offsets, which would otherwise stay live across the solve)
  uint32_t oes = oe;
  asm volatile("" : "+v"(oes));
  const int rows = full ? M : (int)std::min<int64_t>(M, std::max<int64_t>(0, L.N - row0));
#pragma unroll
  for (int r = 0; r < M; ++r)
    if (r < rows) x[oes + (uint32_t)r * se] = X[r];
}

size_t reg_lds_bytes(const Geom &g, const HybridCfg &c) {
  const size_t T = (size_t)c.W * c.P;
  size_t lds = hyb::ring_bytes<double>((int)T, 2);
  if (g.stride == 1) lds += sizeof(double) * T * (c.M + 1);
  return lds;
}

size_t wide_lds_bytes(const HybridCfg &c) {
  const size_t T = (size_t)c.W * c.P;
  return 2 * sizeof(float) * (size_t)c.P * c.M * c.W + hyb::ring_bytes<double>((int)T, 2);
}

bool wide_valid(const Geom &g, const HybridCfg &c) {
  if (g.stride == 1 || g.st0 != 1) return false;
  int maxT = 0, steps = 0;
  if (c.W == 32 && c.M == 32) { maxT = M32Wide<32, 32>::threads; steps = M32Wide<32, 32>::steps; }
  if (c.W == 16 && c.M == 32) { maxT = M32Wide<32, 16>::threads; steps = M32Wide<32, 16>::steps; }
  if (c.W == 16 && c.M == 16) { maxT = M32Wide<16, 16>::threads; steps = M32Wide<16, 16>::steps; }
  if (maxT == 0 || c.P < 1 || c.G < 1) return false;
  const int64_t T = (int64_t)c.W * c.P;
  if ((int64_t)c.P * c.M < g.N || c.P > (1 << steps) || T > maxT || T % 64 != 0) return false;
  // 32-bit element offsets, and x's byte extent below 4 GiB
  const int64_t nb0 = (g.n0 + c.W - 1) / c.W * c.W;
  const int64_t last = (nb0 - 1) + (g.n1 - 1) * g.st1 + ((int64_t)c.P * c.M - 1) * g.stride;
  if ((last + 1) * (int64_t)sizeof(double) > (int64_t)0xFFFFFFFFll) return false;
  return wide_lds_bytes(c) <= M32_LDS_MAX;
}

size_t lds_kernel_bytes(const Geom &g, const HybridCfg &c) {
  const size_t T = (size_t)c.W * c.P;
  const size_t tiles = g.stride == 1 ? sizeof(double) * T * (c.M + 1)  // per-system tiles
                                     : 2 * sizeof(float) * T * c.M;    // a, c [row][thread]
  return tiles + hyb::ring_bytes<double>((int)T, 2);
}

bool lds_valid(const Geom &g, const HybridCfg &c) {
  if (c.M != 8 && c.M != 16) return false;
  if (c.P < 1 || c.W < 1 || c.G < 1) return false;
  if ((int64_t)c.P * c.M < g.N || c.P > (1 << MIXED_MAX_STEPS)) return false;
  if ((int64_t)c.W * c.P > M32_THREADS) return false;
  return lds_kernel_bytes(g, c) <= M32_LDS_MAX;
}

bool reg_valid(const Geom &g, const HybridCfg &c) {
  if (c.M != 8 && c.M != 16) return false;
  if (c.P < 1 || c.W < 1 || c.G < 1) return false;
  if ((int64_t)c.P * c.M < g.N || c.P > (1 << MIXED_MAX_STEPS)) return false;
  if ((int64_t)c.W * c.P > M32_THREADS) return false;
  return reg_lds_bytes(g, c) <= M32_LDS_MAX;
}

template <int M, bool CONTIG, bool ONE>
hipError_t launch_reg_one(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                          const float *b, const float *c, const float *d, double *x,
                          hipStream_t st) {
  const int T = cfg.W * cfg.P;
  const size_t lds = reg_lds_bytes(g, cfg);
  auto kern = mixed32_kernel<M, CONTIG, ONE>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.W, cfg.G,
                     refine, a, b, c, d, x);
  return hipGetLastError();
}

// At most one refinement runs the kernels built for it (fewer registers).
template <int M, bool CONTIG>
hipError_t launch_reg(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                      const float *b, const float *c, const float *d, double *x,
                      hipStream_t st) {
  if (refine <= 1) return launch_reg_one<M, CONTIG, true>(g, cfg, refine, a, b, c, d, x, st);
  if constexpr (M == 8)
    return launch_reg_one<M, CONTIG, false>(g, cfg, refine, a, b, c, d, x, st);
  return hipErrorInvalidValue;  // not general_ok
}

template <int M, int STEPS, bool ONE, bool CONTIG>
hipError_t launch_lds_one(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                          const float *b, const float *c, const float *d, double *x,
                          hipStream_t st) {
  const int T = cfg.W * cfg.P;
  const size_t lds = lds_kernel_bytes(g, cfg);
  auto kern = mixed32_lds_kernel<M, STEPS, ONE, CONTIG>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.W, cfg.G,
                     refine, a, b, c, d, x);
  return hipGetLastError();
}

// Systems of at most 64 lanes take the 6-step PCR (fewer registers).
template <int M, bool CONTIG>
hipError_t launch_lds_c(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                        const float *b, const float *c, const float *d, double *x,
                        hipStream_t st) {
  constexpr int S9 = MIXED_MAX_STEPS;
  if (cfg.P <= 64)
    return refine <= 1 ? launch_lds_one<M, 6, true, CONTIG>(g, cfg, refine, a, b, c, d, x, st)
                       : launch_lds_one<M, 6, false, CONTIG>(g, cfg, refine, a, b, c, d, x, st);
  return refine <= 1 ? launch_lds_one<M, S9, true, CONTIG>(g, cfg, refine, a, b, c, d, x, st)
                     : launch_lds_one<M, S9, false, CONTIG>(g, cfg, refine, a, b, c, d, x, st);
}

template <int M>
hipError_t launch_lds(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                      const float *b, const float *c, const float *d, double *x,
                      hipStream_t st) {
  return g.stride == 1 ? launch_lds_c<M, true>(g, cfg, refine, a, b, c, d, x, st)
                       : launch_lds_c<M, false>(g, cfg, refine, a, b, c, d, x, st);
}

template <int M, int W, bool ONE>
hipError_t launch_wide32_one(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                             const float *b, const float *c, const float *d, double *x,
                             hipStream_t st) {
  const int T = cfg.W * cfg.P;
  const size_t lds = wide_lds_bytes(cfg);
  // plain loads and stores: with nontemporal ones z (512^3) took 0.90 ms
  // against 0.81
  auto kern = mixed32_wide_kernel<M, W, ONE>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.G, refine, a,
                     b, c, d, x);
  return hipGetLastError();
}

template <int M, int W>
hipError_t launch_wide32(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                         const float *b, const float *c, const float *d, double *x,
                         hipStream_t st) {
  if (refine <= 1) return launch_wide32_one<M, W, true>(g, cfg, refine, a, b, c, d, x, st);
  if constexpr (M == 16 && W == 16)
    return launch_wide32_one<M, W, false>(g, cfg, refine, a, b, c, d, x, st);
  return hipErrorInvalidValue;  // not general_ok
}

// Shapes with a kernel for more than one refinement: the others' general
// loop would spill (mixed32_kernel at M = 16, the wide kernel at M = 32) or
// run at one wave per SIMD.
bool general_ok(const HybridCfg &c) {
  if (c.V == 0) return c.M == 8;
  if (c.V == 1) return true;
  if (c.V == 3) return c.M == 16 && c.W == 16;
  return false;
}

}  // namespace

// V = 0: mixed32_kernel (M = 8 or 16, W * P <= 512); V = 1: the kernel
// with a and c in LDS (M = 8 or 16, W * P <= 512); V = 3: the wide
// strided kernel ((M, W) = (32, 32), (32, 16), (16, 16)). More than one
// refinement: general_ok shapes only.
bool mixed32_cfg_valid(const Geom &g, const HybridCfg &c, int refine) {
  if (refine > 1 && !general_ok(c)) return false;
  if (c.V == 3) return wide_valid(g, c);
  if (c.V == 1) return lds_valid(g, c);
  if (c.V != 0) return false;
  return reg_valid(g, c);
}

// Default shapes, from sweeps at 512^3 (times vary by a few percent from
// run to run):
// - contiguous systems: the kernel with a and c in LDS at 8 rows per lane
//   (16 past 4096 rows), workgroups of about one wave (W = 64 / P
//   systems): x 0.61-0.63 ms at one refinement (122 VGPRs, four waves per
//   SIMD; mixed32_kernel at 16 rows per lane 0.65-0.69, at 8 0.67-0.73),
//   0.74 at two (mixed32_kernel at 8 rows per lane 1.14).
// - strided systems, one refinement: the wide kernel at 32 rows per lane,
//   128-byte rows, workgroups dealt to the XCDs in 32 groups (y 0.65, z
//   0.80 ms), else 16 x-neighbours in registers. More: a and c in LDS, 8
//   rows per lane (16 past 4096 rows), 8 x-neighbours (y 0.89, z 1.21 ms
//   at two; the wide kernel at 16 rows per lane 0.99 / 1.28,
//   mixed32_kernel 1.58 / 2.03).
bool mixed32_default_cfg(const Geom &g, int refine, HybridCfg *cfg) {
  HybridCfg c;
  const bool contig = g.stride == 1;
  if (contig || refine > 1) {
    for (int M : {8, 16}) {
      c.M = M;
      c.V = 1;
      c.P = (int)((g.N + M - 1) / M);
      if (contig) {
        c.G = 1;
        c.W = (int)std::max<int64_t>(1, 64 / c.P);
      } else {
        c.G = 8;
        c.W = 8;
        while (c.W > 1 && (int64_t)c.W * c.P > M32_THREADS) c.W /= 2;
      }
      if (mixed32_cfg_valid(g, c, refine)) { *cfg = c; return true; }
    }
    return false;
  }
  c.V = 3;
  c.G = 32;
  c.M = 32;
  c.W = 32;
  c.P = (int)((g.N + c.M - 1) / c.M);
  c.P += c.P & 1;  // whole waves: W * P a multiple of 64
  if (mixed32_cfg_valid(g, c, refine)) { *cfg = c; return true; }
  c.V = 0;
  c.G = 8;
  for (int M : {g.N >= 256 ? 16 : 8, 8, 16}) {
    c.M = M;
    c.P = (int)((g.N + M - 1) / M);
    c.W = 16;
    while (c.W > 1 && (int64_t)c.W * c.P > M32_THREADS) c.W /= 2;
    if (mixed32_cfg_valid(g, c, refine)) { *cfg = c; return true; }
  }
  return false;
}

namespace {

template <bool CONTIG>
__global__ __launch_bounds__(256) void mixed32_promote_kernel(
    Geom g, const float *__restrict__ a, const float *__restrict__ b,
    const float *__restrict__ c, const float *__restrict__ d, double *__restrict__ A,
    double *__restrict__ B, double *__restrict__ C, double *__restrict__ x) {
  const int64_t total = g.nsys() * g.N;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    int64_t k;
    const int64_t i = hyb::batch_elem<CONTIG>(g, e, &k);
    A[i] = (double)a[i];
    B[i] = (double)b[i];
    C[i] = (double)c[i];
    x[i] = (double)d[i];
  }
}

}  // namespace

hipError_t launch_mixed32_promote(const Geom &g, const float *a, const float *b, const float *c,
                                  const float *d, double *A, double *B, double *C, double *x,
                                  hipStream_t st) {
  if (g.stride == 1)
    hipLaunchKernelGGL(mixed32_promote_kernel<true>, dim3(batch_elem_blocks(g)), dim3(256), 0,
                       st, g, a, b, c, d, A, B, C, x);
  else
    hipLaunchKernelGGL(mixed32_promote_kernel<false>, dim3(batch_elem_blocks(g)), dim3(256), 0,
                       st, g, a, b, c, d, A, B, C, x);
  return hipGetLastError();
}

hipError_t launch_mixed32(const Geom &g, const HybridCfg &cfg, int refine, const float *a,
                          const float *b, const float *c, const float *d, double *x,
                          hipStream_t st) {
  const bool contig = g.stride == 1;
  if (cfg.V == 3) {
    if (cfg.W == 32 && cfg.M == 32)
      return launch_wide32<32, 32>(g, cfg, refine, a, b, c, d, x, st);
    if (cfg.W == 16 && cfg.M == 32)
      return launch_wide32<32, 16>(g, cfg, refine, a, b, c, d, x, st);
    if (cfg.W == 16 && cfg.M == 16)
      return launch_wide32<16, 16>(g, cfg, refine, a, b, c, d, x, st);
    return hipErrorInvalidValue;
  }
  if (cfg.V == 1) {
    if (cfg.M == 8) return launch_lds<8>(g, cfg, refine, a, b, c, d, x, st);
    if (cfg.M == 16) return launch_lds<16>(g, cfg, refine, a, b, c, d, x, st);
    return hipErrorInvalidValue;
  }
  switch (cfg.M) {
    case 8: return contig ? launch_reg<8, true>(g, cfg, refine, a, b, c, d, x, st)
                          : launch_reg<8, false>(g, cfg, refine, a, b, c, d, x, st);
    case 16: return contig ? launch_reg<16, true>(g, cfg, refine, a, b, c, d, x, st)
                           : launch_reg<16, false>(g, cfg, refine, a, b, c, d, x, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace trid
