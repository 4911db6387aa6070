// kernels_stream.hip - strided Thomas-PCR hybrid with streamed loads, for wide
// row segments (256 bytes: 32 FP64 or 64 FP32 x-neighbouring systems per
// workgroup).
//
// The plain hybrid (kernels_solve.hip) issues all 4*M loads of a lane before
// its forward sweep, so it needs registers for a, b, c and d of every row;
// 32 FP64 systems of 512 rows would not fit the register file. Here the loads
// run LA rows ahead of the forward sweep (a ring of LA rows) and scheduling
// barriers stop the compiler from hoisting them all to the top, so the live
// set is the sweep's three values per row (al, ga, de) plus the ring. The
// arithmetic is exactly the hybrid's (hybrid_core.hpp phases (i)-(iv)); only
// the order of loads differs.
//
// Why wider segments: a workgroup reads each row of its systems as one
// contiguous segment of W * sizeof(REAL) bytes. For z solves every row lies in
// a different 2 MiB page and for y in a different DRAM row, so fewer, wider
// segments mean fewer translations and activations per byte (a 128-byte
// segment streams at ~4.0 TB/s for z, a 256-byte one at ~5.0 TB/s).

#include <algorithm>
#include <atomic>

#include "hybrid_core.hpp"
#include "trid_internal.h"

namespace trid {
namespace {

constexpr int STREAM_THREADS = 1024;

template <typename REAL, int M, int LA, bool INC>
__global__ __launch_bounds__(STREAM_THREADS) void hybrid_stream_kernel(
    Geom g, int P, int W, int G, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, REAL *__restrict__ d, REAL *__restrict__ u) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<false>(g, P, W, G);
  hyb::Ring ring{smem_raw, (int)(3 * sizeof(REAL) * L.T), 0};
  const int64_t row0 = (int64_t)L.p * M;

  auto load_row = [&](int r, REAL &ar, REAL &br, REAL &cr, REAL &dr) {
    const int64_t row = row0 + r;
    const bool ok = L.active && row < L.N;
    const int64_t o = L.base + row * L.stride;
    ar = (ok && row > 0) ? a[o] : REAL(0);
    br = ok ? b[o] : REAL(1);
    cr = (ok && row < L.N - 1) ? c[o] : REAL(0);
    dr = ok ? d[o] : REAL(0);
  };

  REAL al[M], ga[M], de[M];
  REAL qa[LA], qb[LA], qc[LA], qd[LA];
#pragma unroll
  for (int k = 0; k < LA; ++k) load_row(k, qa[k], qb[k], qc[k], qd[k]);
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int s = r % LA;
    const REAL ar = qa[s], br = qb[s], cr = qc[s], dr = qd[s];
    if (r + LA < M) load_row(r + LA, qa[s], qb[s], qc[s], qd[s]);
    // keep the next row's loads here, behind this row's arithmetic
    __builtin_amdgcn_sched_barrier(0);
    if (r == 0) {
      const REAL inv = REAL(1) / br;
      al[0] = ar * inv;
      ga[0] = cr * inv;
      de[0] = dr * inv;
    } else {
      const REAL inv = REAL(1) / (br - ar * ga[r - 1]);
      al[r] = -ar * al[r - 1] * inv;
      ga[r] = cr * inv;
      de[r] = (dr - ar * de[r - 1]) * inv;
    }
  }
  const REAL al_e = al[M - 1], ga_e = ga[M - 1], de_e = de[M - 1];
  REAL f_d, f_a, f_g;
  hyb::backward_spike<REAL, M>(al, ga, de, f_d, f_a, f_g);
  REAL y, ym;
  hyb::reduced_solve<REAL>(L, ring, al_e, ga_e, de_e, f_d, f_a, f_g, y, ym);
  if (!L.active) return;
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int64_t row = row0 + r;
    if (row < L.N) {
      const REAL x = (r < M - 1) ? de[r] - al[r] * ym - ga[r] * y : y;
      const int64_t o = L.base + row * L.stride;
      if (INC) u[o] += x;
      else d[o] = x;
    }
  }
}

// Lookahead rows: as deep as the register file allows at 1024 threads
// (FP64 M = 16 and FP32 M = 32 hold 96 VGPRs of sweep state).
template <typename REAL, int M>
constexpr int stream_lookahead() {
  return sizeof(REAL) * M >= 128 ? 2 : 4;
}

// Variant with one of the three per-row sweep values (al, then ap) in LDS,
// so that 32 FP64 (64 FP32) systems of up to 16 (32) rows per lane fit one
// workgroup: registers hold ga and de and the load ring, LDS holds
// al[r][thread] (M * T values) and a single-slot exchange buffer. Element
// addresses are 32-bit byte offsets from the array bases (uniform registers),
// which saves a 64-bit address pair per load; the caller's arrays must span
// less than 4 GiB (checked by stream_cfg_valid).
template <typename REAL, int M, int LA, bool INC>
__global__ __launch_bounds__(STREAM_THREADS) void hybrid_lds_kernel(
    Geom g, int P, int W, int G, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, REAL *__restrict__ d, REAL *__restrict__ u) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<false>(g, P, W, G);
  const int T = L.T, t = L.t;
  REAL *al_s = reinterpret_cast<REAL *>(smem_raw);
  hyb::Ring ring{smem_raw + sizeof(REAL) * (size_t)M * T, (int)(3 * sizeof(REAL) * T), 0, 1};
  const int64_t row0 = (int64_t)L.p * M;

  // 32-bit byte offsets from the uniform array bases (extent < 4 GiB)
  const uint32_t ob = (uint32_t)((L.base + row0 * L.stride) * (int64_t)sizeof(REAL));
  const uint32_t os = (uint32_t)(L.stride * (int64_t)sizeof(REAL));
  auto at = [](const REAL *p, uint32_t off) -> REAL {
    return *reinterpret_cast<const REAL *>(reinterpret_cast<const char *>(p) + off);
  };
  auto load_row = [&](int r, REAL &ar, REAL &br, REAL &cr, REAL &dr) {
    const int64_t row = row0 + r;
    const bool ok = L.active && row < L.N;
    const uint32_t o = ob + (uint32_t)r * os;
    ar = (ok && row > 0) ? at(a, o) : REAL(0);
    br = ok ? at(b, o) : REAL(1);
    cr = (ok && row < L.N - 1) ? at(c, o) : REAL(0);
    dr = ok ? at(d, o) : REAL(0);
  };

  REAL ga[M], de[M];
  REAL qa[LA], qb[LA], qc[LA], qd[LA];
  REAL al_prev = 0;
#pragma unroll
  for (int k = 0; k < LA; ++k) load_row(k, qa[k], qb[k], qc[k], qd[k]);
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int s = r % LA;
    const REAL ar = qa[s], br = qb[s], cr = qc[s], dr = qd[s];
    if (r + LA < M) load_row(r + LA, qa[s], qb[s], qc[s], qd[s]);
    __builtin_amdgcn_sched_barrier(0);
    REAL al;
    if (r == 0) {
      const REAL inv = REAL(1) / br;
      al = ar * inv;
      ga[0] = cr * inv;
      de[0] = dr * inv;
    } else {
      const REAL inv = REAL(1) / (br - ar * ga[r - 1]);
      al = -ar * al_prev * inv;
      ga[r] = cr * inv;
      de[r] = (dr - ar * de[r - 1]) * inv;
    }
    al_s[r * T + t] = al;
    al_prev = al;
  }
  const REAL al_e = al_prev, ga_e = ga[M - 1], de_e = de[M - 1];
  // backward sweep; ap replaces al in LDS, gp and dp replace ga and de
  REAL ap_next = M >= 2 ? al_s[(M - 2) * T + t] : REAL(0);
#pragma unroll
  for (int r = M - 3; r >= 0; --r) {
    const REAL gr = ga[r];
    de[r] = de[r] - gr * de[r + 1];
    const REAL ap = al_s[r * T + t] - gr * ap_next;
    al_s[r * T + t] = ap;
    ap_next = ap;
    ga[r] = -gr * ga[r + 1];
  }
  REAL f_d, f_a, f_g;
  if (M >= 2) {
    f_d = de[0]; f_a = ap_next; f_g = ga[0];
  } else {
    f_d = REAL(0); f_a = REAL(0); f_g = REAL(-1);
  }
  REAL y, ym;
  hyb::reduced_solve<REAL>(L, ring, al_e, ga_e, de_e, f_d, f_a, f_g, y, ym);
  if (!L.active) return;
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int64_t row = row0 + r;
    if (row < L.N) {
      const REAL x = (r < M - 1) ? de[r] - al_s[r * T + t] * ym - ga[r] * y : y;
      const uint32_t o = ob + (uint32_t)r * os;
      REAL *dst = reinterpret_cast<REAL *>(reinterpret_cast<char *>(INC ? u : d) + o);
      if (INC) *dst += x;
      else *dst = x;
    }
  }
}

template <typename REAL, int M, bool INC>
hipError_t launch_lds_m(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                        const REAL *c, REAL *d, REAL *u, hipStream_t st) {
  constexpr int LA = 4;
  const int T = cfg.W * cfg.P;
  const size_t lds = sizeof(REAL) * ((size_t)M * T + (size_t)3 * T);
  auto kern = hybrid_lds_kernel<REAL, M, LA, INC>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.W, cfg.G, a,
                     b, c, d, u);
  return hipGetLastError();
}

template <typename REAL, int M, bool INC>
hipError_t launch_stream_m(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                           const REAL *c, REAL *d, REAL *u, hipStream_t st) {
  constexpr int LA = stream_lookahead<REAL, M>();
  const int T = cfg.W * cfg.P;
  const size_t lds = hyb::ring_bytes<REAL>(T, 3);
  auto kern = hybrid_stream_kernel<REAL, M, LA, INC>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.W, cfg.G, a,
                     b, c, d, u);
  return hipGetLastError();
}

}  // namespace

// Rows per lane compiled for the streamed variants: 8 (FP64) and 16 (FP32).
// Twice that spills at their 1024 threads (128 VGPRs); the wide kernel
// (kernels_wide.hip) takes those shapes with `a` in LDS.
bool stream_m_supported(int fp64, int M) { return fp64 ? M == 8 : M == 16; }

template <typename REAL>
bool stream_cfg_valid(const Geom &g, const HybridCfg &c) {
  if (g.stride == 1) return false;  // strided systems only
  if (!stream_m_supported(sizeof(REAL) == 8, c.M)) return false;
  if (c.P < 1 || c.W < 1 || c.G < 1) return false;
  if ((int64_t)c.P * c.M < g.N) return false;
  const int64_t T = (int64_t)c.W * c.P;
  if (T > STREAM_THREADS) return false;
  const size_t lds = c.V == 2 ? sizeof(REAL) * ((size_t)c.M * T + (size_t)3 * T)
                              : hyb::ring_bytes<REAL>((int)T, 3);
  if (c.V == 2) {  // 32-bit byte offsets
    const int64_t nb0 = (g.n0 + c.W - 1) / c.W * c.W;  // inactive lanes address too
    const int64_t last =
        (nb0 - 1) * g.st0 + (g.n1 - 1) * g.st1 + ((int64_t)c.P * c.M - 1) * g.stride;
    if ((last + 1) * (int64_t)sizeof(REAL) > (int64_t)0xFFFFFFFFll) return false;
  }
  return lds <= 160 * 1024;
}

template <typename REAL, bool INC>
hipError_t launch_stream(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                         const REAL *c, REAL *d, REAL *u, hipStream_t st) {
  constexpr int M = sizeof(REAL) == 8 ? 8 : 16;
  if (cfg.M != M) return hipErrorInvalidValue;
  if (cfg.V == 2) return launch_lds_m<REAL, M, INC>(g, cfg, a, b, c, d, u, st);
  return launch_stream_m<REAL, M, INC>(g, cfg, a, b, c, d, u, st);
}

#define TRID_STREAM_INST(REAL)                                                               \
  template bool stream_cfg_valid<REAL>(const Geom &, const HybridCfg &);                      \
  template hipError_t launch_stream<REAL, false>(const Geom &, const HybridCfg &, const REAL *, \
                                                 const REAL *, const REAL *, REAL *, REAL *,   \
                                                 hipStream_t);                                \
  template hipError_t launch_stream<REAL, true>(const Geom &, const HybridCfg &, const REAL *,  \
                                                const REAL *, const REAL *, REAL *, REAL *,    \
                                                hipStream_t);
TRID_STREAM_INST(float)
TRID_STREAM_INST(double)

}  // namespace trid
