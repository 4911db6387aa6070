// kernels_wide.hip - strided Thomas-PCR hybrid with 256-byte row segments.
//
// A workgroup solves W = 256 / sizeof(REAL) x-neighbouring systems (32 FP64,
// 64 FP32), so each row it reads is one contiguous 256-byte segment. For z
// solves every row lies in its own 2 MiB page; twice the bytes per segment
// halves the address translations per byte moved, which is what limits the
// 128-byte (16 FP64 systems) plain kernel.
//
// The plain kernel keeps a, b, c and d of all its rows in registers; at 32
// FP64 systems of 512 rows that is twice the register file. Here the `a`
// operand goes straight from HBM into an LDS tile by the gfx950 global->LDS
// loads (global_load_lds_dword: one wave instruction moves one 256-byte row),
// and the forward and backward sweeps keep their first coefficient (al, then
// ap) in that tile. b, c and d are loaded into registers up front as in the
// plain kernel, so every load of the workgroup is in flight at once.
//
// Arithmetic is the hybrid's (hybrid_core.hpp phases (i)-(iv)); only where
// one operand lives differs.
//
// LDS: tile P*M rows x SEG bytes, plus the PCR ring (3 values per thread per
// slot, two slots when they fit). 512 rows of 256 bytes: 128 KiB + 24 KiB.

#include <algorithm>
#include <atomic>

#include "hybrid_core.hpp"
#include "trid_internal.h"

namespace trid {
namespace {

// SEG: bytes of one row segment (W * sizeof(REAL)), 128 or 256.
template <typename REAL, int SEG>
struct WideShape {
  static constexpr int W = SEG / (int)sizeof(REAL);
};
// Threads per workgroup the register file allows: FP64 M = 32 holds 3 x 64
// VGPRs of b, c, d (512 threads); everything else fits 1024.
template <typename REAL, int M>
struct WideBounds {
  static constexpr int threads = (sizeof(REAL) == 8 && M == 32) ? 512 : 1024;
};
constexpr size_t WIDE_LDS = 160 * 1024;

// PCR exchange slots: one (two barriers per step) for 128-byte rows, whose
// workgroups are small enough for two per CU only with half the ring; for
// 256-byte rows two (one barrier per step) when the LDS allows.
template <typename REAL>
size_t wide_lds_bytes(int M, int P, int seg, int slots) {
  const int T = seg / (int)sizeof(REAL) * P;
  return (size_t)P * M * seg + (size_t)slots * 3 * sizeof(REAL) * (size_t)T;
}
template <typename REAL>
int wide_slots(int M, int P, int seg) {
  if (seg == 128) return 1;
  return wide_lds_bytes<REAL>(M, P, seg, 2) <= WIDE_LDS ? 2 : 1;
}

typedef __attribute__((address_space(3))) void lds_void;

template <typename REAL, int M, int SEG, bool INC>
__global__ __launch_bounds__((WideBounds<REAL, M>::threads)) void hybrid_wide_kernel(
    Geom g, int P, int G, int slots, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, REAL *__restrict__ d, REAL *__restrict__ u) {
  constexpr int W = WideShape<REAL, SEG>::W;
  constexpr int K = 256 / SEG;  // rows per wave-wide global->LDS load
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<false>(g, P, W, G);
  const int T = L.T;
  const int64_t R = (int64_t)P * M;  // rows in the tile (>= N)
  REAL *tile = reinterpret_cast<REAL *>(smem_raw);
  hyb::Ring ring{smem_raw + sizeof(REAL) * (size_t)R * W, (int)(3 * sizeof(REAL) * T), 0,
                 slots};

  // (0) a -> tile[row][0..W) by global->LDS loads. One wave instruction
  // writes 256 contiguous LDS bytes (lane l: bytes 4l..4l+3), i.e. K rows:
  // lane l moves byte 4l % SEG of row (first row + 4l / SEG), which belongs
  // to system first + (4l % SEG) / sizeof(REAL). Wave v moves rows K*v,
  // K*(v + T/64), ...; lanes whose row is >= N or whose system is beyond n0
  // are masked off. Systems are consecutive in memory (st0 == 1).
  {
    const int lane = L.t & 63, wv = L.t >> 6, nw = T >> 6;
    const int sub = (lane * 4) / SEG, byte = (lane * 4) % SEG;
    const int64_t first = L.i0 - L.w;  // x index of the workgroup's system 0
    const bool sys_ok = first + byte / (int)sizeof(REAL) < g.n0;
    const int64_t row_bytes = L.stride * (int64_t)sizeof(REAL);
    const char *src = reinterpret_cast<const char *>(a + first + L.i1 * g.st1) + byte +
                      sub * row_bytes;
    if (sys_ok) {
      for (int64_t row = (int64_t)K * wv; row < L.N; row += (int64_t)K * nw)
        if (row + sub < L.N)
          __builtin_amdgcn_global_load_lds(src + row * row_bytes,
                                           (lds_void *)(tile + row * W), 4, 0, 0);
    }
  }

  // b, c, d of the lane's rows into registers (identity rows past N; c of
  // row N-1 is 0). Addresses are 32-bit byte offsets from the uniform array
  // bases (extent < 4 GiB, checked by wide_cfg_valid), so no 64-bit address
  // pair per row stays live until the stores.
  const int64_t row0 = (int64_t)L.p * M;
  const uint32_t ob = (uint32_t)((L.base + row0 * L.stride) * (int64_t)sizeof(REAL));
  const uint32_t os = (uint32_t)(L.stride * (int64_t)sizeof(REAL));
  auto at = [](const REAL *p, uint32_t off) -> REAL {
    return hyb::ld<true>(reinterpret_cast<const REAL *>(reinterpret_cast<const char *>(p) + off));
  };
  // Lanes whose M rows all lie inside the system (every lane when N = P*M)
  // load without per-row masks.
  const bool full = L.active && row0 + M <= L.N;
  REAL rb[M], rc[M], rd[M];
  if (full) {
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const uint32_t o = ob + (uint32_t)r * os;
      rb[r] = at(b, o);
      rc[r] = at(c, o);
      rd[r] = at(d, o);
    }
    if (row0 + M == L.N) rc[M - 1] = REAL(0);
  } else {
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const int64_t row = row0 + r;
      const bool ok = L.active && row < L.N;
      const uint32_t o = ob + (uint32_t)r * os;
      rb[r] = ok ? at(b, o) : REAL(1);
      rc[r] = (ok && row < L.N - 1) ? at(c, o) : REAL(0);
      rd[r] = ok ? at(d, o) : REAL(0);
    }
  }
  REAL *al_s = tile + row0 * W + L.w;  // al_s[r * W] = this lane's row r
  // a of identity rows (past N, or systems past n0) is 0: the lane clears
  // its own tile slots, which no global->LDS load writes
  if (!full) {
#pragma unroll
    for (int r = 0; r < M; ++r)
      if (!L.active || row0 + r >= L.N) al_s[r * W] = REAL(0);
  }
  // the tile is written by other waves' loads: wait for this wave's, then
  // for every wave
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  // (i)-(iii) with al, then ap, in the tile
  REAL y, ym;
  hyb::sweep_lds_a<REAL, M, W>(L, ring, al_s, row0 == 0, rb, rc, rd, y, ym);
  if (!L.active) return;
  // (iv) substitution. The store offsets are recomputed here (the empty asm
  // hides that they equal the load offsets, which would otherwise be kept
  // live in registers across the whole kernel).
  uint32_t obs = ob;
  asm volatile("" : "+v"(obs));
  const int rows = full ? M : (int)std::min<int64_t>(M, std::max<int64_t>(0, L.N - row0));
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < rows) {
      const REAL x = (r < M - 1) ? rd[r] - al_s[r * W] * ym - rc[r] * y : y;
      REAL *dst = reinterpret_cast<REAL *>(reinterpret_cast<char *>(INC ? u : d) + obs +
                                           (uint32_t)r * os);
      if (INC) *dst += x;
      else hyb::st<true>(dst, x);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <typename REAL, int M, int SEG, bool INC>
hipError_t launch_wide_m(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                         const REAL *c, REAL *d, REAL *u, hipStream_t st) {
  constexpr int W = WideShape<REAL, SEG>::W;
  const int T = W * cfg.P;
  const int slots = wide_slots<REAL>(M, cfg.P, SEG);
  const size_t lds = wide_lds_bytes<REAL>(M, cfg.P, SEG, slots);
  auto kern = hybrid_wide_kernel<REAL, M, SEG, INC>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + W - 1) / W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.G, slots, a,
                     b, c, d, u);
  return hipGetLastError();
}

}  // namespace

// Rows per lane compiled for the wide variant.
bool wide_m_supported(int M) { return M == 16 || M == 32; }

template <typename REAL>
bool wide_cfg_valid(const Geom &g, const HybridCfg &c) {
  const int seg = c.W * (int)sizeof(REAL);
  if (g.stride == 1 || g.st0 != 1) return false;  // x-neighbouring strided systems
  if (!wide_m_supported(c.M) || (seg != 128 && seg != 256) || c.P < 1 || c.G < 1)
    return false;
  if ((int64_t)c.P * c.M < g.N) return false;
  const int W = c.W;
  const int64_t T = (int64_t)W * c.P;
  const int maxT = c.M == 32 ? WideBounds<REAL, 32>::threads : WideBounds<REAL, 16>::threads;
  if (T > maxT || T % 64 != 0) return false;  // whole waves
  // 32-bit byte offsets: every address a lane forms lies below 4 GiB
  const int64_t nb0 = (g.n0 + W - 1) / W * W;
  const int64_t last = (nb0 - 1) + (g.n1 - 1) * g.st1 + ((int64_t)c.P * c.M - 1) * g.stride;
  if ((last + 1) * (int64_t)sizeof(REAL) > (int64_t)0xFFFFFFFFll) return false;
  return wide_lds_bytes<REAL>(c.M, c.P, seg, 1) <= WIDE_LDS;
}

template <typename REAL, bool INC>
hipError_t launch_wide(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                       const REAL *c, REAL *d, REAL *u, hipStream_t st) {
  const bool s256 = cfg.W * (int)sizeof(REAL) == 256;
  if (cfg.M == 16)
    return s256 ? launch_wide_m<REAL, 16, 256, INC>(g, cfg, a, b, c, d, u, st)
                : launch_wide_m<REAL, 16, 128, INC>(g, cfg, a, b, c, d, u, st);
  if (cfg.M == 32)
    return s256 ? launch_wide_m<REAL, 32, 256, INC>(g, cfg, a, b, c, d, u, st)
                : launch_wide_m<REAL, 32, 128, INC>(g, cfg, a, b, c, d, u, st);
  return hipErrorInvalidValue;
}

#define TRID_WIDE_INST(REAL)                                                                  \
  template bool wide_cfg_valid<REAL>(const Geom &, const HybridCfg &);                        \
  template hipError_t launch_wide<REAL, false>(const Geom &, const HybridCfg &, const REAL *,  \
                                               const REAL *, const REAL *, REAL *, REAL *,     \
                                               hipStream_t);                                  \
  template hipError_t launch_wide<REAL, true>(const Geom &, const HybridCfg &, const REAL *,   \
                                              const REAL *, const REAL *, REAL *, REAL *,      \
                                              hipStream_t);
TRID_WIDE_INST(float)
TRID_WIDE_INST(double)

}  // namespace trid
