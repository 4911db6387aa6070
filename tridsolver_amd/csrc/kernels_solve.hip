// kernels_solve.hip - gfx950 kernels of the batched scalar tridiagonal solve.
//
// Two algorithms:
//
// 1. lane_kernel: one Thomas recurrence per lane (wave64 lanes = 64 systems).
//    For strided systems (solvedim 1/2) neighbouring lanes own x-neighbouring
//    systems, so every row load of a wave is one contiguous 64*sizeof(REAL)
//    segment. The forward coefficients c', d' go to a workspace laid out
//    [row][system] (coalesced), d' in place in d for the plain solve. Any
//    length. Arithmetic order of the reference CPU Thomas
//    (scalar/src/cpu/trid_cpu.cpp:312-353): bb = 1/b0, cc = bb*c0, dd = bb*d0;
//    bb = 1/(b - a*cc), cc = bb*c, dd = bb*(d - a*dd); x = d' - c'*x.
//    This is the reference's trid_strided_multidim / trid_linear role
//    (scalar/src/cuda/trid_strided_multidim.hpp:43-189, trid_linear.hpp:40-109).
//
// 2. hybrid_kernel: Thomas-PCR hybrid with the whole system on chip, one HBM
//    read of a,b,c,d and one write of the result. The reference's
//    batchedTrid_{contig,strided}_ker (scalar/src/cuda/trid_thomaspcr.hpp:121-235,
//    308-504) has the same three phases; the reduction here is a different
//    one (one reduced unknown per chunk instead of two) and the reduced system
//    is solved in LDS by every lane of the workgroup at once.
//
//    A system of length N is cut into P chunks of M rows; lane (w, p) owns rows
//    s = p*M .. e = s+M-1 of system w in registers (rows >= N are identity rows
//    a=0, b=1, c=0, d=0; a[0] and c[N-1] are forced to 0).
//    (i) forward sweep with a left spike: every row i of the chunk becomes
//          al_i x_{s-1} + x_i + ga_i x_{i+1} = de_i
//        with inv = 1/(b_i - a_i ga_{i-1}), al_i = -a_i al_{i-1} inv,
//        ga_i = c_i inv, de_i = (d_i - a_i de_{i-1}) inv (row s: divide by b_s);
//    (ii) backward sweep expressing rows s..e-1 through x_{s-1} and x_e:
//          x_i = dp_i - ap_i x_{s-1} - gp_i x_e,
//        dp_i = de_i - ga_i dp_{i+1}, ap_i = al_i - ga_i ap_{i+1},
//        gp_i = -ga_i gp_{i+1}, starting from (dp, ap, gp)_e = (0, 0, -1);
//    (iii) substituting chunk k+1's first row into chunk k's last row gives one
//        equation per chunk in y_k = x_{e_k}:
//          al_e y_{k-1} + (1 - ga_e ap'_s) y_k - ga_e gp'_s y_{k+1}
//                                                   = de_e - ga_e dp'_s
//        (primes: chunk k+1's first row), a P-row tridiagonal system solved
//        by parallel cyclic reduction in LDS, ceil(log2 P) steps;
//    (iv) x_i = dp_i - ap_i y_{k-1} - gp_i y_k inside the chunk, x_e = y_k.
//    Contiguous systems (solvedim 0, ragged) are loaded coalesced (lane p reads
//    element p + P*r) and transposed through a skewed LDS tile so that lane p
//    owns rows p*M..p*M+M-1; the result goes back through the same tile.
//    Strided systems are loaded directly: W x-neighbouring systems per
//    workgroup, W consecutive lanes read W consecutive addresses.
//    kernels_stream.hip and kernels_wide.hip hold strided variants that keep
//    fewer values in registers (the latter, with a staged in LDS by
//    global->LDS loads, is the default for strided systems).
//
// Division is IEEE (v_div_* sequence) as in the reference CPU code.

#include <algorithm>
#include <atomic>
#include <cstdint>

#include "hybrid_core.hpp"
#include "trid_internal.h"

namespace trid {
namespace {

// ---------------------------------------------------------------------------
// 1. Per-lane Thomas
// ---------------------------------------------------------------------------
constexpr int LANE_BLOCK = 256;
constexpr int LANE_K = 8;  // rows fetched per batch of loads

template <typename REAL, bool INC>
__global__ __launch_bounds__(LANE_BLOCK) void lane_kernel(
    Geom g, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, REAL *__restrict__ d, REAL *__restrict__ u,
    REAL *__restrict__ cws, REAL *__restrict__ dws) {
  const int64_t ns = g.nsys();
  const int64_t s = (int64_t)blockIdx.x * LANE_BLOCK + threadIdx.x;
  if (s >= ns) return;
  const int64_t base = (s % g.n0) * g.st0 + (s / g.n0) * g.st1;
  const int64_t N = g.N, str = g.stride;
  a += base; b += base; c += base; d += base;
  if (INC) u += base;
  // d' lives in d for the plain solve (d is overwritten anyway) and in the
  // second workspace for Inc, where d must stay untouched.
  auto dprime = [&](int64_t r) -> REAL & {
    return INC ? dws[r * ns + s] : d[r * str];
  };

  REAL bb = REAL(1) / b[0];
  REAL cc = bb * c[0];
  REAL dd = bb * d[0];
  cws[s] = cc;
  dprime(0) = dd;
  int64_t r = 1;
  for (; r + LANE_K <= N; r += LANE_K) {
    REAL av[LANE_K], bv[LANE_K], cv[LANE_K], dv[LANE_K];
#pragma unroll
    for (int k = 0; k < LANE_K; ++k) {
      const int64_t o = (r + k) * str;
      av[k] = a[o]; bv[k] = b[o]; cv[k] = c[o]; dv[k] = d[o];
    }
#pragma unroll
    for (int k = 0; k < LANE_K; ++k) {
      bb = REAL(1) / (bv[k] - av[k] * cc);
      cc = bb * cv[k];
      dd = bb * (dv[k] - av[k] * dd);
      cws[(r + k) * ns + s] = cc;
      dprime(r + k) = dd;
    }
  }
  for (; r < N; ++r) {
    const int64_t o = r * str;
    const REAL aa = a[o];
    bb = REAL(1) / (b[o] - aa * cc);
    cc = bb * c[o];
    dd = bb * (d[o] - aa * dd);
    cws[r * ns + s] = cc;
    dprime(r) = dd;
  }

  // backward: x_{N-1} = d'_{N-1}; x_r = d'_r - c'_r x_{r+1}
  auto put = [&](int64_t row, REAL x) {
    if (INC) u[row * str] += x;
    else d[row * str] = x;
  };
  put(N - 1, dd);
  r = N - 2;
  for (; r - LANE_K + 1 >= 0; r -= LANE_K) {
    REAL cv[LANE_K], dv[LANE_K];
#pragma unroll
    for (int k = 0; k < LANE_K; ++k) {
      cv[k] = cws[(r - k) * ns + s];
      dv[k] = dprime(r - k);
    }
#pragma unroll
    for (int k = 0; k < LANE_K; ++k) {
      dd = dv[k] - cv[k] * dd;
      put(r - k, dd);
    }
  }
  for (; r >= 0; --r) {
    dd = dprime(r) - cws[r * ns + s] * dd;
    put(r, dd);
  }
}

// ---------------------------------------------------------------------------
// 2. Thomas-PCR hybrid
// ---------------------------------------------------------------------------
template <typename REAL, int M>
struct HybridBounds {
  static constexpr int threads = sizeof(REAL) * M > 64 ? 512 : 1024;
};
template <typename REAL, int M>
constexpr int hybrid_max_threads() {
  return HybridBounds<REAL, M>::threads;
}

template <typename REAL, int M, bool INC, bool CONTIG, bool VEC>
__global__ __launch_bounds__((HybridBounds<REAL, M>::threads)) void hybrid_kernel(
    Geom g, int P, int W, int G, const REAL *__restrict__ a,
    const REAL *__restrict__ b, const REAL *__restrict__ c, REAL *__restrict__ d,
    REAL *__restrict__ u) {
  constexpr int V = 16 / sizeof(REAL);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<CONTIG>(g, P, W, G);
  const size_t tile_bytes = CONTIG ? sizeof(REAL) * (size_t)W * P * (M + 1) : 0;
  REAL *tile = reinterpret_cast<REAL *>(smem_raw) + (size_t)L.w * P * (M + 1);
  hyb::Ring ring{smem_raw + tile_bytes, (int)(3 * sizeof(REAL) * L.T), 0};

  REAL ra[M], rb[M], rc[M], rd[M];
  if constexpr (VEC) {
    hyb::vfetch_abcd<M, V, true>(L, a, b, c, d, ra, rb, rc, rd);
    hyb::vto_chunk<M, V>(L, ra, tile);
    hyb::vto_chunk<M, V>(L, rb, tile);
    hyb::vto_chunk<M, V>(L, rc, tile);
    hyb::vto_chunk<M, V>(L, rd, tile);
  } else {
    hyb::fetch_abcd<M, CONTIG, true>(L, a, b, c, d, ra, rb, rc, rd);
    hyb::to_chunk<M, CONTIG>(L, ra, tile);
    hyb::to_chunk<M, CONTIG>(L, rb, tile);
    hyb::to_chunk<M, CONTIG>(L, rc, tile);
    hyb::to_chunk<M, CONTIG>(L, rd, tile);
  }
  REAL x[M];
  hyb::solve_chunk<REAL, M, CONTIG>(L, ring, ra, rb, rc, rd, x);
  if constexpr (VEC) {
    hyb::vto_fetch<M, V>(L, x, tile);
    hyb::vstore<M, V, INC, true>(L, INC ? u : d, x);
  } else {
    hyb::to_fetch<M, CONTIG>(L, x, tile);
    if (INC) hyb::store_add<M, CONTIG>(L, u, x);
    else hyb::store<M, CONTIG, true>(L, d, x);
  }
}

template <typename REAL, int M, bool INC, bool CONTIG, bool VEC>
hipError_t launch_hybrid_m(const Geom &g, const HybridCfg &cfg, const REAL *a,
                           const REAL *b, const REAL *c, REAL *d, REAL *u,
                           hipStream_t st) {
  const int T = cfg.W * cfg.P;
  size_t lds = sizeof(REAL) * (size_t)6 * T;
  if (CONTIG) lds += sizeof(REAL) * (size_t)cfg.W * cfg.P * (M + 1);
  auto kern = hybrid_kernel<REAL, M, INC, CONTIG, VEC>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t nb0 = (g.n0 + cfg.W - 1) / cfg.W;
  const int64_t blocks = nb0 * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P,
                     cfg.W, cfg.G, a, b, c, d, u);
  return hipGetLastError();
}

// Contiguous systems of 4097-8192 rows, one per 256-thread workgroup, and of
// 8193-16384 rows, one per 512-thread workgroup (variant 4, P = T):
// hyb::long32_solve, 32 rows per lane with `a` moved HBM->LDS in chunk order;
// at FP64 two waves per SIMD (two 256-thread workgroups per CU, or one of
// 512). The workgroups are dealt to the XCDs in G groups as the other
// kernels' are.
template <typename REAL, bool INC, int T>
__global__ __launch_bounds__(T, 512 / T) void hybrid_long32_kernel(
    Geom g, int G, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, REAL *__restrict__ d, REAL *__restrict__ u) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const uint32_t n0 = (uint32_t)g.n0;
  const uint32_t sys = hyb::xcd_remap(blockIdx.x, gridDim.x, G);
  const int64_t base = (int64_t)(sys % n0) * g.st0 + (int64_t)(sys / n0) * g.st1;
  hyb::long32_solve<REAL, INC, T>(hyb::wave_uniform(base), (int)g.N, smem_raw, a, b, c, d, u,
                                  [] {});
}

template <typename REAL, bool INC, int T>
hipError_t launch_long32_t(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                           const REAL *c, REAL *d, REAL *u, hipStream_t st) {
  auto kern = hybrid_long32_kernel<REAL, INC, T>;
  static LdsGrant lds_granted;
  constexpr size_t lds = hyb::long32_lds_bytes<REAL, T>();
  hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((unsigned)(g.n0 * g.n1)), dim3(T), lds, st, g, cfg.G, a, b, c,
                     d, u);
  return hipGetLastError();
}

template <typename REAL, bool INC>
hipError_t launch_long32(const Geom &g, const HybridCfg &cfg, const REAL *a, const REAL *b,
                         const REAL *c, REAL *d, REAL *u, hipStream_t st) {
  if (cfg.P == hyb::LONG32_T_MAX)
    return launch_long32_t<REAL, INC, hyb::LONG32_T_MAX>(g, cfg, a, b, c, d, u, st);
  return launch_long32_t<REAL, INC, hyb::LONG32_T>(g, cfg, a, b, c, d, u, st);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <typename REAL, bool INC>
hipError_t launch_lane(const Geom &g, const REAL *a, const REAL *b, const REAL *c,
                       REAL *d, REAL *u, REAL *cws, REAL *dws, hipStream_t st) {
  const int64_t ns = g.nsys();
  const int64_t blocks = (ns + LANE_BLOCK - 1) / LANE_BLOCK;
  hipLaunchKernelGGL((lane_kernel<REAL, INC>), dim3((unsigned)blocks),
                     dim3(LANE_BLOCK), 0, st, g, a, b, c, d, u, cws, dws);
  return hipGetLastError();
}

template <typename REAL>
static int max_threads_for_m(int M) {
  switch (M) {
    case 1: return hybrid_max_threads<REAL, 1>();
    case 2: return hybrid_max_threads<REAL, 2>();
    case 4: return hybrid_max_threads<REAL, 4>();
    case 8: return hybrid_max_threads<REAL, 8>();
    case 16: return hybrid_max_threads<REAL, 16>();
    default: return 0;
  }
}

template <typename REAL>
bool hybrid_cfg_valid(const Geom &g, const HybridCfg &cfg) {
  if (cfg.V == 1 || cfg.V == 2) return stream_cfg_valid<REAL>(g, cfg);
  if (cfg.V == 3) return wide_cfg_valid<REAL>(g, cfg);
  if (cfg.V == 4) {  // one contiguous system per workgroup: P = 256 for 4097-8192
                     // rows, 512 for 8193-16384
    const bool p256 = cfg.P == hyb::LONG32_T, p512 = cfg.P == hyb::LONG32_T_MAX;
    return g.stride == 1 && cfg.M == hyb::LONG32_M && cfg.W == 1 && cfg.G >= 1 &&
           ((p256 && g.N > 4096) || (p512 && g.N > (int64_t)hyb::LONG32_M * hyb::LONG32_T)) &&
           g.N <= (int64_t)hyb::LONG32_M * cfg.P;
  }
  if (cfg.V != 0) return false;
  const int maxT = max_threads_for_m<REAL>(cfg.M);
  if (maxT == 0 || cfg.P < 1 || cfg.W < 1) return false;
  if ((int64_t)cfg.P * cfg.M < g.N) return false;
  const int64_t T = (int64_t)cfg.W * cfg.P;
  if (T > maxT) return false;
  size_t lds = sizeof(REAL) * (size_t)6 * T;
  if (g.stride == 1) lds += sizeof(REAL) * (size_t)T * (cfg.M + 1);
  return lds <= 160 * 1024;
}

template <typename REAL>
bool hybrid_plain_cfg(const Geom &g, HybridCfg *cfg) {
  const bool contig = g.stride == 1;
  static const int Ms[] = {1, 2, 4, 8, 16};
  if (contig) {
    // One wave (or fewer lanes) per system when possible, 256-thread groups.
    for (int M : Ms) {
      const int64_t P = (g.N + M - 1) / M;
      if (P > 64 && M < 16) continue;
      HybridCfg c;
      c.M = M;
      c.P = (int)P;
      c.W = (int)std::max<int64_t>(1, 256 / P);
      if (hybrid_cfg_valid<REAL>(g, c)) { *cfg = c; return true; }
    }
    return false;
  }
  // Strided: 128 bytes of neighbouring systems per row (W = 16 FP64 / 32
  // FP32); the smallest M whose workgroup fits the preferred size (FP64: 512
  // threads at M = 16), else any valid shape.
  const int pref = sizeof(REAL) == 8 ? 512 : 1024;
  for (int pass = 0; pass < 2; ++pass) {
    for (int W = 128 / (int)sizeof(REAL); W >= 4; W /= 2) {
      for (int M : Ms) {
        HybridCfg c;
        c.M = M;
        c.P = (int)((g.N + M - 1) / M);
        c.W = W;
        if (pass == 0 && (int64_t)c.W * c.P > pref) continue;
        if (hybrid_cfg_valid<REAL>(g, c)) { *cfg = c; return true; }
      }
    }
  }
  return false;
}

template <typename REAL>
bool hybrid_default_cfg(const Geom &g, HybridCfg *cfg) {
  if (g.stride != 1) {
    // Strided systems: the variant with a staged in LDS, shapes measured
    // at 256^3-512^3. Rows at least 64 KiB apart (z) lie in different pages:
    // FP64 takes 256-byte rows (M = 32 above 256 rows) in plain workgroup
    // order, FP32 128-byte rows. Rows close together (y) take 128-byte rows,
    // XCD-grouped.
    HybridCfg c;
    const bool fp64 = sizeof(REAL) == 8;
    const int64_t row_bytes = g.stride * (int64_t)sizeof(REAL);
    const bool far = row_bytes >= (64 << 10);
    c.M = (far && fp64 && g.N > 256) ? 32 : 16;
    c.W = ((far && fp64) ? 256 : 128) / (int)sizeof(REAL);
    const int per_wave = std::max(1, 64 / c.W);
    c.P = (int)((g.N + c.M - 1) / c.M);
    c.P = (c.P + per_wave - 1) / per_wave * per_wave;
    c.G = !far ? 8 : (fp64 || row_bytes >= (1 << 20)) ? 1 : 8;
    c.V = 3;
    if (hybrid_cfg_valid<REAL>(g, c)) {
      *cfg = c;
      return true;
    }
  }
  // Contiguous systems of 4097-8192 rows: one per 256-thread workgroup at 32
  // rows per lane (variant 4), FP64 throughout (4500 / 6000 / 8192 rows: 3.8
  // / 4.3 / 4.7 TB/s against 3.2 / 3.7 / 4.0 with 512 lanes of 16 rows), FP32
  // up to 6144 rows (4.0 / 4.4 against 3.2 / 3.8; at 8192 rows 4.4 against
  // 4.7).
  // FP64 systems of 8193-16384 rows, longer than the other shapes hold, one
  // per 512-thread workgroup.
  if (g.stride == 1 && g.N > 4096 && (sizeof(REAL) == 8 || g.N <= 6144)) {
    HybridCfg c;
    c.M = hyb::LONG32_M;
    c.P = g.N > (int64_t)hyb::LONG32_M * hyb::LONG32_T ? hyb::LONG32_T_MAX : hyb::LONG32_T;
    c.W = 1;
    c.G = sizeof(REAL) == 8 ? 8 : 1;
    c.V = 4;
    if (hybrid_cfg_valid<REAL>(g, c)) {
      *cfg = c;
      return true;
    }
  }
  return hybrid_plain_cfg<REAL>(g, cfg);
}

template <typename REAL, bool INC>
hipError_t launch_hybrid(const Geom &g, const HybridCfg &cfg, const REAL *a,
                         const REAL *b, const REAL *c, REAL *d, REAL *u,
                         hipStream_t st) {
  if (cfg.V == 1 || cfg.V == 2) return launch_stream<REAL, INC>(g, cfg, a, b, c, d, u, st);
  if (cfg.V == 3) return launch_wide<REAL, INC>(g, cfg, a, b, c, d, u, st);
  if (cfg.V == 4) return launch_long32<REAL, INC>(g, cfg, a, b, c, d, u, st);
  const bool contig = g.stride == 1;
  // 16-byte vector access on contiguous systems when every vector lies wholly
  // inside or wholly beyond a row: N and the row pitches multiples of V,
  // 16-byte aligned arrays, M a multiple of V. Only for M <= 4: measured on
  // 1024- and 4096-row FP64 systems, vector loads at M = 8 and 16 run at
  // 3.1-4.2 TB/s against 5.2-5.4 with element loads, while at M = 4
  // (512 rows) they gain 4 %.
  constexpr int V = 16 / sizeof(REAL);
  constexpr int VEC_MAX_M = 4;
  auto al16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool vec = contig && cfg.M <= VEC_MAX_M && cfg.M % V == 0 && g.N % V == 0 &&
                   g.st0 % V == 0 && g.st1 % V == 0 && al16(a) && al16(b) && al16(c) &&
                   al16(d) && (!INC || al16(u));
#define TRID_HYB_CASE(MM)                                                               \
  case MM:                                                                              \
    if (!contig) return launch_hybrid_m<REAL, MM, INC, false, false>(g, cfg, a, b, c, d, u, st); \
    if constexpr (MM % V == 0 && MM <= VEC_MAX_M) {                                     \
      if (vec) return launch_hybrid_m<REAL, MM, INC, true, true>(g, cfg, a, b, c, d, u, st);    \
    }                                                                                   \
    return launch_hybrid_m<REAL, MM, INC, true, false>(g, cfg, a, b, c, d, u, st);
  switch (cfg.M) {
    TRID_HYB_CASE(1)
    TRID_HYB_CASE(2)
    TRID_HYB_CASE(4)
    TRID_HYB_CASE(8)
    TRID_HYB_CASE(16)
    default:
      return hipErrorInvalidValue;
  }
#undef TRID_HYB_CASE
}

#define TRID_INST(REAL)                                                          \
  template hipError_t launch_lane<REAL, false>(const Geom &, const REAL *,       \
      const REAL *, const REAL *, REAL *, REAL *, REAL *, REAL *, hipStream_t);  \
  template hipError_t launch_lane<REAL, true>(const Geom &, const REAL *,        \
      const REAL *, const REAL *, REAL *, REAL *, REAL *, REAL *, hipStream_t);  \
  template hipError_t launch_hybrid<REAL, false>(const Geom &, const HybridCfg &,\
      const REAL *, const REAL *, const REAL *, REAL *, REAL *, hipStream_t);    \
  template hipError_t launch_hybrid<REAL, true>(const Geom &, const HybridCfg &, \
      const REAL *, const REAL *, const REAL *, REAL *, REAL *, hipStream_t);    \
  template bool hybrid_cfg_valid<REAL>(const Geom &, const HybridCfg &);         \
  template bool hybrid_default_cfg<REAL>(const Geom &, HybridCfg *);               \
  template bool hybrid_plain_cfg<REAL>(const Geom &, HybridCfg *);

TRID_INST(float)
TRID_INST(double)

}  // namespace trid
