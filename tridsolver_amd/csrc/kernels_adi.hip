// kernels_adi.hip - the ADI heat-diffusion caller of the batched solve.
//
// The reference app (apps/adi/src/adi_cuda.cu:282-326) runs, per timestep,
//   preproc -> tridXmtsvStridedBatch(solvedim 0) -> (solvedim 1)
//           -> tridXmtsvStridedBatchInc(solvedim 2, u += x),
// with a device synchronisation inside every solve. Here the step is four
// stream-ordered launches with no host synchronisation, so a whole step can be
// captured once into a hipGraph and replayed (tridAdiRun*).
//
// preproc (reference apps/adi/src/preproc.cu:40-90, CPU apps/adi/src/preproc.hpp:172-214):
//   boundary points:  d = 0, a = 0, b = 1, c = 0
//   interior points:  d = lambda * (sum of the 6 neighbours - 6 u),
//                     a = c = -0.5 lambda, b = 1 + lambda.
// The reference writes the same a, b, c into nine arrays (x, y and z); the
// kernel writes each distinct output pointer once, so a caller may pass one
// coefficient set for all three directions (4 arrays written instead of 10).
// Each thread owns one (i, j) column and walks a chunk of k with the three
// planes k-1, k, k+1 of u in registers.

#include <algorithm>

#include "../../include/tridsolver.h"
#include "../../include/tridsolver_adi.h"
#include "../../include/tridsolver_hip.h"
#include "hybrid_core.hpp"
#include "trid_internal.h"
#include "trid_tune.h"

namespace trid {
namespace {

constexpr int PRE_BX = 64, PRE_BY = 4, PRE_KCHUNK = 16;

template <typename REAL>
struct Coef {
  REAL *a, *b, *c;
};

// Right-hand side at an interior point, from u at the point (c) and its six
// neighbours; the one expression both the preproc kernel and the x sweep
// that computes its own right-hand side evaluate, so they agree bit for bit.
template <typename REAL>
__device__ __forceinline__ REAL adi_rhs(REAL lambda, REAL xm, REAL xp, REAL ym, REAL yp,
                                        REAL zm, REAL zp, REAL c) {
  // lambda * (u[-1] + u[+1] + u[-nx] + u[+nx] + u[-nx*ny] + u[+nx*ny] - 6 u)
  return lambda * (xm + xp + ym + yp + zm + zp - REAL(6) * c);
}

template <typename REAL>
__global__ __launch_bounds__(PRE_BX *PRE_BY) void adi_preproc_kernel(
    REAL lambda, const REAL *__restrict__ u, REAL *__restrict__ du, Coef<REAL> c0,
    Coef<REAL> c1, Coef<REAL> c2, int nx, int ny, int nz) {
  const int i = blockIdx.x * PRE_BX + threadIdx.x;
  const int j = blockIdx.y * PRE_BY + threadIdx.y;
  const int k0 = blockIdx.z * PRE_KCHUNK;
  if (i >= nx || j >= ny || k0 >= nz) return;
  const int k1 = min(nz, k0 + PRE_KCHUNK);
  const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
  const bool edge_ij = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
  int64_t ind = (int64_t)k0 * sz + (int64_t)j * sy + i;
  REAL um = 0, uc = u[ind], up = 0;
  if (!edge_ij && k0 > 0) um = u[ind - sz];
  const REAL ca = REAL(-0.5) * lambda, cb = REAL(1) + lambda;
  for (int k = k0; k < k1; ++k, ind += sz) {
    const bool inner = !edge_ij && k > 0 && k < nz - 1;
    if (!edge_ij && k < nz - 1) up = u[ind + sz];
    REAL d, a, b, c;
    if (inner) {
      d = adi_rhs(lambda, u[ind - sx], u[ind + sx], u[ind - sy], u[ind + sy], um, up, uc);
      a = ca;
      b = cb;
      c = ca;
    } else {
      d = 0; a = 0; b = 1; c = 0;
    }
    du[ind] = d;
    if (c0.a) { c0.a[ind] = a; c0.b[ind] = b; c0.c[ind] = c; }
    if (c1.a) { c1.a[ind] = a; c1.b[ind] = b; c1.c[ind] = c; }
    if (c2.a) { c2.a[ind] = a; c2.b[ind] = b; c2.c[ind] = c; }
    um = uc;
    uc = up;
  }
}

template <typename REAL>
tridStatus_t preproc(REAL lambda, const REAL *u, REAL *du, REAL *ax, REAL *bx, REAL *cx,
                     REAL *ay, REAL *by, REAL *cy, REAL *az, REAL *bz, REAL *cz, int nx,
                     int ny, int nz) {
  const bool x_null = !ax && !bx && !cx;  // du only (matrix-free step)
  if (!u || !du || nx < 1 || ny < 1 || nz < 1 || (!x_null && (!ax || !bx || !cx)))
    return TRID_STATUS_INVALID_VALUE;
  // A direction whose three pointers are NULL, or equal to an earlier set,
  // is not written again.
  auto same = [](REAL *p, REAL *q, REAL *r, REAL *P, REAL *Q, REAL *R) {
    return p == P && q == Q && r == R;
  };
  Coef<REAL> c0{ax, bx, cx}, c1{nullptr, nullptr, nullptr}, c2{nullptr, nullptr, nullptr};
  const bool y_null = !ay && !by && !cy, z_null = !az && !bz && !cz;
  if ((!y_null && (!ay || !by || !cy)) || (!z_null && (!az || !bz || !cz)))
    return TRID_STATUS_INVALID_VALUE;
  if (x_null && (!y_null || !z_null)) return TRID_STATUS_INVALID_VALUE;
  if (!y_null && !same(ay, by, cy, ax, bx, cx)) c1 = {ay, by, cy};
  if (!z_null && !same(az, bz, cz, ax, bx, cx) && !same(az, bz, cz, ay, by, cy))
    c2 = {az, bz, cz};
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  dim3 grid((nx + PRE_BX - 1) / PRE_BX, (ny + PRE_BY - 1) / PRE_BY,
            (nz + PRE_KCHUNK - 1) / PRE_KCHUNK);
  hipLaunchKernelGGL(adi_preproc_kernel<REAL>, grid, dim3(PRE_BX, PRE_BY), 0,
                     current_stream(), lambda, u, du, c0, c1, c2, nx, ny, nz);
  return hipGetLastError() == hipSuccess ? TRID_STATUS_SUCCESS
                                         : TRID_STATUS_EXECUTION_FAILED;
}

// ---------------------------------------------------------------------------
// Matrix-free sweeps for the ADI step. The coefficients preproc would store
// are functions of the grid position only (interior: a = c = -0.5 lambda,
// b = 1 + lambda; any boundary point: a = 0, b = 1, c = 0), so each sweep
// computes them from (i, j, k) with the same expressions in the same
// precision, and reads only du (and u for the z sweep): 16 B per point
// per sweep instead of 40 (FP64). The solve is the Thomas-PCR hybrid of
// hybrid_core.hpp, so results equal the array-based step bit for bit.
// ---------------------------------------------------------------------------
template <typename REAL>
struct AdiCoef {
  REAL lambda;
  int nx, ny, nz, sd;  // sd: direction of the sweep
};

// Coefficients of chunk-order rows r = 0..M-1 of the lane's system.
template <typename REAL, int M, bool CONTIG>
__device__ __forceinline__ void adi_coefs(const hyb::Lane &L, const AdiCoef<REAL> &k,
                                          REAL (&ra)[M], REAL (&rb)[M], REAL (&rc)[M]) {
  const REAL ca = REAL(-0.5) * k.lambda, cb = REAL(1) + k.lambda;
  // the contiguous sweep is the x sweep
  const int sd = CONTIG ? 0 : k.sd;
  // coordinates of the system; the row index runs along direction sd
  int ci, cj, ck;
  if (sd == 0) { cj = (int)L.i0; ck = (int)L.i1; ci = 0; }
  else if (sd == 1) { ci = (int)L.i0; ck = (int)L.i1; cj = 0; }
  else { ci = (int)L.i0; cj = (int)L.i1; ck = 0; }
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int64_t row = (int64_t)L.p * M + r;
    int i = ci, j = cj, kk = ck;
    if (sd == 0) i = (int)row;
    else if (sd == 1) j = (int)row;
    else kk = (int)row;
    const bool pad = !L.active || row >= L.N;
    const bool edge = i == 0 || i == k.nx - 1 || j == 0 || j == k.ny - 1 || kk == 0 ||
                      kk == k.nz - 1;
    const bool inner = !pad && !edge;
    ra[r] = inner ? ca : REAL(0);
    rb[r] = inner ? cb : REAL(1);
    rc[r] = inner ? ca : REAL(0);
  }
}

template <typename REAL, int M>
struct AdiBounds {  // same register budget as the array-based hybrid kernel
  static constexpr int threads = sizeof(REAL) * M > 64 ? 512 : 1024;
};

// RHS (x sweep only): the right-hand side is not read from d but computed
// from u, as the preproc kernel does: the line's own u through the LDS tile
// for the x neighbours, the four neighbour lines (j +- 1, k +- 1) loaded
// alongside (coalesced; at 256^3 u stays in the last-level cache). d is
// written only.
template <typename REAL, int M, bool INC, bool CONTIG, bool RHS>
__global__ __launch_bounds__((AdiBounds<REAL, M>::threads)) void adi_solve_kernel(
    Geom g, int P, int W, int G, AdiCoef<REAL> kc, REAL *__restrict__ d,
    REAL *__restrict__ u) {
  static_assert(!RHS || (CONTIG && !INC), "right-hand side from u: x sweep only");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<CONTIG>(g, P, W, G);
  const size_t tile_bytes = CONTIG ? sizeof(REAL) * (size_t)W * P * (M + 1) : 0;
  REAL *tile = reinterpret_cast<REAL *>(smem_raw) + (size_t)L.w * P * (M + 1);
  hyb::Ring ring{smem_raw + tile_bytes, (int)(3 * sizeof(REAL) * L.T), 0};
  REAL ra[M], rb[M], rc[M], rd[M];
  if constexpr (RHS) {
    // Every lane loads all of its rows unconditionally, from addresses kept
    // inside the grid: rows past N read row N-1, lanes past the last line
    // read line 0, and lines on the boundary read their own line as the
    // "neighbours". Those values only feed rows whose right-hand side is 0
    // below, so no load needs a branch around it.
    // Offsets are 32-bit byte offsets from u (the launch checks that the
    // grid is below 4 GiB).
    REAL uc[M], ym[M], yp[M], zm[M], zp[M];
    const int j = (int)L.i0, k = (int)L.i1;  // the x line (j, k)
    const bool line_inner = L.active && j > 0 && j < kc.ny - 1 && k > 0 && k < kc.nz - 1;
    const uint32_t sy = line_inner ? (uint32_t)(kc.nx * sizeof(REAL)) : 0u;
    const uint32_t sz = line_inner ? (uint32_t)((uint32_t)kc.nx * kc.ny * sizeof(REAL)) : 0u;
    const uint32_t lbase = L.active ? (uint32_t)L.base : 0u;
    const int n = (int)L.N;
    const char *ub = reinterpret_cast<const char *>(u);
    auto ld = [ub](uint32_t off) { return *reinterpret_cast<const REAL *>(ub + off); };
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const int row = (int)hyb::fetch_row<M, true>(L, r);
      const uint32_t o = (lbase + (uint32_t)(row < n ? row : n - 1)) * (uint32_t)sizeof(REAL);
      uc[r] = ld(o);
      ym[r] = ld(o - sy);
      yp[r] = ld(o + sy);
      zm[r] = ld(o - sz);
      zp[r] = ld(o + sz);
    }
    // x neighbours: the line's u in the tile, row order
#pragma unroll
    for (int r = 0; r < M; ++r) tile[hyb::skew<M>(L.p + L.P * r)] = uc[r];
    __syncthreads();
    const int last = P * M - 1;
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const int row = L.p + L.P * r;
      const bool inner = line_inner && row > 0 && row < L.N - 1;
      const REAL xm = tile[hyb::skew<M>(row > 0 ? row - 1 : 0)];
      const REAL xp = tile[hyb::skew<M>(row < last ? row + 1 : last)];
      rd[r] = inner ? adi_rhs(kc.lambda, xm, xp, ym[r], yp[r], zm[r], zp[r], uc[r]) : REAL(0);
    }
    __syncthreads();  // the tile is reused below
  } else {
    hyb::fetch<M, CONTIG>(L, d, rd, REAL(0), 0, 0, L.N);
  }
  hyb::to_chunk<M, CONTIG>(L, rd, tile);
  adi_coefs<REAL, M, CONTIG>(L, kc, ra, rb, rc);
  REAL x[M];
  hyb::solve_chunk<REAL, M, CONTIG>(L, ring, ra, rb, rc, rd, x);
  hyb::to_fetch<M, CONTIG>(L, x, tile);
  if constexpr (RHS) {  // 32-bit byte offsets, as the loads
    if (!L.active) return;
    const int n = (int)L.N;
    const uint32_t lbase = (uint32_t)L.base;
    char *db = reinterpret_cast<char *>(d);
    auto st = [&](int r) {
      const uint32_t row = (uint32_t)hyb::fetch_row<M, true>(L, r);
      *reinterpret_cast<REAL *>(db + (lbase + row) * (uint32_t)sizeof(REAL)) = x[r];
    };
    if ((int64_t)P * M == L.N) {  // every row of every lane lies inside the line
#pragma unroll
      for (int r = 0; r < M; ++r) st(r);
    } else {
#pragma unroll
      for (int r = 0; r < M; ++r)
        if ((int)hyb::fetch_row<M, true>(L, r) < n) st(r);
    }
  } else if (INC) {
    hyb::store_add<M, CONTIG>(L, u, x);
  } else {
    hyb::store<M, CONTIG>(L, d, x);
  }
}

// The matrix-free x sweep with every access in chunk order (x lines of
// N = P * 8 rows, P lanes of 8 rows, one wave per workgroup): lane p loads
// its own rows 8p .. 8p+7 of u and of the four neighbour lines as 16-byte
// vectors, takes the x neighbours of its first and last row from lanes p-1
// and p+1 by ds_bpermute, solves (the PCR exchanges are wave-local too) and
// stores its rows as vectors. No LDS tile and no transpose: the values and
// the arithmetic are those of adi_solve_kernel's x sweep, so the results are
// the same bit for bit. For x lines whose rows start 16-byte aligned
// (nx * sizeof(REAL) a multiple of 16) on grids below 4 GiB, at FP32: 8 FP32
// rows are two 16-byte vectors per lane. Measured on the 256^3 matrix-free
// step (replay time per step): FP32 0.105 ms against 0.117-0.12 with the
// LDS-tiled sweep; FP64, four vectors per lane (64-byte strides between
// lanes), 0.197 ms against 0.178, so FP64 keeps the tiled sweep.
constexpr int ADI_DX_M = 8;

template <typename REAL>
__global__ __launch_bounds__(64) void adi_x_direct_kernel(Geom g, int P, int W, int G,
                                                          AdiCoef<REAL> kc,
                                                          REAL *__restrict__ d,
                                                          const REAL *__restrict__ u) {
  constexpr int M = ADI_DX_M;
  constexpr int V = 16 / (int)sizeof(REAL);  // elements per 16-byte vector
  typedef REAL vec __attribute__((ext_vector_type(V)));
  __shared__ __align__(16) unsigned char ring_lds[2 * 3 * sizeof(REAL) * 64];
  const hyb::Lane L = hyb::lane_of<true>(g, P, W, G);
  hyb::Ring ring{ring_lds, (int)(3 * sizeof(REAL) * 64), 0};
  const int j = (int)L.i0, k = (int)L.i1;  // the x line (j, k)
  const bool line_inner = L.active && j > 0 && j < kc.ny - 1 && k > 0 && k < kc.nz - 1;
  const uint32_t sy = line_inner ? (uint32_t)(kc.nx * sizeof(REAL)) : 0u;
  const uint32_t sz = line_inner ? (uint32_t)((uint32_t)kc.nx * kc.ny * sizeof(REAL)) : 0u;
  // lanes past the last line read line 0; their rows are identity rows
  const uint32_t o0 = ((L.active ? (uint32_t)L.base : 0u) + (uint32_t)(L.p * M)) *
                      (uint32_t)sizeof(REAL);
  const char *ub = reinterpret_cast<const char *>(u);
  auto ldv = [ub](uint32_t off) { return *reinterpret_cast<const vec *>(ub + off); };
  REAL uc[M], ym[M], yp[M], zm[M], zp[M];
#pragma unroll
  for (int q = 0; q < M / V; ++q) {
    const uint32_t o = o0 + (uint32_t)(q * 16);
    const vec c0 = ldv(o), c1 = ldv(o - sy), c2 = ldv(o + sy), c3 = ldv(o - sz), c4 = ldv(o + sz);
#pragma unroll
    for (int e = 0; e < V; ++e) {
      uc[q * V + e] = c0[e];
      ym[q * V + e] = c1[e];
      yp[q * V + e] = c2[e];
      zm[q * V + e] = c3[e];
      zp[q * V + e] = c4[e];
    }
  }
  // x neighbours: rows 8p-1 and 8p+8 live in lanes p-1 and p+1
  REAL edge[2] = {uc[M - 1], uc[0]}, nb_lo[2], nb_hi[2];
  hyb::exchange<3, 2>(L, ring, true, edge, L.pstep, L.p > 0, nb_lo, L.pstep, L.p + 1 < L.P,
                      nb_hi);
  REAL rd[M];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int row = L.p * M + r;
    const bool inner = line_inner && row > 0 && row < L.N - 1;
    const REAL xm = r > 0 ? uc[r - 1] : nb_lo[0];
    const REAL xp = r < M - 1 ? uc[r + 1] : nb_hi[1];
    rd[r] = inner ? adi_rhs(kc.lambda, xm, xp, ym[r], yp[r], zm[r], zp[r], uc[r]) : REAL(0);
  }
  REAL ra[M], rb[M], rc[M];
  adi_coefs<REAL, M, true>(L, kc, ra, rb, rc);
  REAL x[M];
  hyb::solve_chunk<REAL, M, true>(L, ring, ra, rb, rc, rd, x);
  if (!L.active) return;
  char *db = reinterpret_cast<char *>(d);
#pragma unroll
  for (int q = 0; q < M / V; ++q) {
    vec v;
#pragma unroll
    for (int e = 0; e < V; ++e) v[e] = x[q * V + e];
    *reinterpret_cast<vec *>(db + o0 + (uint32_t)(q * 16)) = v;
  }
}

// The direct x sweep applies: 8 rows per lane filling the lines exactly, a
// whole line's lanes in the wave, 16-byte aligned rows and arrays.
template <typename REAL>
bool adi_x_direct_ok(const Geom &g, const HybridCfg &cfg, const REAL *d, const REAL *u,
                     int nx) {
  auto al16 = [](const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  return sizeof(REAL) == 4 && cfg.M == ADI_DX_M && (int64_t)cfg.P * ADI_DX_M == g.N &&
         cfg.W * cfg.P <= 64 && ((int64_t)nx * (int64_t)sizeof(REAL)) % 16 == 0 && al16(d) &&
         al16(u);
}

template <typename REAL>
hipError_t launch_adi_x_direct(const Geom &g, const HybridCfg &cfg, const AdiCoef<REAL> &kc,
                               REAL *d, const REAL *u, hipStream_t st) {
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(adi_x_direct_kernel<REAL>, dim3((unsigned)blocks), dim3(cfg.W * cfg.P), 0,
                     st, g, cfg.P, cfg.W, cfg.G, kc, d, u);
  return hipGetLastError();
}

template <typename REAL, int M, bool INC, bool CONTIG, bool RHS = false>
hipError_t launch_adi_solve_m(const Geom &g, const HybridCfg &cfg, const AdiCoef<REAL> &kc,
                              REAL *d, REAL *u, hipStream_t st) {
  const int T = cfg.W * cfg.P;
  size_t lds = hyb::ring_bytes<REAL>(T, 3);
  if (CONTIG) lds += sizeof(REAL) * (size_t)T * (M + 1);
  auto kern = adi_solve_kernel<REAL, M, INC, CONTIG, RHS>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  const int64_t blocks = (g.n0 + cfg.W - 1) / cfg.W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), lds, st, g, cfg.P, cfg.W, cfg.G,
                     kc, d, u);
  return hipGetLastError();
}

// Shape: the one the array-based solve of this geometry uses (tuned or
// default), with plain loads. Systems longer than one workgroup covers are
// not supported here.
template <typename REAL, bool INC>
bool adi_cfg(int dev, const Geom &g, HybridCfg *cfg) {
  if (!tuned_config<REAL, INC>(dev, g, cfg)) return false;
  if (cfg->V != 0) {
    // a variant shape: keep its chunking (M, P: same arithmetic) with plain
    // loads of 16 systems per workgroup (128-byte FP64 rows, 64-byte FP32
    // rows), which suit a sweep that reads only d. Measured on the 256^3
    // matrix-free step over W in {8, 16, 32, 64} and G in {1, 8, 16}: FP32
    // 0.123 ms at W = 16 against 0.134 at 32 (128-byte rows) and 0.143 at 64;
    // FP64 0.198-0.200 ms at W = 16, 0.205 at 8, 0.24 at 32.
    HybridCfg c = *cfg;
    c.V = 0;
    c.W = 16;
    if (hybrid_cfg_valid<REAL>(g, c)) {
      *cfg = c;
      return true;
    }
    return hybrid_plain_cfg<REAL>(g, cfg);
  }
  return hybrid_cfg_valid<REAL>(g, *cfg) || hybrid_plain_cfg<REAL>(g, cfg);
}

// Shape of the ADI x sweeps, shared by the array step's x solve and the
// matrix-free x sweep so that the two steps agree bit for bit. An x line
// fits one workgroup easily; what costs is the PCR across its lanes (a
// division per step and lane), so fewer lanes of more rows than the general
// contiguous default (up to 64 lanes of 4 rows at 256 rows) pay off. Measured
// on 256^3 steps over M in {4, 8, 16}, T in {64, 128, 256}, G in {1, 8}:
// 8 rows per lane, one 64-thread wave per workgroup (matrix-free step FP64
// 0.217 against 0.229-0.233 ms at M = 4; FP32 0.149 against 0.159-0.162).
// G = 1 and 8 time alike; with G = 8 neighbouring x lines run on one XCD,
// so the halo lines the right-hand side reads (j +- 1) hit in its L2
// instead of being fetched again. Re-measured with the branch-free
// right-hand-side loads: M = 16 / T = 64, M = 8 / T = 128, M = 16 / T = 128
// and M = 4 / T = 64 take the FP64 step to 0.206-0.215 ms against 0.200, and
// the FP32 step to 0.134-0.160 ms against 0.133.
template <typename REAL>
struct AdiX {
  static constexpr int M = 8;
  static constexpr int T = 64;
  static constexpr int G = 8;
};

template <typename REAL>
bool adi_x_cfg(const Geom &g, HybridCfg *cfg) {
  HybridCfg c;
  c.M = AdiX<REAL>::M;
  c.P = (int)((g.N + c.M - 1) / c.M);
  c.W = (int)std::max<int64_t>(1, AdiX<REAL>::T / c.P);
  c.G = AdiX<REAL>::G;
  c.V = 0;
  if (g.stride != 1 || !hybrid_cfg_valid<REAL>(g, c)) return false;
  *cfg = c;
  return true;
}

// The x solve of the array step, in the ADI x shape unless the caller has
// forced a shape of its own.
template <typename REAL>
tridStatus_t x_solve(const TridParams *p, const REAL *a, const REAL *b, const REAL *c, REAL *d,
                     const int *dims) {
  Geom g;
  bool empty = false;
  HybridCfg xc;
  const bool shape = forced_config().M == 0 &&
                     make_geom(3, 0, dims, dims, &g, &empty) == TRID_STATUS_SUCCESS &&
                     !empty && adi_x_cfg<REAL>(g, &xc);
  if (shape) tridHipForceHybridConfig(xc.M, xc.P, xc.W, xc.G, xc.V);
  tridStatus_t s;
  if constexpr (sizeof(REAL) == 8) s = tridDmtsvStridedBatch(p, a, b, c, d, 3, 0, dims, dims);
  else s = tridSmtsvStridedBatch(p, a, b, c, d, 3, 0, dims, dims);
  if (shape) tridHipForceHybridConfig(0, 0, 0, 0, 0);
  return s;
}

template <typename REAL, bool INC, bool RHS = false>
tridStatus_t adi_sweep(int dev, REAL lambda, REAL *d, REAL *u, int nx, int ny, int nz,
                       int sd) {
  static_assert(!RHS || !INC, "right-hand side from u: x sweep only");
  const int dims[3] = {nx, ny, nz};
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(3, sd, dims, dims, &g, &empty);
  if (s != TRID_STATUS_SUCCESS || empty) return s;
  HybridCfg cfg;
  if (!(sd == 0 && adi_x_cfg<REAL>(g, &cfg)) && !adi_cfg<REAL, INC>(dev, g, &cfg))
    return TRID_STATUS_INVALID_VALUE;
  // The sweep forms coefficients from grid coordinates, so it is not split
  // into sub-batches: its grid must fit HIP's 2^32 work-items (about 2^35
  // points at 8 rows per lane, beyond a GPU's memory for two arrays).
  if ((g.n0 + cfg.W - 1) / cfg.W * g.n1 > (int64_t)UINT32_MAX / (cfg.W * cfg.P))
    return TRID_STATUS_INVALID_VALUE;
  AdiCoef<REAL> kc{lambda, nx, ny, nz, sd};
  hipStream_t st = current_stream();
  const bool contig = sd == 0;
  hipError_t e;
  if (RHS && !contig) return TRID_STATUS_INTERNAL_ERROR;
  // The right-hand side is computed in the sweep unless FP32 x lines are
  // at 16 rows per lane (they would spill) or the grid is 4 GiB or more (the
  // sweep's loads use 32-bit byte offsets): there preproc writes du and the
  // plain sweep reads it.
  constexpr bool RHS16 = RHS && sizeof(REAL) == 8;
  const bool grid32 =
      (uint64_t)nx * (uint64_t)ny * (uint64_t)nz * sizeof(REAL) < (uint64_t(1) << 32);
  const bool rhs_here = RHS && grid32 && (cfg.M != 16 || RHS16);
  if (RHS && !rhs_here) {
    tridStatus_t sp = preproc<REAL>(lambda, u, d, nullptr, nullptr, nullptr, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr, nx, ny, nz);
    if (sp != TRID_STATUS_SUCCESS) return sp;
  }
  if (contig && rhs_here && adi_x_direct_ok<REAL>(g, cfg, d, u, nx)) {
    e = launch_adi_x_direct<REAL>(g, cfg, kc, d, u, st);
    if (e != hipSuccess) {
      record_hip_error(e);
      return TRID_STATUS_EXECUTION_FAILED;
    }
    return TRID_STATUS_SUCCESS;
  }
#define TRID_ADI_CASE(MM)                                                                  \
  case MM:                                                                                 \
    if (contig && rhs_here)                                                                \
      e = launch_adi_solve_m<REAL, MM, INC, true, (MM == 16 ? RHS16 : RHS)>(g, cfg, kc, d, \
                                                                              u, st);      \
    else if (contig)                                                                       \
      e = launch_adi_solve_m<REAL, MM, INC, true, false>(g, cfg, kc, d, u, st);            \
    else                                                                                   \
      e = launch_adi_solve_m<REAL, MM, INC, false>(g, cfg, kc, d, u, st);                  \
    break;
  switch (cfg.M) {
    TRID_ADI_CASE(1)
    TRID_ADI_CASE(2)
    TRID_ADI_CASE(4)
    TRID_ADI_CASE(8)
    TRID_ADI_CASE(16)
    default:
      return TRID_STATUS_INVALID_VALUE;
  }
#undef TRID_ADI_CASE
  if (e != hipSuccess) {
    record_hip_error(e);
    return TRID_STATUS_EXECUTION_FAILED;
  }
  return TRID_STATUS_SUCCESS;
}

// Matrix-free step: du = solve_x(rhs(u)) (the right-hand side computed in
// the x sweep); du = solve_y(du); u += solve_z(du). 56 B per point at FP64:
// u read + du written, du read + written, du and u read + u written.
template <typename REAL>
tridStatus_t step_fused(REAL lambda, REAL *u, REAL *du, int nx, int ny, int nz) {
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s == TRID_STATUS_SUCCESS && (nx < 1 || ny < 1 || nz < 1)) s = TRID_STATUS_INVALID_VALUE;
  if (s == TRID_STATUS_SUCCESS)
    s = adi_sweep<REAL, false, true>(dev, lambda, du, u, nx, ny, nz, 0);
  if (s == TRID_STATUS_SUCCESS) s = adi_sweep<REAL, false>(dev, lambda, du, u, nx, ny, nz, 1);
  if (s == TRID_STATUS_SUCCESS) s = adi_sweep<REAL, true>(dev, lambda, du, u, nx, ny, nz, 2);
  return s;
}

// One step, stream-ordered, no host synchronisation (ctx->sync forced to 0
// for the solves; the caller decides when to wait).
template <typename REAL>
tridStatus_t step(const TridParams *ctx, REAL lambda, REAL *u, REAL *du, REAL *a,
                  REAL *b, REAL *c, int nx, int ny, int nz);

template <>
tridStatus_t step<double>(const TridParams *ctx, double lambda, double *u, double *du,
                          double *a, double *b, double *c, int nx, int ny, int nz) {
  TridParams p = ctx ? *ctx : TridParams{nullptr, {0, 0, 0}, 0};
  p.sync = 0;
  const int dims[3] = {nx, ny, nz};
  tridStatus_t s = preproc<double>(lambda, u, du, a, b, c, nullptr, nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nx, ny, nz);
  if (s == TRID_STATUS_SUCCESS) s = x_solve<double>(&p, a, b, c, du, dims);
  if (s == TRID_STATUS_SUCCESS) s = tridDmtsvStridedBatch(&p, a, b, c, du, 3, 1, dims, dims);
  if (s == TRID_STATUS_SUCCESS)
    s = tridDmtsvStridedBatchInc(&p, a, b, c, du, u, 3, 2, dims, dims);
  return s;
}

template <>
tridStatus_t step<float>(const TridParams *ctx, float lambda, float *u, float *du,
                         float *a, float *b, float *c, int nx, int ny, int nz) {
  TridParams p = ctx ? *ctx : TridParams{nullptr, {0, 0, 0}, 0};
  p.sync = 0;
  const int dims[3] = {nx, ny, nz};
  tridStatus_t s = preproc<float>(lambda, u, du, a, b, c, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nx, ny, nz);
  if (s == TRID_STATUS_SUCCESS) s = x_solve<float>(&p, a, b, c, du, dims);
  if (s == TRID_STATUS_SUCCESS) s = tridSmtsvStridedBatch(&p, a, b, c, du, 3, 1, dims, dims);
  if (s == TRID_STATUS_SUCCESS)
    s = tridSmtsvStridedBatchInc(&p, a, b, c, du, u, 3, 2, dims, dims);
  return s;
}

template <typename REAL>
tridStatus_t run(const TridParams *ctx, REAL lambda, REAL *u, REAL *du, REAL *a, REAL *b,
                 REAL *c, int nx, int ny, int nz, int iters, int use_graph) {
  // a == NULL selects the matrix-free step
  auto one_step = [&]() {
    return a ? step<REAL>(ctx, lambda, u, du, a, b, c, nx, ny, nz)
             : step_fused<REAL>(lambda, u, du, nx, ny, nz);
  };
  if (iters < 0) return TRID_STATUS_INVALID_VALUE;
  hipStream_t st = current_stream();
  tridStatus_t s = TRID_STATUS_SUCCESS;
  const bool sync = ctx == nullptr || ctx->sync != 0;
  if (!use_graph || iters == 0) {
    for (int it = 0; it < iters && s == TRID_STATUS_SUCCESS; ++it) s = one_step();
  } else {
    // Capture one step on a private stream (the null stream cannot be
    // captured), replay it `iters` times on the caller's stream.
    int dev = 0;
    if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
    hipStream_t cap = nullptr;
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess)
      return TRID_STATUS_EXECUTION_FAILED;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *prev = nullptr;
    tridHipGetStream(&prev);
    if (a) {
      // Solves that take the per-lane path (ctx->opts 1 or 2, or systems
      // longer than one workgroup holds) need a workspace, which cannot be
      // allocated during capture: reserve it for the capture stream now.
      const int dims[3] = {nx, ny, nz};
      size_t need = 0;
      HybridCfg xc;
      Geom gx;
      bool empty = false;
      const bool x_forced = forced_config().M == 0 &&
                            make_geom(3, 0, dims, dims, &gx, &empty) == TRID_STATUS_SUCCESS &&
                            !empty && adi_x_cfg<REAL>(gx, &xc);
      if (!x_forced) need = lane_workspace_need<REAL, false>(dev, ctx, 3, 0, dims, dims);
      need = std::max(need, lane_workspace_need<REAL, false>(dev, ctx, 3, 1, dims, dims));
      need = std::max(need, lane_workspace_need<REAL, true>(dev, ctx, 3, 2, dims, dims));
      void *ws = nullptr;
      if (need > 0) s = get_workspace(dev, cap, need, &ws);
    }
    if (s == TRID_STATUS_SUCCESS) {
      if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        s = TRID_STATUS_EXECUTION_FAILED;
      } else {
        tridHipSetStream(cap);
        tridStatus_t s_step = one_step();
        tridHipSetStream(prev);
        hipError_t e = hipStreamEndCapture(cap, &graph);
        if (s_step != TRID_STATUS_SUCCESS) s = s_step;
        else if (e != hipSuccess) s = TRID_STATUS_EXECUTION_FAILED;
      }
    }
    if (s == TRID_STATUS_SUCCESS &&
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      s = TRID_STATUS_EXECUTION_FAILED;
    for (int it = 0; it < iters && s == TRID_STATUS_SUCCESS; ++it)
      if (hipGraphLaunch(exec, st) != hipSuccess) s = TRID_STATUS_EXECUTION_FAILED;
    // Wait for the replays before the executable graph is destroyed.
    if (exec) {
      if (hipStreamSynchronize(st) != hipSuccess && s == TRID_STATUS_SUCCESS)
        s = TRID_STATUS_EXECUTION_FAILED;
      (void)hipGraphExecDestroy(exec);
    }
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cap);
    drop_stream_workspaces(dev, cap);
  }
  if (s == TRID_STATUS_SUCCESS && sync && hipStreamSynchronize(st) != hipSuccess)
    s = TRID_STATUS_EXECUTION_FAILED;
  return s;
}

}  // namespace
}  // namespace trid

using namespace trid;

EXTERN_C tridStatus_t tridAdiPreprocD(double lambda, const double *u, double *du,
                                      double *ax, double *bx, double *cx, double *ay,
                                      double *by, double *cy, double *az, double *bz,
                                      double *cz, int nx, int ny, int nz) {
  return preproc<double>(lambda, u, du, ax, bx, cx, ay, by, cy, az, bz, cz, nx, ny, nz);
}

EXTERN_C tridStatus_t tridAdiPreprocS(float lambda, const float *u, float *du, float *ax,
                                      float *bx, float *cx, float *ay, float *by,
                                      float *cy, float *az, float *bz, float *cz, int nx,
                                      int ny, int nz) {
  return preproc<float>(lambda, u, du, ax, bx, cx, ay, by, cy, az, bz, cz, nx, ny, nz);
}

EXTERN_C tridStatus_t tridAdiStepD(const TridParams *ctx, double lambda, double *u,
                                   double *du, double *a, double *b, double *c, int nx,
                                   int ny, int nz) {
  tridStatus_t s = step<double>(ctx, lambda, u, du, a, b, c, nx, ny, nz);
  if (s == TRID_STATUS_SUCCESS && (ctx == nullptr || ctx->sync != 0) &&
      hipStreamSynchronize(current_stream()) != hipSuccess)
    s = TRID_STATUS_EXECUTION_FAILED;
  return s;
}

EXTERN_C tridStatus_t tridAdiStepS(const TridParams *ctx, float lambda, float *u, float *du,
                                   float *a, float *b, float *c, int nx, int ny, int nz) {
  tridStatus_t s = step<float>(ctx, lambda, u, du, a, b, c, nx, ny, nz);
  if (s == TRID_STATUS_SUCCESS && (ctx == nullptr || ctx->sync != 0) &&
      hipStreamSynchronize(current_stream()) != hipSuccess)
    s = TRID_STATUS_EXECUTION_FAILED;
  return s;
}

EXTERN_C tridStatus_t tridAdiStepFusedD(double lambda, double *u, double *du, int nx, int ny,
                                        int nz) {
  if (!u || !du) return TRID_STATUS_INVALID_VALUE;
  return step_fused<double>(lambda, u, du, nx, ny, nz);
}

EXTERN_C tridStatus_t tridAdiStepFusedS(float lambda, float *u, float *du, int nx, int ny,
                                        int nz) {
  if (!u || !du) return TRID_STATUS_INVALID_VALUE;
  return step_fused<float>(lambda, u, du, nx, ny, nz);
}

EXTERN_C tridStatus_t tridAdiRunFusedD(const TridParams *ctx, double lambda, double *u,
                                       double *du, int nx, int ny, int nz, int iters,
                                       int use_graph) {
  if (!u || !du) return TRID_STATUS_INVALID_VALUE;
  return run<double>(ctx, lambda, u, du, nullptr, nullptr, nullptr, nx, ny, nz, iters,
                     use_graph);
}

EXTERN_C tridStatus_t tridAdiRunFusedS(const TridParams *ctx, float lambda, float *u, float *du,
                                       int nx, int ny, int nz, int iters, int use_graph) {
  if (!u || !du) return TRID_STATUS_INVALID_VALUE;
  return run<float>(ctx, lambda, u, du, nullptr, nullptr, nullptr, nx, ny, nz, iters,
                    use_graph);
}

EXTERN_C tridStatus_t tridAdiRunD(const TridParams *ctx, double lambda, double *u,
                                  double *du, double *a, double *b, double *c, int nx,
                                  int ny, int nz, int iters, int use_graph) {
  if (!a || !b || !c) return TRID_STATUS_INVALID_VALUE;
  return run<double>(ctx, lambda, u, du, a, b, c, nx, ny, nz, iters, use_graph);
}

EXTERN_C tridStatus_t tridAdiRunS(const TridParams *ctx, float lambda, float *u, float *du,
                                  float *a, float *b, float *c, int nx, int ny, int nz,
                                  int iters, int use_graph) {
  if (!a || !b || !c) return TRID_STATUS_INVALID_VALUE;
  return run<float>(ctx, lambda, u, du, a, b, c, nx, ny, nz, iters, use_graph);
}
