// kernels_long.hip - batched solve of systems longer than one workgroup holds
// (contiguous: above 16384 rows; strided: above 2048 FP64 / 4096 FP32 rows),
// by the Thomas-PCR hybrid with its reduced system solved outside the
// workgroups.
//
// The hybrid (hybrid_core.hpp, kernels_solve.hip) cuts a system into chunks
// of M rows, reduces every chunk to one equation in its last unknown y_k
//     A_k y_{k-1} + y_k + C_k y_{k+1} = D_k,
// solves that reduced system by PCR inside the workgroup and substitutes back.
// A longer system does not fit one workgroup, so here:
//   1. long_reduce_kernel: workgroups take segments of SEG = P * M rows (P
//      lanes per system segment), run phases (i) and (ii) of the hybrid on
//      every chunk and write the chunk's six reduced-form values (its last
//      row's al, ga, de and its first row's dp, ap, gp) -- 6 values per M rows;
//   2. long_build_kernel: the reduced equations A_k, 1, C_k, D_k from chunk
//      k's last row and chunk k+1's first row (the hybrid's phase (iii)
//      coefficients), one thread per chunk;
//   3. the reduced systems (ceil(N / M) rows each) are solved by the batched
//      hybrid itself -- or, when still too long, by this path again;
//   4. long_subst_kernel: the segments are read again, phases (i) and (ii)
//      recomputed (the same arithmetic, so the same values), and
//      x_i = dp_i - ap_i y_{k-1} - gp_i y_k written (or added to u).
// HBM traffic: a, b, c, d read twice (64 B per point at FP64), x written once,
// plus 6 + 4 reals per M rows of reduced records: about 85 B per point
// against the one-workgroup hybrid's 40, and the reduced solve. It replaces
// the per-lane Thomas kernel, whose one recurrence per system leaves the GPU
// idle at a few thousand long systems.
//
// Reduced records and systems live in the caller's workspace, laid out so
// that lanes write consecutive addresses: contiguous systems [system][chunk]
// (the reduced solve is contiguous), strided ones [chunk][system] (the
// reduced solve is strided, x-neighbouring systems adjacent).

#include <algorithm>
#include <atomic>

#include "hybrid_core.hpp"
#include "trid_internal.h"
#include "trid_tune.h"

namespace trid {
namespace {

constexpr int LONG_M = 16;  // rows per lane

// Lanes per system segment and systems per workgroup: contiguous systems one
// system, 512 lanes (8192-row segments); strided ones 128-byte rows of
// x-neighbours, 32 lanes (512-row segments).
template <typename REAL, bool CONTIG>
struct LongShape {
  static constexpr int W = CONTIG ? 1 : 128 / (int)sizeof(REAL);
  static constexpr int P = CONTIG ? 512 : 32;
  static constexpr int T = W * P;
  static constexpr int64_t SEG = (int64_t)P * LONG_M;
};

// The lane of segment `seg` of its system. Chunk k = seg * P + p of the system
// owns rows k*M .. k*M+M-1; L.base and L.N are relative to the segment, so
// the shared fetch order and masks apply inside it.
template <typename REAL, bool CONTIG>
__device__ __forceinline__ hyb::Lane long_lane(const Geom &g, int64_t nseg, int64_t *seg_out,
                                               int64_t *sys_out) {
  using S = LongShape<REAL, CONTIG>;
  hyb::Lane L;
  L.T = S::T;
  L.t = threadIdx.x;
  L.P = S::P;
  L.W = S::W;
  L.p = CONTIG ? L.t : L.t / S::W;
  L.w = CONTIG ? 0 : L.t % S::W;
  L.pstep = CONTIG ? 1 : S::W;
  const int64_t nb0 = (g.n0 + S::W - 1) / S::W;
  const int64_t bid = blockIdx.x;
  const int64_t seg = bid % nseg;
  const int64_t grp = bid / nseg;  // system group: (i0 block, i1)
  const int64_t i0 = (grp % nb0) * S::W + L.w;
  const int64_t i1 = grp / nb0;
  L.active = i0 < g.n0;
  L.i0 = i0;
  L.i1 = i1;
  L.stride = g.stride;
  L.base = i0 * g.st0 + i1 * g.st1 + seg * S::SEG * g.stride;
  const int64_t left = g.N - seg * S::SEG;
  L.N = left < S::SEG ? left : S::SEG;
  *seg_out = seg;
  *sys_out = i0 + g.n0 * i1;
  return L;
}

// a, b, c, d of the lane's chunk in chunk order. a is 0 only at the system's
// first row and c only at its last; rows past N are identity rows.
template <typename REAL, bool CONTIG>
__device__ __forceinline__ void long_fetch(const hyb::Lane &L, bool first_seg, bool last_seg,
                                           const REAL *a, const REAL *b, const REAL *c,
                                           const REAL *d, REAL *tile, REAL (&ra)[LONG_M],
                                           REAL (&rb)[LONG_M], REAL (&rc)[LONG_M],
                                           REAL (&rd)[LONG_M]) {
  constexpr int M = LONG_M;
  // nontemporal: each pass reads the arrays once (8192 strided systems of
  // 16384 rows: 2.18 against 2.29 ms)
  hyb::fetch<M, CONTIG, true>(L, a, ra, REAL(0), 0, first_seg ? 1 : 0, L.N);
  hyb::fetch<M, CONTIG, true>(L, b, rb, REAL(1), 0, 0, L.N);
  hyb::fetch<M, CONTIG, true>(L, c, rc, REAL(0), 0, 0, last_seg ? L.N - 1 : L.N);
  hyb::fetch<M, CONTIG, true>(L, d, rd, REAL(0), 0, 0, L.N);
  hyb::to_chunk<M, CONTIG>(L, ra, tile);
  hyb::to_chunk<M, CONTIG>(L, rb, tile);
  hyb::to_chunk<M, CONTIG>(L, rc, tile);
  hyb::to_chunk<M, CONTIG>(L, rd, tile);
}

// Index of chunk k of system s in the reduced arrays.
template <bool CONTIG>
__device__ __forceinline__ int64_t long_idx(int64_t s, int64_t k, int64_t nsys, int64_t nchunk) {
  return CONTIG ? s * nchunk + k : k * nsys + s;
}

template <typename REAL, bool CONTIG>
__global__ __launch_bounds__((LongShape<REAL, CONTIG>::T)) void long_reduce_kernel(
    Geom g, int64_t nseg, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, const REAL *__restrict__ d, REAL *__restrict__ rec) {
  using S = LongShape<REAL, CONTIG>;
  constexpr int M = LONG_M;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int64_t seg, sys;
  const hyb::Lane L = long_lane<REAL, CONTIG>(g, nseg, &seg, &sys);
  REAL *tile = reinterpret_cast<REAL *>(smem_raw);
  REAL ra[M], rb[M], rc[M], rd[M];
  long_fetch<REAL, CONTIG>(L, seg == 0, seg == nseg - 1, a, b, c, d, tile, ra, rb, rc, rd);
  hyb::forward_spike<REAL, M>(ra, rb, rc, rd);
  const REAL al_e = ra[M - 1], ga_e = rc[M - 1], de_e = rd[M - 1];
  REAL f_d, f_a, f_g;
  hyb::backward_spike<REAL, M>(ra, rc, rd, f_d, f_a, f_g);
  if (!L.active) return;
  const int64_t nsys = g.nsys(), nchunk = nseg * S::P;
  const int64_t R = nsys * nchunk;
  const int64_t i = long_idx<CONTIG>(sys, seg * S::P + L.p, nsys, nchunk);
  rec[i] = al_e;
  rec[R + i] = ga_e;
  rec[2 * R + i] = de_e;
  rec[3 * R + i] = f_d;
  rec[4 * R + i] = f_a;
  rec[5 * R + i] = f_g;
}

// One thread per chunk: A_k, 1, C_k, D_k of the reduced system.
template <typename REAL, bool CONTIG>
__global__ __launch_bounds__(256) void long_build_kernel(int64_t nsys, int64_t nchunk,
                                                         const REAL *__restrict__ rec,
                                                         REAL *__restrict__ ra,
                                                         REAL *__restrict__ rb,
                                                         REAL *__restrict__ rc,
                                                         REAL *__restrict__ rd) {
  const int64_t R = nsys * nchunk;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= R) return;
  const int64_t k = CONTIG ? i % nchunk : i / nsys;
  const int64_t nxt = CONTIG ? i + 1 : i + nsys;  // chunk k+1 of the same system
  const REAL al_e = rec[i], ga_e = rec[R + i], de_e = rec[2 * R + i];
  REAL nd = 0, na = 0, ng = 0;
  if (k + 1 < nchunk) {
    nd = rec[3 * R + nxt];
    na = rec[4 * R + nxt];
    ng = rec[5 * R + nxt];
  }
  // the hybrid's phase (iii) equation, normalised as in hyb::reduced_solve
  REAL A = al_e, Cc = -ga_e * ng, D = de_e - ga_e * nd;
  const REAL inv = REAL(1) / (REAL(1) - ga_e * na);
  A *= inv; Cc *= inv; D *= inv;
  ra[i] = A;
  rb[i] = REAL(1);
  rc[i] = Cc;
  rd[i] = D;
}

template <typename REAL, bool CONTIG, bool INC>
__global__ __launch_bounds__((LongShape<REAL, CONTIG>::T)) void long_subst_kernel(
    Geom g, int64_t nseg, const REAL *__restrict__ a, const REAL *__restrict__ b,
    const REAL *__restrict__ c, REAL *__restrict__ d, REAL *__restrict__ u,
    const REAL *__restrict__ y) {
  using S = LongShape<REAL, CONTIG>;
  constexpr int M = LONG_M;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int64_t seg, sys;
  const hyb::Lane L = long_lane<REAL, CONTIG>(g, nseg, &seg, &sys);
  REAL *tile = reinterpret_cast<REAL *>(smem_raw);
  REAL ra[M], rb[M], rc[M], rd[M];
  long_fetch<REAL, CONTIG>(L, seg == 0, seg == nseg - 1, a, b, c, d, tile, ra, rb, rc, rd);
  hyb::forward_spike<REAL, M>(ra, rb, rc, rd);
  REAL f_d, f_a, f_g;
  hyb::backward_spike<REAL, M>(ra, rc, rd, f_d, f_a, f_g);
  const int64_t nsys = g.nsys(), nchunk = nseg * S::P;
  const int64_t k = seg * S::P + L.p;
  REAL yk = 0, ym = 0;
  if (L.active) {
    yk = y[long_idx<CONTIG>(sys, k, nsys, nchunk)];
    if (k > 0) ym = y[long_idx<CONTIG>(sys, k - 1, nsys, nchunk)];
  }
  REAL x[M];
  hyb::substitute<REAL, M>(ra, rc, rd, yk, ym, x);
  hyb::to_fetch<M, CONTIG>(L, x, tile);
  if (INC) hyb::store_add<M, CONTIG>(L, u, x);
  else hyb::store<M, CONTIG, true>(L, d, x);
}

template <typename REAL, bool CONTIG>
size_t long_lds() {
  using S = LongShape<REAL, CONTIG>;
  return CONTIG ? sizeof(REAL) * (size_t)S::T * (LONG_M + 1) : 0;
}

// Reduced-record reals per level: 6 records + 4 reduced arrays per chunk.
template <typename REAL>
int64_t long_chunks(const Geom &g) {
  const bool contig = g.stride == 1;
  const int64_t seg = contig ? LongShape<REAL, true>::SEG : LongShape<REAL, false>::SEG;
  const int64_t P = contig ? LongShape<REAL, true>::P : LongShape<REAL, false>::P;
  return (g.N + seg - 1) / seg * P;
}

// Geometry of the reduced systems of g (nchunk rows per system).
Geom reduced_geom(const Geom &g, int64_t nchunk) {
  Geom r;
  const int64_t nsys = g.nsys();
  r.N = nchunk;
  if (g.stride == 1) {  // [system][chunk]
    r.stride = 1; r.n0 = nsys; r.n1 = 1; r.st0 = nchunk; r.st1 = nsys * nchunk;
  } else {              // [chunk][system]
    r.stride = nsys; r.n0 = nsys; r.n1 = 1; r.st0 = 1; r.st1 = nsys;
  }
  return r;
}

template <typename REAL, bool CONTIG, bool INC>
hipError_t launch_long_kernels(const Geom &g, int64_t nseg, const REAL *a, const REAL *b,
                               const REAL *c, REAL *d, REAL *u, REAL *rec, REAL *y, bool subst,
                               hipStream_t st) {
  using S = LongShape<REAL, CONTIG>;
  const int64_t nb0 = (g.n0 + S::W - 1) / S::W;
  const int64_t blocks = nseg * nb0 * g.n1;
  if (blocks > (int64_t)UINT32_MAX / S::T) return hipErrorInvalidValue;
  const size_t lds = long_lds<REAL, CONTIG>();
  if (!subst) {
    auto kern = long_reduce_kernel<REAL, CONTIG>;
    static LdsGrant granted;
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, granted);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(S::T), lds, st, g, nseg, a, b, c, d,
                       rec);
  } else {
    auto kern = long_subst_kernel<REAL, CONTIG, INC>;
    static LdsGrant granted;
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, granted);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(S::T), lds, st, g, nseg, a, b, c, d, u,
                       y);
  }
  return hipGetLastError();
}

}  // namespace

template <typename REAL>
size_t long_workspace_bytes(const Geom &g) {
  size_t total = 0;
  Geom cur = g;
  // levels until the reduced systems fit one workgroup of the hybrid
  for (int level = 0; level < 8; ++level) {
    const int64_t nchunk = long_chunks<REAL>(cur);
    total += sizeof(REAL) * (size_t)(10 * nchunk * cur.nsys());
    Geom r = reduced_geom(cur, nchunk);
    HybridCfg cfg;
    if (hybrid_default_cfg<REAL>(r, &cfg)) break;
    cur = r;
  }
  return total;
}

template <typename REAL, bool INC>
hipError_t launch_long(int dev, const Geom &g, const REAL *a, const REAL *b, const REAL *c,
                       REAL *d, REAL *u, void *ws, size_t ws_bytes, hipStream_t st) {
  const bool contig = g.stride == 1;
  const int64_t nseg = contig ? (g.N + LongShape<REAL, true>::SEG - 1) / LongShape<REAL, true>::SEG
                              : (g.N + LongShape<REAL, false>::SEG - 1) / LongShape<REAL, false>::SEG;
  const int64_t nchunk = long_chunks<REAL>(g);
  const int64_t R = nchunk * g.nsys();
  if (ws_bytes < sizeof(REAL) * (size_t)(10 * R)) return hipErrorInvalidValue;
  REAL *rec = static_cast<REAL *>(ws);
  REAL *ra = rec + 6 * R, *rb = ra + R, *rc = rb + R, *rd = rc + R;
  hipError_t e = contig
      ? launch_long_kernels<REAL, true, INC>(g, nseg, a, b, c, d, u, rec, nullptr, false, st)
      : launch_long_kernels<REAL, false, INC>(g, nseg, a, b, c, d, u, rec, nullptr, false, st);
  if (e != hipSuccess) return e;
  {
    const int64_t blocks = (R + 255) / 256;
    if (blocks > (int64_t)UINT32_MAX) return hipErrorInvalidValue;
    if (contig)
      hipLaunchKernelGGL((long_build_kernel<REAL, true>), dim3((unsigned)blocks), dim3(256), 0,
                         st, g.nsys(), nchunk, rec, ra, rb, rc, rd);
    else
      hipLaunchKernelGGL((long_build_kernel<REAL, false>), dim3((unsigned)blocks), dim3(256), 0,
                         st, g.nsys(), nchunk, rec, ra, rb, rc, rd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // the reduced systems: the batched hybrid (tuned shape if one is cached),
  // or this path again with the rest of the workspace
  const Geom r = reduced_geom(g, nchunk);
  HybridCfg cfg;
  if (tuned_config<REAL, false>(dev, r, &cfg) || hybrid_default_cfg<REAL>(r, &cfg)) {
    e = for_each_grid_chunk(r, cfg.W, cfg.W * cfg.P, [&](const Geom &sub, int64_t o) {
      return launch_hybrid<REAL, false>(sub, cfg, ra + o, rb + o, rc + o, rd + o, nullptr, st);
    });
  } else {
    e = launch_long<REAL, false>(dev, r, ra, rb, rc, rd, nullptr, rd + R,
                                 ws_bytes - sizeof(REAL) * (size_t)(10 * R), st);
  }
  if (e != hipSuccess) return e;
  return contig
      ? launch_long_kernels<REAL, true, INC>(g, nseg, a, b, c, d, u, nullptr, rd, true, st)
      : launch_long_kernels<REAL, false, INC>(g, nseg, a, b, c, d, u, nullptr, rd, true, st);
}

#define TRID_LONG_INST(REAL)                                                                  \
  template size_t long_workspace_bytes<REAL>(const Geom &);                                  \
  template hipError_t launch_long<REAL, false>(int, const Geom &, const REAL *, const REAL *, \
                                               const REAL *, REAL *, REAL *, void *, size_t,  \
                                               hipStream_t);                                  \
  template hipError_t launch_long<REAL, true>(int, const Geom &, const REAL *, const REAL *,  \
                                              const REAL *, REAL *, REAL *, void *, size_t,   \
                                              hipStream_t);
TRID_LONG_INST(float)
TRID_LONG_INST(double)

}  // namespace trid
