// hybrid_core.hpp - device building blocks of the Thomas-PCR hybrid kernels
// (plain solve, mixed precision, adjoint, ragged). The algorithm is described
// in kernels_solve.hip; this header holds its phases:
//
//   Lane           which system / chunk a thread owns inside its workgroup
//   fetch / store  masked global access in "fetch order" (strided: the chunk's
//                  own rows p*M+r; contiguous: coalesced rows p + P*r)
//   to_chunk /     transposes between fetch order and chunk order through a
//   to_fetch       skewed LDS tile (contiguous systems only)
//   Ring           two LDS buffers used alternately for the lane exchanges;
//                  one barrier per exchange is enough because a thread that
//                  passed exchange k's barrier knows every thread finished
//                  reading exchange k-1's buffer
//   from_above /   lane exchanges of the reduced system: within one wave by
//   from_below /   ds_bpermute (no LDS store, no barrier) when every system's
//   from_both      lanes lie in one wave, else through the Ring
//   forward_spike / backward_spike / reduced_solve / substitute: phases (i)-(iv)
#pragma once

#include "trid_internal.h"

namespace trid {
namespace hyb {

// Workgroup ids are handed to the XCDs round-robin (id % 8). Dealing ids so
// that group q of G gets a contiguous range of system groups makes
// neighbouring groups -- which read neighbouring 128-byte row segments -- run
// on one XCD at the same time.
// Grid sizes and ids are 32-bit, so is this arithmetic (a 64-bit division
// costs several times the scalar instructions of a 32-bit one).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nb, int G) {
  if (G <= 1) return bid;
  const uint32_t g = (uint32_t)G;
  const uint32_t per = nb / g;
  if (bid >= per * g) return bid;  // tail beyond a multiple of G: identity
  return (bid % g) * per + bid / g;
}

// Skewed tile index of element e of one system: one pad slot per M elements,
// so both the coalesced order (e = p + P*r) and the chunk order (e = p*M + r)
// spread over the LDS banks.
template <int M>
__device__ __forceinline__ int skew(int e) {
  return e + e / M;
}

struct Lane {
  int t;      // thread id in the workgroup
  int p;      // chunk index inside the system, 0..P-1
  int w;      // system index inside the workgroup, 0..W-1
  int P, W, T;
  int pstep;  // thread-id distance between chunk p and p+1 of one system
  int64_t base;  // element offset of the system's first row
  int64_t N;     // system length
  int64_t stride;
  int64_t i0, i1;  // the system's coordinates in the n0 x n1 grid of systems
  bool active;     // the lane belongs to an existing system
};

// Element e of the batch (0 <= e < nsys * N), for elementwise kernels, in an
// order that keeps neighbouring threads on neighbouring addresses: along the
// system for contiguous systems, across x-neighbouring systems for strided
// ones. *k is the row inside the system; the result the element's index.
// (32-bit divisions while e < 2^32, which covers every batch of fewer than
// 2^32 elements; 64-bit ones beyond.)
template <typename U, bool CONTIG>
__device__ __forceinline__ void batch_split(U e, U N, U n0, U *k, U *i0, U *i1) {
  if (CONTIG) {
    const U s = e / N;
    *k = e - s * N;
    *i1 = s / n0;
    *i0 = s - *i1 * n0;
  } else {
    const U r = e / n0;
    *i0 = e - r * n0;
    *i1 = r / N;
    *k = r - *i1 * N;
  }
}

template <bool CONTIG>
__device__ __forceinline__ int64_t batch_elem(const Geom &g, int64_t e, int64_t *k) {
  int64_t i0, i1;
  if ((uint64_t)e >> 32 == 0) {
    uint32_t kk, j0, j1;
    batch_split<uint32_t, CONTIG>((uint32_t)e, (uint32_t)g.N, (uint32_t)g.n0, &kk, &j0, &j1);
    *k = kk; i0 = j0; i1 = j1;
  } else {
    uint64_t kk, j0, j1;
    batch_split<uint64_t, CONTIG>((uint64_t)e, (uint64_t)g.N, (uint64_t)g.n0, &kk, &j0, &j1);
    *k = (int64_t)kk; i0 = (int64_t)j0; i1 = (int64_t)j1;
  }
  return i0 * g.st0 + i1 * g.st1 + *k * g.stride;
}

// Lane of a batched (Geom) launch: system groups of W x-neighbours (strided)
// or W rows (contiguous), P lanes per system.
template <bool CONTIG>
__device__ __forceinline__ Lane lane_of(const Geom &g, int P, int W, int G) {
  Lane L;
  L.T = W * P;
  L.t = threadIdx.x;
  L.P = P;
  L.W = W;
  L.p = CONTIG ? L.t % P : L.t / W;
  L.w = CONTIG ? L.t / P : L.t % W;
  L.pstep = CONTIG ? 1 : W;
  // n0 < 2^31 and one workgroup per W systems: 32-bit block arithmetic
  const uint32_t nb0 = (uint32_t)((g.n0 + W - 1) / W);
  const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x, G);
  const int64_t i0 = (int64_t)(bid % nb0) * W + L.w;
  const int64_t i1 = bid / nb0;
  L.active = i0 < g.n0;
  L.base = i0 * g.st0 + i1 * g.st1;
  L.N = g.N;
  L.stride = g.stride;
  L.i0 = i0;
  L.i1 = i1;
  return L;
}

template <int M, bool CONTIG>
__device__ __forceinline__ int64_t fetch_row(const Lane &L, int r) {
  return CONTIG ? (int64_t)L.p + (int64_t)L.P * r : (int64_t)L.p * M + r;
}

// Element access with the nontemporal hint (NT) or without. NT marks data
// that a kernel reads or writes once. The strided-batch kernels use it:
// 512^3 FP64 x / y / z took 0.89 / 0.98 / 1.14 ms with it against 1.01 /
// 1.04 / 1.19 without. Kernels whose data a caller reuses across calls from
// the 256 MB Infinity Cache (the FP32 ragged solve, the fused ADI steps) ran
// slower with it and keep plain access.
template <bool NT, typename T>
__device__ __forceinline__ T ld(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st(T *p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// v[r] = src[row(r) + shift] for the lane's fetch-order rows; rows >= N,
// inactive lanes and shifted rows outside [lo, hi) get `fill`.
template <int M, bool CONTIG, bool NT = false, typename TI, typename TR>
__device__ __forceinline__ void fetch(const Lane &L, const TI *__restrict__ src, TR (&v)[M],
                                      TR fill, int shift, int64_t lo, int64_t hi) {
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int64_t row = fetch_row<M, CONTIG>(L, r);
    const int64_t sr = row + shift;
    const bool ok = L.active && row < L.N && sr >= lo && sr < hi;
    v[r] = ok ? (TR)ld<NT>(&src[L.base + sr * L.stride]) : fill;
  }
}

// Standard masked loads of a tridiagonal system: a[0] = 0, c[N-1] = 0 and
// identity rows (a=0, b=1, c=0, d=0) beyond the length.
template <int M, bool CONTIG, bool NT = false, typename TI, typename TR>
__device__ __forceinline__ void fetch_abcd(const Lane &L, const TI *a, const TI *b,
                                           const TI *c, const TI *d, TR (&ra)[M],
                                           TR (&rb)[M], TR (&rc)[M], TR (&rd)[M]) {
  fetch<M, CONTIG, NT>(L, a, ra, TR(0), 0, 1, L.N);
  fetch<M, CONTIG, NT>(L, b, rb, TR(1), 0, 0, L.N);
  fetch<M, CONTIG, NT>(L, c, rc, TR(0), 0, 0, L.N - 1);
  fetch<M, CONTIG, NT>(L, d, rd, TR(0), 0, 0, L.N);
}

// dst = v for the lane's rows below N. The lane's base offset passes through
// an empty asm: the row addresses are formed here, instead of the fetch's
// being kept live (in registers) across the solve between the two.
template <int M, bool CONTIG, bool NT = false, typename TR, typename TO>
__device__ __forceinline__ void store(const Lane &L, TO *__restrict__ dst, const TR (&v)[M]) {
  if (!L.active) return;
  const int64_t row0 = fetch_row<M, CONTIG>(L, 0);
  int64_t b = L.base + row0 * L.stride;
  asm volatile("" : "+v"(b));
  const int64_t step = CONTIG ? (int64_t)L.P : L.stride;
#pragma unroll
  for (int r = 0; r < M; ++r)
    if (fetch_row<M, CONTIG>(L, r) < L.N) st<NT>(&dst[b + r * step], (TO)v[r]);
}

template <int M, bool CONTIG, typename TR, typename TO>
__device__ __forceinline__ void store_add(const Lane &L, TO *__restrict__ dst,
                                          const TR (&v)[M]) {
  if (!L.active) return;
  const int64_t row0 = fetch_row<M, CONTIG>(L, 0);
  int64_t b = L.base + row0 * L.stride;
  asm volatile("" : "+v"(b));
  const int64_t step = CONTIG ? (int64_t)L.P : L.stride;
#pragma unroll
  for (int r = 0; r < M; ++r)
    if (fetch_row<M, CONTIG>(L, r) < L.N) dst[b + r * step] += (TO)v[r];
}

// ---- 16-byte vector access for contiguous systems --------------------------
// Fetch order with vectors of V elements: slot q*V + k of lane p holds element
// V*(p + P*q) + k. Used when N and the row pitch are multiples of V and every
// array is 16-byte aligned, so a whole vector is either inside [0, N) or
// beyond it.
template <typename TR, int V>
struct VecOf;
template <>
struct VecOf<double, 2> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct VecOf<float, 4> {
  typedef float type __attribute__((ext_vector_type(4)));
};

template <int M, int V>
__device__ __forceinline__ int64_t vfetch_row(const Lane &L, int q) {
  return (int64_t)V * ((int64_t)L.p + (int64_t)L.P * q);
}

template <int M, int V, bool NT = false, typename TR>
__device__ __forceinline__ void vfetch(const Lane &L, const TR *__restrict__ src, TR (&v)[M],
                                       TR fill, int64_t lo, int64_t hi) {
  typedef typename VecOf<TR, V>::type vec;
#pragma unroll
  for (int q = 0; q < M / V; ++q) {
    const int64_t e0 = vfetch_row<M, V>(L, q);
    if (L.active && e0 < L.N) {
      const vec x = ld<NT>(reinterpret_cast<const vec *>(src + L.base + e0));
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const int64_t e = e0 + k;
        v[q * V + k] = (e >= lo && e < hi) ? x[k] : fill;
      }
    } else {
#pragma unroll
      for (int k = 0; k < V; ++k) v[q * V + k] = fill;
    }
  }
}

template <int M, int V, bool NT = false, typename TR>
__device__ __forceinline__ void vfetch_abcd(const Lane &L, const TR *a, const TR *b,
                                            const TR *c, const TR *d, TR (&ra)[M],
                                            TR (&rb)[M], TR (&rc)[M], TR (&rd)[M]) {
  vfetch<M, V, NT>(L, a, ra, TR(0), 1, L.N);
  vfetch<M, V, NT>(L, b, rb, TR(1), 0, L.N);
  vfetch<M, V, NT>(L, c, rc, TR(0), 0, L.N - 1);
  vfetch<M, V, NT>(L, d, rd, TR(0), 0, L.N);
}

// (NT applies to the plain store; the Inc form reads what it adds to.)
template <int M, int V, bool ADD, bool NT = false, typename TR>
__device__ __forceinline__ void vstore(const Lane &L, TR *__restrict__ dst, const TR (&v)[M]) {
  typedef typename VecOf<TR, V>::type vec;
  if (!L.active) return;
#pragma unroll
  for (int q = 0; q < M / V; ++q) {
    const int64_t e0 = vfetch_row<M, V>(L, q);
    if (e0 < L.N) {
      vec *p = reinterpret_cast<vec *>(dst + L.base + e0);
      vec x;
      if (ADD) x = *p;
#pragma unroll
      for (int k = 0; k < V; ++k) x[k] = ADD ? x[k] + v[q * V + k] : v[q * V + k];
      st<NT && !ADD>(p, x);
    }
  }
}

template <int M, int V, typename TR>
__device__ __forceinline__ void vto_chunk(const Lane &L, TR (&v)[M], TR *tile) {
#pragma unroll
  for (int q = 0; q < M / V; ++q)
#pragma unroll
    for (int k = 0; k < V; ++k) tile[skew<M>((int)vfetch_row<M, V>(L, q) + k)] = v[q * V + k];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < M; ++r) v[r] = tile[L.p * (M + 1) + r];
  __syncthreads();
}

template <int M, int V, typename TR>
__device__ __forceinline__ void vto_fetch(const Lane &L, TR (&v)[M], TR *tile) {
#pragma unroll
  for (int r = 0; r < M; ++r) tile[L.p * (M + 1) + r] = v[r];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < M / V; ++q)
#pragma unroll
    for (int k = 0; k < V; ++k) v[q * V + k] = tile[skew<M>((int)vfetch_row<M, V>(L, q) + k)];
  __syncthreads();
}

// Fetch order -> chunk order (contiguous systems). `tile` is this system's
// P*(M+1)-element LDS tile.
template <int M, bool CONTIG, typename TR>
__device__ __forceinline__ void to_chunk(const Lane &L, TR (&v)[M], TR *tile) {
  if (!CONTIG) return;
#pragma unroll
  for (int r = 0; r < M; ++r) tile[skew<M>(L.p + L.P * r)] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < M; ++r) v[r] = tile[L.p * (M + 1) + r];
  __syncthreads();
}

// Chunk order -> fetch order (contiguous systems).
template <int M, bool CONTIG, typename TR>
__device__ __forceinline__ void to_fetch(const Lane &L, TR (&v)[M], TR *tile) {
  if (!CONTIG) return;
#pragma unroll
  for (int r = 0; r < M; ++r) tile[L.p * (M + 1) + r] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < M; ++r) v[r] = tile[skew<M>(L.p + L.P * r)];
  __syncthreads();
}

// LDS exchange buffers of `slot` bytes. With two slots they alternate and an
// exchange needs one barrier; with one slot (nslot = 1, when LDS is short) the
// reader calls release() to add the barrier that protects the next write.
struct Ring {
  unsigned char *base;
  int slot;
  int cur;
  int nslot = 2;
  template <typename X>
  __device__ __forceinline__ X *next() {
    X *q = reinterpret_cast<X *>(base + (size_t)cur * slot);
    if (nslot == 2) cur ^= 1;
    return q;
  }
  __device__ __forceinline__ void release() const {
    if (nslot == 1) __syncthreads();
  }
};

// Bytes of LDS a ring needs for up to `words` values of type X per thread.
template <typename X>
constexpr size_t ring_bytes(int T, int words) {
  return 2 * sizeof(X) * (size_t)words * T;
}

// (i) forward sweep with left spike; (ra, rc, rd) become (al, ga, de).
template <typename TR, int M>
__device__ __forceinline__ void forward_spike(TR (&ra)[M], const TR (&rb)[M], TR (&rc)[M],
                                              TR (&rd)[M]) {
  TR inv = TR(1) / rb[0];
  ra[0] = ra[0] * inv;
  rc[0] = rc[0] * inv;
  rd[0] = rd[0] * inv;
#pragma unroll
  for (int r = 1; r < M; ++r) {
    const TR ar = ra[r];
    inv = TR(1) / (rb[r] - ar * rc[r - 1]);
    ra[r] = -ar * ra[r - 1] * inv;
    rc[r] = rc[r] * inv;
    rd[r] = (rd[r] - ar * rd[r - 1]) * inv;
  }
}

// (ii) backward sweep over rows 0..M-2: (rd, ra, rc) <- (dp, ap, gp); row M-2
// keeps its forward values. Returns the chunk's first row in reduced form.
template <typename TR, int M>
__device__ __forceinline__ void backward_spike(TR (&ra)[M], TR (&rc)[M], TR (&rd)[M],
                                               TR &f_d, TR &f_a, TR &f_g) {
#pragma unroll
  for (int r = M - 3; r >= 0; --r) {
    const TR gr = rc[r];
    rd[r] = rd[r] - gr * rd[r + 1];
    ra[r] = ra[r] - gr * ra[r + 1];
    rc[r] = -gr * rc[r + 1];
  }
  if (M >= 2) {
    f_d = rd[0]; f_a = ra[0]; f_g = rc[0];
  } else {
    f_d = TR(0); f_a = TR(0); f_g = TR(-1);
  }
}

// ---- lane exchanges ----------------------------------------------------------
// A system's P lanes lie in one wave when the workgroup is one wave, or when
// the system's lanes are consecutive (pstep 1) and P divides 64 (systems then
// start at multiples of P). Such systems exchange values with ds_bpermute:
// wave-local, no LDS store and no workgroup barrier. The values moved are
// the same as through LDS, so results do not depend on the path.
__device__ __forceinline__ bool wave_local(const Lane &L) {
  return L.T <= 64 || (L.pstep == 1 && L.P <= 64 && (L.P & (L.P - 1)) == 0);
}

// v of lane `src` of this wave. ds_bpermute moves 32 bits per lane and reads
// the lane given by bits 7:2 of the byte address, so src is taken modulo 64;
// callers discard values from lanes outside their system.
template <typename X>
__device__ __forceinline__ X wave_read(X v, int src) {
  static_assert(sizeof(X) % 4 == 0, "32-bit words");
  constexpr int NW = sizeof(X) / 4;
  int w[NW];
  __builtin_memcpy(w, &v, sizeof(X));
#pragma unroll
  for (int k = 0; k < NW; ++k) w[k] = __builtin_amdgcn_ds_bpermute(src << 2, w[k]);
  X r;
  __builtin_memcpy(&r, w, sizeof(X));
  return r;
}

// vm[k] = v[k] of thread t - dm (0 unless okm) and vp[k] = v[k] of thread
// t + dp (0 unless okp), for the K values of every thread of the workgroup.
// Every thread calls it (it contains a barrier on the LDS path). SIDES: 1
// below only, 2 above only, 3 both.
template <int SIDES, int K, typename X>
__device__ __forceinline__ void exchange(const Lane &L, Ring &ring, bool wave, const X (&v)[K],
                                         int dm, bool okm, X (&vm)[K], int dp, bool okp,
                                         X (&vp)[K]) {
  if (wave) {
    const int lane = L.t & 63;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (SIDES & 1) {
        const X m = wave_read(v[k], lane - dm);
        vm[k] = okm ? m : X(0);
      }
      if (SIDES & 2) {
        const X q = wave_read(v[k], lane + dp);
        vp[k] = okp ? q : X(0);
      }
    }
    return;
  }
  X *q = ring.next<X>();
#pragma unroll
  for (int k = 0; k < K; ++k) q[k * L.T + L.t] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (SIDES & 1) vm[k] = X(0);
    if (SIDES & 2) vp[k] = X(0);
  }
  if (SIDES & 1) {
    if (okm) {
      const int sl = L.t - dm;
#pragma unroll
      for (int k = 0; k < K; ++k) vm[k] = q[k * L.T + sl];
    }
  }
  if (SIDES & 2) {
    if (okp) {
      const int sl = L.t + dp;
#pragma unroll
      for (int k = 0; k < K; ++k) vp[k] = q[k * L.T + sl];
    }
  }
  ring.release();
}

// K values of chunk p+1 (0 for the last chunk).
template <int K, typename X>
__device__ __forceinline__ void from_above(const Lane &L, Ring &ring, bool wave,
                                           const X (&v)[K], X (&vp)[K]) {
  X unused[K];
  exchange<2, K>(L, ring, wave, v, 0, false, unused, L.pstep, L.p + 1 < L.P, vp);
}

// K values of chunk p-1 (0 for the first chunk).
template <int K, typename X>
__device__ __forceinline__ void from_below(const Lane &L, Ring &ring, bool wave,
                                           const X (&v)[K], X (&vm)[K]) {
  X unused[K];
  exchange<1, K>(L, ring, wave, v, L.pstep, L.p > 0, vm, 0, false, unused);
}

// K values of chunks p-h and p+h (0 outside the system): one PCR step.
template <int K, typename X>
__device__ __forceinline__ void from_both(const Lane &L, Ring &ring, bool wave, int h,
                                          const X (&v)[K], X (&vm)[K], X (&vp)[K]) {
  exchange<3, K>(L, ring, wave, v, h * L.pstep, L.p - h >= 0, vm, h * L.pstep, L.p + h < L.P,
                 vp);
}

// (iii) reduced equation of this chunk from its last-row forward values and
// chunk p+1's first row, then PCR over the P chunks of each system. Returns
// y = x at this chunk's last row and ym = x at the previous chunk's last row.
// WAVE: the kernel's systems are contiguous (pstep 1), so they may lie in one
// wave; strided kernels pass false and keep only the LDS path (the wave path
// would add registers to kernels whose systems span waves).
template <typename TR, bool WAVE = false>
__device__ __forceinline__ void reduced_solve(const Lane &L, Ring &ring, TR al_e, TR ga_e,
                                              TR de_e, TR f_d, TR f_a, TR f_g, TR &y,
                                              TR &ym) {
  const int P = L.P;
  const bool wave = WAVE && wave_local(L);
  TR nx[3];
  {
    const TR f[3] = {f_d, f_a, f_g};
    from_above<3>(L, ring, wave, f, nx);
  }
  const TR nd = nx[0], na = nx[1], ng = nx[2];
  TR A = al_e, Cc = -ga_e * ng, D = de_e - ga_e * nd;
  {
    const TR inv = TR(1) / (TR(1) - ga_e * na);
    A *= inv; Cc *= inv; D *= inv;
  }
  for (int h = 1; h < P; h <<= 1) {
    const TR v[3] = {A, Cc, D};
    TR m[3], q[3];
    from_both<3>(L, ring, wave, h, v, m, q);
    const TR Am = m[0], Cm = m[1], Dm = m[2], Ap = q[0], Cp = q[1], Dp = q[2];
    const TR inv = TR(1) / (TR(1) - A * Cm - Cc * Ap);
    const TR nA = -inv * A * Am;
    const TR nC = -inv * Cc * Cp;
    const TR nD = inv * (D - A * Dm - Cc * Dp);
    A = nA; Cc = nC; D = nD;
  }
  y = D;
  TR yv[1] = {y}, ymv[1];
  from_below<1>(L, ring, wave, yv, ymv);
  ym = ymv[0];
}

// Phases (i)-(iii) with the chunk's `a` kept in LDS: al_s[r * S] holds a of
// the lane's row r (not read for the system's first row: first_chunk). The
// forward sweep overwrites it with al, the backward sweep with ap; rc and rd
// become gp and dp. y and ym as from reduced_solve. The arithmetic is
// solve_chunk's; only where a lives differs. One row's LDS traffic at a time
// (scheduling barriers), so the compiler does not hoist every LDS read.
template <typename TR, int M, int S>
__device__ __forceinline__ void sweep_lds_a(const Lane &L, Ring &ring, TR *al_s,
                                            bool first_chunk, const TR (&rb)[M], TR (&rc)[M],
                                            TR (&rd)[M], TR &y, TR &ym) {
  // (i) forward sweep with left spike; al in the tile, ga in rc, de in rd
  TR al_prev = 0;
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const TR ar = (r > 0 || !first_chunk) ? al_s[r * S] : TR(0);  // a[0] is not used
    TR al;
    if (r == 0) {
      const TR inv = TR(1) / rb[0];
      al = ar * inv;
      rc[0] = rc[0] * inv;
      rd[0] = rd[0] * inv;
    } else {
      const TR inv = TR(1) / (rb[r] - ar * rc[r - 1]);
      al = -ar * al_prev * inv;
      rc[r] = rc[r] * inv;
      rd[r] = (rd[r] - ar * rd[r - 1]) * inv;
    }
    al_s[r * S] = al;
    al_prev = al;
    __builtin_amdgcn_sched_barrier(0);
  }
  const TR al_e = al_prev, ga_e = rc[M - 1], de_e = rd[M - 1];
  // (ii) backward sweep; ap replaces al in the tile, gp and dp replace ga, de
  TR ap_next = M >= 2 ? al_s[(M - 2) * S] : TR(0);
#pragma unroll
  for (int r = M - 3; r >= 0; --r) {
    const TR gr = rc[r];
    rd[r] = rd[r] - gr * rd[r + 1];
    const TR ap = al_s[r * S] - gr * ap_next;
    al_s[r * S] = ap;
    ap_next = ap;
    rc[r] = -gr * rc[r + 1];
    __builtin_amdgcn_sched_barrier(0);
  }
  TR f_d, f_a, f_g;
  if (M >= 2) {
    f_d = rd[0]; f_a = ap_next; f_g = rc[0];
  } else {
    f_d = TR(0); f_a = TR(0); f_g = TR(-1);
  }
  // (iii) reduced system across the P chunks
  reduced_solve<TR>(L, ring, al_e, ga_e, de_e, f_d, f_a, f_g, y, ym);
}

// (iv) x_r = dp_r - ap_r ym - gp_r y; x_{M-1} = y.
template <typename TR, int M>
__device__ __forceinline__ void substitute(const TR (&ra)[M], const TR (&rc)[M],
                                           const TR (&rd)[M], TR y, TR ym, TR (&x)[M]) {
#pragma unroll
  for (int r = 0; r < M - 1; ++r) x[r] = rd[r] - ra[r] * ym - rc[r] * y;
  x[M - 1] = y;
}

// Whole solve of the lane's chunk, given a, b, c, d in chunk order; x in
// chunk order on return. a, c, d are overwritten. WAVE as for reduced_solve.
template <typename TR, int M, bool WAVE = false>
__device__ __forceinline__ void solve_chunk(const Lane &L, Ring &ring, TR (&ra)[M],
                                            const TR (&rb)[M], TR (&rc)[M], TR (&rd)[M],
                                            TR (&x)[M]) {
  forward_spike<TR, M>(ra, rb, rc, rd);
  const TR al_e = ra[M - 1], ga_e = rc[M - 1], de_e = rd[M - 1];
  TR f_d, f_a, f_g;
  backward_spike<TR, M>(ra, rc, rd, f_d, f_a, f_g);
  TR y, ym;
  reduced_solve<TR, WAVE>(L, ring, al_e, ga_e, de_e, f_d, f_a, f_g, y, ym);
  substitute<TR, M>(ra, rc, rd, y, ym, x);
}

// ---- one long system in one workgroup, 32 rows per lane -----------------
// 32 rows per lane, T lanes: systems of 4097-8192 rows in 256 threads, 8193-
// 16384 in 512. b, c and d go to chunk order through the LDS tile; `a` then
// goes from HBM into the same tile by global->LDS loads, in chunk order, and
// stays there through the sweeps (sweep_lds_a), so the lane holds 32 rows of
// b, c, d only (FP64: about 256 VGPRs, two waves per SIMD: two 256-thread
// workgroups per CU or one of 512). The tile is skewed with one pad slot per
// 32 rows (skew<32>, as the transposes use), so lane p reads its rows at
// slots 33p + r. A global->LDS wave instruction writes 256 contiguous LDS
// bytes (lane l: bytes 4l .. 4l+3 from its own source): FP64, rows 32k ..
// 32k+31 to slots 33k ..; FP32, rows 64j .. 64j+63 by two instructions with
// complementary lane masks, lanes 0-31 (chunk 2j) to slots 66j .. and lanes
// 32-63 (chunk 2j+1) from base slot 66j + 1, which puts row 64j + l at slot
// 66j + 1 + l = 33(2j+1) + (l - 32). base and N are the system's first
// element and length (N <= 32 T), the same for every lane: callers pass them
// through wave_uniform / readfirstlane, so they sit in scalar registers and
// rows are 32-bit offsets from base. The arithmetic is solve_chunk's.
constexpr int LONG32_M = 32, LONG32_T = 256, LONG32_T_MAX = 512;

// LDS of long32_solve at T lanes: the tile and the reduced solve's ring
// (FP64 at 512 lanes: 156 KiB)
template <typename REAL, int T = LONG32_T>
constexpr size_t long32_lds_bytes() {
  return sizeof(REAL) * (size_t)T * (LONG32_M + 1) + ring_bytes<REAL>(T, 3);
}
static_assert(long32_lds_bytes<double, LONG32_T_MAX>() <= 160 * 1024, "LDS of one CU");

typedef __attribute__((address_space(3))) void lds_void;

// v of lane 0, which the compiler then holds in scalar registers (a value
// every lane of the wave has).
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// `fetched()` runs once b is in chunk order (after a barrier that follows
// the loads); a persistent caller decodes its next item there.
template <typename REAL, bool INC, int T, typename OnFetched>
__device__ __forceinline__ void long32_solve(int64_t base, int N, unsigned char *smem_raw,
                                             const REAL *__restrict__ a,
                                             const REAL *__restrict__ b,
                                             const REAL *__restrict__ c,
                                             REAL *__restrict__ d, REAL *__restrict__ u,
                                             OnFetched fetched) {
  constexpr int M = LONG32_M;
  static_assert(T % 64 == 0 && T <= LONG32_T_MAX, "whole waves, 32 T rows at most");
  const int t = threadIdx.x;
  Lane L;
  L.P = T; L.W = 1; L.T = T; L.t = t; L.p = t; L.w = 0; L.pstep = 1; L.stride = 1;
  L.active = true; L.base = base; L.N = N; L.i0 = 0; L.i1 = 0;
  REAL *tile = reinterpret_cast<REAL *>(smem_raw);
  Ring ring{smem_raw + sizeof(REAL) * (size_t)T * (M + 1), (int)(3 * sizeof(REAL) * T), 0};
  const REAL *sb = b + base, *sc = c + base, *sd = d + base;
  // b, c, d in fetch order (lane t: rows t + T*r), from rows clamped into
  // the system so that every load is unconditional; the identity rows past N
  // and c of row N-1 are set once the rows are in chunk order, where the
  // lane's rows are row0 .. row0 + 31 (one row index, constant offsets,
  // rather than 32 row indices held from the loads)
  REAL rb[M], rc[M], rd[M];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int row = t + T * r;
    const int rr = row < N ? row : N - 1;
    rb[r] = sb[rr];
    rc[r] = sc[rr];
    rd[r] = sd[rr];
  }
  // transposes through the tile, hyb::to_chunk / to_fetch at P = T, M = 32:
  // fetch-order row t + Tr sits at slot skew<32>(t + Tr) = (t + t/32) +
  // (T + T/32)r, so each lane addresses all its slots from one base with
  // constant offsets (no per-row address registers)
  const int fslot = t + (t >> 5), cslot = t * (M + 1);
  auto to_chunk = [&](REAL (&v)[M]) {
#pragma unroll
    for (int r = 0; r < M; ++r) tile[fslot + (T + T / M) * r] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < M; ++r) v[r] = tile[cslot + r];
    __syncthreads();
  };
  to_chunk(rb);
  fetched();
  to_chunk(rc);
  to_chunk(rd);  // ends with a barrier: the tile is free
  const int row0 = t * M;
  if (row0 + M >= N) {
#pragma unroll
    for (int r = 0; r < M; ++r) {
      rb[r] = row0 + r < N ? rb[r] : REAL(1);
      rc[r] = row0 + r < N - 1 ? rc[r] : REAL(0);
      rd[r] = row0 + r < N ? rd[r] : REAL(0);
    }
  }
  {
    // wave v moves the 256-byte windows k = v, v + T/64, ...; lane l the
    // dword l of window k (row k*E + l*E/64, E rows a window). Row 0's a is
    // not used and rows >= N are masked off. (The thread index passes
    // through an empty asm: the loop's per-lane values are formed here for
    // each item rather than hoisted out of the caller's item loop and held
    // across it.)
    constexpr int E = 256 / (int)sizeof(REAL);  // rows of a 256-byte window
    int tt = t;
    asm volatile("" : "+v"(tt));
    const int lane = tt & 63;
    const int rl = lane * E / 64;
    const char *src = reinterpret_cast<const char *>(a + base) + 4 * lane;
    for (int k = tt >> 6; k * E < N; k += T / 64) {
      const int row = k * E + rl;
      const bool ok = row >= 1 && row < N;
      const char *s = src + k * 256;
      if constexpr (E == M) {  // FP64: the window is chunk k
        if (ok) __builtin_amdgcn_global_load_lds(s, (lds_void *)(tile + k * (M + 1)), 4, 0, 0);
      } else {  // FP32: chunks 2k and 2k+1
        if (ok && lane < 32)
          __builtin_amdgcn_global_load_lds(s, (lds_void *)(tile + 2 * k * (M + 1)), 4, 0, 0);
        if (ok && lane >= 32)
          __builtin_amdgcn_global_load_lds(s, (lds_void *)(tile + 2 * k * (M + 1) + 1), 4, 0, 0);
      }
    }
  }
  REAL *al_s = tile + t * (M + 1);  // this lane's rows, chunk order
  // a of the identity rows past N is 0: the lane clears its own slots, which
  // no load writes
  if (row0 + M > N) {
#pragma unroll
    for (int r = 0; r < M; ++r)
      if (row0 + r >= N) al_s[r] = REAL(0);
  }
  // the tile is written by other waves' loads: wait for this wave's, then
  // for every wave
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  REAL y, ym;
  sweep_lds_a<REAL, M, 1>(L, ring, al_s, row0 == 0, rb, rc, rd, y, ym);
  // (iv) substitution into rd; each lane reads only its own slots, which
  // to_fetch then overwrites first
#pragma unroll
  for (int r = 0; r < M - 1; ++r) rd[r] = rd[r] - al_s[r] * ym - rc[r] * y;
  rd[M - 1] = y;
#pragma unroll
  for (int r = 0; r < M; ++r) tile[cslot + r] = rd[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < M; ++r) rd[r] = tile[fslot + (T + T / M) * r];
  __syncthreads();
  REAL *dst = (INC ? u : d) + base;
  // the store rows formed here from a thread index the compiler cannot
  // match with the loads' (an empty asm), so that the loads' row indices
  // are not held through the solve
  int ts = t;
  asm volatile("" : "+v"(ts));
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int row = ts + T * r;
    if (row < N) {
      // nontemporal stores at 256 lanes (8192 systems of 8192 rows: 0.50
      // against 0.53 ms), plain ones at 512 (16384 rows: 1.14 against 1.18)
      if (INC) dst[row] += rd[r];
      else st<(T == LONG32_T)>(&dst[row], rd[r]);
    }
  }
}

}  // namespace hyb
}  // namespace trid
