// mixed_core.hpp - one chunk of the FP64 solve by FP32 sweeps plus FP64
// iterative refinement (device code shared by the strided and ragged
// kernels). No counterpart in the reference.
//
// The lane holds the FP64 operands B, D of its M rows (chunk order) in
// registers; A and Cc are read through accessors (registers, or an LDS tile
// when the register file cannot hold all four operands).
//   1. the matrix is factored in FP32: the forward spike sweep's multipliers
//      (1/beta_i, ga_i, al_i per row), the chunk's first row in reduced form
//      (backward recurrences, scalars only) and, for every PCR step, the
//      pre-step coefficients (A, C, 1/(1 - A Cm - C Ap)) -- none of these
//      depend on the right-hand side;
//   2. x0 = FP32 solve of (float) d with that factorisation (right-hand-side
//      recurrences only: de_i = (rhs_i - a_i de_{i-1}) / beta_i, the first
//      row's dp_0 by the backward recurrence dp_i = de_i - ga_i dp_{i+1}, PCR
//      on D only, then x_i = de_i - al_i x_{s-1} - ga_i x_{i+1} from the
//      chunk's end);
//   3. `refine` times: r = d - A x in FP64 (neighbour chunks' boundary x via
//      LDS), dx = FP32 solve of (float) r, x += dx in FP64.
// The residual and the accumulation are the only FP64 arithmetic. P, the
// lanes per system, is at most 2^STEPS (STEPS <= MIXED_MAX_STEPS; a caller
// whose workgroups have fewer lanes passes fewer steps and saves registers).

#pragma once

#include "hybrid_core.hpp"

namespace trid {

constexpr int MIXED_MAX_STEPS = 9;  // PCR steps for P <= 512

// Aof(r), Cof(r): the chunk's sub- and super-diagonal at row r (0 where
// masked). The operands are FP64 (TB = double), or FP32 values that define
// the system exactly (TB = float: the FP64 residual promotes them). The ring
// needs slots of 2 doubles per thread. With STORE_INV false the pivots
// 1/beta_i are recomputed by every right-hand-side solve instead of being
// kept in registers (same expression, same values). ONE is for callers that
// never ask for more than one refinement: x0 stays in FP32 registers (its
// FP64 value is exact), and x = x0 + dx is formed in FP64 only at the end --
// the same values as the general loop, with M fewer registers live. FAST
// takes the FP32 reciprocals of the factorisation from the hardware's
// approximate reciprocal (v_rcp_f32, about 1 ulp) instead of a correctly
// rounded division: the factorisation only has to be close to A's, the FP64
// residual and the refinement make up the difference.
template <int M, int STEPS, bool STORE_INV = true, bool ONE = false, bool FAST = false,
          typename TB, typename GetA, typename GetC>
__device__ __forceinline__ void mixed_chunk_acc(const hyb::Lane &L, hyb::Ring &ring, GetA Aof,
                                                const TB (&B)[M], GetC Cof, const TB (&D)[M],
                                                int refine, double (&X)[M]) {
  const int T = L.T, t = L.t, p = L.p, ps = L.pstep, P = L.P;

  // ---- FP32 factorisation ------------------------------------------------
  // Per row: inv = 1/beta, ga and the forward spike al. The backward sweep
  // is kept only for the chunk's first row (f_a, f_g); the substitution runs
  // from the chunk's end on the forward values, so ap and gp are not stored.
  float inv[STORE_INV ? M : 1], ga[M], al[M];
  auto rcp = [](float v) -> float { return FAST ? __builtin_amdgcn_rcpf(v) : 1.0f / v; };
  auto pivot = [&](int r) -> float {  // 1/beta_r, given ga[r-1]
    return r == 0 ? rcp((float)B[0]) : rcp((float)B[r] - (float)Aof(r) * ga[r - 1]);
  };
  {
    const float i0 = pivot(0);
    if (STORE_INV) inv[0] = i0;
    al[0] = (float)Aof(0) * i0;
    ga[0] = (float)Cof(0) * i0;
  }
#pragma unroll
  for (int r = 1; r < M; ++r) {
    const float ar = (float)Aof(r);
    const float ir = pivot(r);
    if (STORE_INV) inv[STORE_INV ? r : 0] = ir;
    al[r] = -ar * al[r - 1] * ir;
    ga[r] = (float)Cof(r) * ir;
  }
  auto inv_of = [&](int r) -> float { return STORE_INV ? inv[STORE_INV ? r : 0] : pivot(r); };
  const float al_e = al[M - 1], ga_e = ga[M - 1];
  float f_a = 0.0f, f_g = -1.0f;
  if (M >= 2) {
    f_a = al[M - 2];
    f_g = ga[M - 2];
#pragma unroll
    for (int r = M - 3; r >= 0; --r) {
      f_a = al[r] - ga[r] * f_a;
      f_g = -ga[r] * f_g;
    }
  }
  float na = 0, ng = 0;
  {
    float *q = ring.next<float>();
    q[t] = f_a; q[T + t] = f_g;
    __syncthreads();
    if (p + 1 < P) { na = q[t + ps]; ng = q[T + t + ps]; }
  }
  const float invB = rcp(1.0f - ga_e * na);
  float RA = al_e * invB, RC = -ga_e * ng * invB;
  float sA[STEPS], sC[STEPS], sR[STEPS];
#pragma unroll
  for (int s = 0; s < STEPS; ++s) {
    const int h = 1 << s;
    sA[s] = RA; sC[s] = RC; sR[s] = 1.0f;
    if (h < P) {
      float *q = ring.next<float>();
      q[t] = RA; q[T + t] = RC;
      __syncthreads();
      float Am = 0, Cm = 0, Ap = 0, Cp = 0;
      if (p - h >= 0) { Am = q[t - h * ps]; Cm = q[T + t - h * ps]; }
      if (p + h < P) { Ap = q[t + h * ps]; Cp = q[T + t + h * ps]; }
      const float rs = rcp(1.0f - RA * Cm - RC * Ap);
      sR[s] = rs;
      RA = -rs * RA * Am;
      RC = -rs * RC * Cp;
    }
  }

  // ---- FP32 right-hand-side solve with the stored factorisation ----------
  auto solve_rhs = [&](float (&v)[M]) {
    v[0] = v[0] * inv_of(0);
#pragma unroll
    for (int r = 1; r < M; ++r) v[r] = (v[r] - (float)Aof(r) * v[r - 1]) * inv_of(r);
    const float de_e = v[M - 1];
    float f_d = 0.0f;
    if (M >= 2) {
      f_d = v[M - 2];
#pragma unroll
      for (int r = M - 3; r >= 0; --r) f_d = v[r] - ga[r] * f_d;
    }
    float nd = 0;
    {
      float *q = ring.next<float>();
      q[t] = f_d;
      __syncthreads();
      if (p + 1 < P) nd = q[t + ps];
    }
    float Dr = (de_e - ga_e * nd) * invB;
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int h = 1 << s;
      if (h < P) {
        float *q = ring.next<float>();
        q[t] = Dr;
        __syncthreads();
        const float Dm = (p - h >= 0) ? q[t - h * ps] : 0.0f;
        const float Dp = (p + h < P) ? q[t + h * ps] : 0.0f;
        Dr = sR[s] * (Dr - sA[s] * Dm - sC[s] * Dp);
      }
    }
    const float y = Dr;
    float ym = 0;
    {
      float *q = ring.next<float>();
      q[t] = y;
      __syncthreads();
      if (p > 0) ym = q[t - ps];
    }
    v[M - 1] = y;
#pragma unroll
    for (int r = M - 2; r >= 0; --r) v[r] = v[r] - al[r] * ym - ga[r] * v[r + 1];
  };

  if constexpr (ONE) {
    float v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) v[r] = (float)D[r];
    solve_rhs(v);
    if (refine > 0) {
      float xprev = 0, xnext = 0;  // neighbour chunks' boundary values of x0
      {
        float *q = ring.next<float>();
        q[t] = v[0]; q[T + t] = v[M - 1];
        __syncthreads();
        if (p > 0) xprev = q[T + t - ps];
        if (p + 1 < P) xnext = q[t + ps];
      }
      float w[M];
#pragma unroll
      for (int r = 0; r < M; ++r) {
        const double xl = r > 0 ? v[r - 1] : xprev;
        const double xr = r < M - 1 ? v[r + 1] : xnext;
        const double res = (double)D[r] - ((double)B[r] * (double)v[r] + (double)Aof(r) * xl +
                                           (double)Cof(r) * xr);
        w[r] = (float)res;
      }
      solve_rhs(w);
#pragma unroll
      for (int r = 0; r < M; ++r) X[r] = (double)v[r] + (double)w[r];
    } else {
#pragma unroll
      for (int r = 0; r < M; ++r) X[r] = (double)v[r];
    }
    return;
  }
  {
    float v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) v[r] = (float)D[r];
    solve_rhs(v);
#pragma unroll
    for (int r = 0; r < M; ++r) X[r] = (double)v[r];
  }
  for (int it = 0; it < refine; ++it) {
    // neighbour chunks' boundary values of x
    double xprev = 0, xnext = 0;
    {
      double *q = ring.next<double>();
      q[t] = X[0]; q[T + t] = X[M - 1];
      __syncthreads();
      if (p > 0) xprev = q[T + t - ps];
      if (p + 1 < P) xnext = q[t + ps];
    }
    float v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const double xl = r > 0 ? X[r - 1] : xprev;
      const double xr = r < M - 1 ? X[r + 1] : xnext;
      const double res = (double)D[r] - ((double)B[r] * X[r] + (double)Aof(r) * xl +
                                         (double)Cof(r) * xr);
      v[r] = (float)res;
    }
    solve_rhs(v);
#pragma unroll
    for (int r = 0; r < M; ++r) X[r] += (double)v[r];
  }
}

// All four operands in registers.
template <int M>
__device__ __forceinline__ void mixed_chunk(const hyb::Lane &L, hyb::Ring &ring,
                                            const double (&A)[M], const double (&B)[M],
                                            const double (&Cc)[M], const double (&D)[M],
                                            int refine, double (&X)[M]) {
  mixed_chunk_acc<M, MIXED_MAX_STEPS, true, false, true>(
      L, ring, [&](int r) { return A[r]; }, B, [&](int r) { return Cc[r]; }, D, refine, X);
}

// Contiguous system (lane p reads rows p + P*r) with A and Cc kept in this
// system's LDS tiles tA, tC of P*(M+1) doubles each: b and d are transposed
// to chunk order through tA first, then a and c are written there skewed and
// stay; lane p reads only its own chunk slots p*(M+1)+r of them, so X goes
// back to fetch order through tA without a barrier before the writes. `ring`
// needs slots of 2 doubles per thread. `fetched()` runs once the loads are
// issued and a barrier has passed (a persistent caller decodes its next
// item there).
// Rows are 32-bit here (p + P*r < P*M <= 8192) and each array is addressed
// from one per-lane pointer, so no 64-bit row index per row stays live (the
// shared fetch's would, through to the store, and spill at this register
// pressure). NT: nontemporal loads and stores (hyb::ld / hyb::st).
// `store_base()` gives L.base again for the stores: a caller whose register
// file is full returns it from an LDS slot, so that no 64-bit base stays
// live across the solve.
template <int M, int STEPS, bool NT, typename OnFetched, typename StoreBase>
__device__ __forceinline__ void mixed_contig_lds(const hyb::Lane &L, hyb::Ring &ring,
                                                 const double *__restrict__ a,
                                                 const double *__restrict__ b,
                                                 const double *__restrict__ c,
                                                 double *__restrict__ d, double *tA, double *tC,
                                                 int refine, OnFetched fetched,
                                                 StoreBase store_base) {
  const int P = L.P, p = L.p, N = L.active ? (int)L.N : 0;
  const int64_t lb = L.base + p;
  double A[M], B[M], Cc[M], D[M];
  {
    const double *pa = a + lb, *pb = b + lb, *pc = c + lb, *pd = d + lb;
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const int row = p + P * r;
      const bool ok = row < N;
      A[r] = (ok && row > 0) ? hyb::ld<NT>(pa + P * r) : 0.0;
      B[r] = ok ? hyb::ld<NT>(pb + P * r) : 1.0;
      Cc[r] = (row < N - 1) ? hyb::ld<NT>(pc + P * r) : 0.0;
      D[r] = ok ? hyb::ld<NT>(pd + P * r) : 0.0;
    }
  }
  hyb::to_chunk<M, true>(L, B, tA);
  fetched();
  hyb::to_chunk<M, true>(L, D, tA);
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int e = hyb::skew<M>(L.p + L.P * r);
    tA[e] = A[r];
    tC[e] = Cc[r];
  }
  __syncthreads();
  const double *ca = tA + L.p * (M + 1), *cc = tC + L.p * (M + 1);
  double X[M];
  mixed_chunk_acc<M, STEPS, true, false, true>(
      L, ring, [&](int r) { return ca[r]; }, B, [&](int r) { return cc[r]; }, D, refine, X);
  hyb::to_fetch<M, true>(L, X, tA);
  // the store base is recomputed behind an empty asm, so that the load
  // addresses are not kept live across the solve
  int64_t sb = store_base() + p;
  asm volatile("" : "+v"(sb));
  double *qd = d + sb;
#pragma unroll
  for (int r = 0; r < M; ++r)
    if (p + P * r < N) hyb::st<NT>(qd + P * r, X[r]);
}

}  // namespace trid
