// mixed_core.hpp - one chunk of the FP64 solve by FP32 sweeps plus FP64
// iterative refinement (device code shared by the strided and ragged
// kernels). No counterpart in the reference.
//
// The lane holds the FP64 operands A, B, Cc, D of its M rows (chunk order).
//   1. the matrix is factored in FP32: the spike sweeps' multipliers
//      (1/beta_i, ga_i, ap_i, gp_i per row) and, for every PCR step, the
//      pre-step coefficients (A, C, 1/(1 - A Cm - C Ap)) -- none of these
//      depend on the right-hand side;
//   2. x0 = FP32 solve of (float) d with that factorisation (right-hand-side
//      recurrences only: de_i = (rhs_i - a_i de_{i-1}) / beta_i, backward
//      dp_i = de_i - ga_i dp_{i+1}, PCR on D only, substitution);
//   3. `refine` times: r = d - A x in FP64 (neighbour chunks' boundary x via
//      LDS), dx = FP32 solve of (float) r, x += dx in FP64.
// The residual and the accumulation are the only FP64 arithmetic. P, the
// lanes per system, is at most 2^MIXED_MAX_STEPS.

#pragma once

#include "hybrid_core.hpp"

namespace trid {

constexpr int MIXED_MAX_STEPS = 9;  // PCR steps for P <= 512

template <int M>
__device__ __forceinline__ void mixed_chunk(const hyb::Lane &L, hyb::Ring &ring,
                                            const double (&A)[M], const double (&B)[M],
                                            const double (&Cc)[M], const double (&D)[M],
                                            int refine, double (&X)[M]) {
  const int T = L.T, t = L.t, p = L.p, ps = L.pstep, P = L.P;

  // ---- FP32 factorisation ------------------------------------------------
  float inv[M], ga[M], ap[M], gp[M];
  inv[0] = 1.0f / (float)B[0];
  ap[0] = (float)A[0] * inv[0];  // forward spike al, becomes ap below
  ga[0] = (float)Cc[0] * inv[0];
#pragma unroll
  for (int r = 1; r < M; ++r) {
    const float ar = (float)A[r];
    inv[r] = 1.0f / ((float)B[r] - ar * ga[r - 1]);
    ap[r] = -ar * ap[r - 1] * inv[r];
    ga[r] = (float)Cc[r] * inv[r];
  }
  const float al_e = ap[M - 1], ga_e = ga[M - 1];
  if (M >= 2) gp[M - 2] = ga[M - 2];
#pragma unroll
  for (int r = M - 3; r >= 0; --r) {
    ap[r] = ap[r] - ga[r] * ap[r + 1];
    gp[r] = -ga[r] * gp[r + 1];
  }
  const float f_a = M >= 2 ? ap[0] : 0.0f, f_g = M >= 2 ? gp[0] : -1.0f;
  float na = 0, ng = 0;
  {
    float *q = ring.next<float>();
    q[t] = f_a; q[T + t] = f_g;
    __syncthreads();
    if (p + 1 < P) { na = q[t + ps]; ng = q[T + t + ps]; }
  }
  const float invB = 1.0f / (1.0f - ga_e * na);
  float RA = al_e * invB, RC = -ga_e * ng * invB;
  float sA[MIXED_MAX_STEPS], sC[MIXED_MAX_STEPS], sR[MIXED_MAX_STEPS];
#pragma unroll
  for (int s = 0; s < MIXED_MAX_STEPS; ++s) {
    const int h = 1 << s;
    sA[s] = RA; sC[s] = RC; sR[s] = 1.0f;
    if (h < P) {
      float *q = ring.next<float>();
      q[t] = RA; q[T + t] = RC;
      __syncthreads();
      float Am = 0, Cm = 0, Ap = 0, Cp = 0;
      if (p - h >= 0) { Am = q[t - h * ps]; Cm = q[T + t - h * ps]; }
      if (p + h < P) { Ap = q[t + h * ps]; Cp = q[T + t + h * ps]; }
      const float rs = 1.0f / (1.0f - RA * Cm - RC * Ap);
      sR[s] = rs;
      RA = -rs * RA * Am;
      RC = -rs * RC * Cp;
    }
  }

  // ---- FP32 right-hand-side solve with the stored factorisation ----------
  auto solve_rhs = [&](float (&v)[M]) {
    v[0] = v[0] * inv[0];
#pragma unroll
    for (int r = 1; r < M; ++r) v[r] = (v[r] - (float)A[r] * v[r - 1]) * inv[r];
    const float de_e = v[M - 1];
#pragma unroll
    for (int r = M - 3; r >= 0; --r) v[r] = v[r] - ga[r] * v[r + 1];
    const float f_d = M >= 2 ? v[0] : 0.0f;
    float nd = 0;
    {
      float *q = ring.next<float>();
      q[t] = f_d;
      __syncthreads();
      if (p + 1 < P) nd = q[t + ps];
    }
    float Dr = (de_e - ga_e * nd) * invB;
#pragma unroll
    for (int s = 0; s < MIXED_MAX_STEPS; ++s) {
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
#pragma unroll
    for (int r = 0; r < M - 1; ++r) v[r] = v[r] - ap[r] * ym - gp[r] * y;
    v[M - 1] = y;
  };

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
      const double res = D[r] - (B[r] * X[r] + A[r] * xl + Cc[r] * xr);
      v[r] = (float)res;
    }
    solve_rhs(v);
#pragma unroll
    for (int r = 0; r < M; ++r) X[r] += (double)v[r];
  }
}

}  // namespace trid
