// kernels_ragged.hip - ragged batch: S independent systems of uneven lengths
// stored back to back (system s = rows offsets[s] .. offsets[s+1]-1 of a, b,
// c, d), solved by persistent workgroups that take work items from a queue.
// No counterpart in the reference, whose API needs equal lengths; per system
// the result is the plain solve's.
//
// Plan (four tiny kernels, all stream-ordered and capturable):
//   reset     zero the class counters, the work counter and the error flag
//   classify  class(s) = log2 of the lanes a system needs, P = next_pow2(ceil(L/M));
//             counted per workgroup in LDS, one global atomic per class
//   prefix    per class: first list slot and first work item; items are
//             ordered longest class first (largest-first scheduling)
//   scatter   list[slot] = s, slots reserved per workgroup and class
// Solve: one persistent launch of T-thread workgroups, one per CU (times
// `per_cu`), over every work item; or, split in two, systems needing more
// than T/2 lanes by that launch and the rest by a second launch of two
// T/2-thread workgroups per CU (one loads while the other computes). Each
// workgroup repeatedly takes the next work item of its range from an atomic
// counter -- W systems of one class, P lanes each, the Thomas-PCR hybrid of
// hybrid_core.hpp -- and leaves when the counter passes the range's last
// item, so every wave exits. Systems longer than M*T set the error flag and
// are skipped.

#include <algorithm>
#include <atomic>

#include "../../include/tridsolver_hip.h"
#include "hybrid_core.hpp"
#include "mixed_core.hpp"
#include "trid_internal.h"

namespace trid {
namespace {

constexpr int RG_CLASSES = 11;  // P = 1, 2, 4, ..., 1024

struct RaggedPlan {
  int counts[RG_CLASSES];
  int cursor[RG_CLASSES];
  int64_t sys_prefix[RG_CLASSES + 1];
  int64_t item_prefix[RG_CLASSES + 1];  // in scheduling order (longest first)
  int work;       // next item of the long-system launch
  int work_short; // next item of the short-system launch
  int error;
  int total_items;
  int long_items;  // items [0, long_items) need more than Tshort lanes
};

__global__ void rg_reset(RaggedPlan *plan) {
  const int t = threadIdx.x;
  if (t < RG_CLASSES) {
    plan->counts[t] = 0;
    plan->cursor[t] = 0;
  }
  if (t == 0) {
    plan->work = 0;
    plan->work_short = 0;
    plan->long_items = 0;
    plan->error = 0;
    plan->total_items = 0;
  }
}

__device__ __forceinline__ int rg_class(int64_t L, int M, int Tmax) {
  if (L <= 0) return -1;
  const int64_t need = (L + M - 1) / M;
  if (need > Tmax) return -2;
  int c = 0;
  while ((int64_t(1) << c) < need) ++c;
  return c;
}

// Counting and scattering aggregate per workgroup in LDS first: one global
// atomic per class and workgroup instead of one per system (systems of one
// class would otherwise serialise on a single counter).
constexpr int RG_PLAN_BLOCK = 256;

__global__ __launch_bounds__(RG_PLAN_BLOCK) void rg_classify(const int64_t *__restrict__ off,
                                                             int64_t S, int M, int Tmax,
                                                             RaggedPlan *plan) {
  __shared__ int cnt[RG_CLASSES];
  const int t = threadIdx.x;
  if (t < RG_CLASSES) cnt[t] = 0;
  __syncthreads();
  const int64_t s = (int64_t)blockIdx.x * RG_PLAN_BLOCK + t;
  if (s < S) {
    const int c = rg_class(off[s + 1] - off[s], M, Tmax);
    if (c >= 0) atomicAdd(&cnt[c], 1);
    else if (c == -2) atomicOr(&plan->error, 1);
  }
  __syncthreads();
  if (t < RG_CLASSES && cnt[t] > 0) atomicAdd(&plan->counts[t], cnt[t]);
}

// Lanes of the workgroups that solve class c: Tmax for systems needing more
// than Tshort lanes, Tshort for the rest.
__host__ __device__ __forceinline__ int rg_threads(int c, int Tmax, int Tshort) {
  return (1 << c) > Tshort ? Tmax : Tshort;
}

__global__ void rg_prefix(RaggedPlan *plan, int Tmax, int Tshort) {
  if (threadIdx.x != 0) return;
  int64_t sp = 0;
  for (int c = 0; c < RG_CLASSES; ++c) {
    plan->sys_prefix[c] = sp;
    sp += plan->counts[c];
  }
  plan->sys_prefix[RG_CLASSES] = sp;
  // item_prefix[k] for scheduling slot k = class RG_CLASSES-1-k
  int64_t ip = 0;
  int long_items = 0;
  for (int k = 0; k < RG_CLASSES; ++k) {
    const int c = RG_CLASSES - 1 - k;
    plan->item_prefix[k] = ip;
    const int W = rg_threads(c, Tmax, Tshort) >> c;
    const int64_t items = W > 0 ? (plan->counts[c] + W - 1) / W : 0;
    ip += items;
    if ((1 << c) > Tshort) long_items += (int)items;
  }
  plan->item_prefix[RG_CLASSES] = ip;
  plan->total_items = (int)ip;
  plan->long_items = long_items;
}

__global__ __launch_bounds__(RG_PLAN_BLOCK) void rg_scatter(const int64_t *__restrict__ off,
                                                            int64_t S, int M, int Tmax,
                                                            RaggedPlan *plan,
                                                            int *__restrict__ list) {
  __shared__ int cnt[RG_CLASSES], first[RG_CLASSES];
  const int t = threadIdx.x;
  if (t < RG_CLASSES) cnt[t] = 0;
  __syncthreads();
  const int64_t s = (int64_t)blockIdx.x * RG_PLAN_BLOCK + t;
  int c = -1, pos = 0;
  if (s < S) {
    c = rg_class(off[s + 1] - off[s], M, Tmax);
    if (c >= 0) pos = atomicAdd(&cnt[c], 1);  // position among this block's class-c systems
  }
  __syncthreads();
  if (t < RG_CLASSES && cnt[t] > 0) first[t] = atomicAdd(&plan->cursor[t], cnt[t]);
  __syncthreads();
  if (c >= 0) list[plan->sys_prefix[c] + first[c] + pos] = (int)s;
}

// One lane's view of work item `item`: the class (P lanes per system, W
// systems), and the lane's system (offset, length) if it has one. The list
// and offset loads depend on each other, so the persistent loops below decode
// the next item while the current one is being solved.
struct RgMeta {
  int cls;
  bool active;
  int N;  // listed systems passed rg_class: N <= M * Tmax
  int64_t base;
};

// The plan's class prefixes, copied into LDS once per workgroup at the start
// of a solve launch (rg_prefix has written them before it): item decoding
// then reads them from LDS, with no chain of dependent global loads.
struct RgPlanView {
  int64_t sys_prefix[RG_CLASSES + 1];
  int64_t item_prefix[RG_CLASSES + 1];
};

// Call with every thread before the launch's first barrier.
__device__ __forceinline__ void rg_view_load(RgPlanView &v, const RaggedPlan *plan, int t) {
  if (t <= RG_CLASSES) {
    v.sys_prefix[t] = plan->sys_prefix[t];
    v.item_prefix[t] = plan->item_prefix[t];
  }
}

__device__ __forceinline__ RgMeta rg_meta(const RgPlanView &v, const int *__restrict__ list,
                                          const int64_t *__restrict__ off, int item, int t,
                                          int Tmax) {
  RgMeta m;
  // scheduling slot k: the last one whose first item is <= item (the
  // prefixes do not decrease, so that is the count of k >= 1 whose first
  // item is <= item; independent LDS reads)
  int k = 0;
#pragma unroll
  for (int j = 1; j < RG_CLASSES; ++j) k += v.item_prefix[j] <= item ? 1 : 0;
  m.cls = RG_CLASSES - 1 - k;
  const int P = 1 << m.cls, W = Tmax >> m.cls;
  const int64_t first = v.sys_prefix[m.cls] + (int64_t)(item - v.item_prefix[k]) * W;
  const int64_t slot = first + t / P;
  m.active = slot < v.sys_prefix[m.cls + 1];
  const int64_t sys = m.active ? list[slot] : 0;
  m.base = m.active ? off[sys] : 0;
  m.N = m.active ? (int)(off[sys + 1] - off[sys]) : 0;
  return m;
}

__device__ __forceinline__ hyb::Lane rg_lane(const RgMeta &m, int t, int Tmax) {
  hyb::Lane L;
  L.P = 1 << m.cls;
  L.W = Tmax >> m.cls;
  L.T = Tmax;
  L.t = t;
  L.p = t % L.P;
  L.w = t / L.P;
  L.pstep = 1;
  L.stride = 1;
  L.active = m.active;
  L.base = m.base;
  L.N = m.N;
  L.i0 = 0;
  L.i1 = 0;
  return L;
}

// Threads per workgroup: 1024 for the M = 4/8 variants, 512 for M = 16 (FP64
// M = 16 needs that register budget; FP32 M = 16 at 1024 threads would spill,
// and 512 threads of 16 rows cover the same 8192 rows with half the lanes).
template <typename REAL, int M>
struct RgBounds {
  static constexpr int threads = M >= 16 ? 512 : 1024;
};

// One persistent launch over a range of items: LONG = the items whose systems
// need more than Tshort lanes (workgroups of T = Tmax threads, one per CU),
// else the rest (T = Tshort, two workgroups per CU, so that one loads while
// the other computes). A single launch plans with Tshort = Tmax: no item is
// long, and the LONG = false launch at T = Tmax takes them all.
// One work item of rg_solve: W systems of one class, P lanes each, in T
// threads. `fetched()` runs once the loads are issued and a barrier has
// passed (the caller decodes its next item there).
template <typename REAL, int M, bool INC, int T, typename OnFetched>
__device__ __forceinline__ void rg_item(const RgMeta &cur, unsigned char *smem_raw,
                                        const REAL *__restrict__ a, const REAL *__restrict__ b,
                                        const REAL *__restrict__ c, REAL *__restrict__ d,
                                        REAL *__restrict__ u, OnFetched fetched) {
  const int t = threadIdx.x;
  const hyb::Lane L = rg_lane(cur, t, T);
  REAL *tile = reinterpret_cast<REAL *>(smem_raw) + (size_t)L.w * L.P * (M + 1);
  hyb::Ring ring{smem_raw + sizeof(REAL) * (size_t)T * (M + 1), (int)(3 * sizeof(REAL) * T), 0};
  REAL ra[M], rb[M], rc[M], rd[M];
  hyb::fetch_abcd<M, true>(L, a, b, c, d, ra, rb, rc, rd);
  hyb::to_chunk<M, true>(L, ra, tile);  // has barriers
  fetched();
  hyb::to_chunk<M, true>(L, rb, tile);
  hyb::to_chunk<M, true>(L, rc, tile);
  hyb::to_chunk<M, true>(L, rd, tile);
  REAL x[M];
  hyb::solve_chunk<REAL, M, true>(L, ring, ra, rb, rc, rd, x);
  hyb::to_fetch<M, true>(L, x, tile);
  if (INC) hyb::store_add<M, true>(L, u, x);
  else hyb::store<M, true>(L, d, x);
}

template <typename REAL, int M, bool INC, int T, bool LONG>
__global__ __launch_bounds__(T) void rg_solve(
    const REAL *__restrict__ a, const REAL *__restrict__ b, const REAL *__restrict__ c,
    REAL *__restrict__ d, REAL *__restrict__ u, const int64_t *__restrict__ off,
    const int *__restrict__ list, RaggedPlan *plan) {
  constexpr int Tmax = T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_next;
  __shared__ RgPlanView s_view;
  rg_view_load(s_view, plan, threadIdx.x);
  const int t = threadIdx.x;
  int *counter = LONG ? &plan->work : &plan->work_short;
  const int lo = LONG ? 0 : plan->long_items;
  const int total = LONG ? plan->long_items : plan->total_items;
  if (t == 0) s_next = lo + atomicAdd(counter, 1);
  __syncthreads();
  int item = s_next;
  RgMeta cur = item < total ? rg_meta(s_view, list, off, item, t, Tmax) : RgMeta{};
  for (;;) {
    __syncthreads();  // every thread has read s_next
    if (item >= total) break;
    // claim the following item now; its id is read after the next barrier
    if (t == 0) s_next = lo + atomicAdd(counter, 1);
    int nxt = total;
    RgMeta next{};
    rg_item<REAL, M, INC, Tmax>(cur, smem_raw, a, b, c, d, u, [&] {
      nxt = s_next;
      if (nxt < total) next = rg_meta(s_view, list, off, nxt, t, Tmax);
    });
    item = nxt;
    cur = next;
  }
}

// Systems needing more than 256 lanes at 16 rows per lane (4097-8192 rows):
// hyb::long32_solve (32 rows per lane in 256 threads, `a` moved HBM->LDS in
// chunk order). The items are single systems (plan class 9 at Tmax 512).
constexpr int RGL_T = hyb::LONG32_T, RGL_PLAN_T = 512;

// LDS of rg_solve_split: the long items' tile and ring (the rest need less).
template <typename REAL>
constexpr size_t rg_split_lds_bytes() {
  return hyb::long32_lds_bytes<REAL>();
}

// The three-launch split's solve (plan at 16 rows per lane with Tmax 512 and
// Tshort 256): one persistent launch of 256-thread workgroups over every
// item (FP64: two per CU). The long items [0, long_items) come first in the
// queue and take hyb::long32_solve at 32 rows per lane; the rest take rg_item at
// 16 rows per lane. One launch, so a workgroup that runs out of long items
// goes straight on to the short ones.
template <typename REAL, bool INC>
__global__ __launch_bounds__(RGL_T, 2) void rg_solve_split(
    const REAL *__restrict__ a, const REAL *__restrict__ b, const REAL *__restrict__ c,
    REAL *__restrict__ d, REAL *__restrict__ u, const int64_t *__restrict__ off,
    const int *__restrict__ list, RaggedPlan *plan) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_next;
  __shared__ RgPlanView s_view;
  rg_view_load(s_view, plan, threadIdx.x);
  const int t = threadIdx.x;
  const int total = plan->total_items, nlong = plan->long_items;
  auto decode = [&](int i) {
    return i < nlong ? rg_meta(s_view, list, off, i, 0, RGL_PLAN_T)
                     : rg_meta(s_view, list, off, i, t, RGL_T);
  };
  if (t == 0) s_next = atomicAdd(&plan->work, 1);
  __syncthreads();
  int item = s_next;
  RgMeta cur = item < total ? decode(item) : RgMeta{};
  for (;;) {
    __syncthreads();  // every thread has read s_next
    if (item >= total) break;
    if (t == 0) s_next = atomicAdd(&plan->work, 1);
    int nxt = total;
    RgMeta next{};
    auto fetched = [&] {
      nxt = s_next;
      if (nxt < total) next = decode(nxt);
    };
    if (item < nlong)
      hyb::long32_solve<REAL, INC, RGL_T>(hyb::wave_uniform(cur.base),
                                   __builtin_amdgcn_readfirstlane(cur.N), smem_raw, a, b, c, d, u,
                                   fetched);
    else rg_item<REAL, 16, INC, RGL_T>(cur, smem_raw, a, b, c, d, u, fetched);
    item = nxt;
    cur = next;
  }
}

// The four planning launches: work items of at most Tmax lanes.
void launch_plan(const int64_t *off, int64_t S, int M, int Tmax, int Tshort, void *ws,
                 hipStream_t st) {
  RaggedPlan *plan = static_cast<RaggedPlan *>(ws);
  int *list = reinterpret_cast<int *>(static_cast<unsigned char *>(ws) + RAGGED_PLAN_HEADER_BYTES);
  hipLaunchKernelGGL(rg_reset, dim3(1), dim3(64), 0, st, plan);
  const unsigned nb = (unsigned)((S + RG_PLAN_BLOCK - 1) / RG_PLAN_BLOCK);
  if (nb > 0) {
    hipLaunchKernelGGL(rg_classify, dim3(nb), dim3(RG_PLAN_BLOCK), 0, st, off, S, M, Tmax,
                       plan);
  }
  hipLaunchKernelGGL(rg_prefix, dim3(1), dim3(64), 0, st, plan, Tmax, Tshort);
  if (nb > 0) {
    hipLaunchKernelGGL(rg_scatter, dim3(nb), dim3(RG_PLAN_BLOCK), 0, st, off, S, M, Tmax,
                       plan, list);
  }
}

template <typename REAL, int M, bool INC>
hipError_t launch_m(const REAL *a, const REAL *b, const REAL *c, REAL *d, REAL *u,
                    const int64_t *off, int64_t S, int per_cu, int launches, void *ws,
                    hipStream_t st, int dev) {
  constexpr int Tmax = RgBounds<REAL, M>::threads;
  constexpr int Tshort = Tmax / 2;
  RaggedPlan *plan = static_cast<RaggedPlan *>(ws);
  int *list = reinterpret_cast<int *>(static_cast<unsigned char *>(ws) + RAGGED_PLAN_HEADER_BYTES);
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int per = std::max(1, per_cu);
  if (launches == 1) {  // every item in one launch of Tmax-thread workgroups
    launch_plan(off, S, M, Tmax, Tmax, ws, st);
    const size_t lds = sizeof(REAL) * (size_t)Tmax * (M + 1) + hyb::ring_bytes<REAL>(Tmax, 3);
    auto kern = rg_solve<REAL, M, INC, Tmax, false>;
    static LdsGrant lds_granted;
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3((unsigned)(cus * per)), dim3(Tmax), lds, st, a, b, c, d, u,
                       off, list, plan);
    return hipGetLastError();
  }
  if (launches == 3 && M != 16) return hipErrorInvalidValue;
  launch_plan(off, S, M, Tmax, Tshort, ws, st);
  if constexpr (M == 16) {
    if (launches == 3) {  // one launch: the long systems at 32 rows per lane, then the rest
      auto kern = rg_solve_split<REAL, INC>;
      static LdsGrant lds_granted;
      constexpr size_t lds = rg_split_lds_bytes<REAL>();
      hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
      if (e != hipSuccess) return e;
      int per_sm = 0;  // resident workgroups per CU
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_sm, reinterpret_cast<const void *>(kern),
                                                       RGL_T, lds) != hipSuccess || per_sm < 1)
        per_sm = 1;
      hipLaunchKernelGGL(kern, dim3((unsigned)(per_sm * cus * per)), dim3(RGL_T), lds, st, a, b,
                         c, d, u, off, list, plan);
      return hipGetLastError();
    }
  }
  if (launches == 2) {  // the long systems first (an empty range exits at once)
    const size_t lds = sizeof(REAL) * (size_t)Tmax * (M + 1) + hyb::ring_bytes<REAL>(Tmax, 3);
    auto kern = rg_solve<REAL, M, INC, Tmax, true>;
    static LdsGrant lds_granted;
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3((unsigned)(cus * per)), dim3(Tmax), lds, st, a, b, c, d, u,
                       off, list, plan);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const size_t lds = sizeof(REAL) * (size_t)Tshort * (M + 1) + hyb::ring_bytes<REAL>(Tshort, 3);
  auto kern = rg_solve<REAL, M, INC, Tshort, false>;
  static LdsGrant lds_granted;
  hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((unsigned)(2 * cus * per)), dim3(Tshort), lds, st, a, b, c, d,
                     u, off, list, plan);
  return hipGetLastError();
}

// FP64 systems by FP32 sweeps plus FP64 refinement (mixed_core.hpp): the
// same work queue, 512-thread workgroups (P <= 2^MIXED_MAX_STEPS lanes).
constexpr int RG_MIXED_THREADS = 512;

template <int M>
__global__ __launch_bounds__(RG_MIXED_THREADS) void rg_solve_mixed(
    const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
    double *__restrict__ d, const int64_t *__restrict__ off, const int *__restrict__ list,
    RaggedPlan *plan, int refine) {
  constexpr int Tmax = RG_MIXED_THREADS;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_next;
  __shared__ RgPlanView s_view;
  rg_view_load(s_view, plan, threadIdx.x);
  const int t = threadIdx.x;
  const int total = plan->total_items;
  if (t == 0) s_next = atomicAdd(&plan->work, 1);
  __syncthreads();
  int item = s_next;
  RgMeta cur = item < total ? rg_meta(s_view, list, off, item, t, Tmax) : RgMeta{};
  for (;;) {
    __syncthreads();  // every thread has read s_next
    if (item >= total) break;
    if (t == 0) s_next = atomicAdd(&plan->work, 1);
    const hyb::Lane L = rg_lane(cur, t, Tmax);
    double *tile = reinterpret_cast<double *>(smem_raw) + (size_t)L.w * L.P * (M + 1);
    hyb::Ring ring{smem_raw + sizeof(double) * (size_t)Tmax * (M + 1),
                   (int)(3 * sizeof(double) * Tmax), 0};
    double A[M], B[M], Cc[M], D[M];
    hyb::fetch_abcd<M, true>(L, a, b, c, d, A, B, Cc, D);
    hyb::to_chunk<M, true>(L, A, tile);
    const int nxt = s_next;
    const RgMeta next = nxt < total ? rg_meta(s_view, list, off, nxt, t, Tmax) : RgMeta{};
    hyb::to_chunk<M, true>(L, B, tile);
    hyb::to_chunk<M, true>(L, Cc, tile);
    hyb::to_chunk<M, true>(L, D, tile);
    double X[M];
    mixed_chunk<M>(L, ring, A, B, Cc, D, refine, X);
    hyb::to_fetch<M, true>(L, X, tile);
    hyb::store<M, true>(L, d, X);
    item = nxt;
    cur = next;
  }
}

// FP64 solution of FP32-stored systems (the arithmetic of kernels_mixed32.hip):
// FP32 a, b, c, d in, FP64 x out (d is not written); 16 rows per lane in
// 512 threads, so systems up to 8192 rows; ONE: at most one refinement. The
// LDS tile holds doubles for x's transpose; the FP32 operands go through its
// first half. AC_LDS keeps a and c in the tile (its two FP32 halves) for the
// solve, 32 registers a lane fewer: the general refinement loop needs that
// to fit the register file.
constexpr int RG_M32_M = 16;

// T threads, LONG as in rg_solve (the range of a two-launch split).
template <bool ONE, bool AC_LDS, int T, bool LONG>
__global__ __launch_bounds__(T) void rg_solve_mixed32(
    const float *__restrict__ a, const float *__restrict__ b, const float *__restrict__ c,
    const float *__restrict__ d, double *__restrict__ x, const int64_t *__restrict__ off,
    const int *__restrict__ list, RaggedPlan *plan, int refine) {
  constexpr int M = RG_M32_M, Tmax = T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_next;
  __shared__ RgPlanView s_view;
  rg_view_load(s_view, plan, threadIdx.x);
  const int t = threadIdx.x;
  int *counter = LONG ? &plan->work : &plan->work_short;
  const int lo = LONG ? 0 : plan->long_items;
  const int total = LONG ? plan->long_items : plan->total_items;
  if (t == 0) s_next = lo + atomicAdd(counter, 1);
  __syncthreads();
  int item = s_next;
  RgMeta cur = item < total ? rg_meta(s_view, list, off, item, t, Tmax) : RgMeta{};
  for (;;) {
    __syncthreads();  // every thread has read s_next
    if (item >= total) break;
    if (t == 0) s_next = lo + atomicAdd(counter, 1);
    const hyb::Lane L = rg_lane(cur, t, Tmax);
    double *dtile = reinterpret_cast<double *>(smem_raw) + (size_t)L.w * L.P * (M + 1);
    float *ftile = reinterpret_cast<float *>(dtile);
    hyb::Ring ring{smem_raw + sizeof(double) * (size_t)Tmax * (M + 1),
                   (int)(2 * sizeof(double) * Tmax), 0};
    float A[M], B[M], Cc[M], D[M];
    hyb::fetch_abcd<M, true>(L, a, b, c, d, A, B, Cc, D);
    hyb::to_chunk<M, true>(L, B, ftile);  // has barriers: s_next is visible below
    const int nxt = s_next;
    const RgMeta next = nxt < total ? rg_meta(s_view, list, off, nxt, t, Tmax) : RgMeta{};
    hyb::to_chunk<M, true>(L, D, ftile);
    double X[M];
    if constexpr (AC_LDS) {
      float *tC = ftile + L.P * (M + 1);
#pragma unroll
      for (int r = 0; r < M; ++r) {
        const int e = hyb::skew<M>(L.p + L.P * r);
        ftile[e] = A[r];
        tC[e] = Cc[r];
      }
      __syncthreads();
      const float *ca = ftile + L.p * (M + 1), *cc = tC + L.p * (M + 1);
      mixed_chunk_acc<M, MIXED_MAX_STEPS, false, ONE, true>(
          L, ring, [&](int r) { return ca[r]; }, B, [&](int r) { return cc[r]; }, D, refine, X);
      __syncthreads();  // a and c are read to the end before x overwrites the tile
    } else {
      hyb::to_chunk<M, true>(L, A, ftile);
      hyb::to_chunk<M, true>(L, Cc, ftile);
      mixed_chunk_acc<M, MIXED_MAX_STEPS, false, ONE, true>(
          L, ring, [&](int r) { return A[r]; }, B, [&](int r) { return Cc[r]; }, D, refine, X);
    }
    hyb::to_fetch<M, true>(L, X, dtile);
    hyb::store<M, true>(L, x, X);
    item = nxt;
    cur = next;
  }
}

template <bool ONE, bool AC_LDS, int T, bool LONG>
hipError_t launch_mixed32_range(const float *a, const float *b, const float *c, const float *d,
                                double *x, const int64_t *off, RaggedPlan *plan, int *list,
                                unsigned blocks, int refine, hipStream_t st) {
  const size_t lds = sizeof(double) * (size_t)T * (RG_M32_M + 1) + hyb::ring_bytes<double>(T, 2);
  auto kern = rg_solve_mixed32<ONE, AC_LDS, T, LONG>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(T), lds, st, a, b, c, d, x, off, list, plan,
                     refine);
  return hipGetLastError();
}

// One launch of 512-thread workgroups, or two: the systems needing more than
// 256 lanes in 512-thread workgroups, then the rest in 256-thread ones, two
// per CU.
template <bool ONE, bool AC_LDS>
hipError_t launch_mixed32_one(const float *a, const float *b, const float *c, const float *d,
                              double *x, const int64_t *off, int64_t S, int per_cu,
                              int launches, int refine, void *ws, hipStream_t st, int dev) {
  constexpr int Tmax = RG_MIXED_THREADS, Tshort = Tmax / 2;
  RaggedPlan *plan = static_cast<RaggedPlan *>(ws);
  int *list = reinterpret_cast<int *>(static_cast<unsigned char *>(ws) + RAGGED_PLAN_HEADER_BYTES);
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const unsigned per = (unsigned)std::max(1, per_cu);
  if (launches == 1) {  // no long range: the LONG = false launch takes every item
    launch_plan(off, S, RG_M32_M, Tmax, Tmax, ws, st);
    return launch_mixed32_range<ONE, AC_LDS, Tmax, false>(a, b, c, d, x, off, plan, list,
                                                          cus * per, refine, st);
  }
  launch_plan(off, S, RG_M32_M, Tmax, Tshort, ws, st);
  hipError_t e = launch_mixed32_range<ONE, AC_LDS, Tmax, true>(a, b, c, d, x, off, plan, list,
                                                               cus * per, refine, st);
  if (e != hipSuccess) return e;
  return launch_mixed32_range<ONE, AC_LDS, Tshort, false>(a, b, c, d, x, off, plan, list,
                                                          2 * cus * per, refine, st);
}

template <int M>
hipError_t launch_mixed_m(const double *a, const double *b, const double *c, double *d,
                          const int64_t *off, int64_t S, int per_cu, int refine, void *ws,
                          hipStream_t st, int dev) {
  constexpr int Tmax = RG_MIXED_THREADS;
  RaggedPlan *plan = static_cast<RaggedPlan *>(ws);
  int *list = reinterpret_cast<int *>(static_cast<unsigned char *>(ws) + RAGGED_PLAN_HEADER_BYTES);
  launch_plan(off, S, M, Tmax, Tmax, ws, st);  // one launch: no long range
  const size_t lds =
      sizeof(double) * (size_t)Tmax * (M + 1) + hyb::ring_bytes<double>(Tmax, 3);
  auto kern = rg_solve_mixed<M>;
  static LdsGrant lds_granted;
  {
    hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
    if (e != hipSuccess) return e;
  }
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  hipLaunchKernelGGL(kern, dim3((unsigned)(cus * std::max(1, per_cu))), dim3(Tmax), lds, st,
                     a, b, c, d, off, list, plan, refine);
  return hipGetLastError();
}

// M = 16 with the FP64 a and c operands in LDS (mixed_contig_lds): 8192 rows
// in 512 threads. One launch of 512-thread workgroups, one per CU, or the
// FP64 solve's split: systems needing more than 256 lanes by that launch, the
// rest by 256-thread workgroups, two per CU (LDS 2 tiles x T*(M+1) doubles +
// 2 ring slots of 2 doubles per thread: 152 KiB at 512 threads, 76 KiB at
// 256).
constexpr int RG_MIXED_LDS_M = 16;

template <int T>
constexpr size_t rg_mixed_lds_bytes() {
  return 2 * sizeof(double) * (size_t)T * (RG_MIXED_LDS_M + 1) + hyb::ring_bytes<double>(T, 2);
}

// Two waves per SIMD in both shapes (the 256-thread workgroups run two per
// CU), i.e. at most 256 VGPRs.
template <int T, bool LONG>
__global__ __launch_bounds__(T, 2) void rg_solve_mixed_lds(
    const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
    double *__restrict__ d, const int64_t *__restrict__ off, const int *__restrict__ list,
    RaggedPlan *plan, int refine) {
  constexpr int M = RG_MIXED_LDS_M;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_next;
  __shared__ RgPlanView s_view;
  // each thread's system offset of the current item, read again for the
  // stores: with the next item decoded during the solve, holding both in
  // registers across it would spill
  __shared__ int64_t s_base[T];
  rg_view_load(s_view, plan, threadIdx.x);
  const int t = threadIdx.x;
  int *counter = LONG ? &plan->work : &plan->work_short;
  const int lo = LONG ? 0 : plan->long_items;
  const int total = LONG ? plan->long_items : plan->total_items;
  if (t == 0) s_next = lo + atomicAdd(counter, 1);
  __syncthreads();
  int item = s_next;
  RgMeta cur = item < total ? rg_meta(s_view, list, off, item, t, T) : RgMeta{};
  double *tiles = reinterpret_cast<double *>(smem_raw);
  hyb::Ring ring{smem_raw + 2 * sizeof(double) * (size_t)T * (M + 1),
                 (int)(2 * sizeof(double) * T), 0};
  for (;;) {
    __syncthreads();  // every thread has read s_next
    if (item >= total) break;
    // claim the following item now; its id is read after the first transpose
    if (t == 0) s_next = lo + atomicAdd(counter, 1);
    s_base[t] = cur.base;
    const hyb::Lane L = rg_lane(cur, t, T);
    double *tA = tiles + (size_t)L.w * L.P * (M + 1);
    double *tC = tA + (size_t)T * (M + 1);
    int nxt = 0;
    RgMeta next{};
    // at most T lanes per system: log2(T) PCR steps
    mixed_contig_lds<M, T == 512 ? 9 : 8, false>(L, ring, a, b, c, d, tA, tC, refine, [&] {
      nxt = s_next;
      if (nxt < total) next = rg_meta(s_view, list, off, nxt, t, T);
    }, [&] { return s_base[t]; });
    item = nxt;
    cur = next;
  }
}

template <int T, bool LONG>
hipError_t launch_rg_mixed_lds(const double *a, const double *b, const double *c, double *d,
                               const int64_t *off, int *list, RaggedPlan *plan, unsigned blocks,
                               int refine, hipStream_t st) {
  auto kern = rg_solve_mixed_lds<T, LONG>;
  static LdsGrant lds_granted;
  hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), rg_mixed_lds_bytes<T>(),
                            lds_granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(T), rg_mixed_lds_bytes<T>(), st, a, b, c, d, off,
                     list, plan, refine);
  return hipGetLastError();
}

hipError_t launch_mixed_lds(const double *a, const double *b, const double *c, double *d,
                            const int64_t *off, int64_t S, int per_cu, int launches, int refine,
                            void *ws, hipStream_t st, int dev) {
  constexpr int Tmax = 512, Tshort = 256;
  RaggedPlan *plan = static_cast<RaggedPlan *>(ws);
  int *list = reinterpret_cast<int *>(static_cast<unsigned char *>(ws) + RAGGED_PLAN_HEADER_BYTES);
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int per = std::max(1, per_cu);
  if (launches == 1) {  // no long range: the LONG = false launch takes every item
    launch_plan(off, S, RG_MIXED_LDS_M, Tmax, Tmax, ws, st);
    return launch_rg_mixed_lds<Tmax, false>(a, b, c, d, off, list, plan,
                                            (unsigned)(cus * per), refine, st);
  }
  launch_plan(off, S, RG_MIXED_LDS_M, Tmax, Tshort, ws, st);
  hipError_t e = launch_rg_mixed_lds<Tmax, true>(a, b, c, d, off, list, plan,
                                                 (unsigned)(cus * per), refine, st);
  if (e != hipSuccess) return e;
  return launch_rg_mixed_lds<Tshort, false>(a, b, c, d, off, list, plan,
                                            (unsigned)(2 * cus * per), refine, st);
}

}  // namespace

int ragged_mixed32_max_len() { return RG_M32_M * RG_MIXED_THREADS; }

hipError_t launch_ragged_mixed32(const float *a, const float *b, const float *c, const float *d,
                                 double *x, const int64_t *off, int64_t S, int per_cu,
                                 int launches, int refine, void *ws, hipStream_t st, int dev) {
  if (launches != 1 && launches != 2) return hipErrorInvalidValue;
  return refine <= 1 ? launch_mixed32_one<true, false>(a, b, c, d, x, off, S, per_cu, launches,
                                                       refine, ws, st, dev)
                     : launch_mixed32_one<false, true>(a, b, c, d, x, off, S, per_cu, launches,
                                                       refine, ws, st, dev);
}

int ragged_mixed_max_len(int M) {
  if (M == RG_MIXED_LDS_M) return M * 512;
  return (M == 4 || M == 8) ? M * RG_MIXED_THREADS : 0;
}

hipError_t launch_ragged_mixed(int M, const double *a, const double *b, const double *c,
                               double *d, const int64_t *off, int64_t S, int per_cu,
                               int launches, int refine, void *ws, hipStream_t st, int dev) {
  if (launches != 1 && launches != 2) return hipErrorInvalidValue;
  switch (M) {
    case 4: return launch_mixed_m<4>(a, b, c, d, off, S, per_cu, refine, ws, st, dev);
    case 8: return launch_mixed_m<8>(a, b, c, d, off, S, per_cu, refine, ws, st, dev);
    case RG_MIXED_LDS_M:
      return launch_mixed_lds(a, b, c, d, off, S, per_cu, launches, refine, ws, st, dev);
    default: return hipErrorInvalidValue;
  }
}

size_t ragged_workspace_bytes(int64_t S) {
  return RAGGED_PLAN_HEADER_BYTES + sizeof(int) * (size_t)std::max<int64_t>(S, 1);
}

int ragged_max_len(int fp64, int M) {
  (void)fp64;  // the same shapes at both precisions
  if (M != 4 && M != 8 && M != 16) return 0;
  return M * (M >= 16 ? 512 : 1024);
}

template <typename REAL, bool INC>
hipError_t launch_ragged(int M, const REAL *a, const REAL *b, const REAL *c, REAL *d, REAL *u,
                         const int64_t *off, int64_t S, int per_cu, int launches, void *ws,
                         hipStream_t st, int dev) {
  static_assert(sizeof(RaggedPlan) <= RAGGED_PLAN_HEADER_BYTES, "plan header");
  if (launches < 1 || launches > 3) return hipErrorInvalidValue;
  switch (M) {
    case 4: return launch_m<REAL, 4, INC>(a, b, c, d, u, off, S, per_cu, launches, ws, st, dev);
    case 8: return launch_m<REAL, 8, INC>(a, b, c, d, u, off, S, per_cu, launches, ws, st, dev);
    case 16:
      return launch_m<REAL, 16, INC>(a, b, c, d, u, off, S, per_cu, launches, ws, st, dev);
    default: return hipErrorInvalidValue;
  }
}

// Reads the plan's error flag (after the stream has completed).
hipError_t ragged_error_flag(void *ws, int *flag) {
  RaggedPlan *plan = static_cast<RaggedPlan *>(ws);
  return hipMemcpy(flag, &plan->error, sizeof(int), hipMemcpyDeviceToHost);
}

#define TRID_RG_INST(REAL, INC)                                                              \
  template hipError_t launch_ragged<REAL, INC>(int, const REAL *, const REAL *, const REAL *, \
                                               REAL *, REAL *, const int64_t *, int64_t, int, \
                                               int, void *, hipStream_t, int);
TRID_RG_INST(float, false)
TRID_RG_INST(float, true)
TRID_RG_INST(double, false)
TRID_RG_INST(double, true)

}  // namespace trid
