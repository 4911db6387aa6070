// kernels_util.hip - measurement helpers of libtridhip.so:
//   stream_copy_kernel   a plain streaming copy whose bytes moved are known
//                        exactly. It gives the HBM bandwidth the solver's
//                        roofline fraction is quoted against, and the
//                        reference kernel that calibrates rocprofv3's
//                        FETCH_SIZE / WRITE_SIZE counters; its shape (threads,
//                        16-byte vectors per thread, non-temporal or plain
//                        accesses) can be chosen, so the bandwidth quoted is the
//                        best of a sweep;
//   access_probe_kernel  the loads and stores of the strided (y / z) FP64 solve
//                        with 128- or 256-byte row segments (kernels_wide.hip:
//                        a by global->LDS loads into a [row][system] tile, b, c,
//                        d into registers, d written back), with one add per
//                        value in place of the solve: the bandwidth that access
//                        pattern allows, the ceiling the solve's rate is
//                        compared with. access_probe_x_kernel does the same
//                        for contiguous (x) systems. The probes' b, c, d and
//                        stores are nontemporal, as the solves' are.
#include "../../include/tridsolver_hip.h"
#include <algorithm>

#include "hybrid_core.hpp"
#include "trid_internal.h"

namespace trid {
namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int COPY_BLOCK = 256;
constexpr int COPY_UNROLL = 4;  // 16 KiB per workgroup, all loads before stores

// U 16-byte vectors per thread, all loads before the stores; NT: non-temporal
// loads and stores.
template <int U, bool NT>
__global__ __launch_bounds__(1024) void stream_copy_kernel(const u32x4 *__restrict__ src,
                                                           u32x4 *__restrict__ dst, size_t n16) {
  // the workgroup's first vector (uniform) and how many of its vectors exist:
  // per-thread indices stay 32-bit
  const size_t first = (size_t)blockIdx.x * (blockDim.x * U);
  const size_t left = n16 - first;
  const uint32_t lim = left < (size_t)blockDim.x * U ? (uint32_t)left : blockDim.x * U;
  const u32x4 *s = src + first;
  u32x4 *t = dst + first;
  u32x4 v[U];
  if (lim == blockDim.x * U) {  // every workgroup but the last: no masks
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t i = threadIdx.x + (uint32_t)k * blockDim.x;
      v[k] = NT ? __builtin_nontemporal_load(s + i) : s[i];
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t i = threadIdx.x + (uint32_t)k * blockDim.x;
      if (NT) __builtin_nontemporal_store(v[k], t + i);
      else t[i] = v[k];
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint32_t i = threadIdx.x + (uint32_t)k * blockDim.x;
    if (i < lim) v[k] = NT ? __builtin_nontemporal_load(s + i) : s[i];
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint32_t i = threadIdx.x + (uint32_t)k * blockDim.x;
    if (i < lim) {
      if (NT) __builtin_nontemporal_store(v[k], t + i);
      else t[i] = v[k];
    }
  }
}

template <int U, bool NT>
hipError_t launch_copy(const void *src, void *dst, size_t n16, int block, hipStream_t st) {
  const size_t per_block = (size_t)block * U;
  const size_t blocks = (n16 + per_block - 1) / per_block;
  if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
  hipLaunchKernelGGL((stream_copy_kernel<U, NT>), dim3((unsigned)blocks), dim3(block), 0, st,
                     static_cast<const u32x4 *>(src), static_cast<u32x4 *>(dst), n16);
  return hipGetLastError();
}

typedef __attribute__((address_space(3))) void lds_void;

// FP64 only. W = SEG / 8 x-neighbouring systems per workgroup, P lanes per
// system (W * P whole waves), lane (w, p) holds rows p*M .. p*M+M-1 of its
// system, as in hybrid_wide_kernel; LDS: the P*M x SEG tile of a.
template <int M, int SEG>
__global__ __launch_bounds__(M == 32 ? 512 : 1024) void access_probe_kernel(
    Geom g, int P, int G, const double *__restrict__ a, const double *__restrict__ b,
    const double *__restrict__ c, double *__restrict__ d) {
  constexpr int W = SEG / 8;
  constexpr int K = 256 / SEG;  // rows per wave-wide global->LDS load
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const hyb::Lane L = hyb::lane_of<false>(g, P, W, G);
  const int T = L.T;
  double *tile = reinterpret_cast<double *>(smem_raw);
  {  // a -> tile[row][0..W): lane l of a wave writes LDS bytes 4l..4l+3 of K rows
    const int lane = L.t & 63, wv = L.t >> 6, nw = T >> 6;
    const int sub = (lane * 4) / SEG, byte = (lane * 4) % SEG;
    const int64_t first = L.i0 - L.w;
    const bool sys_ok = first + byte / 8 < g.n0;
    const int64_t row_bytes = L.stride * 8;
    const char *src = reinterpret_cast<const char *>(a + first + L.i1 * g.st1) + byte +
                      sub * row_bytes;
    if (sys_ok) {
      for (int64_t row = (int64_t)K * wv; row < L.N; row += (int64_t)K * nw)
        if (row + sub < L.N)
          __builtin_amdgcn_global_load_lds(src + row * row_bytes,
                                           (lds_void *)(tile + row * W), 4, 0, 0);
    }
  }
  const int64_t row0 = (int64_t)L.p * M;
  const uint32_t ob = (uint32_t)((L.base + row0 * L.stride) * 8);
  const uint32_t os = (uint32_t)(L.stride * 8);
  auto at = [](const double *p, uint32_t off) -> double {
    return hyb::ld<true>(reinterpret_cast<const double *>(reinterpret_cast<const char *>(p) + off));
  };
  const int rows = L.active ? (int)std::min<int64_t>(M, std::max<int64_t>(0, L.N - row0)) : 0;
  double rb[M], rc[M], rd[M];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const uint32_t o = ob + (uint32_t)r * os;
    const bool ok = r < rows;
    rb[r] = ok ? at(b, o) : 0.0;
    rc[r] = ok ? at(c, o) : 0.0;
    rd[r] = ok ? at(d, o) : 0.0;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  const double *al = tile + row0 * W + L.w;
  uint32_t obs = ob;
  asm volatile("" : "+v"(obs));
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < rows) {
      double *dst = reinterpret_cast<double *>(reinterpret_cast<char *>(d) + obs +
                                               (uint32_t)r * os);
      hyb::st<true>(dst, rb[r] + rc[r] + rd[r] + al[r * W]);
    }
  }
}

template <int M, int SEG>
hipError_t launch_probe(const Geom &g, int P, int G, const double *a, const double *b,
                        const double *c, double *d, hipStream_t st) {
  constexpr int W = SEG / 8;
  const size_t lds = (size_t)P * M * SEG;
  auto kern = access_probe_kernel<M, SEG>;
  static LdsGrant lds_granted;
  hipError_t e = ensure_lds(reinterpret_cast<const void *>(kern), lds, lds_granted);
  if (e != hipSuccess) return e;
  const int64_t blocks = (g.n0 + W - 1) / W * g.n1;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(W * P), lds, st, g, P, G, a, b, c, d);
  return hipGetLastError();
}

// Contiguous systems (the x solve's pattern): W systems of P lanes in a
// 256-thread workgroup, lane p holds rows p + P*r (r < M) of its system,
// coalesced; reads a, b, c, d and writes d = a + b + c + d.
template <int M>
__global__ __launch_bounds__(256) void access_probe_x_kernel(Geom g, int P, int W, int G,
                                                             const double *__restrict__ a,
                                                             const double *__restrict__ b,
                                                             const double *__restrict__ c,
                                                             double *__restrict__ d) {
  const hyb::Lane L = hyb::lane_of<true>(g, P, W, G);
  double v[M];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    const int64_t row = hyb::fetch_row<M, true>(L, r);
    const int64_t e = L.base + row;
    v[r] = (L.active && row < L.N)
               ? hyb::ld<true>(a + e) + hyb::ld<true>(b + e) + hyb::ld<true>(c + e) +
                     hyb::ld<true>(d + e)
               : 0.0;
  }
  hyb::store<M, true, true>(L, d, v);
}

template <int M>
hipError_t launch_probe_x(const Geom &g, int P, int W, int G, const double *a, const double *b,
                          const double *c, double *d, hipStream_t st) {
  const int64_t blocks = (g.n0 + W - 1) / W * g.n1;
  hipLaunchKernelGGL(access_probe_x_kernel<M>, dim3((unsigned)blocks), dim3(W * P), 0, st, g, P,
                     W, G, a, b, c, d);
  return hipGetLastError();
}

}  // namespace
}  // namespace trid

EXTERN_C tridStatus_t tridHipStreamCopyShape(const void *src, void *dst, size_t bytes,
                                             int threads, int vectors, int nontemporal) {
  if (!src || !dst || bytes % 16 != 0 || (reinterpret_cast<uintptr_t>(src) % 16) ||
      (reinterpret_cast<uintptr_t>(dst) % 16))
    return TRID_STATUS_INVALID_VALUE;
  if (threads < 64 || threads > 1024 || threads % 64 != 0) return TRID_STATUS_INVALID_VALUE;
  if (vectors != 1 && vectors != 2 && vectors != 4) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  tridStatus_t s = trid::check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (bytes == 0) return TRID_STATUS_SUCCESS;
  const size_t n16 = bytes / 16;
  hipStream_t st = trid::current_stream();
  hipError_t e;
  const bool nt = nontemporal != 0;
  switch (vectors) {
    case 1: e = nt ? trid::launch_copy<1, true>(src, dst, n16, threads, st)
                   : trid::launch_copy<1, false>(src, dst, n16, threads, st); break;
    case 2: e = nt ? trid::launch_copy<2, true>(src, dst, n16, threads, st)
                   : trid::launch_copy<2, false>(src, dst, n16, threads, st); break;
    case 4: e = nt ? trid::launch_copy<4, true>(src, dst, n16, threads, st)
                   : trid::launch_copy<4, false>(src, dst, n16, threads, st); break;
    default: return TRID_STATUS_INVALID_VALUE;
  }
  return e == hipSuccess ? TRID_STATUS_SUCCESS : TRID_STATUS_EXECUTION_FAILED;
}

EXTERN_C tridStatus_t tridHipStreamCopy(const void *src, void *dst, size_t bytes) {
  return tridHipStreamCopyShape(src, dst, bytes, trid::COPY_BLOCK, trid::COPY_UNROLL, 1);
}

EXTERN_C tridStatus_t tridHipAccessProbe(const double *a, const double *b, const double *c,
                                         double *d, int solvedim, const int *dims, int seg_bytes,
                                         int M, int G) {
  if (!a || !b || !c || !d || !dims || solvedim < 0 || solvedim > 2 || G < 1)
    return TRID_STATUS_INVALID_VALUE;
  if (solvedim == 0 ? (seg_bytes != 0 || (M != 4 && M != 8 && M != 16))
                    : ((seg_bytes != 128 && seg_bytes != 256) || (M != 16 && M != 32)))
    return TRID_STATUS_INVALID_VALUE;
  trid::Geom g;
  bool empty = false;
  tridStatus_t s = trid::make_geom(3, solvedim, dims, dims, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  int dev = 0;
  if ((s = trid::check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  if (empty) return TRID_STATUS_SUCCESS;
  if (solvedim == 0) {  // contiguous: P lanes of M rows per system, 256-thread workgroups
    const int64_t P = (g.N + M - 1) / M;
    if (P > 256 || g.n0 * g.n1 > (int64_t)0x7fffffff) return TRID_STATUS_INVALID_VALUE;
    const int W = (int)std::max<int64_t>(1, 256 / P);
    hipStream_t st = trid::current_stream();
    hipError_t e;
    if (M == 4) e = trid::launch_probe_x<4>(g, (int)P, W, G, a, b, c, d, st);
    else if (M == 8) e = trid::launch_probe_x<8>(g, (int)P, W, G, a, b, c, d, st);
    else e = trid::launch_probe_x<16>(g, (int)P, W, G, a, b, c, d, st);
    return e == hipSuccess ? TRID_STATUS_SUCCESS : TRID_STATUS_EXECUTION_FAILED;
  }
  const int W = seg_bytes / 8;
  const int per_wave = std::max(1, 64 / W);
  int P = (int)((g.N + M - 1) / M);
  P = (P + per_wave - 1) / per_wave * per_wave;
  const int64_t T = (int64_t)W * P;
  // the kernel's register budget, one LDS tile, 32-bit byte offsets
  if (T > (M == 32 ? 512 : 1024) || (int64_t)P * M * seg_bytes > 160 * 1024 ||
      (int64_t)dims[0] * dims[1] * dims[2] * 8 > (int64_t)0xFFFFFFFFll)
    return TRID_STATUS_INVALID_VALUE;
  hipStream_t st = trid::current_stream();
  hipError_t e;
  if (M == 16) e = seg_bytes == 256 ? trid::launch_probe<16, 256>(g, P, G, a, b, c, d, st)
                                    : trid::launch_probe<16, 128>(g, P, G, a, b, c, d, st);
  else e = seg_bytes == 256 ? trid::launch_probe<32, 256>(g, P, G, a, b, c, d, st)
                            : trid::launch_probe<32, 128>(g, P, G, a, b, c, d, st);
  return e == hipSuccess ? TRID_STATUS_SUCCESS : TRID_STATUS_EXECUTION_FAILED;
}
