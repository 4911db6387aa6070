// trid_tune.hip - shape selection and the autotuner.
//
// Every call asks tuned_config() (or its mixed / ragged counterparts) for the
// kernel shape of its geometry. The answer comes from a per-device cache keyed
// by the exact geometry; on a miss it is the measured default of
// hybrid_default_cfg(), or -- in autotune-on-first-use mode, outside stream
// capture -- the result of tuning that geometry now.
//
// Tuning a geometry: scratch arrays of the same extent are filled on the GPU
// with a diagonally dominant pattern (the caller's arrays are never touched),
// and every candidate is timed on a private stream with hipEvents (one warm-up,
// median of three): the per-lane Thomas kernel and every valid hybrid shape
//   M in {1,2,4,8,16} rows per lane (the Thomas / PCR split: M rows by Thomas,
//   log2(P) PCR steps across the P = ceil(N/M) lanes),
//   W systems per workgroup, G XCD grouping in {1, 8, 16},
// skipping strided shapes whose rows are narrower than 64 bytes, plus, for
// strided systems, the streamed variants and the variant with `a` staged in
// LDS (128- and 256-byte rows, M in {16, 32}, G in {1, 2, 4, 8, 16, 32}). The
// fastest is cached (the fastest four of a first pass are timed again, median
// of 9). The adjoint is tuned the same way over its own kernel's shapes.
// The cache can be saved to and loaded from a text file.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/tridsolver_hip.h"
#include "trid_tune.h"

namespace trid {
namespace {

enum Kind : int { KIND_SOLVE = 0, KIND_MIXED = 1, KIND_RAGGED = 2, KIND_MIXED32 = 3, KIND_ADJOINT = 4 };

struct Key {
  int dev, kind, fp64;
  int inc;  // the Inc solve; for KIND_MIXED32: more than one refinement
  int64_t N, stride, n0, n1, st0, st1;
  bool operator<(const Key &o) const {
    return std::tie(dev, kind, fp64, inc, N, stride, n0, n1, st0, st1) <
           std::tie(o.dev, o.kind, o.fp64, o.inc, o.N, o.stride, o.n0, o.n1, o.st0, o.st1);
  }
};

struct Choice {
  int algo;  // 1 lane, 3 hybrid (solve); 2 mixed; 5 ragged (cfg.M, cfg.G = per_cu,
             // cfg.V = solve launches); 10 FP32-stored mixed; 4 adjoint
  HybridCfg cfg;
  float ms;
};

std::mutex g_tmu;
std::map<Key, Choice> g_cache;
int g_mode = -1;  // -1: read TRIDSOLVER_AUTOTUNE on first use

int autotune_mode() {
  std::lock_guard<std::mutex> lk(g_tmu);
  if (g_mode < 0) {
    const char *e = std::getenv("TRIDSOLVER_AUTOTUNE");
    g_mode = (e && std::atoi(e) > 0) ? 1 : 0;
  }
  return g_mode;
}

Key key_of(int dev, int kind, int fp64, int inc, const Geom &g) {
  return Key{dev, kind, fp64, inc, g.N, g.stride, g.n0, g.n1, g.st0, g.st1};
}

bool lookup(const Key &k, Choice *c) {
  std::lock_guard<std::mutex> lk(g_tmu);
  auto it = g_cache.find(k);
  if (it == g_cache.end()) return false;
  *c = it->second;
  return true;
}

void store(const Key &k, const Choice &c) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_cache[k] = c;
}

bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess) return true;
  return cap != hipStreamCaptureStatusNone;
}

// Deterministic diagonally dominant fill: a, c in (-1, 1), b = 2 + |a| + |c| + t,
// d (and u) in (-1, 1), from an integer hash of the element index.
__device__ __forceinline__ double hash01(uint64_t i, uint64_t salt) {
  uint64_t x = i * 0x9E3779B97F4A7C15ull + salt;
  x ^= x >> 31; x *= 0xBF58476D1CE4E5B9ull; x ^= x >> 29; x *= 0x94D049BB133111EBull;
  x ^= x >> 32;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

template <typename REAL>
__global__ void tune_fill(REAL *a, REAL *b, REAL *c, REAL *d, REAL *u, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double av = 2 * hash01(i, 1) - 1, cv = 2 * hash01(i, 2) - 1;
  a[i] = (REAL)av;
  c[i] = (REAL)cv;
  b[i] = (REAL)(2 + fabs(av) + fabs(cv) + hash01(i, 3));
  d[i] = (REAL)(2 * hash01(i, 4) - 1);
  if (u) u[i] = (REAL)(2 * hash01(i, 5) - 1);
}

size_t extent_of(const Geom &g) {
  return (size_t)((g.n0 - 1) * g.st0 + (g.n1 - 1) * g.st1 + (g.N - 1) * g.stride + 1);
}

struct Scratch {
  void *mem = nullptr;
  size_t n = 0;
  ~Scratch() {
    if (mem) (void)hipFree(mem);
  }
};

// Times fn() on st: one warm-up, then the median of `reps` runs.
template <typename F>
float time_on(hipStream_t st, F fn, int reps = 3) {
  constexpr int MAX_REPS = 9;
  reps = std::min(std::max(reps, 1), MAX_REPS);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return INFINITY;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return INFINITY;
  }
  float t[MAX_REPS];
  for (int k = 0; k < MAX_REPS; ++k) t[k] = INFINITY;
  if (fn() == hipSuccess && hipStreamSynchronize(st) == hipSuccess) {
    for (int k = 0; k < reps; ++k) {
      (void)hipEventRecord(e0, st);
      if (fn() != hipSuccess) break;
      (void)hipEventRecord(e1, st);
      if (hipEventSynchronize(e1) != hipSuccess) break;
      (void)hipEventElapsedTime(&t[k], e0, e1);
    }
  }
  (void)hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::sort(t, t + reps);
  return t[reps / 2];  // median
}

template <typename REAL>
std::vector<HybridCfg> candidates(const Geom &g) {
  std::vector<HybridCfg> out;
  static const int Ms[] = {1, 2, 4, 8, 16};
  static const int Ws[] = {1, 2, 4, 8, 16, 32, 64};
  static const int Gs[] = {1, 8, 16};
  for (int M : Ms)
    for (int W : Ws)
      for (int G : Gs) {
        HybridCfg c;
        c.M = M;
        c.P = (int)((g.N + M - 1) / M);
        c.W = W;
        c.G = G;
        if (g.stride != 1 && (int64_t)W * (int64_t)sizeof(REAL) < 64 && W < g.n0) continue;
        if ((int64_t)W > g.n0 && W > 1) continue;
        if ((int64_t)c.W * c.P < 32 && c.W < g.n0) continue;  // under-filled workgroups
        if (hybrid_cfg_valid<REAL>(g, c)) out.push_back(c);
        if (g.stride != 1) {  // streamed strided variants of the same shape
          for (int V : {1, 2}) {
            c.V = V;
            if (hybrid_cfg_valid<REAL>(g, c)) out.push_back(c);
          }
        }
      }
  if (g.stride == 1 && g.N > 4096) {  // one system of up to 8192 (256 lanes) or 16384
                                       // rows (512 lanes) per workgroup, a in LDS
    for (int P : {256, 512})
      for (int G : {1, 8}) {
        HybridCfg c;
        c.M = 32; c.P = P; c.W = 1; c.G = G; c.V = 4;
        if (hybrid_cfg_valid<REAL>(g, c)) out.push_back(c);
      }
  }
  if (g.stride != 1) {  // wide variant: 128- or 256-byte rows, a in LDS; finer XCD groupings
    for (int seg : {128, 256})
      for (int M : {16, 32})
        for (int G : {1, 2, 4, 8, 16, 32}) {
          HybridCfg c;
          c.M = M;
          c.W = seg / (int)sizeof(REAL);
          const int per_wave = std::max(1, 64 / c.W);  // whole waves
          c.P = (int)((g.N + M - 1) / M);
          c.P = (c.P + per_wave - 1) / per_wave * per_wave;
          c.G = G;
          c.V = 3;
          if (hybrid_cfg_valid<REAL>(g, c)) out.push_back(c);
        }
  }
  return out;
}

template <typename REAL, bool INC>
tridStatus_t tune_solve(int dev, const Geom &g, Choice *best) {
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return TRID_STATUS_EXECUTION_FAILED;
  const size_t n = extent_of(g);
  const int arrays = INC ? 5 : 4;
  Scratch s;
  if (hipMalloc(&s.mem, sizeof(REAL) * n * arrays) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(st);
    return TRID_STATUS_ALLOC_FAILED;
  }
  REAL *a = static_cast<REAL *>(s.mem), *b = a + n, *c = b + n, *d = c + n;
  REAL *u = INC ? d + n : nullptr;
  hipLaunchKernelGGL(tune_fill<REAL>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a,
                     b, c, d, u, n);
  best->algo = 0;
  best->ms = INFINITY;
  // every candidate with a median of 3, then the fastest few again with a
  // median of 9: shapes within a few percent of each other are told apart
  // by the second pass, not by one noisy launch
  std::vector<std::pair<float, HybridCfg>> timed;
  for (const HybridCfg &cfg : candidates<REAL>(g)) {
    const float ms = time_on(st, [&] { return launch_hybrid<REAL, INC>(g, cfg, a, b, c, d, u, st); });
    if (std::isfinite(ms)) timed.emplace_back(ms, cfg);
  }
  std::sort(timed.begin(), timed.end(),
            [](const std::pair<float, HybridCfg> &x, const std::pair<float, HybridCfg> &y) {
              return x.first < y.first;
            });
  constexpr size_t FINALISTS = 4;
  for (size_t k = 0; k < timed.size() && k < FINALISTS; ++k) {
    const HybridCfg &cfg = timed[k].second;
    const float ms =
        time_on(st, [&] { return launch_hybrid<REAL, INC>(g, cfg, a, b, c, d, u, st); }, 9);
    if (ms < best->ms) {
      best->ms = ms;
      best->algo = 3;
      best->cfg = cfg;
    }
  }
  // systems longer than one workgroup holds: the long-system path
  {
    HybridCfg one;
    const size_t lwb = long_workspace_bytes<REAL>(g);
    void *lws = nullptr;
    if (!hybrid_default_cfg<REAL>(g, &one) && lwb < (size_t(16) << 30) &&
        hipMalloc(&lws, lwb) == hipSuccess) {
      const float ms = time_on(st, [&] {
        return launch_long<REAL, INC>(dev, g, a, b, c, d, u, lws, lwb, st);
      });
      if (ms < best->ms) {
        best->ms = ms;
        best->algo = 12;
      }
      (void)hipStreamSynchronize(st);
      (void)hipFree(lws);
    } else {
      (void)hipGetLastError();
    }
  }
  // per-lane Thomas, when its workspace fits comfortably
  const size_t wsb = sizeof(REAL) * (size_t)g.nsys() * (size_t)g.N * (INC ? 2 : 1);
  void *ws = nullptr;
  if (wsb < (size_t(16) << 30) && hipMalloc(&ws, wsb) == hipSuccess) {
    REAL *cws = static_cast<REAL *>(ws);
    REAL *dws = INC ? cws + (size_t)g.nsys() * (size_t)g.N : nullptr;
    const float ms = time_on(st, [&] { return launch_lane<REAL, INC>(g, a, b, c, d, u, cws, dws, st); });
    if (ms < best->ms) {
      best->ms = ms;
      best->algo = 1;
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(ws);
  } else {
    (void)hipGetLastError();
  }
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  return best->algo ? TRID_STATUS_SUCCESS : TRID_STATUS_EXECUTION_FAILED;
}

tridStatus_t tune_mixed(int dev, const Geom &g, Choice *best) {
  (void)dev;
  {
    // no mixed kernel shape holds these systems: the call runs the FP64
    // batched solve, nothing to tune here (algo 15, no time)
    HybridCfg any;
    if (!mixed_default_cfg(g, &any)) {
      best->algo = 15;
      best->ms = 0.0f;
      return TRID_STATUS_SUCCESS;
    }
  }
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return TRID_STATUS_EXECUTION_FAILED;
  const size_t n = extent_of(g);
  Scratch s;
  if (hipMalloc(&s.mem, sizeof(double) * n * 4) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(st);
    return TRID_STATUS_ALLOC_FAILED;
  }
  double *a = static_cast<double *>(s.mem), *b = a + n, *c = b + n, *d = c + n;
  hipLaunchKernelGGL(tune_fill<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a,
                     b, c, d, (double *)nullptr, n);
  best->algo = 0;
  best->ms = INFINITY;
  // M = 4, 8: every operand in registers; M = 16: a and c in LDS (V = 3,
  // P rounded up to whole waves for strided systems)
  static const int Ms[] = {4, 8, 16};
  static const int Ws[] = {1, 2, 4, 8, 16, 32};
  static const int Gs[] = {1, 8, 16};
  for (int M : Ms)
    for (int W : Ws)
      for (int G : Gs) {
        HybridCfg cfg;
        cfg.M = M;
        cfg.P = (int)((g.N + M - 1) / M);
        cfg.W = W;
        cfg.G = G;
        if (M == 16) {
          cfg.V = 3;
          if (g.stride != 1 && W < 64) cfg.P = (cfg.P + 64 / W - 1) / (64 / W) * (64 / W);
        }
        if ((int64_t)W > g.n0 && W > 1) continue;
        if (!mixed_cfg_valid(g, cfg)) continue;
        const float ms = time_on(st, [&] { return launch_mixed(g, cfg, 2, a, b, c, d, st); });
        if (ms < best->ms) {
          best->ms = ms;
          best->algo = 2;
          best->cfg = cfg;
        }
      }
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  return best->algo ? TRID_STATUS_SUCCESS : TRID_STATUS_EXECUTION_FAILED;
}

// The FP32-stored solve (tridDmtsvStridedBatchMixedF32) at `refine`
// refinements (1, or 2 for the shapes of more than one): contiguous systems
// the register kernel and the kernel with a and c in LDS over M x workgroup
// size x G; strided ones the wide kernel (128- and 64-byte rows) over G, and
// the register kernel at 16 and 8 rows per lane. Shapes without a kernel for
// `refine` are skipped (mixed32_cfg_valid).
tridStatus_t tune_mixed32(int dev, const Geom &g, int refine, Choice *best) {
  (void)dev;
  {
    // no shape holds these systems: the call promotes them and runs the FP64
    // batched solve, nothing to tune here (algo 16, no time)
    HybridCfg any;
    if (!mixed32_default_cfg(g, refine, &any)) {
      best->algo = 16;
      best->ms = 0.0f;
      return TRID_STATUS_SUCCESS;
    }
  }
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return TRID_STATUS_EXECUTION_FAILED;
  const size_t n = extent_of(g);
  Scratch s;
  if (hipMalloc(&s.mem, sizeof(float) * n * 4 + sizeof(double) * n + 256) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(st);
    return TRID_STATUS_ALLOC_FAILED;
  }
  float *a = static_cast<float *>(s.mem), *b = a + n, *c = b + n, *d = c + n;
  double *x = reinterpret_cast<double *>(
      (reinterpret_cast<uintptr_t>(d + n) + 255) & ~uintptr_t(255));
  hipLaunchKernelGGL(tune_fill<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b,
                     c, d, (float *)nullptr, n);
  best->algo = 0;
  best->ms = INFINITY;
  std::vector<HybridCfg> cands;
  static const int Gs[] = {1, 8, 16, 32};
  if (g.stride == 1) {  // V = 1: a and c kept in LDS
    for (int V : {0, 1})
      for (int M : {8, 16})
        for (int W : {1, 2, 4, 8, 16})
          for (int G : {1, 8}) {
            HybridCfg cfg;
            cfg.M = M; cfg.P = (int)((g.N + M - 1) / M); cfg.W = W; cfg.G = G; cfg.V = V;
            if ((int64_t)W > g.n0 && W > 1) continue;
            cands.push_back(cfg);
          }
  } else {
    const int wide[][2] = {{32, 32}, {32, 16}, {16, 16}};  // {M, W}
    for (const auto &mw : wide)
      for (int G : Gs) {
        HybridCfg cfg;
        cfg.M = mw[0]; cfg.W = mw[1]; cfg.G = G; cfg.V = 3;
        const int per_wave = 64 / cfg.W;
        cfg.P = (int)((g.N + cfg.M - 1) / cfg.M);
        cfg.P = (cfg.P + per_wave - 1) / per_wave * per_wave;
        cands.push_back(cfg);
      }
    for (int V : {0, 1}) {  // V = 1 (a and c in LDS) for more than one refinement
      if (V == 1 && refine <= 1) continue;
      for (int M : {8, 16})
        for (int W : {4, 8, 16}) {
          HybridCfg cfg;
          cfg.M = M; cfg.P = (int)((g.N + M - 1) / M); cfg.W = W; cfg.G = 8; cfg.V = V;
          cands.push_back(cfg);
        }
    }
  }
  // median of 3 for every candidate, then median of 9 for the fastest few
  // (the z shapes at XCD grouping 16 and 32 lie within a few percent)
  std::vector<std::pair<float, HybridCfg>> timed;
  for (const HybridCfg &cfg : cands) {
    if (!mixed32_cfg_valid(g, cfg, refine)) continue;
    const float ms =
        time_on(st, [&] { return launch_mixed32(g, cfg, refine, a, b, c, d, x, st); });
    if (std::isfinite(ms)) timed.emplace_back(ms, cfg);
  }
  std::sort(timed.begin(), timed.end(),
            [](const std::pair<float, HybridCfg> &p, const std::pair<float, HybridCfg> &q) {
              return p.first < q.first;
            });
  constexpr size_t FINALISTS = 4;
  for (size_t k = 0; k < timed.size() && k < FINALISTS; ++k) {
    const HybridCfg &cfg = timed[k].second;
    const float ms =
        time_on(st, [&] { return launch_mixed32(g, cfg, refine, a, b, c, d, x, st); }, 9);
    if (ms < best->ms) {
      best->ms = ms;
      best->algo = 10;
      best->cfg = cfg;
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  return best->algo ? TRID_STATUS_SUCCESS : TRID_STATUS_EXECUTION_FAILED;
}

}  // namespace

template <typename REAL, bool INC>
int tuned_solve_algo(int dev, const Geom &g, HybridCfg *cfg) {
  const Key k = key_of(dev, KIND_SOLVE, sizeof(REAL) == 8, INC, g);
  Choice c;
  if (!lookup(k, &c) && autotune_mode() == 1 && !capturing(current_stream())) {
    if (tune_solve<REAL, INC>(dev, g, &c) == TRID_STATUS_SUCCESS) store(k, c);
    else c.algo = 0;
  } else if (!lookup(k, &c)) {
    c.algo = 0;
  }
  if (c.algo == 3 && hybrid_cfg_valid<REAL>(g, c.cfg)) {
    *cfg = c.cfg;
    return 3;
  }
  if (c.algo == 1 || c.algo == 12) return c.algo;
  return hybrid_default_cfg<REAL>(g, cfg) ? 3 : 12;
}

template <typename REAL, bool INC>
bool tuned_config(int dev, const Geom &g, HybridCfg *cfg) {
  return tuned_solve_algo<REAL, INC>(dev, g, cfg) == 3;
}

bool mixed_tuned_config(int dev, const Geom &g, HybridCfg *cfg) {
  const Key k = key_of(dev, KIND_MIXED, 1, 0, g);
  Choice c;
  if (!lookup(k, &c) && autotune_mode() == 1 && !capturing(current_stream())) {
    if (tune_mixed(dev, g, &c) == TRID_STATUS_SUCCESS) store(k, c);
    else c.algo = 0;
  } else if (!lookup(k, &c)) {
    c.algo = 0;
  }
  if (c.algo == 2 && mixed_cfg_valid(g, c.cfg)) {
    *cfg = c.cfg;
    return true;
  }
  return mixed_default_cfg(g, cfg);
}

// Two tuned shapes per geometry: for at most one refinement (timed at 1) and
// for more (timed at 2; the key's second flag).
bool mixed32_tuned_config(int dev, const Geom &g, int refine, HybridCfg *cfg) {
  const bool general = refine > 1;
  const Key k = key_of(dev, KIND_MIXED32, 0, general, g);
  Choice c;
  if (!lookup(k, &c) && autotune_mode() == 1 && !capturing(current_stream())) {
    if (tune_mixed32(dev, g, general ? 2 : 1, &c) == TRID_STATUS_SUCCESS) store(k, c);
    else c.algo = 0;
  } else if (!lookup(k, &c)) {
    c.algo = 0;
  }
  if (c.algo == 10 && mixed32_cfg_valid(g, c.cfg, refine)) {
    *cfg = c.cfg;
    return true;
  }
  return mixed32_default_cfg(g, refine, cfg);
}

// Times every valid adjoint shape (M rows per lane, W systems per
// workgroup, XCD grouping G) on scratch data of g's extent, with (grads) or
// without the coefficient gradients; the fastest four again, median of 9.
template <typename REAL>
tridStatus_t tune_adjoint(int dev, const Geom &g, bool grads, Choice *best) {
  (void)dev;
  {
    // no adjoint kernel shape holds these systems: the call takes the
    // transposed batched solve, nothing to tune (algo 14, no time)
    HybridCfg any;
    if (!adjoint_default_cfg<REAL>(g, &any)) {
      best->algo = 14;
      best->ms = 0.0f;
      return TRID_STATUS_SUCCESS;
    }
  }
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return TRID_STATUS_EXECUTION_FAILED;
  const size_t n = extent_of(g);
  const int arrays = grads ? 9 : 6;  // a b c x w lam (ga gb gc)
  Scratch s;
  if (hipMalloc(&s.mem, sizeof(REAL) * n * arrays) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(st);
    return TRID_STATUS_ALLOC_FAILED;
  }
  REAL *a = static_cast<REAL *>(s.mem), *b = a + n, *c = b + n, *x = c + n, *w = x + n;
  REAL *lam = w + n;
  REAL *ga = grads ? lam + n : nullptr, *gb = grads ? lam + 2 * n : nullptr;
  REAL *gc = grads ? lam + 3 * n : nullptr;
  hipLaunchKernelGGL(tune_fill<REAL>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a,
                     b, c, x, w, n);
  best->algo = 0;
  best->ms = INFINITY;
  static const int Ms[] = {1, 2, 4, 8, 16};
  static const int Ws[] = {1, 2, 4, 8, 16, 32, 64};
  static const int Gs[] = {1, 8, 16};
  std::vector<std::pair<float, HybridCfg>> timed;
  for (int M : Ms)
    for (int W : Ws)
      for (int G : Gs) {
        HybridCfg cfg;
        cfg.M = M;
        cfg.P = (int)((g.N + M - 1) / M);
        cfg.W = W;
        cfg.G = G;
        if (g.stride != 1 && (int64_t)W * (int64_t)sizeof(REAL) < 64 && W < g.n0) continue;
        if ((int64_t)W > g.n0 && W > 1) continue;
        if ((int64_t)cfg.W * cfg.P < 32 && cfg.W < g.n0) continue;
        if (!adjoint_cfg_valid<REAL>(g, cfg)) continue;
        const float ms = time_on(st, [&] {
          return launch_adjoint<REAL>(g, cfg, a, b, c, x, w, lam, ga, gb, gc, st);
        });
        if (std::isfinite(ms)) timed.emplace_back(ms, cfg);
      }
  std::sort(timed.begin(), timed.end(),
            [](const std::pair<float, HybridCfg> &p, const std::pair<float, HybridCfg> &q) {
              return p.first < q.first;
            });
  for (size_t k = 0; k < timed.size() && k < 4; ++k) {
    const HybridCfg &cfg = timed[k].second;
    const float ms = time_on(st, [&] {
      return launch_adjoint<REAL>(g, cfg, a, b, c, x, w, lam, ga, gb, gc, st);
    }, 9);
    if (ms < best->ms) {
      best->ms = ms;
      best->algo = 4;
      best->cfg = cfg;
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  return best->algo ? TRID_STATUS_SUCCESS : TRID_STATUS_EXECUTION_FAILED;
}

template <typename REAL>
bool adjoint_tuned_config(int dev, const Geom &g, bool grads, HybridCfg *cfg) {
  const Key k = key_of(dev, KIND_ADJOINT, sizeof(REAL) == 8, grads ? 1 : 0, g);
  Choice c;
  if (!lookup(k, &c) && autotune_mode() == 1 && !capturing(current_stream())) {
    if (tune_adjoint<REAL>(dev, g, grads, &c) == TRID_STATUS_SUCCESS) store(k, c);
    else c.algo = 0;
  } else if (!lookup(k, &c)) {
    c.algo = 0;
  }
  if (c.algo == 4 && adjoint_cfg_valid<REAL>(g, c.cfg)) {
    *cfg = c.cfg;
    return true;
  }
  return adjoint_default_cfg<REAL>(g, cfg);
}

template <typename REAL>
void ragged_tuned_config(int dev, int64_t max_len, int *M, int *per_cu, int *launches) {
  Geom g{max_len, 1, 0, 0, 0, 0};
  const Key k = key_of(dev, KIND_RAGGED, sizeof(REAL) == 8, 0, g);
  Choice c;
  if (lookup(k, &c) && c.algo == 5 && ragged_max_len(sizeof(REAL) == 8, c.cfg.M) >= max_len) {
    *M = c.cfg.M;
    *per_cu = c.cfg.G;
    *launches = (c.cfg.V >= 1 && c.cfg.V <= 3) ? c.cfg.V : 1;
    return;
  }
  // untuned: where systems may exceed 4096 rows, 16 rows per lane -- FP64
  // in the split launch (systems over 4096 rows at 32 rows per lane), FP32
  // in one launch of two 512-thread workgroups per CU (the tuned choices on
  // the 40,000-system batch of lengths up to 8192: FP64 0.45 ms against
  // 0.53 in one launch; FP32 0.27 against 0.30 split and 0.35 at one
  // workgroup per CU); else 8 rows per lane in one launch
  *M = max_len > 4096 ? 16 : 8;
  *per_cu = (max_len > 4096 && sizeof(REAL) == 4) ? 2 : 1;
  *launches = (max_len > 4096 && sizeof(REAL) == 8) ? 3 : 1;
}

// Tunes the ragged kernel on the caller's offsets with scratch values.
template <typename REAL>
tridStatus_t tune_ragged(int dev, const int64_t *offsets, int64_t nsys, int64_t max_len) {
  int64_t total = 0;
  if (hipMemcpy(&total, offsets + nsys, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
    return TRID_STATUS_EXECUTION_FAILED;
  if (total <= 0) return TRID_STATUS_SUCCESS;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return TRID_STATUS_EXECUTION_FAILED;
  Scratch s;
  const size_t n = (size_t)total;
  if (hipMalloc(&s.mem, sizeof(REAL) * n * 4 + ragged_workspace_bytes(nsys) + 256) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(st);
    return TRID_STATUS_ALLOC_FAILED;
  }
  REAL *a = static_cast<REAL *>(s.mem), *b = a + n, *c = b + n, *d = c + n;
  void *ws = reinterpret_cast<void *>(
      (reinterpret_cast<uintptr_t>(d + n) + 255) & ~uintptr_t(255));
  hipLaunchKernelGGL(tune_fill<REAL>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b,
                     c, d, (REAL *)nullptr, n);
  Choice best;
  best.algo = 0;
  best.ms = INFINITY;
  // One persistent solve launch, and for batches with systems over 4096 rows
  // the split (one launch of 256-thread workgroups: those systems at 32 rows
  // per lane, then the rest at 16). The two-launch split
  // (tridHipForceRaggedConfig) measured within 2-3 % of one launch on 40,000
  // systems of 8-8192 rows and is not timed here.
  struct Cand { int M, per_cu, launches; };
  std::vector<Cand> cands;
  for (int M : {4, 8, 16})
    for (int per_cu : {1, 2}) cands.push_back({M, per_cu, 1});
  if (max_len > 4096) cands.push_back({16, 1, 3});
  for (const Cand &cd : cands) {
    if (ragged_max_len(sizeof(REAL) == 8, cd.M) < max_len) continue;
    const float ms = time_on(st, [&] {
      return launch_ragged<REAL, false>(cd.M, a, b, c, d, nullptr, offsets, nsys, cd.per_cu,
                                        cd.launches, ws, st, dev);
    });
    if (ms < best.ms) {
      best.ms = ms;
      best.algo = 5;
      best.cfg.M = cd.M;
      best.cfg.G = cd.per_cu;
      best.cfg.V = cd.launches;
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  if (!best.algo) return TRID_STATUS_EXECUTION_FAILED;
  Geom g{max_len, 1, 0, 0, 0, 0};
  store(key_of(dev, KIND_RAGGED, sizeof(REAL) == 8, 0, g), best);
  return TRID_STATUS_SUCCESS;
}

template int tuned_solve_algo<float, false>(int, const Geom &, HybridCfg *);
template int tuned_solve_algo<float, true>(int, const Geom &, HybridCfg *);
template int tuned_solve_algo<double, false>(int, const Geom &, HybridCfg *);
template int tuned_solve_algo<double, true>(int, const Geom &, HybridCfg *);
template bool tuned_config<float, false>(int, const Geom &, HybridCfg *);
template bool tuned_config<float, true>(int, const Geom &, HybridCfg *);
template bool tuned_config<double, false>(int, const Geom &, HybridCfg *);
template bool tuned_config<double, true>(int, const Geom &, HybridCfg *);
template void ragged_tuned_config<float>(int, int64_t, int *, int *, int *);
template void ragged_tuned_config<double>(int, int64_t, int *, int *, int *);
template bool adjoint_tuned_config<float>(int, const Geom &, bool, HybridCfg *);
template bool adjoint_tuned_config<double>(int, const Geom &, bool, HybridCfg *);

}  // namespace trid

using namespace trid;

EXTERN_C tridStatus_t tridHipSetAutotuneMode(int mode) {
  if (mode < 0 || mode > 1) return TRID_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_tmu);
  g_mode = mode;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipAutotune(int fp64, int inc, int ndim, int solvedim, const int *dims,
                                      const int *pads, float *best_ms) {
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  if (empty) return TRID_STATUS_SUCCESS;
  Choice c;
  if (fp64)
    s = inc ? tune_solve<double, true>(dev, g, &c) : tune_solve<double, false>(dev, g, &c);
  else
    s = inc ? tune_solve<float, true>(dev, g, &c) : tune_solve<float, false>(dev, g, &c);
  if (s != TRID_STATUS_SUCCESS) return s;
  store(key_of(dev, KIND_SOLVE, fp64 ? 1 : 0, inc ? 1 : 0, g), c);
  if (best_ms) *best_ms = c.ms;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipAutotuneMixed(int ndim, int solvedim, const int *dims,
                                           const int *pads, float *best_ms) {
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  if (empty) return TRID_STATUS_SUCCESS;
  Choice c;
  if ((s = tune_mixed(dev, g, &c)) != TRID_STATUS_SUCCESS) return s;
  store(key_of(dev, KIND_MIXED, 1, 0, g), c);
  if (best_ms) *best_ms = c.ms;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipAutotuneMixedF32(int ndim, int solvedim, const int *dims,
                                              const int *pads, float *best_ms) {
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  if (empty) return TRID_STATUS_SUCCESS;
  Choice c, c2;
  if ((s = tune_mixed32(dev, g, 1, &c)) != TRID_STATUS_SUCCESS) return s;
  if ((s = tune_mixed32(dev, g, 2, &c2)) != TRID_STATUS_SUCCESS) return s;
  store(key_of(dev, KIND_MIXED32, 0, 0, g), c);
  store(key_of(dev, KIND_MIXED32, 0, 1, g), c2);
  if (best_ms) *best_ms = c.ms;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipAutotuneAdjoint(int fp64, int grads, int ndim, int solvedim,
                                             const int *dims, const int *pads, float *best_ms) {
  Geom g;
  bool empty = false;
  tridStatus_t s = make_geom(ndim, solvedim, dims, pads, &g, &empty);
  if (s != TRID_STATUS_SUCCESS) return s;
  int dev = 0;
  if ((s = check_device(&dev)) != TRID_STATUS_SUCCESS) return s;
  if (empty) return TRID_STATUS_SUCCESS;
  Choice c;
  s = fp64 ? tune_adjoint<double>(dev, g, grads != 0, &c)
           : tune_adjoint<float>(dev, g, grads != 0, &c);
  if (s != TRID_STATUS_SUCCESS) return s;
  store(key_of(dev, KIND_ADJOINT, fp64 ? 1 : 0, grads ? 1 : 0, g), c);
  if (best_ms) *best_ms = c.ms;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipAutotuneRagged(int fp64, const int64_t *offsets, int64_t nsys,
                                            int64_t max_len) {
  if (!offsets || nsys < 0 || max_len < 1) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  tridStatus_t s = check_device(&dev);
  if (s != TRID_STATUS_SUCCESS) return s;
  if (nsys == 0) return TRID_STATUS_SUCCESS;
  return fp64 ? tune_ragged<double>(dev, offsets, nsys, max_len)
              : tune_ragged<float>(dev, offsets, nsys, max_len);
}

// Text format, one entry per line:
//   kind fp64 inc N stride n0 n1 st0 st1 algo M P W G V ms
// (the device is not stored: entries load onto the current device).
EXTERN_C tridStatus_t tridHipAutotuneSave(const char *path) {
  if (!path) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TRID_STATUS_NOT_INITIALIZED;
  FILE *f = std::fopen(path, "w");
  if (!f) return TRID_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_tmu);
  std::fprintf(f, "# tridsolver_amd autotune cache v2\n");
  for (const auto &kv : g_cache) {
    const Key &k = kv.first;
    const Choice &c = kv.second;
    if (k.dev != dev) continue;
    std::fprintf(f, "%d %d %d %lld %lld %lld %lld %lld %lld %d %d %d %d %d %d %.6f\n", k.kind,
                 k.fp64, k.inc, (long long)k.N, (long long)k.stride, (long long)k.n0,
                 (long long)k.n1, (long long)k.st0, (long long)k.st1, c.algo, c.cfg.M, c.cfg.P,
                 c.cfg.W, c.cfg.G, c.cfg.V, c.ms);
  }
  std::fclose(f);
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridHipAutotuneLoad(const char *path) {
  if (!path) return TRID_STATUS_INVALID_VALUE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TRID_STATUS_NOT_INITIALIZED;
  FILE *f = std::fopen(path, "r");
  if (!f) return TRID_STATUS_INVALID_VALUE;
  char line[512];
  tridStatus_t s = TRID_STATUS_SUCCESS;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    Key k{};
    Choice c{};
    long long N, stride, n0, n1, st0, st1;
    if (std::sscanf(line, "%d %d %d %lld %lld %lld %lld %lld %lld %d %d %d %d %d %d %f", &k.kind,
                    &k.fp64, &k.inc, &N, &stride, &n0, &n1, &st0, &st1, &c.algo, &c.cfg.M,
                    &c.cfg.P, &c.cfg.W, &c.cfg.G, &c.cfg.V, &c.ms) != 16) {
      s = TRID_STATUS_INVALID_VALUE;
      break;
    }
    k.dev = dev;
    k.N = N; k.stride = stride; k.n0 = n0; k.n1 = n1; k.st0 = st0; k.st1 = st1;
    std::lock_guard<std::mutex> lk(g_tmu);
    g_cache[k] = c;
  }
  std::fclose(f);
  return s;
}

EXTERN_C tridStatus_t tridHipAutotuneClear(void) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_cache.clear();
  return TRID_STATUS_SUCCESS;
}
