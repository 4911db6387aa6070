// trid_internal.h - host-side declarations shared by the C ABI layer and the
// kernel translation units of libtridhip.so. Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/trid_common.h"

namespace trid {

// Shape of one batched solve, after the (ndim, solvedim, dims, pads) of the
// C API have been resolved (the reference does the same per solvedim in
// scalar/src/cpu/trid_cpu.cpp:369-454):
//   system s = i0 + n0 * i1 starts at element i0 * st0 + i1 * st1 and its
//   element n is at that base + n * stride.
// For solvedim 1 and 2, i0 is the x index (st0 = 1), so neighbouring lanes
// read neighbouring addresses. For solvedim 0, i0 is y (st0 = pads[0]) and the
// system itself is contiguous (stride = 1).
struct Geom {
  int64_t N;       // system length (dims[solvedim])
  int64_t stride;  // element stride inside a system
  int64_t n0, n1;  // systems = n0 * n1
  int64_t st0, st1;
  __host__ __device__ int64_t nsys() const { return n0 * n1; }
};

// Tuning of the Thomas-PCR hybrid for one call: M rows per lane, P lanes per
// system, W systems per workgroup (block = W * P threads), and G, the number
// of groups the workgroups are dealt into so that consecutive system groups
// run on the same XCD at the same time (workgroups are dispatched to the 8
// XCDs round-robin; G = 1 keeps the plain order).
struct HybridCfg {
  int M = 0;
  int P = 0;
  int W = 0;
  int G = 8;
};

enum class Algo : int { Auto = 0, Lane = 1, Hybrid = 3 };

// ---- launchers (kernels_solve.hip) --------------------------------------
// Per-lane Thomas, any length. cws: N * nsys reals of workspace; dws: same
// size, only needed for Inc (the plain solve keeps d' in d itself).
template <typename REAL, bool INC>
hipError_t launch_lane(const Geom &g, const REAL *a, const REAL *b, const REAL *c,
                       REAL *d, REAL *u, REAL *cws, REAL *dws, hipStream_t st);

// Thomas-PCR hybrid; contiguous systems (stride == 1) use the LDS-staged
// loader, strided systems load per lane with x-adjacent lanes.
template <typename REAL, bool INC>
hipError_t launch_hybrid(const Geom &g, const HybridCfg &cfg, const REAL *a,
                         const REAL *b, const REAL *c, REAL *d, REAL *u,
                         hipStream_t st);

// ---- shared host helpers (trid_api.hip) ---------------------------------
hipStream_t current_stream();
tridStatus_t check_device(int *dev);
tridStatus_t get_workspace(int dev, hipStream_t st, size_t bytes, void **out);

// Picks a valid hybrid configuration for g, or returns false if the systems
// are too long for one workgroup.
template <typename REAL>
bool hybrid_default_cfg(const Geom &g, HybridCfg *cfg);
template <typename REAL>
bool hybrid_cfg_valid(const Geom &g, const HybridCfg &cfg);

}  // namespace trid
