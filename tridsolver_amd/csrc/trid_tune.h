// trid_tune.h - choice of the hybrid shape for a call (autotuner cache).
#pragma once

#include "trid_internal.h"

namespace trid {

// The algorithm for g on device dev: 3 the hybrid (*cfg filled), 1 the
// per-lane Thomas kernel, 12 the long-system path (systems longer than one
// workgroup holds). A cached tuning result decides, else the defaults.
template <typename REAL, bool INC>
int tuned_solve_algo(int dev, const Geom &g, HybridCfg *cfg);
// tuned_solve_algo(...) == 3.
template <typename REAL, bool INC>
bool tuned_config(int dev, const Geom &g, HybridCfg *cfg);

// Shape of the fused FP32/FP64-refinement kernel for g.
bool mixed_tuned_config(int dev, const Geom &g, HybridCfg *cfg);

// Shape of the FP32-stored solve for g at `refine` refinements.
bool mixed32_tuned_config(int dev, const Geom &g, int refine, HybridCfg *cfg);

// Shape of the adjoint solve for g, with (grads) or without the coefficient
// gradients; false if the systems are longer than one workgroup holds.
template <typename REAL>
bool adjoint_tuned_config(int dev, const Geom &g, bool grads, HybridCfg *cfg);

// Rows per lane M, workgroups per CU and solve launches (1, or 2 with the
// short systems apart) of the ragged persistent kernel for systems up to
// max_len long.
template <typename REAL>
void ragged_tuned_config(int dev, int64_t max_len, int *M, int *per_cu, int *launches);

}  // namespace trid
