// trid_tune.h - choice of the hybrid shape for a call (autotuner cache).
#pragma once

#include "trid_internal.h"

namespace trid {

// Fills *cfg with the shape to use for g on device dev and returns true, or
// returns false when the per-lane Thomas path should run instead.
template <typename REAL, bool INC>
bool tuned_config(int dev, const Geom &g, HybridCfg *cfg);

}  // namespace trid
