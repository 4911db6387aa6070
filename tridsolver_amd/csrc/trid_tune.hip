// trid_tune.hip - shape selection for the Thomas-PCR hybrid.
#include "trid_tune.h"

namespace trid {

template <typename REAL, bool INC>
bool tuned_config(int dev, const Geom &g, HybridCfg *cfg) {
  (void)dev;
  return hybrid_default_cfg<REAL>(g, cfg);
}

bool mixed_tuned_config(int dev, const Geom &g, HybridCfg *cfg) {
  (void)dev;
  return mixed_default_cfg(g, cfg);
}

template <typename REAL>
void ragged_tuned_config(int dev, int64_t max_len, int *M, int *per_cu) {
  (void)dev;
  *per_cu = 1;
  *M = 8;
  if (sizeof(REAL) == 8 && max_len > ragged_max_len(1, 8)) *M = 16;
}

template void ragged_tuned_config<float>(int, int64_t, int *, int *);
template void ragged_tuned_config<double>(int, int64_t, int *, int *);

template bool tuned_config<float, false>(int, const Geom &, HybridCfg *);
template bool tuned_config<float, true>(int, const Geom &, HybridCfg *);
template bool tuned_config<double, false>(int, const Geom &, HybridCfg *);
template bool tuned_config<double, true>(int, const Geom &, HybridCfg *);

}  // namespace trid
