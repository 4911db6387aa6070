// trid_tune.hip - shape selection for the Thomas-PCR hybrid.
#include "trid_tune.h"

namespace trid {

template <typename REAL, bool INC>
bool tuned_config(int dev, const Geom &g, HybridCfg *cfg) {
  (void)dev;
  return hybrid_default_cfg<REAL>(g, cfg);
}

template bool tuned_config<float, false>(int, const Geom &, HybridCfg *);
template bool tuned_config<float, true>(int, const Geom &, HybridCfg *);
template bool tuned_config<double, false>(int, const Geom &, HybridCfg *);
template bool tuned_config<double, true>(int, const Geom &, HybridCfg *);

}  // namespace trid
