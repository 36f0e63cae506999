// Fused KEEP + viscous stage kernel (placeholder until the tiled kernel lands).
#include "kernels.h"

namespace amr {

bool fused_supported(const amr_scheme_t&, int, int) { return false; }

hipError_t launch_fused_stage(const amr_scheme_t&, const FusedConfig&, const TileDev*, int, const BoxDev*,
                              const Closures&, const double*, const StageCoef&, const Fab*, const Fab*,
                              const Fab*, hipStream_t) {
  return hipErrorInvalidValue;
}

}  // namespace amr
