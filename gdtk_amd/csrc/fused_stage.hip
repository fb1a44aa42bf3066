// fused_stage.hip -- fused LDS-tiled stage kernel (placeholder until the tiled kernel lands).
#include <hip/hip_runtime.h>

#include "launchers.hpp"

namespace gdtk {

hipError_t launch_fused_stage(const BlockDev&, const IdealGasDev&, int, int, double, double, const CellLimits&,
                              hipStream_t) {
    return hipErrorNotSupported;
}

}  // namespace gdtk
