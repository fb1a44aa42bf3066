// launchers.hpp -- host-side launch wrappers of the block kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "block.hpp"
#include "gdtk_device.hpp"

namespace gdtk {

// Cell-validity limits and the low-temperature decode fallback
// (src/eilmer/globalconfig.d:79-86, 1005-1006).
struct CellLimits {
    double low_T = 200.0;
    double max_vel = 30000.0;
    double min_T = 0.0;
    double max_T = 50000.0;
};

hipError_t launch_ghost_fill(const BlockDev& b, double* fs, hipStream_t s);
hipError_t launch_face_flux(const BlockDev& b, const IdealGasDev& gas, int flux_calc, double eps, hipStream_t s);
hipError_t launch_face_flux_dir(const BlockDev& b, const IdealGasDev& gas, int flux_calc, double eps, int d,
                                hipStream_t s);
hipError_t launch_cell_update(const BlockDev& b, const IdealGasDev& gas, int stage, double dt, const CellLimits& lim,
                              hipStream_t s);
int dt_partial_count(const BlockDev& b);
hipError_t launch_dt(const BlockDev& b, double dt_current, double cfl, double* part, double* out3, hipStream_t s);
hipError_t launch_encode(const BlockDev& b, double* U, hipStream_t s);
hipError_t launch_decode(const BlockDev& b, const IdealGasDev& gas, double* U, double* fs, double low_T,
                         hipStream_t s);
// Fused LDS-tiled stage kernel (fused_stage.hip); returns hipErrorNotSupported for
// shapes it does not handle.
void set_fused_kchunk(int kc);
hipError_t launch_fused_stage(const BlockDev& b, const IdealGasDev& gas, int stage, int flux_calc, double dt,
                              double eps, const CellLimits& lim, hipStream_t s);

}  // namespace gdtk
