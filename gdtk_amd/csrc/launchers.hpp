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
// Ghost cells of all blocks in one launch (exchange and the other ghost-cell effects);
// mode 1 exchanges the cell lengths of the exchange sides instead.
hipError_t launch_ghost_batch(const BlockDev* tab, const SideDev* sides, int nsides, int64_t total, int mode,
                              hipStream_t s);
// Fused LDS-tiled stage kernel over all blocks (fused_stage.hip): tiles of one block for
// k chunks of kc planes, the persistent grid size for the device, and the launch (dt is
// read from device memory).
int fused_tiles(const BlockDev& b, int kc);
int fused_batch_grid(int dims, int flux_calc);
hipError_t launch_fused_batch(const StageBatch& sb, int dims, int stage, int flux_calc, const double* dtp,
                              double eps, const CellLimits& lim, const IdealGasDev& gas, int grid, hipStream_t s);
hipError_t launch_clear_u32(unsigned* p, int n, hipStream_t s);
hipError_t launch_set_f64(double* p, double v, hipStream_t s);

}  // namespace gdtk
