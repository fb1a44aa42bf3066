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

// Block time-step estimates of blocks [0, nblocks) of tab: one launch of count lanes (count
// a multiple of 256; start[q] = block q's first launch index, start[nblocks] the end, base =
// the launch index of the first lane) writing one partial per workgroup into part (2 per
// workgroup), then one workgroup per block: out[3q..3q+2] = {min dt_local, max cfl_local,
// last signal}; last_signal holds one double per block.
hipError_t launch_dt_batch(const BlockDev* tab, const int64_t* start, int nblocks, int64_t base, int64_t count,
                           int dims, double dt_current, double cfl, double* part, double* last_signal, double* out,
                           hipStream_t s);
hipError_t launch_encode(const BlockDev& b, double* U, hipStream_t s);
hipError_t launch_decode(const BlockDev& b, const IdealGasDev& gas, double* U, double* fs, double low_T,
                         hipStream_t s);
// Face-per-thread kernels over all blocks: the faces of every direction in one launch
// (FaceBatch), then the cell update (dt read from device memory; start[q] = first launch
// index of block q, a multiple of 256; total = launch size).  numerics: 0 Eilmer 4, 1 Eilmer 5
// (lmr) limiter smoothing and AUSMDV sound speed.
hipError_t launch_face_flux_batch(const BlockDev* tab, const FaceBatch& fb, int nblocks, int dims, int flux_calc,
                                 int numerics, double eps, const IdealGasDev& gas, hipStream_t s);
hipError_t launch_cell_update_batch(const BlockDev* tab, const int64_t* start, int nblocks, int64_t total, int dims,
                                   int stage, const double* dtp, const CellLimits& lim, const IdealGasDev& gas,
                                   hipStream_t s);
// Multi-species thermally perfect gas (flux_species.hip), nsp = 2 or 5: the same batched
// face and cell-update launches, and one block's encode/decode at start-up.
struct ChemDev;
bool species_supported(int nsp);
hipError_t launch_face_flux_species(const BlockDev* tab, const FaceBatch& fb, int nblocks, int dims, int flux_calc,
                                   int numerics, double eps, int nsp, const ChemDev* cd, hipStream_t s);
hipError_t launch_face_flux_species_lmr(const BlockDev* tab, const FaceBatch& fb, int nblocks, int dims, int flux_calc,
                                       double eps, int nsp, const ChemDev* cd, int64_t total, hipStream_t s);
hipError_t launch_cell_update_species(const BlockDev* tab, const int64_t* start, int nblocks, int64_t total, int dims,
                                     int stage, const double* dtp, const CellLimits& lim, int nsp, const ChemDev* cd,
                                     hipStream_t s);
hipError_t launch_encode_species(const BlockDev& b, double* U, hipStream_t s);
hipError_t launch_decode_species(const BlockDev& b, const ChemDev* cd, double* U, double* fs, double low_T,
                                 hipStream_t s);
// Operator-split chemistry: every block's cells to dense arrays (rho, T, u, p [n], massf
// [nsp][n]) and the reactor's results back (flow state and U[0] species densities); the
// scatter counts failed cells into nfail.  coff[q] = first dense index of block q.
hipError_t launch_cells_gather(const BlockDev* tab, const int64_t* start, const int64_t* coff, int nblocks,
                               int64_t total, int nsp, int64_t n, double* rho, double* T, double* u, double* p,
                               double* massf, hipStream_t s);
hipError_t launch_cells_scatter(const BlockDev* tab, const int64_t* start, const int64_t* coff, int nblocks,
                                int64_t total, int nsp, int64_t n, const double* T_o, const double* u_o,
                                const double* p_o, const double* massf_o, const double* dtc_o, double* dtc,
                                const int* status, unsigned* nfail, hipStream_t s);
hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t s);
// Ghost cells of all blocks in one launch (exchange and the other ghost-cell effects);
// mode 1 exchanges the cell lengths of the exchange sides instead; mode 2 also copies the
// shock indicators.
// Shock detection over all blocks' interior cells (PJ detector, strict rule applied by
// the adaptive flux kernels), launched like the cell update.
hipError_t launch_ghost_batch(const BlockDev* tab, const SideDev* sides, int nsides, int64_t total, int mode,
                              hipStream_t s);
hipError_t launch_shock_detect(const BlockDev* tab, const FaceBatch& fb, const int64_t* cstart, int nblocks,
                               int64_t ctotal, int dims, double comp_tol, double shear_tol, int damped_outflow,
                               hipStream_t s);
// Fused LDS-tiled stage kernel over all blocks (fused_stage.hip): tiles of one block for
// k chunks of kc planes, the persistent grid size for the device, and the launch (dt is
// read from device memory).
// Tiles are tx x (256 / tx) cells, tx in {16, 32, 64}.
int fused_tiles(const BlockDev& b, int kc, int tx);
bool fused_tile_x_ok(int tx);
int fused_batch_grid(int dims, int flux_calc, int tx);
hipError_t launch_fused_batch(const StageBatch& sb, int dims, int stage, int flux_calc, const double* dtp,
                              double eps, const CellLimits& lim, const IdealGasDev& gas, int grid, int tx,
                              hipStream_t s);
hipError_t launch_clear_u32(unsigned* p, int n, hipStream_t s);
// One launch copying n doubles from src to dst for every segment: start = the segment's
// first chunk of 1024 doubles (prefix sums of ceil(n / 1024)); src and dst 16-byte aligned.
struct CopySeg {
    const double* src;
    double* dst;
    int64_t n, start;
};
hipError_t launch_copy_segments(const CopySeg* segs, int nseg, int64_t nchunks, hipStream_t s);
// Flow-solution records of one block (flow_io.hip): pack the flow state (fs_in) into nv
// doubles per active cell, and unpack records into the flow-state planes fs.
hipError_t launch_pack_flow(const BlockDev& b, const double* pos, const double* mu, const double* kc, const double* dtc,
                            double* rec, int nv, hipStream_t s);
hipError_t launch_unpack_flow(const BlockDev& b, const double* rec, double* fs, double* mu, double* kc, int nv,
                              hipStream_t s);
hipError_t launch_set_f64(double* p, double v, hipStream_t s);

}  // namespace gdtk
