// chem_launch.hpp -- host-side launchers of the chemistry kernels (chem_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "chem_device.hpp"

namespace gdtk {

constexpr int CHEM_BLOCK = 256;   // threads per workgroup of every chemistry kernel
constexpr int CHEM_WAVES_PER_BLOCK = CHEM_BLOCK / 64;
// Workgroups of an integration launch over n cells are at most max(ceil(n / CHEM_BLOCK), 8)
// (the persistent grid is capped there too), so the per-wave statistics need at most this
// many waves.
inline int chem_stat_waves(int n) {
    const int g = (n + CHEM_BLOCK - 1) / CHEM_BLOCK;
    return CHEM_WAVES_PER_BLOCK * (g > 8 ? g : 8);
}
constexpr int QUEUE_STRIDE = 32;  // queue counters 128 bytes apart

// Per-cell flag written by the setup kernel.  CF_SKIP: a cell the caller's mask takes out
// of this call (neither read nor written).
// CF_HEAVY: to integrate, taken from the heavy list (first) rather than the cell ranges.
enum ChemFlag : int { CF_INTEGRATE = 0, CF_FROZEN = 1, CF_FAILED = 2, CF_SKIP = 3, CF_HEAVY = 4 };
constexpr int QUEUE_COUNTERS = 10;  // 8 per-XCD ranges, heavy next, heavy length

// Device work buffers of one chemistry call over n cells (species-major arrays).
struct ChemWork {
    double* kfkb;       // [2*nr][n]  forward then backward rate constants
    double* conc0;      // [nsp][n]   initial concentrations
    double* h0;         // [n]        first substep size
    double* hest;       // [n]        step estimate from the initial state (retry's first step)
    double* conc;       // [nsp][n]   final concentrations of integrated cells
    int* flag;          // [n]        ChemFlag
    int* heavy;         // [n]        cells flagged CF_HEAVY, in the order setup listed them
    int* st;            // [n]        0 integrated, -1 failed (integrated cells)
    int* nsub;          // [n]        substeps per cell
    unsigned* queue;    // [10*32]    per-XCD next-cell counters, then the heavy list's next
                        //            entry and length, 128 B apart
    unsigned long long* stats;  // [4][waves] per integration wave: substeps, corrector iterations,
                                //            max substeps, step attempts (plain stores, no atomics)
    double* Tend;       // [n]        tight temperature coupling: T and u as the last subcycle's
    double* uend;       // [n]        update left them (the start of the final T(u) update)
};

// Launch shape of the integration kernel.
struct ChemTuning {
    int persistent = -1;    // 1: persistent lanes refilled from the work queue; 0: one lane per cell;
                            // -1: persistent when the batch exceeds 2048 cells per CU
    int chunk = 64;         // cells a wave takes from the queue per grab
    int refill_period = 1;  // substep iterations between refills of finished lanes
    int blocks_per_cu = 0;  // persistent grid: workgroups per CU (0 = occupancy limit)
    int heavy_ratio = 4;    // setup lists cells with dt > heavy_ratio * first substep (0 = none)
};

enum ChemPhase : int { CHEM_PHASE_SETUP = 0, CHEM_PHASE_INTEGRATE = 1, CHEM_PHASE_FINISH = 2, CHEM_PHASE_TPG = 3 };

// Arguments of one kernel phase (the fields a phase does not use stay zero).
struct ChemCall {
    int phase;
    const ChemDev* cd;
    int nr, n, mode, method, tight, dense;
    double dt;
    const double *rho, *T, *u, *p, *massf, *dtc_in;
    const int* mask;  // setup: cells with mask[i] != 0 are skipped (may be null)
    double *T_out, *u_out, *p_out, *massf_out, *dtc_out;
    double *rho_io, *p_io, *T_io, *u_io, *a_io;
    int* status;
    ChemWork w;
    ChemTuning tu;
    int num_cu;
    int* waves_out;  // integrate: the waves launched
    hipStream_t s;
};

// Instantiated per species count in chem_nsp<N>.hip.
template <int NSP>
hipError_t chem_run(const ChemCall& c);
extern template hipError_t chem_run<1>(const ChemCall&);
extern template hipError_t chem_run<2>(const ChemCall&);
extern template hipError_t chem_run<3>(const ChemCall&);
extern template hipError_t chem_run<4>(const ChemCall&);
extern template hipError_t chem_run<5>(const ChemCall&);
extern template hipError_t chem_run<6>(const ChemCall&);
extern template hipError_t chem_run<7>(const ChemCall&);
extern template hipError_t chem_run<8>(const ChemCall&);

bool chem_supported(int nsp, int nr);
// Zero the queue counters (a fill kernel, so the call may be captured).
hipError_t launch_chem_clear(const ChemWork& w, hipStream_t s);
// dense: the mechanism's lists are all in ascending species order (ChemDev::all_sorted).
hipError_t launch_chem_setup(const ChemDev* cd, int nsp, int nr, int dense, int n, double dt, const double* rho,
                             const double* T, const double* massf, const double* dtc, const int* mask,
                             int heavy_ratio, const ChemWork& w, hipStream_t s);
// *waves: the waves launched, each of which writes its statistics entries.
hipError_t launch_chem_integrate(const ChemDev* cd, int nsp, int nr, int dense, int method, int tight, int n, double dt,
                                 const double* rho, const double* T, const double* u, const double* dtc_in,
                                 double* dtc_out, const ChemWork& w, const ChemTuning& tu, int num_cu, int* waves,
                                 hipStream_t s);
hipError_t launch_chem_finish(const ChemDev* cd, int nsp, int n, const double* rho, const double* T,
                              const double* u, const double* p, const double* massf, const double* dtc_in,
                              double* T_out, double* u_out, double* p_out, double* massf_out, double* dtc_out,
                              int* status, const ChemWork& w, hipStream_t s);
hipError_t launch_tpg_update(const ChemDev* cd, int nsp, int mode, int n, double* rho, double* p, double* T,
                             double* u, double* a, const double* massf, int* status, hipStream_t s);

}  // namespace gdtk
