// surrogate_launch.hpp -- launchers of the learned chemistry step (surrogate_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "chem_device.hpp"

namespace gdtk {

constexpr int SURR_MAX_FEAT = CH_MAX_SP + 3;
constexpr int SURR_MAX_ELEM = 4;

// Feature scaling, output scaling, element matrix and fallback thresholds (kernel argument).
struct SurrCfg {
    double x_mean[SURR_MAX_FEAT], x_inv_std[SURR_MAX_FEAT];
    double y_scale[CH_MAX_SP];
    double x_abs_max, disagree_tol, T_frozen;
    double elem[SURR_MAX_ELEM][CH_MAX_SP];
    int n_elem;
    double route_ratio, route_T;
};

// Zero k counters (a fill kernel, so the call may be captured).
hipError_t launch_surr_clear(unsigned* c, int k, hipStream_t s);
hipError_t launch_surr_route(int nsp, int n, const double* T, const double* dtc, double dt, const SurrCfg& cfg,
                             int* path, int* taken, int* net_idx, unsigned* net_count, hipStream_t s);
hipError_t launch_surr_features(int nsp, int n, const int* net_idx, const unsigned* net_count, const double* rho,
                                const double* T, const double* massf, double dt, const SurrCfg& cfg, float* x,
                                hipStream_t s);
hipError_t launch_surr_apply(int nsp, int n, const int* net_idx, const unsigned* net_count, const float* x,
                             const float* ya, const float* yb, const double* massf, const SurrCfg& cfg,
                             double* massf_new, int* path, int* taken, unsigned* fb_count, double* disagree,
                             hipStream_t s);
hipError_t launch_surr_finish(int nsp, int n, const int* net_idx, const unsigned* net_count, const double* rho,
                              const double* T, const double* u, const double* p, const double* massf, const double* dtc,
                              const double* massf_new, const ChemDev* cd, double* T_out, double* u_out, double* p_out,
                              double* massf_out, double* dtc_out, int* status, int* path, hipStream_t s);
hipError_t launch_surr_project(int nsp, int n, const double* Y0, const double* d0, const SurrCfg& cfg, double* d,
                               int* ok, hipStream_t s);

}  // namespace gdtk
