// surrogate_kernels.hip -- the learned chemistry step around the BF16 network (no
// counterpart in the reference).
//
// Routing: a cell at or below T_frozen is the chemistry step's (left unchanged); of the others,
// the cells whose integration would be long (the interval spans more than route_ratio of the
// cell's last suggested substep, or, with no suggestion yet, T above route_T) go to the
// networks, compacted into a list by an atomic counter, and the rest to the alpha-QSS
// integrator.  Per network cell: features (log T, log rho, Y[nsp], log dt, standardised) ->
// two networks -> mean mass-fraction change -> exact element-conservation projection with an
// active set that keeps every mass fraction >= 0 -> T, u, p by the gas model's Newton T(u)
// (as the reference's thermochemical_increment ends, fvcell.d:906-925).  A network cell whose
// features lie outside the trained range, whose two networks disagree, or whose projection
// cannot be made feasible goes back to the integrator.  The integrator then runs over all
// cells with a mask of the cells the networks answered, so no count returns to the host.
#include <hip/hip_runtime.h>

#include "chem_device.hpp"
#include "surrogate_launch.hpp"

namespace gdtk {

// Active-set least-squares projection: the change d closest to the network's prediction
// d0 with E d = 0 (every element's mass conserved; the rows of E sum to the species' unit
// mass, so total mass is conserved too) and Y0 + d >= 0.  Species driven below zero are
// fixed at d = -Y0 and the rest re-projected, at most NSP rounds.  Returns false if the
// free species cannot carry the element balance.
template <int NSP>
__device__ __forceinline__ bool project(const SurrCfg& cfg, const double* Y0, const double* d0, double* d) {
    constexpr int ME = SURR_MAX_ELEM;
    const int ne = cfg.n_elem;
    bool fixed[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) fixed[s] = false;
    for (int round = 0; round <= NSP; ++round) {
        // G lam = rhs with G = E_U E_U^T (free species U), rhs = E (d0 on U, -Y0 on fixed)
        double G[ME][ME], rhs[ME], L[ME][ME], y[ME], lam[ME];
#pragma unroll
        for (int e = 0; e < ME; ++e) {
            rhs[e] = 0.0;
            y[e] = 0.0;
            lam[e] = 0.0;
#pragma unroll
            for (int f = 0; f < ME; ++f) {
                G[e][f] = (e == f) ? 1.0 : 0.0;
                L[e][f] = 0.0;
            }
        }
#pragma unroll
        for (int e = 0; e < ME; ++e) {
            if (e >= ne) continue;
#pragma unroll
            for (int s = 0; s < NSP; ++s) rhs[e] += cfg.elem[e][s] * (fixed[s] ? -Y0[s] : d0[s]);
#pragma unroll
            for (int f = 0; f < ME; ++f) {
                if (f >= ne) continue;
                double gsum = 0.0;
#pragma unroll
                for (int s = 0; s < NSP; ++s)
                    if (!fixed[s]) gsum += cfg.elem[e][s] * cfg.elem[f][s];
                G[e][f] = gsum;
            }
        }
        // Cholesky G = L L^T (unused rows keep the identity)
#pragma unroll
        for (int c = 0; c < ME; ++c) {
            double diag = G[c][c];
#pragma unroll
            for (int k = 0; k < c; ++k) diag -= L[c][k] * L[c][k];
            if (!(diag > 1.0e-14)) return false;
            L[c][c] = sqrt(diag);
#pragma unroll
            for (int r = c + 1; r < ME; ++r) {
                double v = G[r][c];
#pragma unroll
                for (int k = 0; k < c; ++k) v -= L[r][k] * L[c][k];
                L[r][c] = v / L[c][c];
            }
        }
#pragma unroll
        for (int r = 0; r < ME; ++r) {
            double v = rhs[r];
#pragma unroll
            for (int k = 0; k < r; ++k) v -= L[r][k] * y[k];
            y[r] = v / L[r][r];
        }
#pragma unroll
        for (int r = ME - 1; r >= 0; --r) {
            double v = y[r];
#pragma unroll
            for (int k = r + 1; k < ME; ++k) v -= L[k][r] * lam[k];
            lam[r] = v / L[r][r];
        }
        bool clean = true;
#pragma unroll
        for (int s = 0; s < NSP; ++s) {
            if (fixed[s]) {
                d[s] = -Y0[s];
                continue;
            }
            double corr = 0.0;
#pragma unroll
            for (int e = 0; e < ME; ++e)
                if (e < ne) corr += cfg.elem[e][s] * lam[e];
            d[s] = d0[s] - corr;
            if (Y0[s] + d[s] < 0.0) {
                fixed[s] = true;
                clean = false;
            }
        }
        if (clean) return true;
    }
    return false;
}

// path: 0 network, 1 integrator, 2 frozen (T <= T_frozen: the integrator's call leaves it
// unchanged, fvcell.d:870), 3 failed (network composition without a temperature).  taken:
// 1 where the networks answer the cell (the integrator's mask).
template <int NSP>
__global__ __launch_bounds__(256) void k_surr_route(int n, const double* __restrict__ T,
                                                    const double* __restrict__ dtc, double dt, SurrCfg cfg,
                                                    int* __restrict__ path, int* __restrict__ taken,
                                                    int* __restrict__ net_idx, unsigned* __restrict__ net_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double Ti = T[i];
    if (Ti <= cfg.T_frozen) {
        path[i] = 2;
        taken[i] = 0;
        return;
    }
    const double d = dtc[i];
    const bool stiff = (cfg.route_ratio <= 0.0) || ((d > 0.0) ? (dt > cfg.route_ratio * d) : (Ti > cfg.route_T));
    if (!stiff) {
        path[i] = 1;
        taken[i] = 0;
        return;
    }
    path[i] = 0;
    taken[i] = 1;
    net_idx[atomicAdd(net_count, 1u)] = i;
}

// Features of the listed cells, x [nsp + 3][n] in list order (entries past the count unused).
template <int NSP>
__global__ __launch_bounds__(256) void k_surr_features(int n, const int* __restrict__ net_idx,
                                                       const unsigned* __restrict__ net_count,
                                                       const double* __restrict__ rho,
                                                       const double* __restrict__ T,
                                                       const double* __restrict__ massf, double dt, SurrCfg cfg,
                                                       float* __restrict__ x) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int)*net_count) return;
    const int i = net_idx[j];
    double f[NSP + 3];
    f[0] = log(T[i]);
    f[1] = log(rho[i]);
#pragma unroll
    for (int s = 0; s < NSP; ++s) f[2 + s] = massf[(size_t)s * n + i];
    f[NSP + 2] = log(dt);
#pragma unroll
    for (int k = 0; k < NSP + 3; ++k) x[(size_t)k * n + j] = (float)((f[k] - cfg.x_mean[k]) * cfg.x_inv_std[k]);
}

// Network cells (list order j): the trained-range and agreement tests, the projection; a
// cell that fails one goes back to the integrator (path 1, taken 0).
template <int NSP>
__global__ __launch_bounds__(256) void k_surr_apply(int n, const int* __restrict__ net_idx,
                                                    const unsigned* __restrict__ net_count,
                                                    const float* __restrict__ x, const float* __restrict__ ya,
                                                    const float* __restrict__ yb, const double* __restrict__ massf,
                                                    SurrCfg cfg, double* __restrict__ massf_new, int* __restrict__ path,
                                                    int* __restrict__ taken, unsigned* __restrict__ fb_count,
                                                    double* __restrict__ disagree) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int)*net_count) return;
    const int i = net_idx[j];
    bool fallback = false;
#pragma unroll
    for (int k = 0; k < NSP + 3; ++k)
        if (!(fabsf(x[(size_t)k * n + j]) <= (float)cfg.x_abs_max)) fallback = true;
    double Y0[NSP], d0[NSP], d[NSP];
    double dis = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        Y0[s] = massf[(size_t)s * n + i];
        const double a = ya[(size_t)s * n + j], b = yb[(size_t)s * n + j];
        d0[s] = 0.5 * (a + b) * cfg.y_scale[s];
        dis = fmax(dis, fabs(a - b) * cfg.y_scale[s]);
    }
    disagree[j] = dis;
    if (!(dis <= cfg.disagree_tol)) fallback = true;
    if (!fallback && !project<NSP>(cfg, Y0, d0, d)) fallback = true;
    if (fallback) {
        path[i] = 1;
        taken[i] = 0;
        atomicAdd(fb_count, 1u);
        return;
    }
#pragma unroll
    for (int s = 0; s < NSP; ++s) massf_new[(size_t)s * n + j] = Y0[s] + d[s];
}

// T, u, p of the network cells by Newton T(u) at the new composition (therm_perf_gas.d).
// A composition that admits no temperature is the reference's failure mode of
// thermochemical_increment (update_thermo_from_rhou throws after the chemistry, fvcell.d:
// 906-925): the cell is reported failed (path 3, status -1, state unchanged, dt_chem -1)
// without running the integrator on it.
template <int NSP>
__global__ __launch_bounds__(256) void k_surr_finish(int n, const int* __restrict__ net_idx,
                                                     const unsigned* __restrict__ net_count,
                                                     const double* __restrict__ rho, const double* T,
                                                     const double* u, const double* p_in, const double* massf_in,
                                                     const double* dtc_in, const double* __restrict__ massf_new,
                                                     const ChemDev* __restrict__ cd, double* T_out, double* u_out,
                                                     double* p_out, double* massf_out, double* dtc_out, int* status,
                                                     int* __restrict__ path) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int)*net_count) return;
    const int i = net_idx[j];
    if (path[i] != 0) return;
    double mf[NSP];
    double Tn = T[i], un = u[i];
    const double dc = dtc_in[i];
#pragma unroll
    for (int s = 0; s < NSP; ++s) mf[s] = massf_new[(size_t)s * n + j];
    if (!tpg_update_temperature<NSP>(*cd, mf, Tn, un)) {
        path[i] = 3;
        const double pv = p_in[i];
        T_out[i] = T[i];
        u_out[i] = u[i];
        p_out[i] = pv;
        dtc_out[i] = -1.0;
#pragma unroll
        for (int s = 0; s < NSP; ++s) {
            const double m0 = massf_in[(size_t)s * n + i];
            massf_out[(size_t)s * n + i] = m0;
        }
        status[i] = -1;
        return;
    }
    const double Rm = tpg_Rmix<NSP>(*cd, mf);
    p_out[i] = rho[i] * Rm * Tn;
    T_out[i] = Tn;
    u_out[i] = un;
    dtc_out[i] = dc;
#pragma unroll
    for (int s = 0; s < NSP; ++s) massf_out[(size_t)s * n + i] = mf[s];
    status[i] = 0;
}

// The projection alone, one thread per cell, species-major arrays: d = the projected change
// of d0 at Y0 and ok = 1, or ok = 0 and d = 0 where the projection is infeasible.
template <int NSP>
__global__ __launch_bounds__(256) void k_surr_project(int n, const double* __restrict__ Y0,
                                                      const double* __restrict__ d0, SurrCfg cfg,
                                                      double* __restrict__ d, int* __restrict__ ok) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double y[NSP], dd[NSP], r[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        y[s] = Y0[(size_t)s * n + i];
        dd[s] = d0[(size_t)s * n + i];
    }
    const bool good = project<NSP>(cfg, y, dd, r);
#pragma unroll
    for (int s = 0; s < NSP; ++s) d[(size_t)s * n + i] = good ? r[s] : 0.0;
    ok[i] = good ? 1 : 0;
}

__global__ void k_surr_clear(unsigned* c, int k) {
    if ((int)threadIdx.x < k) c[threadIdx.x] = 0u;
}

static inline unsigned g256(int n) { return (unsigned)((n + 255) / 256); }

#define GDTK_SURR_NSP(F)  \
    switch (nsp) {        \
    case 1: F(1)          \
    case 2: F(2)          \
    case 3: F(3)          \
    case 4: F(4)          \
    case 5: F(5)          \
    case 6: F(6)          \
    case 7: F(7)          \
    case 8: F(8)          \
    default: return hipErrorInvalidValue; \
    }

hipError_t launch_surr_clear(unsigned* c, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_surr_clear, dim3(1), dim3(64), 0, s, c, k);
    return hipGetLastError();
}

hipError_t launch_surr_route(int nsp, int n, const double* T, const double* dtc, double dt, const SurrCfg& cfg,
                             int* path, int* taken, int* net_idx, unsigned* net_count, hipStream_t s) {
    if (n <= 0) return hipSuccess;
#define F(N)                                                                                                         \
    hipLaunchKernelGGL((k_surr_route<N>), dim3(g256(n)), dim3(256), 0, s, n, T, dtc, dt, cfg, path, taken, net_idx, \
                       net_count);                                                                                   \
    return hipGetLastError();
    GDTK_SURR_NSP(F)
#undef F
}

hipError_t launch_surr_features(int nsp, int n, const int* net_idx, const unsigned* net_count, const double* rho,
                                const double* T, const double* massf, double dt, const SurrCfg& cfg, float* x,
                                hipStream_t s) {
    if (n <= 0) return hipSuccess;
#define F(N)                                                                                                         \
    hipLaunchKernelGGL((k_surr_features<N>), dim3(g256(n)), dim3(256), 0, s, n, net_idx, net_count, rho, T, massf, \
                       dt, cfg, x);                                                                                  \
    return hipGetLastError();
    GDTK_SURR_NSP(F)
#undef F
}

hipError_t launch_surr_apply(int nsp, int n, const int* net_idx, const unsigned* net_count, const float* x,
                             const float* ya, const float* yb, const double* massf, const SurrCfg& cfg,
                             double* massf_new, int* path, int* taken, unsigned* fb_count, double* disagree,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
#define F(N)                                                                                                          \
    hipLaunchKernelGGL((k_surr_apply<N>), dim3(g256(n)), dim3(256), 0, s, n, net_idx, net_count, x, ya, yb, massf,  \
                       cfg, massf_new, path, taken, fb_count, disagree);                                              \
    return hipGetLastError();
    GDTK_SURR_NSP(F)
#undef F
}

hipError_t launch_surr_finish(int nsp, int n, const int* net_idx, const unsigned* net_count, const double* rho,
                              const double* T, const double* u, const double* p, const double* massf, const double* dtc,
                              const double* massf_new, const ChemDev* cd, double* T_out, double* u_out, double* p_out,
                              double* massf_out, double* dtc_out, int* status, int* path, hipStream_t s) {
    if (n <= 0) return hipSuccess;
#define F(N)                                                                                                          \
    hipLaunchKernelGGL((k_surr_finish<N>), dim3(g256(n)), dim3(256), 0, s, n, net_idx, net_count, rho, T, u, p,     \
                       massf, dtc, massf_new, cd, T_out, u_out, p_out, massf_out, dtc_out, status, path);             \
    return hipGetLastError();
    GDTK_SURR_NSP(F)
#undef F
}

hipError_t launch_surr_project(int nsp, int n, const double* Y0, const double* d0, const SurrCfg& cfg, double* d,
                               int* ok, hipStream_t s) {
    if (n <= 0) return hipSuccess;
#define F(N)                                                                                                   \
    hipLaunchKernelGGL((k_surr_project<N>), dim3(g256(n)), dim3(256), 0, s, n, Y0, d0, cfg, d, ok);           \
    return hipGetLastError();
    GDTK_SURR_NSP(F)
#undef F
}

}  // namespace gdtk
