// chem_kernels.hip -- entry points of the chemistry kernels: the queue/statistics clear
// kernel and the dispatch of each phase to the instantiation for the mechanism's species
// count (chem_nsp<N>.hip, compiled as separate translation units).
#include <hip/hip_runtime.h>

#include "chem_launch.hpp"

namespace gdtk {

__global__ void k_chem_clear(unsigned* queue) {
    int t = threadIdx.x;
    if (t < QUEUE_COUNTERS) queue[t * QUEUE_STRIDE] = 0u;
}

bool chem_supported(int nsp, int nr) { return nsp >= 1 && nsp <= CH_MAX_SP && nr >= 0 && nr <= CH_MAX_R; }

hipError_t launch_chem_clear(const ChemWork& w, hipStream_t s) {
    hipLaunchKernelGGL(k_chem_clear, dim3(1), dim3(64), 0, s, w.queue);
    return hipGetLastError();
}

static hipError_t run_nsp(int nsp, const ChemCall& c) {
    if (c.n <= 0) return hipSuccess;
    switch (nsp) {
    case 1: return chem_run<1>(c);
    case 2: return chem_run<2>(c);
    case 3: return chem_run<3>(c);
    case 4: return chem_run<4>(c);
    case 5: return chem_run<5>(c);
    case 6: return chem_run<6>(c);
    case 7: return chem_run<7>(c);
    case 8: return chem_run<8>(c);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_chem_setup(const ChemDev* cd, int nsp, int nr, int dense, int n, double dt, const double* rho,
                             const double* T, const double* massf, const double* dtc, const int* mask,
                             int heavy_ratio, const ChemWork& w, hipStream_t s) {
    ChemCall c{};
    c.phase = CHEM_PHASE_SETUP;
    c.cd = cd; c.nr = nr; c.dense = dense; c.n = n; c.dt = dt;
    c.rho = rho; c.T = T; c.massf = massf; c.dtc_in = dtc; c.mask = mask; c.tu.heavy_ratio = heavy_ratio;
    c.w = w; c.s = s;
    return run_nsp(nsp, c);
}

hipError_t launch_chem_integrate(const ChemDev* cd, int nsp, int nr, int dense, int method, int tight, int n, double dt,
                                 const double* rho, const double* T, const double* u, const double* dtc_in,
                                 double* dtc_out, const ChemWork& w, const ChemTuning& tu, int num_cu, int* waves,
                                 hipStream_t s) {
    ChemCall c{};
    c.phase = CHEM_PHASE_INTEGRATE;
    c.cd = cd; c.nr = nr; c.dense = dense; c.method = method; c.tight = tight; c.n = n; c.dt = dt;
    c.rho = rho; c.T = T; c.u = u; c.dtc_in = dtc_in; c.dtc_out = dtc_out;
    c.w = w; c.tu = tu; c.num_cu = num_cu; c.waves_out = waves; c.s = s;
    return run_nsp(nsp, c);
}

hipError_t launch_chem_finish(const ChemDev* cd, int nsp, int n, const double* rho, const double* T,
                              const double* u, const double* p, const double* massf, const double* dtc_in,
                              double* T_out, double* u_out, double* p_out, double* massf_out, double* dtc_out,
                              int* status, const ChemWork& w, hipStream_t s) {
    ChemCall c{};
    c.phase = CHEM_PHASE_FINISH;
    c.cd = cd; c.n = n;
    c.rho = rho; c.T = T; c.u = u; c.p = p; c.massf = massf; c.dtc_in = dtc_in;
    c.T_out = T_out; c.u_out = u_out; c.p_out = p_out; c.massf_out = massf_out; c.dtc_out = dtc_out;
    c.status = status; c.w = w; c.s = s;
    return run_nsp(nsp, c);
}

hipError_t launch_tpg_update(const ChemDev* cd, int nsp, int mode, int n, double* rho, double* p, double* T,
                             double* u, double* a, const double* massf, int* status, hipStream_t s) {
    ChemCall c{};
    c.phase = CHEM_PHASE_TPG;
    c.cd = cd; c.mode = mode; c.n = n;
    c.rho_io = rho; c.p_io = p; c.T_io = T; c.u_io = u; c.a_io = a; c.massf = massf;
    c.status = status; c.s = s;
    return run_nsp(nsp, c);
}

}  // namespace gdtk
