// tt.hip -- the two-temperature air reactor of the C ABI (include/gdtk_hip.h): the kernel,
// one lane per cell running FVCell.thermochemical_increment (src/eilmer/fvcell.d:858-972)
// around TwoTemperatureAirKinetics.opCall (src/kinetics/two_temperature_air_kinetics.d:
// 162-358), and the handle that builds its tables the way the reference's constructors do
// (two_temperature_air.d:158-220, two_temperature_air_kinetics.d:55-75, 387-439).
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/gdtk_hip.h"
#include "chem_host.hpp"
#include "host_common.hpp"
#include "tt_device.hpp"

namespace gdtk {

constexpr int TT_BLOCK = 64;

// Cells with T <= T_frozen are left as they are.  A reactor failure restores the input state
// and retries with dt_chem = -1; a second failure, or a Tv iteration that does not converge,
// fails the cell (status -1, the input state written back).  Outputs may alias inputs.
template <int NSP, int NRM>
__global__ __launch_bounds__(TT_BLOCK) void k_tt_step(const ChemDev* __restrict__ cdp, const TTDev* __restrict__ tdp,
                                                     int n, double dt, const double* rho, const double* T,
                                                     const double* Tv, const double* u, const double* uv,
                                                     const double* p, const double* massf, const double* dtc,
                                                     double* T_out, double* Tv_out, double* u_out, double* uv_out,
                                                     double* p_out, double* massf_out, double* dtc_out, int* status,
                                                     int* nsub, int* natt) {
    const int i = blockIdx.x * TT_BLOCK + threadIdx.x;
    if (i >= n) return;
    const ChemDev& cd = *cdp;
    const TTDev& td = *tdp;
    TTCell<NSP> Q0;
    Q0.rho = rho[i];
    Q0.p = p[i];
    Q0.T = T[i];
    Q0.Tv = Tv[i];
    Q0.u = u[i];
    Q0.uv = uv[i];
#pragma unroll
    for (int s = 0; s < NSP; ++s) Q0.massf[s] = massf[(size_t)s * n + i];
    double dc = dtc[i];
    TTCell<NSP> Q = Q0;
    TTCounts cnt{0, 0};
    int st = 0;
    if (Q0.T > cd.T_frozen) {
        int r = tt_opcall<NSP, NRM>(cd, td, Q, dt, dc, cnt);
        if (r == 0) {
            Q = Q0;
            dc = -1.0;
            r = tt_opcall<NSP, NRM>(cd, td, Q, dt, dc, cnt);
        }
        if (r != 1 || !tt_update_from_rhou<NSP>(cd, td, Q)) {
            st = -1;
            Q = Q0;
        }
    }
    T_out[i] = Q.T;
    Tv_out[i] = Q.Tv;
    u_out[i] = Q.u;
    uv_out[i] = Q.uv;
    p_out[i] = Q.p;
#pragma unroll
    for (int s = 0; s < NSP; ++s) massf_out[(size_t)s * n + i] = Q.massf[s];
    dtc_out[i] = dc;
    status[i] = st;
    nsub[i] = cnt.substeps;
    natt[i] = cnt.attempts;
}

struct TT {
    ChemDev host{};
    TTDev thost{};
    ChemDev* dev = nullptr;
    TTDev* tdev = nullptr;
    DevArr<int> nsub, natt;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int last_n = 0;
    bool have_events = false;
    ~TT() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (dev) (void)hipFree(dev);
        if (tdev) (void)hipFree(tdev);
    }
};

static std::mutex g_tt_mu;
static std::vector<std::unique_ptr<TT>> g_tts;

static TT& get_tt(int h) {
    std::lock_guard<std::mutex> lk(g_tt_mu);
    if (h < 0 || h >= (int)g_tts.size() || !g_tts[h]) throw Error(GDTK_ERR_ARG, "bad two-temperature handle");
    return *g_tts[h];
}

}  // namespace gdtk

using namespace gdtk;

extern "C" {

int gdtk_hip_tt_create(const gdtk_tpg_gas* gas, const gdtk_mechanism* mech, const gdtk_tt_params* prm, int* tt) {
    return guarded([&] {
        if (!gas || !mech || !prm || !tt) throw Error(GDTK_ERR_ARG, "null argument");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw Error(GDTK_ERR_HIP, "no HIP device");
        auto t = std::make_unique<TT>();
        build_chemdev(*gas, *mech, t->host, true);
        ChemDev& cd = t->host;
        if (cd.nsp != 5)
            throw Error(GDTK_ERR_UNSUPPORTED, "the two-temperature reactor is built for the 5-species neutral air model");
        if (cd.nr > 8) throw Error(GDTK_ERR_UNSUPPORTED, "the two-temperature reactor takes at most 8 reactions");
        if (cd.method != GDTK_ODE_RKF)
            throw Error(GDTK_ERR_UNSUPPORTED, "the two-temperature reactor runs the RKF chemistry step");
        if (!cd.all_sorted)
            throw Error(GDTK_ERR_UNSUPPORTED, "reaction lists must be in ascending species order (as prep-chem writes them)");
        for (int ir = 0; ir < cd.nr; ++ir)
            if (cd.r[ir].f_from_eq || cd.r[ir].b_from_eq)
                throw Error(GDTK_ERR_UNSUPPORTED, "the two-temperature reactor needs explicit forward and backward rates");
        if (!(prm->abs_tol > 0.0) || !(prm->rel_tol >= 0.0)) throw Error(GDTK_ERR_ARG, "bad energy step tolerances");
        TTDev& td = t->thost;
        const double T_REF = TT_T_REF;
        for (int s = 0; s < cd.nsp; ++s) {
            const double R = cd.R[s];
            td.molecular[s] = prm->molecular[s] ? 1 : 0;
            td.del_hf[s] = cea_h(cd.cea[s], 298.15);
            td.Cp_tr_rot[s] = (5. / 2.) * R;
            if (td.molecular[s]) td.Cp_tr_rot[s] += R;
            td.sA[s] = td.Cp_tr_rot[s] * T_REF - td.del_hf[s];
            td.sB[s] = td.Cp_tr_rot[s] - R;
            td.A_MW[s] = td.molecular[s] ? prm->A_MW[s] : 0.0;
            if (td.molecular[s] && !(td.A_MW[s] > 0.0)) throw Error(GDTK_ERR_ARG, "A_MW must be positive for molecules");
            td.mpart[s] = cd.M[s] / TT_AVOGADRO;
            for (int c = 0; c < cd.nsp; ++c) {
                double mu = (cd.M[s] * cd.M[c]) / (cd.M[s] + cd.M[c]);
                mu *= 1000.0;
                td.mu_q[s][c] = std::pow(mu, 0.25);
            }
        }
        td.abs_tol = prm->abs_tol;
        td.rel_tol = prm->rel_tol;
        HIPCHK(hipMalloc(&t->dev, sizeof(ChemDev)));
        HIPCHK(hipMemcpy(t->dev, &t->host, sizeof(ChemDev), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&t->tdev, sizeof(TTDev)));
        HIPCHK(hipMemcpy(t->tdev, &t->thost, sizeof(TTDev), hipMemcpyHostToDevice));
        for (auto& e : t->ev) HIPCHK(hipEventCreate(&e));
        std::lock_guard<std::mutex> lk(g_tt_mu);
        g_tts.push_back(std::move(t));
        *tt = (int)g_tts.size() - 1;
        return GDTK_OK;
    });
}

int gdtk_hip_tt_destroy(int tt) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_tt_mu);
        if (tt < 0 || tt >= (int)g_tts.size() || !g_tts[tt]) throw Error(GDTK_ERR_ARG, "bad two-temperature handle");
        if (g_tts[tt]->have_events) HIPCHK(hipEventSynchronize(g_tts[tt]->ev[1]));
        g_tts[tt].reset();
        return GDTK_OK;
    });
}

int gdtk_hip_tt_step_device(int tt, int n, const double* rho, const double* T, const double* Tv, const double* u,
                            const double* uv, const double* p, const double* massf, const double* dt_chem, double dt,
                            double* T_out, double* Tv_out, double* u_out, double* uv_out, double* p_out,
                            double* massf_out, double* dt_chem_out, int* status, void* stream) {
    return guarded([&] {
        TT& t = get_tt(tt);
        if (n < 0) throw Error(GDTK_ERR_ARG, "negative cell count");
        if (!(dt > 0.0)) throw Error(GDTK_ERR_ARG, "dt must be positive");
        if (n > 0 && (!rho || !T || !Tv || !u || !uv || !p || !massf || !dt_chem || !T_out || !Tv_out || !u_out ||
                      !uv_out || !p_out || !massf_out || !dt_chem_out || !status))
            throw Error(GDTK_ERR_ARG, "null array");
        hipStream_t s = static_cast<hipStream_t>(stream);
        t.nsub.reserve((size_t)std::max(n, 1));
        t.natt.reserve((size_t)std::max(n, 1));
        HIPCHK(hipEventRecord(t.ev[0], s));
        if (n > 0) {
            const dim3 grid((unsigned)((n + TT_BLOCK - 1) / TT_BLOCK));
            if (t.host.nr <= 6)
                hipLaunchKernelGGL((k_tt_step<5, 6>), grid, dim3(TT_BLOCK), 0, s, t.dev, t.tdev, n, dt, rho, T, Tv, u,
                                   uv, p, massf, dt_chem, T_out, Tv_out, u_out, uv_out, p_out, massf_out, dt_chem_out,
                                   status, t.nsub.p, t.natt.p);
            else
                hipLaunchKernelGGL((k_tt_step<5, 8>), grid, dim3(TT_BLOCK), 0, s, t.dev, t.tdev, n, dt, rho, T, Tv, u,
                                   uv, p, massf, dt_chem, T_out, Tv_out, u_out, uv_out, p_out, massf_out, dt_chem_out,
                                   status, t.nsub.p, t.natt.p);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(t.ev[1], s));
        t.last_n = n;
        t.have_events = true;
        return GDTK_OK;
    });
}

int gdtk_hip_tt_stats(int tt, double* stats, int nstats) {
    return guarded([&] {
        TT& t = get_tt(tt);
        if (!t.have_events) throw Error(GDTK_ERR_ARG, "no two-temperature step has run");
        if (!stats || nstats < 0) throw Error(GDTK_ERR_ARG, "bad output");
        HIPCHK(hipEventSynchronize(t.ev[1]));
        std::vector<int> sub((size_t)t.last_n), att((size_t)t.last_n);
        if (t.last_n > 0) {
            HIPCHK(hipMemcpy(sub.data(), t.nsub.p, sub.size() * sizeof(int), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(att.data(), t.natt.p, att.size() * sizeof(int), hipMemcpyDeviceToHost));
        }
        double tot = 0.0, mx = 0.0, ta = 0.0;
        for (int i = 0; i < t.last_n; ++i) {
            tot += sub[i];
            mx = std::max(mx, (double)sub[i]);
            ta += att[i];
        }
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, t.ev[0], t.ev[1]));
        const double v[5] = {tot, mx, ta, (double)ms, (double)t.last_n};
        const int m = std::min(nstats, 5);
        for (int k = 0; k < m; ++k) stats[k] = v[k];
        return m;
    });
}

}  // extern "C"
