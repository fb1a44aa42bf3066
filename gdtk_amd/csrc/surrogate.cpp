// surrogate.cpp -- host side of the learned chemistry step (C ABI in include/gdtk_hip.h).
//
// gdtk_hip_surrogate_step_device is a drop-in for gdtk_hip_chem_step_device (the same
// per-cell outputs) that answers most cells with two trained networks and an exact
// element-conservation projection, and hands the rest (outside the trained range, networks
// in disagreement, infeasible projection, failed T(u) iteration) to the alpha-QSS
// integrator of the chemistry handle.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/gdtk_hip.h"
#include "chem_host.hpp"
#include "host_common.hpp"
#include "mlp_host.hpp"
#include "surrogate_launch.hpp"

namespace gdtk {

static_assert(GDTK_SURR_MAX_FEAT == SURR_MAX_FEAT && GDTK_SURR_MAX_ELEM == SURR_MAX_ELEM, "record sizes");

struct Surrogate {
    int chem = -1, net_a = -1, net_b = -1;
    int nsp = 0;
    SurrCfg cfg{};
    DevArr<float> x, ya, yb;
    DevArr<double> massf_new, disagree;
    DevArr<int> path, taken, net_idx;
    DevArr<unsigned> counts;  // {network cells, cells sent back to the integrator}
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int last_n = 0;
    bool have = false;

    ~Surrogate() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    // Everything stays on the device: the network cells are compacted by an atomic counter
    // that the kernels read, and the integrator runs over all cells with the mask of the
    // cells the networks answered.  Nothing is read back to the host.
    void step(int n, double dt, const double* rho, const double* T, const double* u, const double* p,
              const double* massf, const double* dtc, double* T_out, double* u_out, double* p_out, double* massf_out,
              double* dtc_out, int* status, int* path_out, hipStream_t s) {
        if (n < 0) throw Error(GDTK_ERR_ARG, "negative cell count");
        if (!(dt > 0.0)) throw Error(GDTK_ERR_ARG, "dt must be positive");
        Chem& ch = get_chem(chem);
        Mlp& ma = get_mlp(net_a);
        Mlp& mb = get_mlp(net_b);
        const size_t N = (size_t)std::max(n, 1);
        x.reserve((size_t)(nsp + 3) * N);
        ya.reserve((size_t)nsp * N);
        yb.reserve((size_t)nsp * N);
        massf_new.reserve((size_t)nsp * N);
        disagree.reserve(N);
        path.reserve(N);
        taken.reserve(N);
        net_idx.reserve(N);
        counts.reserve(2);
        int* pth = path_out ? path_out : path.p;
        HIPCHK(hipEventRecord(ev[0], s));
        HIPCHK(launch_surr_clear(counts.p, 2, s));
        HIPCHK(launch_surr_route(nsp, n, T, dtc, dt, cfg, pth, taken.p, net_idx.p, counts.p, s));
        HIPCHK(launch_surr_features(nsp, n, net_idx.p, counts.p, rho, T, massf, dt, cfg, x.p, s));
        mlp_infer(ma, n, x.p, ya.p, s, counts.p);
        mlp_infer(mb, n, x.p, yb.p, s, counts.p);
        HIPCHK(hipEventRecord(ev[1], s));
        HIPCHK(launch_surr_apply(nsp, n, net_idx.p, counts.p, x.p, ya.p, yb.p, massf, cfg, massf_new.p, pth, taken.p,
                                 counts.p + 1, disagree.p, s));
        HIPCHK(launch_surr_finish(nsp, n, net_idx.p, counts.p, rho, T, u, p, massf, dtc, massf_new.p, ch.dev, T_out,
                                  u_out, p_out, massf_out, dtc_out, status, pth, s));
        HIPCHK(hipEventRecord(ev[2], s));
        ch.step_device(n, dt, rho, T, u, p, massf, dtc, T_out, u_out, p_out, massf_out, dtc_out, status, s, nullptr,
                       taken.p);
        HIPCHK(hipEventRecord(ev[3], s));
        last_n = n;
        have = true;
    }

    // {cells, network cells sent back to the integrator, ms routing + features + networks,
    //  ms projection + T(u), ms integrator, max network disagreement, network cells} of the
    //  last step (waits for it).
    int stats(double* out, int nout) {
        if (!have) throw Error(GDTK_ERR_ARG, "no surrogate step has run");
        HIPCHK(hipEventSynchronize(ev[3]));
        float ms[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) HIPCHK(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
        unsigned c[2] = {0, 0};
        HIPCHK(hipMemcpy(c, counts.p, sizeof(c), hipMemcpyDeviceToHost));
        double dmax = 0.0;
        if (c[0]) {
            std::vector<double> d((size_t)c[0]);
            HIPCHK(hipMemcpy(d.data(), disagree.p, d.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (double v : d)
                if (v > dmax) dmax = v;
        }
        const double v[7] = {(double)last_n, (double)c[1], ms[0], ms[1], ms[2], dmax, (double)c[0]};
        const int m = std::min(nout, 7);
        for (int k = 0; k < m; ++k) out[k] = v[k];
        return m;
    }
};

static std::mutex g_surr_mu;
static std::vector<std::unique_ptr<Surrogate>> g_surrs;

static Surrogate& get_surr(int h) {
    std::lock_guard<std::mutex> lk(g_surr_mu);
    if (h < 0 || h >= (int)g_surrs.size() || !g_surrs[h]) throw Error(GDTK_ERR_ARG, "bad surrogate handle");
    return *g_surrs[h];
}

}  // namespace gdtk

using namespace gdtk;

extern "C" {

int gdtk_hip_surrogate_create(int chem, int net_a, int net_b, const gdtk_surrogate_cfg* cfg, int* h) {
    return guarded([&] {
        if (!cfg || !h) throw Error(GDTK_ERR_ARG, "null argument");
        Chem& ch = get_chem(chem);
        Mlp& ma = get_mlp(net_a);
        Mlp& mb = get_mlp(net_b);
        const int nsp = ch.nsp;
        if (ma.n_in != nsp + 3 || mb.n_in != nsp + 3 || ma.n_out != nsp || mb.n_out != nsp)
            throw Error(GDTK_ERR_ARG, "networks must map nsp + 3 features to nsp outputs");
        if (cfg->n_elem < 1 || cfg->n_elem > SURR_MAX_ELEM) throw Error(GDTK_ERR_ARG, "n_elem must be 1..4");
        if (!(cfg->x_abs_max > 0.0) || !(cfg->disagree_tol >= 0.0)) throw Error(GDTK_ERR_ARG, "bad thresholds");
        auto s = std::make_unique<Surrogate>();
        s->chem = chem;
        s->net_a = net_a;
        s->net_b = net_b;
        s->nsp = nsp;
        for (int k = 0; k < SURR_MAX_FEAT; ++k) {
            s->cfg.x_mean[k] = cfg->x_mean[k];
            s->cfg.x_inv_std[k] = cfg->x_inv_std[k];
        }
        for (int k = 0; k < CH_MAX_SP; ++k) s->cfg.y_scale[k] = cfg->y_scale[k];
        s->cfg.x_abs_max = cfg->x_abs_max;
        s->cfg.disagree_tol = cfg->disagree_tol;
        s->cfg.T_frozen = ch.host.T_frozen;
        s->cfg.n_elem = cfg->n_elem;
        for (int e = 0; e < SURR_MAX_ELEM; ++e)
            for (int k = 0; k < CH_MAX_SP; ++k) s->cfg.elem[e][k] = cfg->elem[e][k];
        s->cfg.route_ratio = cfg->route_ratio;
        s->cfg.route_T = cfg->route_T;
        for (auto& e : s->ev) HIPCHK(hipEventCreate(&e));
        std::lock_guard<std::mutex> lk(g_surr_mu);
        g_surrs.push_back(std::move(s));
        *h = (int)g_surrs.size() - 1;
        return GDTK_OK;
    });
}

int gdtk_hip_surrogate_destroy(int h) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_surr_mu);
        if (h < 0 || h >= (int)g_surrs.size() || !g_surrs[h]) throw Error(GDTK_ERR_ARG, "bad surrogate handle");
        HIPCHK(hipDeviceSynchronize());
        g_surrs[h].reset();
        return GDTK_OK;
    });
}

int gdtk_hip_surrogate_step_device(int h, int n, const double* rho, const double* T, const double* u, const double* p,
                                   const double* massf, const double* dt_chem, double dt, double* T_out, double* u_out,
                                   double* p_out, double* massf_out, double* dt_chem_out, int* status, int* path,
                                   void* stream) {
    return guarded([&] {
        Surrogate& s = get_surr(h);
        if (n > 0 && (!rho || !T || !u || !p || !massf || !dt_chem || !T_out || !u_out || !p_out || !massf_out ||
                      !dt_chem_out || !status))
            throw Error(GDTK_ERR_ARG, "null array");
        s.step(n, dt, rho, T, u, p, massf, dt_chem, T_out, u_out, p_out, massf_out, dt_chem_out, status, path,
               static_cast<hipStream_t>(stream));
        return GDTK_OK;
    });
}

int gdtk_hip_surrogate_stats(int h, double* stats, int nstats) {
    return guarded([&] {
        if (!stats || nstats < 1) throw Error(GDTK_ERR_ARG, "null stats");
        return get_surr(h).stats(stats, nstats);
    });
}

int gdtk_hip_surrogate_project(int nsp, int n_elem, const double* elem, int n, const double* Y0, const double* d0,
                               double* d, int* ok, void* stream) {
    return guarded([&] {
        if (nsp < 1 || nsp > CH_MAX_SP) throw Error(GDTK_ERR_ARG, "nsp must be 1..8");
        if (n_elem < 1 || n_elem > SURR_MAX_ELEM) throw Error(GDTK_ERR_ARG, "n_elem must be 1..4");
        if (n < 0) throw Error(GDTK_ERR_ARG, "negative cell count");
        if (!elem || (n > 0 && (!Y0 || !d0 || !d || !ok))) throw Error(GDTK_ERR_ARG, "null array");
        SurrCfg cfg{};
        cfg.n_elem = n_elem;
        for (int e = 0; e < n_elem; ++e)
            for (int k = 0; k < nsp; ++k) cfg.elem[e][k] = elem[e * nsp + k];
        const hipStream_t s = static_cast<hipStream_t>(stream);
        HIPCHK(launch_surr_project(nsp, n, Y0, d0, cfg, d, ok, s));
        return GDTK_OK;
    });
}

}  // extern "C"
