// chem_host.hpp -- the chemistry handle of the C ABI (defined in chem.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "chem_device.hpp"
#include "chem_launch.hpp"
#include "host_common.hpp"

namespace gdtk {

// Gas record -> the gas part of a device table (chem.cpp).
void build_tpg_gas(const gdtk_tpg_gas& g, ChemDev& cd);
// Gas and mechanism records -> a device table, with the reference readers' checks; Park
// rate constants are accepted for the two-temperature reactor only (chem.cpp).
void build_chemdev(const gdtk_tpg_gas& g, const gdtk_mechanism& m, ChemDev& cd, bool two_temperature);

struct Chem {
    ChemDev host{};
    ChemDev* dev = nullptr;
    int nsp = 0, nr = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    ChemTuning tuning;
    DevArr<double> kfkb, conc0, h0, hest, conc, Tend, uend;
    DevArr<int> flag, heavy, st, nsub;
    DevArr<unsigned> queue;
    DevArr<unsigned long long> stats;
    // Staging of the host-array entry points.
    DevArr<double> s_rho, s_T, s_u, s_p, s_massf, s_dtc, s_a;
    DevArr<int> s_status;
    // Last call: cells, its work buffers (the handle's own or a caller's), event stream.
    int last_n = 0, last_waves = 0;
    ChemWork last_w{};
    hipStream_t last_stream = nullptr;
    bool have_events = false;

    ~Chem() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (dev) (void)hipFree(dev);
        if (stream) (void)hipStreamDestroy(stream);
    }

    ChemWork work(int n) {
        const size_t N = (size_t)std::max(n, 1);
        kfkb.reserve(2 * (size_t)std::max(nr, 1) * N);
        conc0.reserve((size_t)nsp * N);
        conc.reserve((size_t)nsp * N);
        h0.reserve(N);
        hest.reserve(N);
        flag.reserve(N);
        heavy.reserve(N);
        st.reserve(N);
        nsub.reserve(N);
        queue.reserve(QUEUE_COUNTERS * QUEUE_STRIDE);
        stats.reserve(4 * (size_t)chem_stat_waves(n));
        Tend.reserve(host.tight ? N : 1);
        uend.reserve(host.tight ? N : 1);
        return ChemWork{kfkb.p, conc0.p, h0.p, hest.p, conc.p, flag.p, heavy.p, st.p, nsub.p, queue.p, stats.p, Tend.p,
                        uend.p};
    }

    // Bytes of the work buffers of one call over n cells when they are laid out in one
    // region (carve_work), each buffer starting on a 256-byte boundary.
    size_t work_bytes(int n) const {
        const size_t N = (size_t)std::max(n, 1);
        auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
        return up(4 * (size_t)chem_stat_waves(n) * sizeof(unsigned long long)) +
               up(QUEUE_COUNTERS * QUEUE_STRIDE * sizeof(unsigned)) + 4 * up(N * sizeof(int)) +
               (host.tight ? 4 : 2) * up(N * sizeof(double)) + 2 * up((size_t)nsp * N * sizeof(double)) +
               up(2 * (size_t)std::max(nr, 1) * N * sizeof(double));
    }

    // The work buffers of a call over n cells carved out of one device region of at least
    // work_bytes(n) bytes (256-byte aligned).
    ChemWork carve_work(int n, void* region) const {
        const size_t N = (size_t)std::max(n, 1);
        auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
        char* q = static_cast<char*>(region);
        auto take = [&](size_t b) {
            char* r = q;
            q += up(b);
            return r;
        };
        ChemWork w;
        w.stats = reinterpret_cast<unsigned long long*>(take(4 * (size_t)chem_stat_waves(n) * sizeof(unsigned long long)));
        w.queue = reinterpret_cast<unsigned*>(take(QUEUE_COUNTERS * QUEUE_STRIDE * sizeof(unsigned)));
        w.flag = reinterpret_cast<int*>(take(N * sizeof(int)));
        w.heavy = reinterpret_cast<int*>(take(N * sizeof(int)));
        w.st = reinterpret_cast<int*>(take(N * sizeof(int)));
        w.nsub = reinterpret_cast<int*>(take(N * sizeof(int)));
        w.h0 = reinterpret_cast<double*>(take(N * sizeof(double)));
        w.hest = reinterpret_cast<double*>(take(N * sizeof(double)));
        w.conc0 = reinterpret_cast<double*>(take((size_t)nsp * N * sizeof(double)));
        w.conc = reinterpret_cast<double*>(take((size_t)nsp * N * sizeof(double)));
        w.kfkb = reinterpret_cast<double*>(take(2 * (size_t)std::max(nr, 1) * N * sizeof(double)));
        w.Tend = host.tight ? reinterpret_cast<double*>(take(N * sizeof(double))) : nullptr;
        w.uend = host.tight ? reinterpret_cast<double*>(take(N * sizeof(double))) : nullptr;
        return w;
    }

    // The three kernels on device arrays; outputs may alias inputs.  The work buffers are the
    // handle's own (grown as needed) unless the caller brings them (wk).  Cells with
    // mask[i] != 0 (mask may be null) are neither read nor written: another path serves them.
    void step_device(int n, double dt, const double* rho, const double* T, const double* u, const double* p,
                     const double* massf, const double* dtc, double* T_out, double* u_out, double* p_out,
                     double* massf_out, double* dtc_out, int* status, hipStream_t s, const ChemWork* wk = nullptr,
                     const int* mask = nullptr) {
        if (n < 0) throw Error(GDTK_ERR_ARG, "negative cell count");
        if (!(dt > 0.0)) throw Error(GDTK_ERR_ARG, "dt must be positive");
        ChemWork w = wk ? *wk : work(n);
        // Below two cells per lane of a full chip (4 waves per SIMD) the one-lane-per-cell
        // launch starts every cell at once; the queue pays off on larger batches.
        ChemTuning tu = tuning;
        if (tu.persistent < 0) tu.persistent = ((long long)n > 2048LL * std::max(num_cu, 1)) ? 1 : 0;
        HIPCHK(hipEventRecord(ev[0], s));
        HIPCHK(launch_chem_clear(w, s));
        HIPCHK(launch_chem_setup(dev, nsp, nr, host.all_sorted, n, dt, rho, T, massf, dtc, mask,
                                 tu.persistent ? tu.heavy_ratio : 0, w, s));
        HIPCHK(hipEventRecord(ev[1], s));
        int waves = 0;
        HIPCHK(launch_chem_integrate(dev, nsp, nr, host.all_sorted, host.method, host.tight, n, dt, rho, T, u, dtc,
                                     dtc_out, w, tu, num_cu, &waves, s));
        HIPCHK(hipEventRecord(ev[2], s));
        HIPCHK(launch_chem_finish(dev, nsp, n, rho, T, u, p, massf, dtc, T_out, u_out, p_out, massf_out, dtc_out,
                                  status, w, s));
        HIPCHK(hipEventRecord(ev[3], s));
        last_n = n;
        last_waves = waves;
        last_w = w;
        last_stream = s;
        have_events = true;
    }

    // stats: {total substeps, total corrector iterations, max substeps, step attempts,
    //         ms setup, ms integrate, ms finish, cells}
    int read_stats(double* out, int nout) {
        if (!have_events) throw Error(GDTK_ERR_ARG, "no chemistry step has run");
        HIPCHK(hipEventSynchronize(ev[3]));
        unsigned long long h[4] = {0, 0, 0, 0};
        if (last_waves > 0) {
            std::vector<unsigned long long> part(4 * (size_t)last_waves);
            HIPCHK(hipMemcpy(part.data(), last_w.stats, part.size() * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
            for (int k = 0; k < last_waves; ++k) {
                h[0] += part[4 * k + 0];
                h[1] += part[4 * k + 1];
                h[2] = std::max(h[2], part[4 * k + 2]);
                h[3] += part[4 * k + 3];
            }
        }
        float ms[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) HIPCHK(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
        const double v[8] = {(double)h[0], (double)h[1], (double)h[2], (double)h[3],
                             ms[0], ms[1], ms[2], (double)last_n};
        int m = std::min(nout, 8);
        for (int k = 0; k < m; ++k) out[k] = v[k];
        return m;
    }
};

Chem& get_chem(int h);

}  // namespace gdtk
