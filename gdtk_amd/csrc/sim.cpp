// sim.cpp -- host side of the MI355X explicit update: block table, time-step control
// and the C ABI declared in include/gdtk_hip.h.
//
// The host logic restates the reference's control flow around the GPU kernels:
//   determine_time_step_size      src/eilmer/simcore_gasdynamic_step.d:60-159
//   FluidBlock dt estimate         src/eilmer/fluidblock.d:987-1084 (reduction on device)
//   explicit pc update + retries   src/eilmer/simcore_gasdynamic_step.d:906-1575
//   time loop                      src/eilmer/simcore.d:1014-1377
//   initial encode/decode          src/eilmer/simcore.d:326-330
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gdtk_hip.h"
#include "block.hpp"
#include "gdtk_device.hpp"
#include "geometry.hpp"
#include "host_common.hpp"
#include "launchers.hpp"

namespace gdtk {

thread_local std::string g_last_error;

struct DevBuf {
    double* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        n = count;
        HIPCHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double)));
        HIPCHK(hipMemset(p, 0, std::max<size_t>(count, 1) * sizeof(double)));
        HIPCHK(hipDeviceSynchronize());  // the fill is on the null stream
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct Block {
    int id;
    gdtk_block_desc desc;
    int dims, nic, njc, nkc, ncq;
    BlockDev dev{};
    DevBuf fs[2], U[2], dUdt0, len, vol, fgeom[3], flux[3], dtpart, dtout;
    int* counters = nullptr;
    int cur = 0;   // which fs/U copy is current (input of the next stage)
    ~Block() {
        if (counters) (void)hipFree(counters);
    }
    int64_t ncell() const { return (int64_t)nic * njc * nkc; }
};

struct Sim {
    gdtk_sim_config cfg;
    IdealGasDev gas;
    CellLimits lim;
    std::vector<std::unique_ptr<Block>> blocks;
    hipStream_t stream = nullptr;
    // SimState (simcore.d)
    double time = 0.0, dt_global = 0.0, dt_allow = 0.0, cfl_max = 0.0;
    int step = 0, attempts_last = 0;
    // Profiling: HIP events around each kernel launch of the stage, per kernel kind
    // (0 face flux, 1 cell update, 2 fused stage, 3 ghost fill), on the library's stream.
    bool profile = false;
    static constexpr int NKIND = 4;
    double prof_ms[NKIND] = {0, 0, 0, 0};
    long long prof_launches[NKIND] = {0, 0, 0, 0};
    double prof_items[NKIND] = {0, 0, 0, 0};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Stage variant chosen by the autotuner: -1 not tuned (use cfg.use_fused), 0 face-per-
    // thread, 1 fused.
    int variant = -1;

    template <class L>
    void timed(int kind, double items, L&& launch) {
        if (!profile) {
            HIPCHK(launch());
            return;
        }
        HIPCHK(hipEventRecord(ev0, stream));
        HIPCHK(launch());
        HIPCHK(hipEventRecord(ev1, stream));
        HIPCHK(hipEventSynchronize(ev1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        prof_ms[kind] += ms;
        prof_launches[kind] += 1;
        prof_items[kind] += items;
    }

    ~Sim() {
        blocks.clear();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int ncq() const { return cfg.dimensions == 3 ? 5 : 4; }

    void refresh_dev(Block& b) {
        BlockDev& d = b.dev;
        d.fs_in = b.fs[b.cur].p;
        d.fs_out = b.fs[1 - b.cur].p;
        d.U0 = b.U[0].p;
        d.Unew = b.U[1].p;
        d.dUdt0 = b.dUdt0.p;
    }

    void add_block(const gdtk_block_desc& desc, int* id) {
        if (desc.nic < 2 || desc.njc < 2 || (cfg.dimensions == 3 && desc.nkc < 2) ||
            (cfg.dimensions == 2 && desc.nkc != 1))
            throw Error(GDTK_ERR_ARG, "block extents must be >= 2 cells per active direction (nkc == 1 in 2-D)");
        if (!desc.vertices) throw Error(GDTK_ERR_ARG, "null vertices");
        for (int s = 0; s < 6; ++s)
            if (desc.bc_kind[s] < 0 || desc.bc_kind[s] > GDTK_BC_EXCHANGE) throw Error(GDTK_ERR_ARG, "bad bc kind");
        auto b = std::make_unique<Block>();
        b->id = (int)blocks.size();
        b->desc = desc;
        b->dims = cfg.dimensions;
        b->nic = desc.nic; b->njc = desc.njc; b->nkc = desc.nkc;
        b->ncq = ncq();
        BlockGeometry geo;
        try {
            compute_block_geometry(cfg.dimensions, desc.nic, desc.njc, desc.nkc, desc.vertices, geo);
        } catch (std::exception& e) {
            throw Error(GDTK_ERR_ARG, e.what());
        }
        BlockDev& d = b->dev;
        d.dims = cfg.dimensions;
        d.nic = desc.nic; d.njc = desc.njc; d.nkc = desc.nkc;
        d.NI = desc.nic + 2 * NGHOST;
        d.NJ = desc.njc + 2 * NGHOST;
        d.NK = (cfg.dimensions == 3) ? desc.nkc + 2 * NGHOST : 1;
        d.npad = (int64_t)d.NI * d.NJ * d.NK;
        d.ncq = b->ncq;
        const size_t np = (size_t)d.npad;
        for (int q = 0; q < 2; ++q) b->fs[q].alloc(np * FS_NBASE);
        for (int q = 0; q < 2; ++q) b->U[q].alloc(np * b->ncq);
        b->dUdt0.alloc(np * b->ncq);
        // Padded lengths and volume, ghosts mirrored (sfluidblock.d:884-1000).
        std::vector<double> hlen(3 * np, 0.0), hvol(np, 0.0);
        const int nic = desc.nic, njc = desc.njc, nkc = desc.nkc;
        auto cidx = [&](int i, int j, int k) { return ((size_t)k * njc + j) * nic + i; };
        const int kg = (cfg.dimensions == 3) ? NGHOST : 0;
        auto pad = [&](int i, int j, int k) {
            return (size_t)(i + NGHOST) + (size_t)d.NI * ((size_t)(j + NGHOST) + (size_t)d.NJ * (size_t)(k + kg));
        };
        auto setlen = [&](size_t pc, size_t c) {
            hlen[pc] = geo.iLen[c];
            hlen[pc + np] = geo.jLen[c];
            hlen[pc + 2 * np] = geo.kLen[c];
        };
        for (int k = 0; k < nkc; ++k)
            for (int j = 0; j < njc; ++j)
                for (int i = 0; i < nic; ++i) {
                    setlen(pad(i, j, k), cidx(i, j, k));
                    hvol[pad(i, j, k)] = geo.volume[cidx(i, j, k)];
                }
        for (int k = 0; k < nkc; ++k)
            for (int j = 0; j < njc; ++j)
                for (int n = 0; n < NGHOST; ++n) {
                    setlen(pad(-1 - n, j, k), cidx(n, j, k));
                    setlen(pad(nic + n, j, k), cidx(nic - 1 - n, j, k));
                }
        for (int k = 0; k < nkc; ++k)
            for (int i = 0; i < nic; ++i)
                for (int n = 0; n < NGHOST; ++n) {
                    setlen(pad(i, -1 - n, k), cidx(i, n, k));
                    setlen(pad(i, njc + n, k), cidx(i, njc - 1 - n, k));
                }
        if (cfg.dimensions == 3)
            for (int j = 0; j < njc; ++j)
                for (int i = 0; i < nic; ++i)
                    for (int n = 0; n < NGHOST; ++n) {
                        setlen(pad(i, j, -1 - n), cidx(i, j, n));
                        setlen(pad(i, j, nkc + n), cidx(i, j, nkc - 1 - n));
                    }
        b->len.alloc(3 * np);
        b->vol.alloc(np);
        HIPCHK(hipMemcpy(b->len.p, hlen.data(), hlen.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b->vol.p, hvol.data(), hvol.size() * sizeof(double), hipMemcpyHostToDevice));
        int nd = (cfg.dimensions == 3) ? 3 : 2;
        for (int q = 0; q < 3; ++q) {
            for (int e = 0; e < 3; ++e) d.nf[q][e] = geo.nf[q][e];
            d.nface[q] = (q < nd) ? (int64_t)geo.nf[q][0] * geo.nf[q][1] * geo.nf[q][2] : 0;
            if (q < nd) {
                b->fgeom[q].alloc(geo.face[q].size());
                HIPCHK(hipMemcpy(b->fgeom[q].p, geo.face[q].data(), geo.face[q].size() * sizeof(double),
                                 hipMemcpyHostToDevice));
                b->flux[q].alloc((size_t)d.nface[q] * b->ncq);
            }
            d.fgeom[q] = b->fgeom[q].p;
            d.flux[q] = b->flux[q].p;
        }
        d.len = b->len.p;
        d.vol = b->vol.p;
        for (int s = 0; s < 6; ++s) {
            d.bc_kind[s] = desc.bc_kind[s];
            if (desc.bc_kind[s] == GDTK_BC_INFLOW_SUPERSONIC) {
                // FlowState(p, T, vel): update_thermo_from_pT + update_sound_speed
                // (ideal_gas.d:89-97, 136-143).
                double p = desc.inflow[s][0], T = desc.inflow[s][1];
                if (!(T > 0.0) || !(p > 0.0)) throw Error(GDTK_ERR_ARG, "inflow state needs positive p and T");
                double rho = p / (T * gas.R);
                double u = gas.Cv * T;
                double a = std::sqrt(gas.gamma * gas.R * T);
                double row[8] = {rho, desc.inflow[s][2], desc.inflow[s][3], desc.inflow[s][4], u, p, T, a};
                for (int f = 0; f < 8; ++f) d.inflow[s][f] = row[f];
            }
            if (desc.bc_kind[s] == GDTK_BC_EXCHANGE) throw Error(GDTK_ERR_UNSUPPORTED, "exchange BC not built yet");
        }
        HIPCHK(hipMalloc(&b->counters, 4 * sizeof(int)));
        HIPCHK(hipMemset(b->counters, 0, 4 * sizeof(int)));
        d.counters = b->counters;
        b->dtpart.alloc(2 * (size_t)dt_partial_count(d) + 2);
        b->dtout.alloc(4);
        refresh_dev(*b);
        *id = b->id;
        blocks.push_back(std::move(b));
    }

    Block& blk(int id) {
        if (id < 0 || id >= (int)blocks.size()) throw Error(GDTK_ERR_ARG, "bad block id");
        return *blocks[id];
    }

    void set_initial(int bid, const double* p, const double* T, const double* vx, const double* vy, const double* vz) {
        Block& b = blk(bid);
        BlockDev& d = b.dev;
        const size_t np = (size_t)d.npad;
        std::vector<double> h(np * FS_NBASE, 0.0);
        const int kg = (cfg.dimensions == 3) ? NGHOST : 0;
        for (int k = 0; k < b.nkc; ++k)
            for (int j = 0; j < b.njc; ++j)
                for (int i = 0; i < b.nic; ++i) {
                    size_t c = ((size_t)k * b.njc + j) * b.nic + i;
                    size_t pc = (size_t)(i + NGHOST) + (size_t)d.NI * ((size_t)(j + NGHOST) + (size_t)d.NJ * (size_t)(k + kg));
                    double pp = p[c], TT = T[c];
                    if (!(TT > 0.0) || !(pp > 0.0)) throw Error(GDTK_ERR_ARG, "initial state needs positive p and T");
                    // FlowState(p, T, vel) (ideal_gas.d:89-97, 136-143)
                    double rho = pp / (TT * gas.R);
                    double u = gas.Cv * TT;
                    double a = std::sqrt(gas.gamma * gas.R * TT);
                    double row[8] = {rho, vx[c], vy[c], vz[c], u, pp, TT, a};
                    for (int f = 0; f < 8; ++f) h[pc + f * np] = row[f];
                }
        b.cur = 0;
        refresh_dev(b);
        HIPCHK(hipMemcpy(b.fs[0].p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemsetAsync(b.counters, 0, 4 * sizeof(int), stream));
        HIPCHK(launch_encode(b.dev, b.U[0].p, stream));
        HIPCHK(launch_decode(b.dev, gas, b.U[0].p, b.fs[0].p, lim.low_T, stream));
        int cnt[4];
        HIPCHK(hipMemcpyAsync(cnt, b.counters, 4 * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (cnt[0] > 0) throw Error(GDTK_ERR_BAD_CELL, "bad cell while initializing flow");
    }

    void get_field(int bid, int field, double* out) {
        Block& b = blk(bid);
        BlockDev& d = b.dev;
        const size_t np = (size_t)d.npad;
        const double* src = nullptr;
        if (field >= GDTK_FIELD_DUDT0 && field < GDTK_FIELD_DUDT0 + b.ncq) src = b.dUdt0.p + (field - GDTK_FIELD_DUDT0) * np;
        else if (field >= GDTK_FIELD_U0 && field < GDTK_FIELD_U0 + b.ncq) src = b.U[0].p + (field - GDTK_FIELD_U0) * np;
        else if (field >= 0 && field < FS_NBASE) src = b.fs[b.cur].p + field * np;
        else if (field == GDTK_FIELD_VOLUME) src = b.vol.p;
        else if (field >= GDTK_FIELD_ILENGTH && field <= GDTK_FIELD_KLENGTH) src = b.len.p + (field - GDTK_FIELD_ILENGTH) * np;
        else throw Error(GDTK_ERR_ARG, "bad field");
        std::vector<double> h(np);
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipMemcpy(h.data(), src, np * sizeof(double), hipMemcpyDeviceToHost));
        const int kg = (cfg.dimensions == 3) ? NGHOST : 0;
        for (int k = 0; k < b.nkc; ++k)
            for (int j = 0; j < b.njc; ++j)
                for (int i = 0; i < b.nic; ++i) {
                    size_t c = ((size_t)k * b.njc + j) * b.nic + i;
                    size_t pc = (size_t)(i + NGHOST) + (size_t)d.NI * ((size_t)(j + NGHOST) + (size_t)d.NJ * (size_t)(k + kg));
                    out[c] = h[pc];
                }
    }

    bool use_fused() const { return variant >= 0 ? variant == 1 : cfg.use_fused != 0; }

    // Launch one stage on every block; returns without synchronising.
    void launch_stage(int stage, double dt) {
        const double eps = 1.0e-12;  // epsilon_van_albada, globalconfig.d:1079
        for (auto& bp : blocks) {
            Block& b = *bp;
            refresh_dev(b);
            const double nc = (double)b.ncell();
            timed(3, nc, [&] { return launch_ghost_fill(b.dev, b.fs[b.cur].p, stream); });
            if (use_fused()) {
                timed(2, nc, [&] { return launch_fused_stage(b.dev, gas, stage, cfg.flux_calculator, dt, eps, lim, stream); });
            } else {
                const int nd = (b.dims == 3) ? 3 : 2;
                for (int d = 0; d < nd; ++d)
                    timed(0, (double)b.dev.nface[d],
                          [&] { return launch_face_flux_dir(b.dev, gas, cfg.flux_calculator, eps, d, stream); });
                timed(1, nc, [&] { return launch_cell_update(b.dev, gas, stage, dt, lim, stream); });
            }
        }
    }

    // Autotuner: times both stages of every candidate launch shape on the current flow
    // state (stage 1 reads fs and U0 and writes the other fs copy and dUdt0; stage 2 then
    // reads the same fs and writes fs_out and U[1]: the state the next step starts from,
    // fs[cur] and U[0], is untouched) and keeps the fastest.  Candidates: face-per-thread
    // kernels, and the fused kernel with k chunks of 4..64 planes.
    std::vector<double> autotune(double dt, int reps) {
        std::vector<double> result;
        const int kcs[] = {4, 8, 16, 32, 64};
        bool saved_profile = profile;
        profile = false;
        // Stage 2 on the same input writes only the scratch copies (fs_out, U[1]), so
        // a stage-1 + stage-2 pair can be timed without advancing the solution.
        auto time_variant = [&](int var, int kc) {
            variant = var;
            if (kc > 0) set_fused_kchunk(kc);
            launch_stage(1, dt);  // warm-up
            HIPCHK(hipEventRecord(ev0, stream));
            for (int r = 0; r < reps; ++r) {
                launch_stage(1, dt);
                launch_stage(2, dt);
            }
            HIPCHK(hipEventRecord(ev1, stream));
            HIPCHK(hipEventSynchronize(ev1));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            return (double)ms / (2 * reps);
        };
        double best = time_variant(0, 0);
        int best_var = 0, best_kc = 0;
        result.push_back(0); result.push_back(0); result.push_back(best);
        if (cfg.dimensions == 3) {
            for (int kc : kcs) {
                double t = time_variant(1, kc);
                result.push_back(1); result.push_back(kc); result.push_back(t);
                if (t < best) { best = t; best_var = 1; best_kc = kc; }
            }
        } else {
            double t = time_variant(1, 1);
            result.push_back(1); result.push_back(1); result.push_back(t);
            if (t < best) { best = t; best_var = 1; best_kc = 1; }
        }
        variant = best_var;
        if (best_var == 1) set_fused_kchunk(best_kc);
        HIPCHK(hipStreamSynchronize(stream));
        profile = saved_profile;
        result.insert(result.begin(), {(double)best_var, (double)best_kc, best});
        return result;
    }

    // Runs one stage on all blocks and reports (decode failures, invalid cells).
    void stage_checked(int stage, double dt, int& n_fail, int& n_invalid) {
        for (auto& bp : blocks) HIPCHK(hipMemsetAsync(bp->counters, 0, 4 * sizeof(int), stream));
        launch_stage(stage, dt);
        n_fail = 0;
        n_invalid = 0;
        std::vector<int> cnt(4 * blocks.size());
        for (size_t q = 0; q < blocks.size(); ++q)
            HIPCHK(hipMemcpyAsync(&cnt[4 * q], blocks[q]->counters, 4 * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (size_t q = 0; q < blocks.size(); ++q) {
            n_fail += cnt[4 * q];
            n_invalid = std::max(n_invalid, cnt[4 * q + 1]);
        }
        // The stage output becomes the current flow state.
        for (auto& bp : blocks) bp->cur = 1 - bp->cur;
    }

    // Public single stage (tests): stage 1 leaves fs = decode(U1), stage 2 fs = decode(U2);
    // U[0] is not swapped.
    int stage_public(int stage, double dt) {
        if (stage != 1 && stage != 2) throw Error(GDTK_ERR_ARG, "stage must be 1 or 2");
        int nf, ni;
        stage_checked(stage, dt, nf, ni);
        if (nf > 0) return GDTK_ERR_BAD_CELL;
        if (ni > cfg.max_invalid_cells) return GDTK_ERR_TOO_MANY_BAD_CELLS;
        return GDTK_OK;
    }

    void block_dt(Block& b, double dt_current, double cfl, bool check_cfl, double out[3]) {
        refresh_dev(b);
        HIPCHK(launch_dt(b.dev, dt_current, cfl, b.dtpart.p, b.dtout.p, stream));
        double h[4];
        HIPCHK(hipMemcpyAsync(h, b.dtout.p, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        double dt_allow = h[0], cflm = h[1], last_signal = h[2];
        const double cfl_allow = 1.2, cfl_adjust = 0.5;  // fluidblock.d:1011, 1002
        if (check_cfl && (cflm < 0.0 || cflm > cfl_allow)) {
            cflm = cfl_adjust * cfl_allow;
            dt_allow = cflm / last_signal;
        }
        out[0] = dt_allow;
        out[1] = cflm;
        out[2] = 0.0;
    }

    // simcore_gasdynamic_step.d:60-159
    void determine_time_step_size() {
        if ((step % cfg.cfl_count) != 0) return;
        double da = std::numeric_limits<double>::max(), cm = 0.0;
        for (auto& bp : blocks) {
            double r[3];
            block_dt(*bp, dt_global, cfg.cfl_value, step > 0, r);
            da = std::min(da, r[0]);
            cm = std::max(cm, r[1]);
        }
        dt_allow = da;
        cfl_max = cm;
        if (step == 0) dt_allow = std::fmin(cfg.dt_init, dt_allow);
        if (dt_allow <= dt_global) {
            dt_global = dt_allow;
        } else {
            dt_global = std::min(dt_global * 1.5, dt_allow);
            dt_global = std::min(dt_global, cfg.dt_max);
        }
    }

    // simcore_gasdynamic_step.d:906-1575 for the pc scheme.
    void gasdynamic_step() {
        double t0 = time;
        int attempt = 0;
        bool failed = false;
        do {
            ++attempt;
            failed = false;
            for (int s = 1; s <= 2; ++s) {
                int nf, ni;
                stage_checked(s, dt_global, nf, ni);
                if (nf > 0) {
                    failed = true;
                    dt_global = dt_global * 0.2;
                    // Restart from U[0]: restore the flow state by decoding U[0]
                    // (the reference retries from whatever state its cells hold).
                    for (auto& bp : blocks) {
                        Block& b = *bp;
                        b.cur = 0;
                        refresh_dev(b);
                        HIPCHK(launch_decode(b.dev, gas, b.U[0].p, b.fs[0].p, lim.low_T, stream));
                    }
                    break;
                }
                if (ni > cfg.max_invalid_cells)
                    throw Error(GDTK_ERR_TOO_MANY_BAD_CELLS, "Too many bad cells during explicit gasdynamic update.");
            }
        } while (failed && attempt < cfg.max_attempts_for_step);
        attempts_last = attempt;
        if (failed) throw Error(GDTK_ERR_STEP_FAILED, "Explicit update failed after max attempts; giving up.");
        // swap(U[0], U[end]) (simcore_gasdynamic_step.d:1557-1562); after two stages the
        // current flow state is again in fs[0] when cur toggled twice.
        for (auto& bp : blocks) {
            std::swap(bp->U[0].p, bp->U[1].p);
            if (bp->cur != 0) {
                std::swap(bp->fs[0].p, bp->fs[1].p);
                bp->cur = 0;
            }
            refresh_dev(*bp);
        }
        time = t0 + dt_global;
    }

    int run(int max_steps, double max_time) {
        int taken = 0;
        while (true) {
            determine_time_step_size();
            gasdynamic_step();
            step += 1;
            ++taken;
            if (time >= max_time) break;
            if (taken >= max_steps) break;
        }
        return taken;
    }

    void state(double* st) const {
        if (!st) return;
        st[0] = time; st[1] = dt_global; st[2] = step; st[3] = cfl_max; st[4] = dt_allow; st[5] = attempts_last;
    }
};

static std::mutex g_mu;
static std::vector<std::unique_ptr<Sim>> g_sims;

static Sim& get_sim(int h) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (h < 0 || h >= (int)g_sims.size() || !g_sims[h]) throw Error(GDTK_ERR_ARG, "bad simulation handle");
    return *g_sims[h];
}

}  // namespace gdtk

using namespace gdtk;

extern "C" {

const char* gdtk_hip_last_error(void) { return g_last_error.c_str(); }

int gdtk_hip_sim_create(const gdtk_sim_config* cfg, int* sim) {
    return guarded([&] {
        if (!cfg || !sim) throw Error(GDTK_ERR_ARG, "null argument");
        if (cfg->dimensions != 2 && cfg->dimensions != 3) throw Error(GDTK_ERR_ARG, "dimensions must be 2 or 3");
        if (cfg->flux_calculator != GDTK_FLUX_AUSMDV && cfg->flux_calculator != GDTK_FLUX_EFM)
            throw Error(GDTK_ERR_UNSUPPORTED, "flux calculator not implemented");
        if (cfg->interpolation_order != 2) throw Error(GDTK_ERR_UNSUPPORTED, "only interpolation_order 2 is implemented");
        if (!(cfg->mMass > 0.0) || !(cfg->gamma > 1.0)) throw Error(GDTK_ERR_ARG, "bad gas parameters");
        if (cfg->cfl_count < 1 || cfg->max_attempts_for_step < 1) throw Error(GDTK_ERR_ARG, "bad step control");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw Error(GDTK_ERR_HIP, "no HIP device");
        auto s = std::make_unique<Sim>();
        s->cfg = *cfg;
        // ideal_gas.d:66-70
        double R = 8.31451 / cfg->mMass;
        s->gas.R = R;
        s->gas.Cv = R / (cfg->gamma - 1.0);
        s->gas.Cvinv = 1.0 / s->gas.Cv;
        s->gas.gamma = cfg->gamma;
        s->dt_global = cfg->dt_init;
        HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&s->ev0));
        HIPCHK(hipEventCreate(&s->ev1));
        std::lock_guard<std::mutex> lk(g_mu);
        g_sims.push_back(std::move(s));
        *sim = (int)g_sims.size() - 1;
        return GDTK_OK;
    });
}

int gdtk_hip_sim_destroy(int sim) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_mu);
        if (sim < 0 || sim >= (int)g_sims.size() || !g_sims[sim]) throw Error(GDTK_ERR_ARG, "bad simulation handle");
        (void)hipStreamSynchronize(g_sims[sim]->stream);
        g_sims[sim].reset();
        return GDTK_OK;
    });
}

int gdtk_hip_sim_add_block(int sim, const gdtk_block_desc* desc, int* block_id) {
    return guarded([&] {
        if (!desc || !block_id) throw Error(GDTK_ERR_ARG, "null argument");
        get_sim(sim).add_block(*desc, block_id);
        return GDTK_OK;
    });
}

int gdtk_hip_block_set_initial(int sim, int block, const double* p, const double* T, const double* vx,
                               const double* vy, const double* vz) {
    return guarded([&] {
        if (!p || !T || !vx || !vy || !vz) throw Error(GDTK_ERR_ARG, "null argument");
        get_sim(sim).set_initial(block, p, T, vx, vy, vz);
        return GDTK_OK;
    });
}

int gdtk_hip_block_get_field(int sim, int block, int field, double* out) {
    return guarded([&] {
        if (!out) throw Error(GDTK_ERR_ARG, "null argument");
        get_sim(sim).get_field(block, field, out);
        return GDTK_OK;
    });
}

int gdtk_hip_stage(int sim, int stage, double dt) {
    return guarded([&] { return get_sim(sim).stage_public(stage, dt); });
}

int gdtk_hip_block_dt(int sim, int block, double dt_current, double cfl_value, int check_cfl, double* out) {
    return guarded([&] {
        if (!out) throw Error(GDTK_ERR_ARG, "null argument");
        Sim& s = get_sim(sim);
        s.block_dt(s.blk(block), dt_current, cfl_value, check_cfl != 0, out);
        return GDTK_OK;
    });
}

int gdtk_hip_sim_run(int sim, int max_steps, double max_time, double* state) {
    Sim* sp = nullptr;
    int rc = guarded([&] {
        sp = &get_sim(sim);
        return sp->run(max_steps, max_time);
    });
    if (sp) sp->state(state);
    return rc;
}

int gdtk_hip_sim_step_fixed_dt(int sim, double dt, double* state) {
    Sim* sp = nullptr;
    int rc = guarded([&] {
        sp = &get_sim(sim);
        sp->dt_global = dt;
        sp->gasdynamic_step();
        sp->step += 1;
        return GDTK_OK;
    });
    if (sp) sp->state(state);
    return rc;
}

int gdtk_hip_sim_set_dt(int sim, double dt_global, int step) {
    return guarded([&] {
        Sim& s = get_sim(sim);
        s.dt_global = dt_global;
        s.step = step;
        return GDTK_OK;
    });
}

int gdtk_hip_sim_profile(int sim, int enable) {
    return guarded([&] {
        Sim& s = get_sim(sim);
        s.profile = enable != 0;
        for (int q = 0; q < Sim::NKIND; ++q) {
            s.prof_ms[q] = 0.0;
            s.prof_launches[q] = 0;
            s.prof_items[q] = 0.0;
        }
        return GDTK_OK;
    });
}

int gdtk_hip_sim_kernel_stats(int sim, double* stats, int nstats) {
    return guarded([&] {
        Sim& s = get_sim(sim);
        for (int k = 0; k < Sim::NKIND; ++k) {
            double v[3] = {(double)s.prof_launches[k], s.prof_ms[k],
                           s.prof_launches[k] ? s.prof_items[k] / s.prof_launches[k] : 0.0};
            for (int q = 0; q < 3; ++q)
                if (3 * k + q < nstats) stats[3 * k + q] = v[q];
        }
        return GDTK_OK;
    });
}

int gdtk_hip_sim_autotune(int sim, double dt, int reps, double* out, int nout) {
    return guarded([&] {
        if (reps < 1) throw Error(GDTK_ERR_ARG, "reps must be >= 1");
        std::vector<double> r = get_sim(sim).autotune(dt, reps);
        for (int q = 0; q < nout && q < (int)r.size(); ++q) out[q] = r[q];
        return (int)r.size();
    });
}

int gdtk_hip_sim_set_tuning(int sim, int key, int value) {
    return guarded([&] {
        (void)get_sim(sim);
        if (key == GDTK_TUNE_KCHUNK) {
            if (value < 1 || value > 4096) throw Error(GDTK_ERR_ARG, "k chunk out of range");
            set_fused_kchunk(value);
            return GDTK_OK;
        }
        throw Error(GDTK_ERR_ARG, "unknown tuning key");
    });
}

}  // extern "C"
