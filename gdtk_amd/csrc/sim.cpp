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
#include "chem_device.hpp"
#include "chem_host.hpp"
#include "flow_io.hpp"
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
    DevBuf fs[2], U[2], dUdt0, len, vol, fgeom[3], flux[3], fs_save;
    DevArr<unsigned char> shock;  // padded cells (adaptive flux calculator)
    DevArr<unsigned char> fshock;  // faces of the three directions, one after another
    // Flow-file fields the update does not touch: cell centroids (3 planes of active cells)
    // and the gas transport coefficients mu, k carried through an inviscid run (allocated
    // when set or read from a file; written as 0 otherwise).
    DevArr<double> pos, mu, kc;
    int* counters = nullptr;
    ~Block() {
        if (counters) (void)hipFree(counters);
    }
    int64_t ncell() const { return (int64_t)nic * njc * nkc; }
};

// Counters of the batched stage launches, per stage s = 1, 2: 8 per-XCD tile-queue
// counters (TILE_QSTRIDE apart), then 4 cell counters per block ([0] decode failures,
// [1] invalid cells).
struct StageCounters {
    int nblocks = 0;
    int region() const { return 8 * TILE_QSTRIDE + 4 * nblocks; }
    int queue(int stage) const { return (stage - 1) * region(); }
    int cells(int stage, int blk) const { return (stage - 1) * region() + 8 * TILE_QSTRIDE + 4 * blk; }
};

struct Sim {
    gdtk_sim_config cfg;
    IdealGasDev gas;
    CellLimits lim;
    // Multi-species thermally perfect gas (gdtk_hip_sim_set_gas_tpg): species count and the
    // gas table (species gas constants, CEA curves) on host and device.
    int nsp = 1;
    ChemDev tpg{};
    DevArr<ChemDev> tpg_dev;
    // Operator-split chemistry (gdtk_hip_sim_set_chemistry): the reactor handle, or a
    // surrogate handle serving it, dense per-cell arrays and each cell's dt_chem.
    int chem_h = -1, surr_h = -1;
    int64_t ncell_all = 0;
    DevArr<double> cin, cout_, dtc;
    DevArr<int> cstatus;
    DevArr<int64_t> coff;
    DevArr<unsigned> cfail;
    double chem_ms = 0.0;
    long long chem_calls = 0;
    hipEvent_t cev0 = nullptr, cev1 = nullptr;
    // The failed-cell count of the last chemistry step, copied to pinned host memory behind
    // the step and checked at the next point where the host waits for the stream anyway.
    unsigned* cfail_host = nullptr;
    bool chem_pending = false;
    std::vector<std::unique_ptr<Block>> blocks;
    hipStream_t stream = nullptr;
    // Which flow-state copy is the input of the next stage, and which U copy holds U[0]
    // (the reference swaps U[0] and U[2] after a step, simcore_gasdynamic_step.d:1557-1562).
    int cur = 0, u0 = 0;
    // Batched launches over all blocks, set up once all blocks exist: per-block device
    // records for each (stage, cur, u0), the ghost-side table, tile prefix sums, counters,
    // the device copy of dt and the captured steps (one per u0).
    bool finalized = false;
    DevArr<BlockDev> tabs;
    DevArr<SideDev> sides;
    int nsides = 0;
    int64_t nghost = 0;
    DevArr<int> tile_start;
    int total_tiles = 0, kchunk = 16, tile_x = 64, tiles_kc = -1, tiles_tx = -1;
    DevArr<unsigned> ctr;
    StageCounters sc;
    DevArr<double> dt_dev;
    int grid = 0, wg_per_cu = 0;
    bool use_graph = true;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    long long graph_sig[2] = {-1, -1};
    // Face-per-thread launches over all blocks: first launch index of each block's faces
    // (per direction) and cells, aligned to the 256-thread workgroup.
    DevArr<int64_t> fstart[3], cstart;
    std::vector<int64_t> cstart_h;
    int64_t ftotal[3] = {0, 0, 0}, ctotal = 0;
    // Block time-step estimates: one partial per workgroup of the cell launch, each block's
    // last signal, and the per-block results {dt, cfl, last signal}.
    DevArr<double> dt_part, dt_last, dt_out;
    // Step-start copy of every block's flow state (multi-species gas): segments per
    // flow-state copy (cur = 0, 1), one launch, part of the captured step.
    DevArr<CopySeg> save_segs[2];
    int64_t save_chunks = 0;
    // SimState (simcore.d)
    double time = 0.0, dt_global = 0.0, dt_allow = 0.0, cfl_max = 0.0;
    int step = 0, attempts_last = 0;
    // Profiling: HIP events around each kernel launch of the stage, per kernel kind
    // (0 face flux, 1 cell update, 2 fused stage, 3 ghost fill, 4 shock detection), on the
    // library's stream.
    bool profile = false;
    static constexpr int NKIND = 5;
    double prof_ms[NKIND] = {0, 0, 0, 0, 0};
    long long prof_launches[NKIND] = {0, 0, 0, 0, 0};
    double prof_items[NKIND] = {0, 0, 0, 0, 0};
    // Shock-detector options (globalconfig.d:1061-1131) of the adaptive flux calculator.
    gdtk_shock_detector sdet{-0.30, 0.20, 1, 0, 0, 1000000000};
    int numerics = GDTK_NUMERICS_EILMER4;
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
        drop_graphs();
        if (cev0) (void)hipEventDestroy(cev0);
        if (cev1) (void)hipEventDestroy(cev1);
        if (cfail_host) (void)hipHostFree(cfail_host);
        blocks.clear();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int ncq() const { return (cfg.dimensions == 3 ? 5 : 4) + (nsp > 1 ? nsp : 0); }
    int nfs() const { return FS_BASE_PLANES + (nsp > 1 ? 2 * nsp : 0); }

    // FlowState(p, T, vel, massf) of the thermally perfect mixture: update_thermo_from_pT
    // and update_sound_speed (therm_perf_gas.d:230-234, 394-404; perf_gas_mix_eos.d:43-76),
    // as one row of flow-state planes.
    void tpg_state_pT(double p, double T, const double* vel, const double* massf, double* row) const {
        double Rm = 0.0, u = 0.0, Cp = 0.0, Cv = 0.0;
        for (int i = 0; i < nsp; ++i) Rm += massf[i] * tpg.R[i];
        const double rho = p / (Rm * T);
        for (int i = 0; i < nsp; ++i) u += massf[i] * (cea_h(tpg.cea[i], T) - tpg.R[i] * T);
        for (int i = 0; i < nsp; ++i) Cp += massf[i] * cea_Cp(tpg.cea[i], T);
        for (int i = 0; i < nsp; ++i) Cv += massf[i] * (cea_Cp(tpg.cea[i], T) - tpg.R[i]);
        const double gamma = Cp / Cv;
        double Rs = 0.0;
        for (int i = 0; i < nsp; ++i) Rs += massf[i] * tpg.R[i];
        const double a = std::sqrt(gamma * (Rs * T));
        const double base[8] = {rho, vel[0], vel[1], vel[2], u, p, T, a};
        for (int f = 0; f < FS_BASE_PLANES; ++f) row[f] = base[f];
        for (int i = 0; i < nsp; ++i) {
            row[FS_BASE_PLANES + i] = massf[i];
            row[FS_BASE_PLANES + nsp + i] = rho * massf[i];
        }
    }

    void set_gas_tpg(const gdtk_tpg_gas& g) {
        if (!blocks.empty()) throw Error(GDTK_ERR_ARG, "the gas must be set before blocks are added");
        if (!species_supported(g.nsp)) throw Error(GDTK_ERR_UNSUPPORTED, "multi-species flux path: 2 or 5 species");
        build_tpg_gas(g, tpg);
        nsp = g.nsp;
        tpg_dev.alloc(1);
        HIPCHK(hipMemcpy(tpg_dev.p, &tpg, sizeof(ChemDev), hipMemcpyHostToDevice));
    }

    void refresh_dev(Block& b) {
        BlockDev& d = b.dev;
        d.fs_in = b.fs[cur].p;
        d.fs_out = b.fs[1 - cur].p;
        d.U0 = b.U[u0].p;
        d.Unew = b.U[1 - u0].p;
        d.dUdt0 = b.dUdt0.p;
        d.counters = b.counters;
    }

    void drop_graphs() {
        for (int q = 0; q < 2; ++q) {
            if (graph[q]) (void)hipGraphExecDestroy(graph[q]);
            graph[q] = nullptr;
            graph_sig[q] = -1;
        }
    }

    bool adaptive() const { return flux_is_adaptive(cfg.flux_calculator); }

    // Whether this step's stage 1 detects shocks (simcore_gasdynamic_step.d:1033-1035).
    bool detect_now() const {
        return adaptive() && (!sdet.frozen_shock_detector || sdet.shock_detector_freeze_step > step);
    }

    // What a captured step depends on besides the buffers: the kernel variant, the fused
    // kernel's launch shape, whether stage 1 detects shocks and the numerics.
    long long launch_signature() const {
        return (use_fused() ? 1000000LL + 1000LL * kchunk + tile_x + 100000000LL * wg_per_cu : 0LL) +
               (detect_now() ? 10000000000LL : 0LL) + 100000000000LL * numerics;
    }

    void build_starts() {
        auto aligned = [](int64_t n) { return (n + 255) / 256 * 256; };
        const size_t nb = blocks.size();
        const int nd = (cfg.dimensions == 3) ? 3 : 2;
        for (int d = 0; d < 3; ++d) {
            std::vector<int64_t> st(nb + 1, 0);
            if (d < nd)
                for (size_t q = 0; q < nb; ++q) st[q + 1] = st[q] + aligned(blocks[q]->dev.nface[d]);
            ftotal[d] = st[nb];
            fstart[d].alloc(st.size());
            HIPCHK(hipMemcpy(fstart[d].p, st.data(), st.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        }
        std::vector<int64_t> st(nb + 1, 0);
        for (size_t q = 0; q < nb; ++q) st[q + 1] = st[q] + aligned(blocks[q]->ncell());
        ctotal = st[nb];
        cstart.alloc(st.size());
        HIPCHK(hipMemcpy(cstart.p, st.data(), st.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        cstart_h = st;
        dt_part.alloc(2 * (size_t)(ctotal / 256));
        dt_last.alloc(nb);
        dt_out.alloc(3 * nb);
    }

    // Face-flux arrays of the face-per-thread kernels, allocated the first time they run.
    void ensure_flux_all() {
        bool grew = false;
        for (auto& bp : blocks) {
            Block& b = *bp;
            const int nd = (b.dims == 3) ? 3 : 2;
            for (int q = 0; q < nd; ++q)
                if (!b.flux[q].p) {
                    b.flux[q].alloc((size_t)b.dev.nface[q] * b.ncq);
                    b.dev.flux[q] = b.flux[q].p;
                    grew = true;
                }
        }
        if (grew) {
            build_tabs();
            drop_graphs();
        }
    }

    static int tab_index(int stage, int c, int u) { return ((stage - 1) * 2 + c) * 2 + u; }
    const BlockDev* tab(int stage, int c, int u) const {
        return tabs.p + (size_t)tab_index(stage, c, u) * blocks.size();
    }

    // Exchange connections (full_face_copy.d:140-695, check_cell_mapping): the source
    // side must exist and hold the same number of cells, in the orientation's arrangement;
    // every ghost cell's source must lie inside the source block (the mapping is affine
    // in the in-face position, so its corners decide).
    void check_exchange(const Block& b, int side) const {
        const gdtk_block_desc& d = b.desc;
        const int nsd = (cfg.dimensions == 3) ? 6 : 4;
        const int ob = d.other_block[side], os = d.other_face[side], oo = d.other_orientation[side];
        if (ob < 0 || ob >= (int)blocks.size()) throw Error(GDTK_ERR_ARG, "exchange: bad other block");
        if (os < 0 || os >= nsd) throw Error(GDTK_ERR_ARG, "exchange: bad other face");
        if (oo < 0 || oo > 3 || (cfg.dimensions == 2 && oo != 0)) throw Error(GDTK_ERR_ARG, "exchange: bad orientation");
        const Block& o = *blocks[ob];
        int A, B, oA, oB;
        side_extents(side, b.nic, b.njc, b.nkc, A, B);
        side_extents(os, o.nic, o.njc, o.nkc, oA, oB);
        const bool turned = (cfg.dimensions == 3) && (oo % 2 == 1);
        const bool ok = (cfg.dimensions == 2) ? (A == oA) : (turned ? (A == oB && B == oA) : (A == oA && B == oB));
        if (!ok) throw Error(GDTK_ERR_ARG, "exchange: the two faces hold different numbers of cells");
        for (int ca = 0; ca < 2; ++ca)
            for (int cb = 0; cb < 2; ++cb)
                for (int n = 0; n < NGHOST; ++n) {
                    int i, j, k;
                    ffc_source(cfg.dimensions, side, os, oo, ca ? A - 1 : 0, cb ? B - 1 : 0, n, o.nic, o.njc, o.nkc, i,
                               j, k);
                    if (i < 0 || i >= o.nic || j < 0 || j >= o.njc || k < 0 || k >= o.nkc)
                        throw Error(GDTK_ERR_ARG, "exchange: mapping leaves the source block");
                }
    }

    // Tile list of the fused stage launch for k chunks of kc planes and tiles tx cells wide,
    // and the persistent grid for that kernel.
    void build_tiles(int kc, int tx) {
        std::vector<int> ts(blocks.size() + 1, 0);
        for (size_t q = 0; q < blocks.size(); ++q) {
            const int64_t t = (int64_t)ts[q] + fused_tiles(blocks[q]->dev, kc, tx);
            if (t > (int64_t)1 << 30) throw Error(GDTK_ERR_UNSUPPORTED, "too many tiles for one launch");
            ts[q + 1] = (int)t;
        }
        total_tiles = ts.back();
        tile_start.alloc(ts.size());
        HIPCHK(hipMemcpy(tile_start.p, ts.data(), ts.size() * sizeof(int), hipMemcpyHostToDevice));
        tiles_kc = kc;
        tiles_tx = tx;
        grid = fused_batch_grid(cfg.dimensions, cfg.flux_calculator, tx);
        if (grid < 8) throw Error(GDTK_ERR_HIP, "cannot size the stage launch");
        drop_graphs();
    }

    void ensure_tiles() {
        if (tiles_kc != kchunk || tiles_tx != tile_x) build_tiles(kchunk, tile_x);
    }

    void build_tabs() {
        const size_t nb = blocks.size();
        std::vector<BlockDev> h(8 * nb);
        for (int st = 1; st <= 2; ++st)
            for (int c = 0; c < 2; ++c)
                for (int u = 0; u < 2; ++u)
                    for (size_t q = 0; q < nb; ++q) {
                        Block& b = *blocks[q];
                        BlockDev d = b.dev;
                        d.fs_in = b.fs[c].p;
                        d.fs_out = b.fs[1 - c].p;
                        d.U0 = b.U[u].p;
                        d.Unew = b.U[1 - u].p;
                        d.dUdt0 = b.dUdt0.p;
                        d.counters = reinterpret_cast<int*>(ctr.p + sc.cells(st, (int)q));
                        h[(size_t)tab_index(st, c, u) * nb + q] = d;
                    }
        tabs.alloc(h.size());
        HIPCHK(hipMemcpy(tabs.p, h.data(), h.size() * sizeof(BlockDev), hipMemcpyHostToDevice));
    }

    // Sets up the batched launches once every block exists (the reference does the same
    // after building all blocks: set_up_cell_mapping, exchange_geometry).
    void finalize() {
        if (finalized) return;
        if (blocks.empty()) throw Error(GDTK_ERR_ARG, "no blocks");
        const int nsd = (cfg.dimensions == 3) ? 6 : 4;
        std::vector<SideDev> hs;
        int64_t start = 0;
        bool any_exchange = false;
        for (auto& bp : blocks) {
            Block& b = *bp;
            for (int side = 0; side < nsd; ++side) {
                SideDev sd{};
                sd.blk = b.id;
                sd.side = side;
                sd.kind = b.desc.bc_kind[side];
                sd.oblk = sd.oside = -1;
                sd.orient = 0;
                if (sd.kind == GDTK_BC_EXCHANGE) {
                    check_exchange(b, side);
                    sd.oblk = b.desc.other_block[side];
                    sd.oside = b.desc.other_face[side];
                    sd.orient = b.desc.other_orientation[side];
                    any_exchange = true;
                }
                side_extents(side, b.nic, b.njc, b.nkc, sd.na, sd.nb);
                sd.start = start;
                start += (int64_t)sd.na * sd.nb * NGHOST;
                hs.push_back(sd);
            }
        }
        nsides = (int)hs.size();
        nghost = start;
        sides.alloc(hs.size());
        HIPCHK(hipMemcpy(sides.p, hs.data(), hs.size() * sizeof(SideDev), hipMemcpyHostToDevice));
        sc.nblocks = (int)blocks.size();
        ctr.alloc((size_t)2 * sc.region());
        dt_dev.alloc(1);
        for (auto& bp : blocks) refresh_dev(*bp);
        build_tabs();
        build_starts();
        build_tiles(kchunk, tile_x);
        // Geometry exchange: ghost cells of exchange sides take their source cells' lengths.
        if (any_exchange) {
            HIPCHK(launch_ghost_batch(tab(1, 0, 0), sides.p, nsides, nghost, 1, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        finalized = true;
    }

    int launch_grid() const {
        if (wg_per_cu <= 0) return grid;
        int dev = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        return std::max(8, ncu * wg_per_cu);
    }

    // Batched stage over all blocks: ghost cells (exchange + BCs), then the fused stage
    // kernel, or the face-per-thread kernels (one launch for the faces of every direction,
    // then the cell update); dt is read from dt_dev.  Counters of the stage must be clear.
    void launch_batched_stage(int stage, int c, int u) {
        const BlockDev* t = tab(stage, c, u);
        const double nc = (double)total_cells();
        const double eps = 1.0e-12;  // epsilon_van_albada, globalconfig.d:1079
        const int nb = (int)blocks.size();
        // Stage 1 of a step that detects shocks: the ghost fill also copies the indicators,
        // then detect_shocks (simcore_gasdynamic_step.d:1030-1037).
        const bool detect = (stage == 1) && detect_now();
        const int nd = (cfg.dimensions == 3) ? 3 : 2;
        FaceBatch fb{};
        int64_t nfaces = 0;
        for (int d = 0; d < 3; ++d) {
            fb.start[d] = fstart[d].p;
            fb.total[d] = (d < nd) ? ftotal[d] : 0;
            if (d < nd) nfaces += nfaces_total(d);
        }
        timed(3, (double)nghost,
              [&] { return launch_ghost_batch(t, sides.p, nsides, nghost, detect ? 2 : 0, stream); });
        if (detect)
            timed(4, nc, [&] {
                return launch_shock_detect(t, fb, cstart.p, nb, ctotal, cfg.dimensions, sdet.compression_tolerance,
                                           sdet.shear_tolerance, sdet.damped_outflow, stream);
            });
        if (!use_fused()) {
            timed(0, (double)nfaces, [&] {
                if (nsp > 1)
                    return launch_face_flux_species(t, fb, nb, cfg.dimensions, cfg.flux_calculator, numerics, eps,
                                                    nsp, tpg_dev.p, stream);
                return launch_face_flux_batch(t, fb, nb, cfg.dimensions, cfg.flux_calculator, numerics, eps, gas,
                                              stream);
            });
            timed(1, nc, [&] {
                if (nsp > 1)
                    return launch_cell_update_species(t, cstart.p, nb, ctotal, cfg.dimensions, stage, dt_dev.p, lim,
                                                      nsp, tpg_dev.p, stream);
                return launch_cell_update_batch(t, cstart.p, nb, ctotal, cfg.dimensions, stage, dt_dev.p, lim, gas,
                                                stream);
            });
            return;
        }
        StageBatch sb;
        sb.tab = t;
        sb.tile_start = tile_start.p;
        sb.queue = ctr.p + sc.queue(stage);
        sb.nblocks = nb;
        sb.total_tiles = total_tiles;
        sb.per_xcd = (total_tiles + 7) / 8;
        sb.KC = kchunk;
        timed(2, nc, [&] {
            return launch_fused_batch(sb, cfg.dimensions, stage, cfg.flux_calculator, dt_dev.p, eps, lim, gas,
                                      launch_grid(), tile_x, stream);
        });
    }

    int64_t nfaces_total(int d) const {
        int64_t n = 0;
        for (auto& bp : blocks) n += bp->dev.nface[d];
        return n;
    }

    int64_t total_cells() const {
        int64_t n = 0;
        for (auto& bp : blocks) n += bp->ncell();
        return n;
    }

    // Reads the cell counters of the given stages: (decode failures summed over blocks,
    // invalid cells of the worst block) per stage.
    void read_counters(int s_lo, int s_hi, int nf[3], int ni[3]) {
        std::vector<unsigned> h((size_t)2 * sc.region());
        HIPCHK(hipMemcpyAsync(h.data(), ctr.p, h.size() * sizeof(unsigned), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int st = s_lo; st <= s_hi; ++st) {
            nf[st] = 0;
            ni[st] = 0;
            for (int q = 0; q < sc.nblocks; ++q) {
                nf[st] += (int)h[sc.cells(st, q)];
                ni[st] = std::max(ni[st], (int)h[sc.cells(st, q) + 1]);
            }
        }
    }

    // The whole pc step of all blocks as one captured graph (per U parity): clear the
    // counters, then ghost cells + stage kernels for stage 1 and stage 2.
    hipGraphExec_t step_graph(int u) {
        const long long sig = launch_signature();
        if (graph[u] && graph_sig[u] == sig) return graph[u];
        if (graph[u]) {
            (void)hipGraphExecDestroy(graph[u]);
            graph[u] = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        bool saved = profile;
        profile = false;
        try {
            HIPCHK(launch_clear_u32(ctr.p, 2 * sc.region(), stream));
            launch_save(0);
            launch_batched_stage(1, 0, u);
            launch_batched_stage(2, 1, u);
        } catch (...) {
            profile = saved;
            hipGraph_t dummy = nullptr;
            (void)hipStreamEndCapture(stream, &dummy);
            if (dummy) (void)hipGraphDestroy(dummy);
            throw;
        }
        profile = saved;
        HIPCHK(hipStreamEndCapture(stream, &g));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
        graph[u] = ex;
        graph_sig[u] = sig;
        return ex;
    }

    void add_block(const gdtk_block_desc& desc, int* id) {
        if (finalized) throw Error(GDTK_ERR_ARG, "blocks cannot be added after the first stage");
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
        d.nsp = nsp;
        d.nfs = nfs();
        const size_t np = (size_t)d.npad;
        for (int q = 0; q < 2; ++q) b->fs[q].alloc(np * (size_t)d.nfs);
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
        b->pos.alloc(geo.pos.size());
        HIPCHK(hipMemcpy(b->pos.p, geo.pos.data(), geo.pos.size() * sizeof(double), hipMemcpyHostToDevice));
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
            }
            d.fgeom[q] = b->fgeom[q].p;
            d.flux[q] = nullptr;  // face-flux arrays: allocated when the face-per-thread kernels run
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
                if (nsp > 1) {
                    tpg_state_pT(p, T, &desc.inflow[s][2], desc.inflow_massf[s], d.inflow[s]);
                    continue;
                }
                double rho = p / (T * gas.R);
                double u = gas.Cv * T;
                double a = std::sqrt(gas.gamma * gas.R * T);
                double row[8] = {rho, desc.inflow[s][2], desc.inflow[s][3], desc.inflow[s][4], u, p, T, a};
                for (int f = 0; f < 8; ++f) d.inflow[s][f] = row[f];
            }
        }
        HIPCHK(hipMalloc(&b->counters, 4 * sizeof(int)));
        HIPCHK(hipMemset(b->counters, 0, 4 * sizeof(int)));
        d.counters = b->counters;
        d.shock = nullptr;
        for (int q = 0; q < 3; ++q) d.fshock[q] = nullptr;
        if (adaptive()) {
            b->shock.alloc(np);  // zero: the initial flow state's S (flowstate.d:93)
            d.shock = b->shock.p;
            b->fshock.alloc((size_t)(d.nface[0] + d.nface[1] + d.nface[2]));
            int64_t off = 0;
            for (int q = 0; q < 3; ++q) {
                d.fshock[q] = b->fshock.p + off;
                off += d.nface[q];
            }
        }
        refresh_dev(*b);
        *id = b->id;
        blocks.push_back(std::move(b));
    }

    Block& blk(int id) {
        if (id < 0 || id >= (int)blocks.size()) throw Error(GDTK_ERR_ARG, "bad block id");
        return *blocks[id];
    }

    // Initial flow state of a block's cells (simcore.d:326-330): FlowState(p, T, vel[,
    // massf]) per cell, U[0] = encode(fs), fs = decode(U[0]).  massf is species-major
    // [nsp][ncell] for a multi-species gas and ignored otherwise.
    void set_initial(int bid, const double* p, const double* T, const double* vx, const double* vy, const double* vz,
                     const double* massf) {
        Block& b = blk(bid);
        BlockDev& d = b.dev;
        const size_t np = (size_t)d.npad;
        if (nsp > 1 && !massf) throw Error(GDTK_ERR_ARG, "a multi-species gas needs initial mass fractions");
        std::vector<double> h(np * (size_t)d.nfs, 0.0);
        const int kg = (cfg.dimensions == 3) ? NGHOST : 0;
        const size_t ncell = (size_t)b.ncell();
        double row[FS_MAX_PLANES], mf[CH_MAX_SP];
        for (int k = 0; k < b.nkc; ++k)
            for (int j = 0; j < b.njc; ++j)
                for (int i = 0; i < b.nic; ++i) {
                    size_t c = ((size_t)k * b.njc + j) * b.nic + i;
                    size_t pc = (size_t)(i + NGHOST) + (size_t)d.NI * ((size_t)(j + NGHOST) + (size_t)d.NJ * (size_t)(k + kg));
                    double pp = p[c], TT = T[c];
                    if (!(TT > 0.0) || !(pp > 0.0)) throw Error(GDTK_ERR_ARG, "initial state needs positive p and T");
                    if (nsp > 1) {
                        const double vel[3] = {vx[c], vy[c], vz[c]};
                        for (int q = 0; q < nsp; ++q) mf[q] = massf[q * ncell + c];
                        tpg_state_pT(pp, TT, vel, mf, row);
                    } else {
                        // FlowState(p, T, vel) (ideal_gas.d:89-97, 136-143)
                        double rho = pp / (TT * gas.R);
                        double u = gas.Cv * TT;
                        double a = std::sqrt(gas.gamma * gas.R * TT);
                        const double base[8] = {rho, vx[c], vy[c], vz[c], u, pp, TT, a};
                        for (int f = 0; f < 8; ++f) row[f] = base[f];
                    }
                    for (int f = 0; f < d.nfs; ++f) h[pc + f * np] = row[f];
                }
        refresh_dev(b);
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipMemcpy(b.fs[cur].p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        if (b.shock.p) HIPCHK(hipMemsetAsync(b.shock.p, 0, b.shock.n, stream));
        cold_start(b);
    }

    // The cold start of init_simulation (simcore.d:318-336) on a block whose flow-state planes
    // hold the initial state: U[0] = encode(fs), fs = decode(U[0]), dt_chem = -1.
    void cold_start(Block& b) {
        HIPCHK(hipMemsetAsync(b.counters, 0, 4 * sizeof(int), stream));
        if (dtc.p) {
            int64_t off = 0;
            for (int q = 0; q < b.id; ++q) off += blocks[q]->ncell();
            HIPCHK(launch_fill_f64(dtc.p + off, b.ncell(), -1.0, stream));
        }
        if (nsp > 1) {
            HIPCHK(launch_encode_species(b.dev, b.U[u0].p, stream));
            HIPCHK(launch_decode_species(b.dev, tpg_dev.p, b.U[u0].p, b.fs[cur].p, lim.low_T, stream));
        } else {
            HIPCHK(launch_encode(b.dev, b.U[u0].p, stream));
            HIPCHK(launch_decode(b.dev, gas, b.U[u0].p, b.fs[cur].p, lim.low_T, stream));
        }
        int cnt[4];
        HIPCHK(hipMemcpyAsync(cnt, b.counters, 4 * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (cnt[0] > 0) throw Error(GDTK_ERR_BAD_CELL, "bad cell while initializing flow");
    }

    // Species names of the flow file's mass-fraction variables: given, or "air" (the
    // ideal-air model's species) for a single-species gas.
    std::vector<std::string> species_list(const char* names) const {
        std::vector<std::string> v = split_names(names);
        if (v.empty() && nsp == 1) v.push_back("air");
        if ((int)v.size() != nsp) throw Error(GDTK_ERR_ARG, "one species name per species is needed");
        return v;
    }

    // write_solution (fluidblockio_old.d:388-447) of one block: the current flow state packed
    // into cell records on the device, copied to the host and written.
    void write_flow(int bid, const char* fn, int format, const char* label, double sim_time, const char* species) {
        Block& b = blk(bid);
        const std::vector<std::string> vars = flow_variable_list(nsp, species_list(species));
        const int nv = (int)vars.size();
        const int64_t n = b.ncell();
        DevArr<double> rec;
        rec.alloc((size_t)n * nv);
        int64_t off = 0;
        for (int q = 0; q < bid; ++q) off += blocks[q]->ncell();
        refresh_dev(b);
        HIPCHK(launch_pack_flow(b.dev, b.pos.p, b.mu.p, b.kc.p, dtc.p ? dtc.p + off : nullptr, rec.p, nv, stream));
        std::vector<double> h((size_t)n * nv);
        HIPCHK(hipMemcpyAsync(h.data(), rec.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        try {
            write_flow_file(fn, format, label ? label : "", sim_time, cfg.dimensions, b.nic, b.njc,
                            cfg.dimensions == 3 ? b.nkc : 1, vars, h.data(), n);
        } catch (std::exception& e) {
            throw Error(GDTK_ERR_ARG, e.what());
        }
    }

    // read_legacy_solution (fluidblockio_old.d:156-273) of one block followed by the cold
    // start (simcore.d:299-336): the file's rho, vel, p, a, u, T, mass fractions, S, mu and k
    // become the cell states, then U[0] = encode(fs), fs = decode(U[0]) and dt_chem = -1.
    // Returns the file's sim_time.
    double read_flow(int bid, const char* fn, int format, const char* species) {
        Block& b = blk(bid);
        const std::vector<std::string> vars = flow_variable_list(nsp, species_list(species));
        const int nv = (int)vars.size();
        std::vector<double> h;
        double t = 0.0;
        try {
            t = read_flow_file(fn, format, cfg.dimensions, b.nic, b.njc, b.nkc, vars, h);
        } catch (std::exception& e) {
            throw Error(GDTK_ERR_ARG, e.what());
        }
        const int64_t n = b.ncell();
        for (int64_t c = 0; c < n; ++c) {
            const double* r = &h[(size_t)c * nv];
            if (!(r[4] > 0.0) || !(r[nv - 1] > 0.0)) throw Error(GDTK_ERR_ARG, "flow file: cell with rho or T <= 0");
        }
        DevArr<double> rec;
        rec.alloc(h.size());
        if (!b.mu.p) b.mu.alloc((size_t)n);
        if (!b.kc.p) b.kc.alloc((size_t)n);
        refresh_dev(b);
        HIPCHK(hipMemcpy(rec.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(launch_unpack_flow(b.dev, rec.p, b.fs[cur].p, b.mu.p, b.kc.p, nv, stream));
        cold_start(b);
        return t;
    }

    // Per-cell transport coefficients written to the flow files (fs.gas.mu, k; an inviscid
    // run carries them unchanged from the initial state).
    void set_transport(int bid, const double* mu_h, const double* k_h) {
        Block& b = blk(bid);
        const size_t n = (size_t)b.ncell();
        if (!b.mu.p) b.mu.alloc(n);
        if (!b.kc.p) b.kc.alloc(n);
        HIPCHK(hipMemcpy(b.mu.p, mu_h, n * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.kc.p, k_h, n * sizeof(double), hipMemcpyHostToDevice));
    }

    void get_field(int bid, int field, double* out) {
        Block& b = blk(bid);
        BlockDev& d = b.dev;
        if (field == GDTK_FIELD_DT_CHEM) {
            // the cell's dt_chem (fvcell.d:78), kept with the chemistry's cell arrays
            if (!ncell_all) throw Error(GDTK_ERR_ARG, "no chemistry step has run");
            check_chem();
            int64_t off = 0;
            for (int q = 0; q < bid; ++q) off += blocks[q]->ncell();
            HIPCHK(hipStreamSynchronize(stream));
            HIPCHK(hipMemcpy(out, dtc.p + off, (size_t)b.ncell() * sizeof(double), hipMemcpyDeviceToHost));
            return;
        }
        const size_t np = (size_t)d.npad;
        const double* src = nullptr;
        if (field >= GDTK_FIELD_DUDT0 && field < GDTK_FIELD_DUDT0 + b.ncq) src = b.dUdt0.p + (field - GDTK_FIELD_DUDT0) * np;
        else if (field >= GDTK_FIELD_U0 && field < GDTK_FIELD_U0 + b.ncq) src = b.U[u0].p + (field - GDTK_FIELD_U0) * np;
        else if (field >= 0 && field < FS_BASE_PLANES) src = b.fs[cur].p + field * np;
        else if (nsp > 1 && field >= GDTK_FIELD_MASSF && field < GDTK_FIELD_MASSF + nsp)
            src = b.fs[cur].p + (FS_BASE_PLANES + field - GDTK_FIELD_MASSF) * np;
        else if (nsp > 1 && field >= GDTK_FIELD_RHO_S && field < GDTK_FIELD_RHO_S + nsp)
            src = b.fs[cur].p + (FS_BASE_PLANES + nsp + field - GDTK_FIELD_RHO_S) * np;
        else if (field == GDTK_FIELD_VOLUME) src = b.vol.p;
        else if (field >= GDTK_FIELD_ILENGTH && field <= GDTK_FIELD_KLENGTH) src = b.len.p + (field - GDTK_FIELD_ILENGTH) * np;
        else if (field != GDTK_FIELD_SHOCK || !b.shock.p) throw Error(GDTK_ERR_ARG, "bad field");
        std::vector<double> h(np);
        HIPCHK(hipStreamSynchronize(stream));
        if (src) {
            HIPCHK(hipMemcpy(h.data(), src, np * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            std::vector<unsigned char> hs(np);
            HIPCHK(hipMemcpy(hs.data(), b.shock.p, np, hipMemcpyDeviceToHost));
            for (size_t q = 0; q < np; ++q) h[q] = hs[q];
        }
        const int kg = (cfg.dimensions == 3) ? NGHOST : 0;
        for (int k = 0; k < b.nkc; ++k)
            for (int j = 0; j < b.njc; ++j)
                for (int i = 0; i < b.nic; ++i) {
                    size_t c = ((size_t)k * b.njc + j) * b.nic + i;
                    size_t pc = (size_t)(i + NGHOST) + (size_t)d.NI * ((size_t)(j + NGHOST) + (size_t)d.NJ * (size_t)(k + kg));
                    out[c] = h[pc];
                }
    }

    // The fused stage kernel serves the single-species gas with the AUSMDV or EFM
    // calculator and Eilmer 4 numerics; a multi-species gas, the Hanel and adaptive
    // calculators and Eilmer 5 numerics run the face-per-thread kernels.
    bool fusable() const {
        return nsp == 1 && cfg.flux_calculator <= GDTK_FLUX_EFM && numerics == GDTK_NUMERICS_EILMER4;
    }
    bool use_fused() const {
        if (!fusable()) return false;
        return variant >= 0 ? variant == 1 : cfg.use_fused != 0;
    }

    // Prepares the launches of the current variant.
    void prepare_launches() {
        finalize();
        if (use_fused()) ensure_tiles();
        else ensure_flux_all();
    }

    // Launches one stage on every block (counters must be clear); returns without
    // synchronising.  dt is written to dt_dev first.
    void launch_stage(int stage, double dt) {
        prepare_launches();
        HIPCHK(launch_set_f64(dt_dev.p, dt, stream));
        launch_batched_stage(stage, cur, u0);
    }

    void clear_stage_counters(int stage) {
        finalize();
        HIPCHK(launch_clear_u32(ctr.p + sc.queue(stage), sc.region(), stream));
    }

    // Autotuner: times both stages of every candidate launch shape on the current flow
    // state (stage 1 reads fs and U0 and writes the other fs copy and dUdt0; stage 2 then
    // reads the same fs and writes fs_out and U[1]: the state the next step starts from,
    // fs[cur] and U[0], is untouched) and keeps the fastest.  Candidates: face-per-thread
    // kernels, and the fused kernel for tiles 64, 32 and 16 cells wide with k chunks of
    // 8..64 planes (3-D).  Returns {best variant, best k chunk, best ms, best tile width,
    // then (variant, k chunk, ms, tile width) per candidate}.
    std::vector<double> autotune(double dt, int reps) {
        std::vector<double> result;
        const int kcs[] = {8, 16, 32, 64};
        const int txs[] = {64, 32, 16};
        bool saved_profile = profile;
        profile = false;
        // Stage 1 may detect shocks: the indicators are restored afterwards, so that the next
        // step's ghost fill copies the same values as without tuning.
        std::vector<std::unique_ptr<DevArr<unsigned char>>> shock_keep;
        if (detect_now()) {
            finalize();
            for (auto& bp : blocks) {
                shock_keep.emplace_back(new DevArr<unsigned char>());
                shock_keep.back()->alloc(bp->shock.n);
                HIPCHK(hipMemcpyAsync(shock_keep.back()->p, bp->shock.p, bp->shock.n, hipMemcpyDeviceToDevice, stream));
            }
        }
        auto time_variant = [&](int var, int kc, int tx) {
            variant = var;
            if (kc > 0) kchunk = kc;
            if (tx > 0) tile_x = tx;
            clear_stage_counters(1);
            launch_stage(1, dt);  // warm-up
            HIPCHK(hipEventRecord(ev0, stream));
            for (int r = 0; r < reps; ++r) {
                clear_stage_counters(1);
                launch_stage(1, dt);
                clear_stage_counters(2);
                launch_stage(2, dt);
            }
            HIPCHK(hipEventRecord(ev1, stream));
            HIPCHK(hipEventSynchronize(ev1));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            return (double)ms / (2 * reps);
        };
        double best = time_variant(0, 0, 0);
        int best_var = 0, best_kc = 0, best_tx = 0;
        auto add = [&](int var, int kc, int tx, double t) {
            result.push_back(var); result.push_back(kc); result.push_back(t); result.push_back(tx);
            if (t < best) { best = t; best_var = var; best_kc = kc; best_tx = tx; }
        };
        result.push_back(0); result.push_back(0); result.push_back(best); result.push_back(0);
        for (int tx : txs) {
            if (!fusable()) break;  // fused: one species, AUSMDV or EFM, Eilmer 4 numerics
            if (cfg.dimensions == 3) {
                for (int kc : kcs) add(1, kc, tx, time_variant(1, kc, tx));
            } else {
                add(1, 1, tx, time_variant(1, 1, tx));
            }
        }
        variant = best_var;
        if (best_var == 1) {
            kchunk = best_kc;
            tile_x = best_tx;
        }
        for (size_t q = 0; q < shock_keep.size(); ++q)
            HIPCHK(hipMemcpyAsync(blocks[q]->shock.p, shock_keep[q]->p, shock_keep[q]->n, hipMemcpyDeviceToDevice,
                                  stream));
        HIPCHK(hipStreamSynchronize(stream));
        profile = saved_profile;
        result.insert(result.begin(), {(double)best_var, (double)best_kc, best, (double)best_tx});
        return result;
    }

    // Runs one stage on all blocks and reports (decode failures, invalid cells).
    void stage_checked(int stage, double dt, int& n_fail, int& n_invalid) {
        clear_stage_counters(stage);
        launch_stage(stage, dt);
        int nf[3], ni[3];
        read_counters(stage, stage, nf, ni);
        n_fail = nf[stage];
        n_invalid = ni[stage];
        // The stage output becomes the current flow state.
        cur = 1 - cur;
        for (auto& bp : blocks) refresh_dev(*bp);
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

    // Time-step estimates of blocks [q0, q0 + n) from the current flow state, one launch;
    // results at dt_out[3 q ...].
    void launch_dt(int q0, int n, double dt_current, double cfl) {
        finalize();
        const int64_t base = cstart_h[q0], count = cstart_h[q0 + n] - base;
        HIPCHK(launch_dt_batch(tab(1, cur, u0) + q0, cstart.p + q0, n, base, count, cfg.dimensions, dt_current, cfl,
                               dt_part.p + 2 * (base / 256), dt_last.p + q0, dt_out.p + 3 * q0, stream));
    }

    // fluidblock.d:987-1084 from the reduced {dt_allow, cfl_max, last signal}.
    static void finish_block_dt(const double h[3], bool check_cfl, double out[3]) {
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

    void block_dt(Block& b, double dt_current, double cfl, bool check_cfl, double out[3]) {
        launch_dt(b.id, 1, dt_current, cfl);
        double h[3];
        HIPCHK(hipMemcpyAsync(h, dt_out.p + 3 * b.id, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        finish_block_dt(h, check_cfl, out);
    }

    // simcore_gasdynamic_step.d:60-159: every block's estimate in one launch, read back
    // with one copy and one synchronisation.
    void determine_time_step_size() {
        if ((step % cfg.cfl_count) != 0) return;
        const size_t nb = blocks.size();
        std::vector<double> h(3 * nb);
        launch_dt(0, (int)nb, dt_global, cfg.cfl_value);
        HIPCHK(hipMemcpyAsync(h.data(), dt_out.p, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        double da = std::numeric_limits<double>::max(), cm = 0.0;
        for (size_t q = 0; q < nb; ++q) {
            double r[3];
            finish_block_dt(&h[3 * q], step > 0, r);
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

    void set_chemistry(int chem, int surrogate) {
        if (nsp < 2) throw Error(GDTK_ERR_ARG, "chemistry needs a multi-species gas (gdtk_hip_sim_set_gas_tpg)");
        Chem& c = get_chem(chem);
        if (c.nsp != nsp) throw Error(GDTK_ERR_ARG, "the chemistry's species differ from the gas's");
        chem_h = chem;
        surr_h = surrogate;
        if (!cev0) {
            HIPCHK(hipEventCreate(&cev0));
            HIPCHK(hipEventCreate(&cev1));
            HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&cfail_host), sizeof(unsigned)));
            *cfail_host = 0;
        }
    }

    // chemistry_step (simcore.d:1619-1638) for the full_T_full_R splitting: every cell's
    // thermochemical_increment over dt on the device (the batched reactor, or the surrogate
    // that serves it), dt_chem kept per cell (initially -1, simcore.d:335).  A failed cell
    // ends the run, as the reference's FlowSolverException does.
    void chemistry_step(double dt) {
        if (chem_h < 0) return;
        finalize();
        const int nb = (int)blocks.size();
        if (!ncell_all) {
            std::vector<int64_t> off(nb + 1, 0);
            for (int q = 0; q < nb; ++q) off[q + 1] = off[q] + blocks[q]->ncell();
            ncell_all = off[nb];
            coff.alloc(off.size());
            HIPCHK(hipMemcpy(coff.p, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
            const size_t n = (size_t)ncell_all;
            cin.alloc((4 + (size_t)nsp) * n);
            cout_.alloc((4 + (size_t)nsp) * n);
            dtc.alloc(n);
            cstatus.alloc(n);
            cfail.alloc(1);
            HIPCHK(launch_fill_f64(dtc.p, ncell_all, -1.0, stream));
        }
        const int64_t n = ncell_all;
        double *rho = cin.p, *T = cin.p + n, *u = cin.p + 2 * n, *p = cin.p + 3 * n, *mf = cin.p + 4 * n;
        double *To = cout_.p, *uo = cout_.p + n, *po = cout_.p + 2 * n, *mfo = cout_.p + 3 * n,
               *dto = cout_.p + (3 + nsp) * n;
        const BlockDev* t = tab(1, cur, u0);
        HIPCHK(hipEventRecord(cev0, stream));
        HIPCHK(launch_cells_gather(t, cstart.p, coff.p, nb, ctotal, nsp, n, rho, T, u, p, mf, stream));
        int rc;
        if (surr_h >= 0) {
            rc = gdtk_hip_surrogate_step_device(surr_h, (int)n, rho, T, u, p, mf, dtc.p, dt, To, uo, po, mfo, dto,
                                                cstatus.p, nullptr, stream);
        } else {
            rc = gdtk_hip_chem_step_device(chem_h, (int)n, rho, T, u, p, mf, dtc.p, dt, To, uo, po, mfo, dto,
                                           cstatus.p, stream);
        }
        if (rc < 0) throw Error(rc, g_last_error);
        HIPCHK(launch_clear_u32(cfail.p, 1, stream));
        HIPCHK(launch_cells_scatter(t, cstart.p, coff.p, nb, ctotal, nsp, n, To, uo, po, mfo, dto, dtc.p, cstatus.p,
                                    cfail.p, stream));
        HIPCHK(hipMemcpyAsync(cfail_host, cfail.p, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipEventRecord(cev1, stream));
        chem_pending = true;
    }

    // The outcome of the last chemistry step: its time, and the reference's fatal
    // FlowSolverException (fvcell.d:923) if a cell failed.  Called where the host has waited
    // for the stream anyway (after the next step's flux attempt) and before a call returns.
    void check_chem() {
        if (!chem_pending) return;
        chem_pending = false;
        HIPCHK(hipEventSynchronize(cev1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, cev0, cev1));
        chem_ms += ms;
        ++chem_calls;
        if (*cfail_host > 0) throw Error(GDTK_ERR_STEP_FAILED, "The thermochemical_increment() failed.");
    }

    // Restart of a failed attempt from U[0]: the flow state is restored by decoding U[0]
    // (the reference retries from whatever state its cells hold).
    // For a multi-species gas the decode iterates from the cell's temperature, so decoding
    // U[0] again would not return the same state bit for bit: the flow state the step
    // started from is kept in fs_save instead and copied back.
    void restore_from_U0() {
        if (nsp > 1) {
            for (auto& bp : blocks) {
                Block& b = *bp;
                HIPCHK(hipMemcpyAsync(b.fs[0].p, b.fs_save.p, b.fs_save.n * sizeof(double), hipMemcpyDeviceToDevice,
                                      stream));
            }
            cur = 0;
            for (auto& bp : blocks) refresh_dev(*bp);
            return;
        }
        cur = 0;
        for (auto& bp : blocks) {
            Block& b = *bp;
            refresh_dev(b);
            HIPCHK(launch_decode(b.dev, gas, b.U[u0].p, b.fs[0].p, lim.low_T, stream));
        }
    }

    // The save buffers and the copy segments, made once (never during a capture).
    void ensure_save() {
        if (nsp <= 1 || save_chunks > 0) return;
        const int nb = (int)blocks.size();
        std::vector<CopySeg> h[2];
        int64_t start = 0;
        for (auto& bp : blocks) {
            Block& b = *bp;
            b.fs_save.alloc((size_t)b.dev.npad * b.dev.nfs);
            for (int c = 0; c < 2; ++c) h[c].push_back(CopySeg{b.fs[c].p, b.fs_save.p, (int64_t)b.fs_save.n, start});
            start += ((int64_t)b.fs_save.n + 1023) / 1024;
        }
        for (int c = 0; c < 2; ++c) {
            save_segs[c].alloc(nb);
            HIPCHK(hipMemcpy(save_segs[c].p, h[c].data(), nb * sizeof(CopySeg), hipMemcpyHostToDevice));
        }
        save_chunks = start;
    }

    // Keeps the flow state the step starts from (fs[cur]) for a retry (restore_from_U0).
    void launch_save(int c) {
        if (nsp <= 1) return;
        HIPCHK(launch_copy_segments(save_segs[c].p, (int)blocks.size(), save_chunks, stream));
    }

    // One attempt of the pc step as the captured graph: both stages run, then the counters
    // decide as the reference's stage-by-stage checks would (a failed stage 1 makes stage
    // 2's results irrelevant; the retry restarts from U[0], which the attempt leaves
    // untouched).  Returns true when the attempt failed.
    bool graph_attempt() {
        prepare_launches();
        hipGraphExec_t ex = step_graph(u0);
        HIPCHK(launch_set_f64(dt_dev.p, dt_global, stream));
        HIPCHK(hipGraphLaunch(ex, stream));
        int nf[3], ni[3];
        read_counters(1, 2, nf, ni);
        for (int st = 1; st <= 2; ++st) {
            if (nf[st] > 0) return true;
            if (ni[st] > cfg.max_invalid_cells)
                throw Error(GDTK_ERR_TOO_MANY_BAD_CELLS, "Too many bad cells during explicit gasdynamic update.");
        }
        return false;
    }

    // simcore_gasdynamic_step.d:906-1575 for the pc scheme.
    void gasdynamic_step() {
        double t0 = time;
        const bool graphed = use_graph && !profile && cur == 0;
        ensure_save();
        if (!graphed) launch_save(cur);  // the captured step saves as its first node
        int attempt = 0;
        bool failed = false;
        do {
            ++attempt;
            failed = false;
            if (graphed) {
                failed = graph_attempt();
                if (failed) {
                    dt_global = dt_global * 0.2;
                    restore_from_U0();
                }
                continue;
            }
            for (int s = 1; s <= 2; ++s) {
                int nf, ni;
                stage_checked(s, dt_global, nf, ni);
                if (nf > 0) {
                    failed = true;
                    dt_global = dt_global * 0.2;
                    restore_from_U0();
                    break;
                }
                if (ni > cfg.max_invalid_cells)
                    throw Error(GDTK_ERR_TOO_MANY_BAD_CELLS, "Too many bad cells during explicit gasdynamic update.");
            }
        } while (failed && attempt < cfg.max_attempts_for_step);
        attempts_last = attempt;
        check_chem();  // the previous step's chemistry, long done: the attempt has waited
        if (failed) throw Error(GDTK_ERR_STEP_FAILED, "Explicit update failed after max attempts; giving up.");
        // swap(U[0], U[end]) (simcore_gasdynamic_step.d:1557-1562).  The two stages leave
        // the flow state in the copy the step started from.
        u0 = 1 - u0;
        for (auto& bp : blocks) refresh_dev(*bp);
        time = t0 + dt_global;
    }

    int run(int max_steps, double max_time) {
        int taken = 0;
        while (true) {
            determine_time_step_size();
            gasdynamic_step();
            chemistry_step(dt_global);
            step += 1;
            ++taken;
            if (time >= max_time) break;
            if (taken >= max_steps) break;
        }
        check_chem();
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
        if (cfg->flux_calculator < GDTK_FLUX_AUSMDV || cfg->flux_calculator > GDTK_FLUX_ADAPTIVE_HANEL_AUSMDV)
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
        get_sim(sim).set_initial(block, p, T, vx, vy, vz, nullptr);
        return GDTK_OK;
    });
}

int gdtk_hip_block_set_initial_massf(int sim, int block, const double* p, const double* T, const double* vx,
                                     const double* vy, const double* vz, const double* massf) {
    return guarded([&] {
        if (!p || !T || !vx || !vy || !vz || !massf) throw Error(GDTK_ERR_ARG, "null argument");
        get_sim(sim).set_initial(block, p, T, vx, vy, vz, massf);
        return GDTK_OK;
    });
}

int gdtk_hip_sim_set_chemistry(int sim, int chem, int surrogate) {
    return guarded([&] {
        get_sim(sim).set_chemistry(chem, surrogate);
        return GDTK_OK;
    });
}

int gdtk_hip_sim_chem_stats(int sim, double* stats, int nstats) {
    return guarded([&] {
        Sim& s = get_sim(sim);
        s.check_chem();
        double queued = 0.0;
        if (s.surr_h >= 0 && s.chem_calls > 0) {
            double sv[2];
            if (gdtk_hip_surrogate_stats(s.surr_h, sv, 2) == 2) queued = sv[1];
        }
        const double v[3] = {(double)s.chem_calls, s.chem_ms, queued};
        int m = std::min(nstats, 3);
        for (int k = 0; k < m; ++k) stats[k] = v[k];
        return m;
    });
}

int gdtk_hip_sim_set_gas_tpg(int sim, const gdtk_tpg_gas* gas) {
    return guarded([&] {
        if (!gas) throw Error(GDTK_ERR_ARG, "null argument");
        Sim& s = get_sim(sim);
        s.set_gas_tpg(*gas);
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
        sp->chemistry_step(sp->dt_global);
        sp->step += 1;
        sp->check_chem();
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

int gdtk_hip_sim_get_state(int sim, double* state) {
    return guarded([&] {
        if (!state) throw Error(GDTK_ERR_ARG, "null argument");
        get_sim(sim).state(state);
        return GDTK_OK;
    });
}

int gdtk_hip_sim_set_time(int sim, double time) {
    return guarded([&] {
        get_sim(sim).time = time;
        return GDTK_OK;
    });
}

int gdtk_hip_block_write_flow(int sim, int block, const char* filename, int format, const char* label,
                              double sim_time, const char* species_names) {
    return guarded([&] {
        if (!filename) throw Error(GDTK_ERR_ARG, "null file name");
        if (format != GDTK_FLOW_GZIPTEXT && format != GDTK_FLOW_RAWBINARY) throw Error(GDTK_ERR_ARG, "bad flow format");
        get_sim(sim).write_flow(block, filename, format, label, sim_time, species_names);
        return GDTK_OK;
    });
}

int gdtk_hip_block_read_flow(int sim, int block, const char* filename, int format, const char* species_names,
                             double* sim_time) {
    return guarded([&] {
        if (!filename) throw Error(GDTK_ERR_ARG, "null file name");
        if (format != GDTK_FLOW_GZIPTEXT && format != GDTK_FLOW_RAWBINARY) throw Error(GDTK_ERR_ARG, "bad flow format");
        const double t = get_sim(sim).read_flow(block, filename, format, species_names);
        if (sim_time) *sim_time = t;
        return GDTK_OK;
    });
}

int gdtk_hip_block_set_transport(int sim, int block, const double* mu, const double* k) {
    return guarded([&] {
        if (!mu || !k) throw Error(GDTK_ERR_ARG, "null argument");
        get_sim(sim).set_transport(block, mu, k);
        return GDTK_OK;
    });
}

int gdtk_hip_flow_nvars(int nsp) {
    return guarded([&] {
        if (nsp < 1 || nsp > GDTK_MAX_SPECIES) throw Error(GDTK_ERR_ARG, "bad species count");
        return 17 + nsp + (nsp > 1 ? 1 : 0);
    });
}

int gdtk_hip_flow_write_records(const char* filename, int format, const char* label, double sim_time, int dims,
                                int nic, int njc, int nkc, int nsp, const char* species_names, const double* records) {
    return guarded([&] {
        if (!filename || !records || nsp < 1 || nsp > GDTK_MAX_SPECIES || nic < 1 || njc < 1 || nkc < 1)
            throw Error(GDTK_ERR_ARG, "bad argument");
        std::vector<std::string> names = split_names(species_names);
        if (names.empty() && nsp == 1) names.push_back("air");
        try {
            write_flow_file(filename, format, label ? label : "", sim_time, dims, nic, njc, nkc,
                            flow_variable_list(nsp, names), records, (int64_t)nic * njc * nkc);
        } catch (std::exception& e) {
            throw Error(GDTK_ERR_ARG, e.what());
        }
        return GDTK_OK;
    });
}

int gdtk_hip_flow_read_records(const char* filename, int format, int dims, int nic, int njc, int nkc, int nsp,
                               const char* species_names, double* records, double* sim_time) {
    return guarded([&] {
        if (!filename || !records || nsp < 1 || nsp > GDTK_MAX_SPECIES || nic < 1 || njc < 1 || nkc < 1)
            throw Error(GDTK_ERR_ARG, "bad argument");
        std::vector<std::string> names = split_names(species_names);
        if (names.empty() && nsp == 1) names.push_back("air");
        std::vector<double> rec;
        double t = 0.0;
        try {
            t = read_flow_file(filename, format, dims, nic, njc, nkc, flow_variable_list(nsp, names), rec);
        } catch (std::exception& e) {
            throw Error(GDTK_ERR_ARG, e.what());
        }
        std::copy(rec.begin(), rec.end(), records);
        if (sim_time) *sim_time = t;
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

int gdtk_hip_sim_set_shock_detector(int sim, const gdtk_shock_detector* opts) {
    return guarded([&] {
        Sim& s = get_sim(sim);
        if (!opts) throw Error(GDTK_ERR_ARG, "null argument");
        // detect_shock_points throws for a positive tolerance (fluidblock.d:484)
        if (opts->compression_tolerance > 0.0) throw Error(GDTK_ERR_ARG, "compression_tolerance should be negative!");
        if (!opts->strict_shock_detector)
            throw Error(GDTK_ERR_UNSUPPORTED, "only the strict shock detector is implemented");
        s.sdet = *opts;
        s.drop_graphs();  // the tolerances are launch arguments of the captured steps
        return GDTK_OK;
    });
}

int gdtk_hip_sim_set_numerics(int sim, int numerics) {
    return guarded([&] {
        Sim& s = get_sim(sim);
        if (numerics != GDTK_NUMERICS_EILMER4 && numerics != GDTK_NUMERICS_LMR)
            throw Error(GDTK_ERR_ARG, "unknown numerics");
        s.numerics = numerics;
        s.drop_graphs();
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

int gdtk_hip_exchange_source(int dims, int face, int other_face, int orientation, int a, int b, int n, int nic,
                             int njc, int nkc, int* ijk) {
    return guarded([&] {
        if (!ijk) throw Error(GDTK_ERR_ARG, "null argument");
        const int nsd = (dims == 3) ? 6 : 4;
        if ((dims != 2 && dims != 3) || face < 0 || face >= nsd || other_face < 0 || other_face >= nsd ||
            orientation < 0 || orientation > 3 || n < 0 || n >= NGHOST)
            throw Error(GDTK_ERR_ARG, "bad exchange description");
        int i, j, k;
        ffc_source(dims, face, other_face, orientation, a, b, n, nic, njc, nkc, i, j, k);
        ijk[0] = i;
        ijk[1] = j;
        ijk[2] = k;
        return GDTK_OK;
    });
}

int gdtk_hip_sim_set_tuning(int sim, int key, int value) {
    return guarded([&] {
        Sim& s = get_sim(sim);
        if (key == GDTK_TUNE_KCHUNK) {
            if (value < 1 || value > 4096) throw Error(GDTK_ERR_ARG, "k chunk out of range");
            s.kchunk = value;
            return GDTK_OK;
        }
        if (key == GDTK_TUNE_GRAPH) {
            s.use_graph = value != 0;
            return GDTK_OK;
        }
        if (key == GDTK_TUNE_WG_PER_CU) {
            if (value < 0 || value > 32) throw Error(GDTK_ERR_ARG, "workgroups per CU out of range");
            s.wg_per_cu = value;
            s.drop_graphs();
            return GDTK_OK;
        }
        if (key == GDTK_TUNE_TILE_X) {
            if (!fused_tile_x_ok(value)) throw Error(GDTK_ERR_ARG, "tile width must be 16, 32 or 64");
            s.tile_x = value;
            return GDTK_OK;
        }
        if (key == GDTK_TUNE_VARIANT) {
            if (value < -1 || value > 1) throw Error(GDTK_ERR_ARG, "variant must be -1, 0 or 1");
            s.variant = value;
            return GDTK_OK;
        }
        throw Error(GDTK_ERR_ARG, "unknown tuning key");
    });
}

}  // extern "C"
