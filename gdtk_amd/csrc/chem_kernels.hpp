// chem_kernels.hpp -- operator-split finite-rate chemistry on gfx950, one lane per cell.
//
// FVCell.thermochemical_increment (src/eilmer/fvcell.d:858-972) for a batch of cells is
// split into three kernels so that the divergent, long-running part runs alone:
//   k_chem_setup      massf2conc, rate constants at the cell temperature and the first
//                     substep size (chemistry_update.d:140-160, reaction_mechanism.d:50-63)
//   k_chem_integrate  the ChemistryUpdate subcycle controller with the alpha-QSS or RKF
//                     step (chemistry_update.d:162-326), written as a per-lane state
//                     machine that advances one ODE step attempt per loop trip; the retry
//                     with dt_chem = -1 of fvcell.d:899-905 is its second pass
//   k_chem_finish     conc2massf and the final renormalisation, then
//                     update_thermo_from_rhou (fvcell.d:906-925)
// Cells differ in substep count by orders of magnitude.  In the persistent variant the
// integration grid is sized to the occupancy limit and a lane whose cell is done takes the
// next one from a work queue, so the waves stay full until the queue drains.  The queue is
// split into one contiguous cell range per XCD (workgroup blockIdx % 8), each with its own
// counter; a wave takes `chunk` cells per atomic grab, and turns to the other XCDs' ranges
// when its own is exhausted.
#pragma once
#include <hip/hip_runtime.h>

#include "chem_device.hpp"
#include "chem_launch.hpp"

namespace gdtk {

template <int NSP, int NRM, bool DENSE>
__global__ __launch_bounds__(CHEM_BLOCK) void k_chem_setup(const ChemDev* __restrict__ cdp, int n, double dt,
                                                           const double* __restrict__ rho,
                                                           const double* __restrict__ T,
                                                           const double* __restrict__ massf,
                                                           const double* __restrict__ dtc,
                                                           const int* __restrict__ mask, int heavy_ratio,
                                                           ChemWork w) {
    const int i = blockIdx.x * CHEM_BLOCK + threadIdx.x;
    if (i >= n) return;
    if (mask && mask[i] != 0) {
        w.flag[i] = CF_SKIP;
        return;
    }
    const ChemDev& cd = *cdp;
    const double Ti = T[i];
    // fvcell.d:870: frozen below T_frozen
    if (Ti <= cd.T_frozen) {
        w.flag[i] = CF_FROZEN;
        return;
    }
    const double r = rho[i];
    // massf2conc throws on a bad density (gas_model.d:238-253); the CEA curves throw on a
    // NaN temperature.  Both attempts of fvcell.d then fail.
    if (!(r > 0.0) || isnan(Ti)) {
        w.flag[i] = CF_FAILED;
        return;
    }
    double conc[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        conc[s] = massf[(size_t)s * n + i] * r / cd.M[s];
        if (conc[s] < CH_MIN_MOLES) conc[s] = 0.0;
        w.conc0[(size_t)s * n + i] = conc[s];
    }
    double kf[NRM], kb[NRM];
    eval_rate_constants<NSP, NRM>(cd, Ti, kf, kb);
    const int nr = cd.nr;
#pragma unroll
    for (int ir = 0; ir < NRM; ++ir) {
        if (ir >= nr) continue;
        w.kfkb[(size_t)ir * n + i] = kf[ir];
        w.kfkb[(size_t)(nr + ir) * n + i] = kb[ir];
    }
    // chemistry_update.d:152-158.  The estimate is also the first step of the retry with
    // dt_chem = -1 (fvcell.d:899-905), which starts from the same state.
    const double d = dtc[i];
    const double est = estimate_step_size<NSP, NRM, DENSE>(cd, kf, kb, conc);
    double h;
    if (d > dt) h = dt;
    else if (d <= 0.0) h = est;
    else h = d;
    w.h0[i] = h;
    w.hest[i] = est;
    // Cells whose integration is predicted long go to the heavy list, one atomic per wave.
    const bool heavy = heavy_ratio > 0 && dt > (double)heavy_ratio * h;
    w.flag[i] = heavy ? CF_HEAVY : CF_INTEGRATE;
    const unsigned long long hb = __ballot(heavy);
    if (hb != 0ull) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((long long)hb) - 1;
        unsigned base = 0u;
        if (lane == leader) base = atomicAdd(&w.queue[9 * QUEUE_STRIDE], (unsigned)__popcll(hb));
        base = __shfl(base, leader);
        if (heavy) w.heavy[base + (unsigned)__popcll(hb & ((1ull << lane) - 1ull))] = i;
    }
}

// Per-lane state of one cell inside the subcycle controller; T and u are the gas state's
// temperature and internal energy, which only the tight temperature coupling changes.
template <int NSP, int NRM>
struct CellRun {
    int cell = -1;
    int pass, cycle, attempt, nsub, iters;
    double rho, t, h, dtSave, dtSug, T, u;
    double kf[NRM], kb[NRM], c0[NSP];
};

template <int NSP, int NRM, bool TIGHT>
__device__ __forceinline__ void load_cell(const ChemDev& cd, int n, int i, const double* __restrict__ rho,
                                          const double* __restrict__ T, const double* __restrict__ u,
                                          const double* __restrict__ dtc_in, const ChemWork& w,
                                          CellRun<NSP, NRM>& c) {
    const int nr = cd.nr;
    c.cell = i;
    c.rho = rho[i];
    if (TIGHT) {
        c.T = T[i];
        c.u = u[i];
    }
#pragma unroll
    for (int ir = 0; ir < NRM; ++ir) {
        c.kf[ir] = (ir < nr) ? w.kfkb[(size_t)ir * n + i] : 0.0;
        c.kb[ir] = (ir < nr) ? w.kfkb[(size_t)(nr + ir) * n + i] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < NSP; ++s) c.c0[s] = w.conc0[(size_t)s * n + i];
    c.h = w.h0[i];
    c.dtSug = dtc_in[i];
    c.pass = 0;
    c.cycle = 0;
    c.attempt = 0;
    c.t = 0.0;
    c.dtSave = 0.0;
    c.nsub = 0;
    c.iters = 0;
}

// TIGHT: tightTempCoupling (chemistry_update.d:265-269): after every successful subcycle the
// mass fractions of the new concentrations, T from u by Newton (started at the gas state's
// T, which keeps the result of the previous update), and the rate constants at that T.  A
// failed T(u) iteration there raises GasModelException in the reference, which the
// thermochemical_increment of fvcell.d:894-905 does not catch: the cell fails without the
// retry.
template <int NSP, int NRM, int METHOD, bool TIGHT, bool DENSE>
__global__ __launch_bounds__(CHEM_BLOCK) void k_chem_integrate(const ChemDev* __restrict__ cdp, int n, double dt,
                                                               const double* __restrict__ rho,
                                                               const double* __restrict__ T_in,
                                                               const double* __restrict__ u_in,
                                                               const double* __restrict__ dtc_in,
                                                               double* __restrict__ dtc_out, ChemWork w,
                                                               int persistent, int chunk, int refill_period) {
    const ChemDev& cd = *cdp;
    const double DT_INCREASE_PERCENT = 10.0, DT_REDUCTION_FACTOR = 10.0, H_MIN = 1.0e-15,
                 ALLOWABLE_MASSF_ERROR = 1.0e-3;
    const int lane = threadIdx.x & 63;
    const unsigned long long lane_bit = 1ull << lane;
    CellRun<NSP, NRM> c;
    unsigned long long acc_sub = 0, acc_it = 0, max_sub = 0, acc_att = 0;

    bool exhausted = true;
    int w_next = 0, w_end = 0;
    int from_heavy = 0;      // [w_next, w_end) indexes the heavy list (wave-uniform)
    unsigned drained = 0u;   // lane 0: ranges seen empty (bits 0-7 per XCD, bit 8 heavy list)
    const int xcd = blockIdx.x & 7;
    if (persistent) {
        exhausted = false;
    } else {
        const int g = blockIdx.x * CHEM_BLOCK + threadIdx.x;
        if (g < n && (w.flag[g] == CF_INTEGRATE || w.flag[g] == CF_HEAVY))
            load_cell<NSP, NRM, TIGHT>(cd, n, g, rho, T_in, u_in, dtc_in, w, c);
    }

    for (int trip = 0;; ++trip) {
        // Refill finished lanes from the queue (wave-uniform control flow).
        if (!exhausted) {
            unsigned long long busy = __ballot(c.cell >= 0);
            if (refill_period <= 1 || (trip % refill_period) == 0 || busy == 0ull) {
                unsigned long long need = __ballot(c.cell < 0);
                while (need != 0ull) {
                    if (w_next >= w_end) {
                        int base = -1, end = 0, heavy = 0;
                        if (lane == 0) {
                            // The heavy list first, then the own XCD's range, then the others'.
                            // A list or range seen drained is not probed again; probing one
                            // costs a load, grabbing from it a contended atomic.
                            if (!(drained & 256u)) {
                                const unsigned len = __hip_atomic_load(&w.queue[9 * QUEUE_STRIDE], __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_AGENT);
                                const unsigned seen = __hip_atomic_load(&w.queue[8 * QUEUE_STRIDE], __ATOMIC_RELAXED,
                                                                        __HIP_MEMORY_SCOPE_AGENT);
                                const unsigned got =
                                    (seen < len) ? atomicAdd(&w.queue[8 * QUEUE_STRIDE], (unsigned)chunk) : len;
                                if (got < len) {
                                    base = (int)got;
                                    end = (int)((got + (unsigned)chunk < len) ? got + (unsigned)chunk : len);
                                    heavy = 1;
                                } else {
                                    drained |= 256u;
                                }
                            }
                            for (int j = 0; j < 8 && base < 0; ++j) {
                                const int q = (xcd + j) & 7;
                                if (drained & (1u << q)) continue;
                                const long long lo = (long long)n * q / 8, hi = (long long)n * (q + 1) / 8;
                                const unsigned seen = __hip_atomic_load(&w.queue[q * QUEUE_STRIDE], __ATOMIC_RELAXED,
                                                                        __HIP_MEMORY_SCOPE_AGENT);
                                const unsigned got = (lo + (long long)seen < hi)
                                                         ? atomicAdd(&w.queue[q * QUEUE_STRIDE], (unsigned)chunk)
                                                         : (unsigned)(hi - lo);
                                if (lo + (long long)got < hi) {
                                    base = (int)(lo + got);
                                    end = (int)((lo + got + chunk < hi) ? lo + got + chunk : hi);
                                } else {
                                    drained |= 1u << q;
                                }
                            }
                        }
                        base = __shfl(base, 0);
                        end = __shfl(end, 0);
                        from_heavy = __shfl(heavy, 0);
                        if (base < 0) {
                            exhausted = true;
                            break;
                        }
                        w_next = base;
                        w_end = end;
                    }
                    const int avail = w_end - w_next;
                    const int k = __popcll(need);
                    const int take = (k < avail) ? k : avail;
                    const int rank = __popcll(need & (lane_bit - 1ull));
                    if ((need & lane_bit) && rank < take) {
                        if (from_heavy) {
                            const int i = w.heavy[w_next + rank];
                            load_cell<NSP, NRM, TIGHT>(cd, n, i, rho, T_in, u_in, dtc_in, w, c);
                        } else {
                            const int i = w_next + rank;
                            if (w.flag[i] == CF_INTEGRATE)
                                load_cell<NSP, NRM, TIGHT>(cd, n, i, rho, T_in, u_in, dtc_in, w, c);
                        }
                    }
                    w_next += take;
                    need = __ballot(c.cell < 0);
                }
            }
        }
        if (__ballot(c.cell >= 0) == 0ull) break;
        if (c.cell < 0) continue;

        // One ODE step attempt of ChemistryUpdate.opCall (chemistry_update.d:162-277).
        ++acc_att;
        if (c.attempt == 0) {
            c.dtSave = c.dtSug;
            c.h = ((dt - c.t) < c.h) ? (dt - c.t) : c.h;
        }
        double y[NSP];
        bool ok;
        if (METHOD == 1) {
            ok = rkf_step<NSP, NRM, DENSE>(cd, c.kf, c.kb, c.c0, c.h, y, c.dtSug);
        } else {
            ok = alpha_qss_step<NSP, NRM, DENSE>(cd, c.kf, c.kb, c.c0, c.h, y, c.dtSug, c.iters);
        }
        bool passes = true;
        if (ok) {
            double mf[NSP];
            const double massfTotal = conc_to_massf<NSP>(cd, y, c.rho, mf);
            if (fabs(massfTotal - 1.0) > ALLOWABLE_MASSF_ERROR) passes = false;
        }
        int outcome = 0;  // 0 continue, 1 done, -1 this pass failed
        if (ok && passes) {
            c.t += c.h;
#pragma unroll
            for (int s = 0; s < NSP; ++s) c.c0[s] = y[s];
            const double hMax = c.h * (1.0 + DT_INCREASE_PERCENT / 100.0);
            if (c.dtSug > c.h) c.h = (hMax < c.dtSug) ? hMax : c.dtSug;
            ++c.nsub;
            bool fatal = false;
            if (TIGHT) {
                double mf[NSP];
                (void)conc_to_massf<NSP>(cd, c.c0, c.rho, mf);
                if (tpg_update_temperature<NSP>(cd, mf, c.T, c.u)) eval_rate_constants<NSP, NRM>(cd, c.T, c.kf, c.kb);
                else fatal = true;
            }
            if (fatal) {
                c.pass = 1;  // no retry
                c.dtSug = -1.0;
                outcome = -1;
            } else if (c.t >= dt) {
                if (c.cycle == 0) c.dtSave = c.dtSug;
                outcome = 1;
            } else {
                ++c.cycle;
                c.attempt = 0;
                if (c.cycle == cd.maxSubcycles) outcome = -1;
            }
        } else {
            c.h /= DT_REDUCTION_FACTOR;
            if (c.h < H_MIN) {
                outcome = -1;
            } else {
                ++c.attempt;
                if (c.attempt == cd.maxAttempts) outcome = -1;
            }
        }
        if (outcome == -1 && c.pass == 0) {
            // fvcell.d:899-905: second attempt from the initial state with dt_chem = -1
            const int i = c.cell;
#pragma unroll
            for (int s = 0; s < NSP; ++s) c.c0[s] = w.conc0[(size_t)s * n + i];
            if (TIGHT) {
                // the gas state and rate constants of the first attempt's start
                c.T = T_in[i];
                c.u = u_in[i];
                const int nr = cd.nr;
#pragma unroll
                for (int ir = 0; ir < NRM; ++ir) {
                    if (ir >= nr) continue;
                    c.kf[ir] = w.kfkb[(size_t)ir * n + i];
                    c.kb[ir] = w.kfkb[(size_t)(nr + ir) * n + i];
                }
            }
            c.pass = 1;
            c.dtSug = -1.0;
            c.h = w.hest[i];
            c.cycle = 0;
            c.attempt = 0;
            c.t = 0.0;
            c.dtSave = 0.0;
            outcome = 0;
        }
        if (outcome != 0) {
            const int i = c.cell;
            if (outcome == 1) {
#pragma unroll
                for (int s = 0; s < NSP; ++s) w.conc[(size_t)s * n + i] = c.c0[s];
                dtc_out[i] = c.dtSave;
                w.st[i] = 0;
                if (TIGHT) {
                    w.Tend[i] = c.T;
                    w.uend[i] = c.u;
                }
            } else {
                dtc_out[i] = c.dtSug;
                w.st[i] = -1;
            }
            w.nsub[i] = c.nsub;
            acc_sub += (unsigned long long)c.nsub;
            acc_it += (unsigned long long)c.iters;
            max_sub = ((unsigned long long)c.nsub > max_sub) ? (unsigned long long)c.nsub : max_sub;
            c.cell = -1;
        }
    }
    // Wave totals of the statistics, stored per wave (the host reduces them when asked):
    // thousands of same-address atomics would serialise at the end of every launch.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        acc_sub += __shfl_xor(acc_sub, off);
        acc_it += __shfl_xor(acc_it, off);
        acc_att += __shfl_xor(acc_att, off);
        const unsigned long long o = __shfl_xor(max_sub, off);
        max_sub = (o > max_sub) ? o : max_sub;
    }
    if (lane == 0) {
        const int wave = blockIdx.x * CHEM_WAVES_PER_BLOCK + (threadIdx.x >> 6);
        w.stats[4 * wave + 0] = acc_sub;
        w.stats[4 * wave + 1] = acc_it;
        w.stats[4 * wave + 2] = max_sub;
        w.stats[4 * wave + 3] = acc_att;
    }
}

template <int NSP>
__global__ __launch_bounds__(CHEM_BLOCK, 3) void k_chem_finish(const ChemDev* __restrict__ cdp, int n,
                                                            const double* __restrict__ rho, const double* T,
                                                            const double* u, const double* p, const double* massf,
                                                            const double* dtc_in, double* T_out, double* u_out,
                                                            double* p_out, double* massf_out, double* dtc_out,
                                                            int* status, ChemWork w) {
    const int i = blockIdx.x * CHEM_BLOCK + threadIdx.x;
    if (i >= n) return;
    const ChemDev& cd = *cdp;
    const int f = w.flag[i];
    if (f == CF_SKIP) return;
    double Ti = T[i], ui = u[i], pi = p[i];
    double mf[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) mf[s] = massf[(size_t)s * n + i];
    int stat = 0;
    if (f == CF_FROZEN) {
        dtc_out[i] = dtc_in[i];
    } else if (f == CF_FAILED) {
        dtc_out[i] = -1.0;
        stat = -1;
    } else if (w.st[i] != 0) {
        stat = -1;
    } else {
        double conc[NSP], mfn[NSP];
#pragma unroll
        for (int s = 0; s < NSP; ++s) conc[s] = w.conc[(size_t)s * n + i];
        const double r = rho[i];
        // chemistry_update.d:290-294
        const double massfTotal = conc_to_massf<NSP>(cd, conc, r, mfn);
#pragma unroll
        for (int s = 0; s < NSP; ++s) mfn[s] /= massfTotal;
        // update_thermo_from_rhou (therm_perf_gas.d:120-135): T from u by Newton, then p.  With
        // the tight coupling the gas state's T and u are those of the last subcycle's update.
        double Tn = Ti, un = ui;
        if (cd.tight) {
            Tn = w.Tend[i];
            un = w.uend[i];
        }
        if (tpg_update_temperature<NSP>(cd, mfn, Tn, un)) {
            const double Rm = tpg_Rmix<NSP>(cd, mfn);
            pi = r * Rm * Tn;
            Ti = Tn;
            ui = un;
#pragma unroll
            for (int s = 0; s < NSP; ++s) mf[s] = mfn[s];
        } else {
            stat = -1;
        }
    }
    T_out[i] = Ti;
    u_out[i] = ui;
    p_out[i] = pi;
#pragma unroll
    for (int s = 0; s < NSP; ++s) massf_out[(size_t)s * n + i] = mf[s];
    status[i] = stat;
}

// Gas-model state updates (therm_perf_gas.d:104-160, 394-404): mode 0 pT, 1 rhou, 2 rhoT,
// then the sound speed.  status -1 where the temperature iteration fails.
template <int NSP>
__global__ __launch_bounds__(CHEM_BLOCK) void k_tpg_update(const ChemDev* __restrict__ cdp, int mode, int n,
                                                           double* rho, double* p, double* T, double* u, double* a,
                                                           const double* __restrict__ massf, int* status) {
    const int i = blockIdx.x * CHEM_BLOCK + threadIdx.x;
    if (i >= n) return;
    const ChemDev& cd = *cdp;
    double mf[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) mf[s] = massf[(size_t)s * n + i];
    int stat = 0;
    if (mode == 0) {
        const double Rm = tpg_Rmix<NSP>(cd, mf);
        const double denom = Rm * T[i];
        rho[i] = p[i] / denom;
        u[i] = tpg_energy<NSP>(cd, mf, T[i]);
    } else if (mode == 1) {
        double Tn = T[i], un = u[i];
        const bool ok = tpg_update_temperature<NSP>(cd, mf, Tn, un);
        T[i] = Tn;
        u[i] = un;
        if (ok) {
            const double Rm = tpg_Rmix<NSP>(cd, mf);
            p[i] = rho[i] * Rm * Tn;
        } else {
            stat = -1;
        }
    } else {
        u[i] = tpg_energy<NSP>(cd, mf, T[i]);
        const double Rm = tpg_Rmix<NSP>(cd, mf);
        p[i] = rho[i] * Rm * T[i];
    }
    if (stat == 0) a[i] = tpg_sound_speed<NSP>(cd, mf, T[i]);
    status[i] = stat;
}

static inline unsigned grid_for(int n) { return (unsigned)((n + CHEM_BLOCK - 1) / CHEM_BLOCK); }

struct SetupL {
    const ChemDev* cd; int n; double dt; const double *rho, *T, *massf, *dtc; const int* mask; int heavy_ratio;
    ChemWork w; hipStream_t s;
    template <int NSP, int NRM, bool DENSE>
    hipError_t run() {
        hipLaunchKernelGGL((k_chem_setup<NSP, NRM, DENSE>), dim3(grid_for(n)), dim3(CHEM_BLOCK), 0, s, cd, n, dt, rho,
                           T, massf, dtc, mask, heavy_ratio, w);
        return hipGetLastError();
    }
};

struct IntegrateL {
    const ChemDev* cd; int n; double dt; const double* rho; const double *T, *u; const double* dtc_in;
    double* dtc_out; ChemWork w; ChemTuning tu; int num_cu; int method; int tight; int* waves_out; hipStream_t s;
    template <int NSP, int NRM, int METHOD, bool TIGHT, bool DENSE>
    hipError_t launch() {
        unsigned grid = grid_for(n);
        if (tu.persistent) {
            int per_cu = tu.blocks_per_cu;
            if (per_cu <= 0) {
                hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, reinterpret_cast<const void*>(&k_chem_integrate<NSP, NRM, METHOD, TIGHT, DENSE>), CHEM_BLOCK,
                    0);
                if (e != hipSuccess) return e;
                if (per_cu < 1) per_cu = 1;
            }
            const unsigned pgrid = (unsigned)(per_cu * (num_cu > 0 ? num_cu : 1));
            grid = pgrid < grid ? pgrid : grid;
            if (grid < 8) grid = 8;  // one workgroup per XCD range at least
        }
        if (waves_out) *waves_out = (int)grid * CHEM_WAVES_PER_BLOCK;
        const int chunk = tu.chunk > 0 ? tu.chunk : 64;
        hipLaunchKernelGGL((k_chem_integrate<NSP, NRM, METHOD, TIGHT, DENSE>), dim3(grid), dim3(CHEM_BLOCK), 0, s, cd,
                           n, dt, rho, T, u, dtc_in, dtc_out, w, tu.persistent, chunk, tu.refill_period);
        return hipGetLastError();
    }
    template <int NSP, int NRM, bool DENSE>
    hipError_t run() {
        if (method == 1) return launch<NSP, NRM, 1, false, DENSE>();  // RKF with tight coupling is refused at set-up
        return tight ? launch<NSP, NRM, 0, true, DENSE>() : launch<NSP, NRM, 0, false, DENSE>();
    }
};

struct FinishL {
    const ChemDev* cd; int n; const double *rho, *T, *u, *p, *massf, *dtc_in; double *T_out, *u_out, *p_out,
        *massf_out, *dtc_out; int* status; ChemWork w; hipStream_t s;
    template <int NSP>
    hipError_t run() {
        hipLaunchKernelGGL((k_chem_finish<NSP>), dim3(grid_for(n)), dim3(CHEM_BLOCK), 0, s, cd, n, rho, T, u, p,
                           massf, dtc_in, T_out, u_out, p_out, massf_out, dtc_out, status, w);
        return hipGetLastError();
    }
};

struct TpgL {
    const ChemDev* cd; int mode, n; double *rho, *p, *T, *u, *a; const double* massf; int* status; hipStream_t s;
    template <int NSP>
    hipError_t run() {
        hipLaunchKernelGGL((k_tpg_update<NSP>), dim3(grid_for(n)), dim3(CHEM_BLOCK), 0, s, cd, mode, n, rho, p, T, u,
                           a, massf, status);
        return hipGetLastError();
    }
};

// One phase of a chemistry call for NSP species.  Mechanisms in dense form get the
// branch-free rate walk with the reaction bound 5 for the 5-species, 5-reaction air
// mechanism, otherwise 8 or 32; the others the list walk with the bound 32.
template <int NSP>
hipError_t chem_run(const ChemCall& c) {
    switch (c.phase) {
    case CHEM_PHASE_SETUP: {
        SetupL l{c.cd, c.n, c.dt, c.rho, c.T, c.massf, c.dtc_in, c.mask, c.tu.heavy_ratio, c.w, c.s};
        if (!c.dense) return l.template run<NSP, 32, false>();
        if (NSP == 5 && c.nr == 5) return l.template run<NSP, 5, true>();
        if (c.nr <= 8) return l.template run<NSP, 8, true>();
        return l.template run<NSP, 32, true>();
    }
    case CHEM_PHASE_INTEGRATE: {
        IntegrateL l{c.cd,    c.n,  c.dt,     c.rho,  c.T,      c.u,        c.dtc_in, c.dtc_out,
                     c.w,     c.tu, c.num_cu, c.method, c.tight, c.waves_out, c.s};
        if (!c.dense) return l.template run<NSP, 32, false>();
        if (NSP == 5 && c.nr == 5) return l.template run<NSP, 5, true>();
        if (c.nr <= 8) return l.template run<NSP, 8, true>();
        return l.template run<NSP, 32, true>();
    }
    case CHEM_PHASE_FINISH: {
        FinishL l{c.cd, c.n, c.rho, c.T, c.u, c.p, c.massf, c.dtc_in, c.T_out, c.u_out, c.p_out, c.massf_out,
                  c.dtc_out, c.status, c.w, c.s};
        return l.template run<NSP>();
    }
    case CHEM_PHASE_TPG: {
        TpgL l{c.cd, c.mode, c.n, c.rho_io, c.p_io, c.T_io, c.u_io, c.a_io, c.massf, c.status, c.s};
        return l.template run<NSP>();
    }
    default: return hipErrorInvalidValue;
    }
}

}  // namespace gdtk
