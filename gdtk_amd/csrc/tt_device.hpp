// tt_device.hpp -- two-temperature air on gfx950: the TwoTemperatureAir gas
// (src/gas/two_temperature_air.d, neutral species) and the TwoTemperatureAirKinetics reactor
// (src/kinetics/two_temperature_air_kinetics.d), one lane per cell.  FP64, no FMA
// contraction; every expression keeps the reference's operation order.  The mechanism is in
// dense form (ascending lists, as prep-chem writes them), with explicit backward rates.
#pragma once
#include <hip/hip_runtime.h>

#include "chem_device.hpp"

namespace gdtk {

constexpr double TT_T_REF = 298.15;                  // two_temperature_air.d:24
constexpr double TT_AVOGADRO = 6.02214e23;           // physical_constants.d:18
constexpr double TT_BOLTZMANN = 1.380658e-23;        // physical_constants.d:17
constexpr double TT_SMALL_MOLE_FRACTION = 1.0e-15;   // gas_model.d:42
constexpr double TT_PI = 3.141592653589793;          // to!double(PI)

// Per-species constants of the gas and the relaxation model, computed on the host as the
// reference's constructors do (two_temperature_air.d:186-205, kinetics :387-439).
struct TTDev {
    int molecular[CH_MAX_SP];
    double Cp_tr_rot[CH_MAX_SP], del_hf[CH_MAX_SP];
    double sA[CH_MAX_SP], sB[CH_MAX_SP];      // Cp_tr_rot T_REF - del_hf, Cp_tr_rot - R
    double A_MW[CH_MAX_SP];                   // Millikan-White A (0 for atoms)
    double mu_q[CH_MAX_SP][CH_MAX_SP];        // pow(mu_ij in g/mol, 0.25)
    double mpart[CH_MAX_SP];                  // M / N_A (kg)
    double abs_tol, rel_tol;                  // energy step
};

// GasState of a cell: T_modes[0] = Tv, u_modes[0] = uv.
template <int NSP>
struct TTCell {
    double rho, p, T, Tv, u, uv;
    double massf[NSP];
};

// two_temperature_air.d:573-583, one species at Tve with log(Tve) given
__device__ __forceinline__ double tt_ev(const ChemDev& cd, const TTDev& td, int s, double Tve, double logTve) {
    const double h_at_Tve = cea_h(cd.cea[s], Tve, logTve);
    return h_at_Tve - td.Cp_tr_rot[s] * (Tve - TT_T_REF) - td.del_hf[s];
}

// :584-591 and :673-680 at one temperature (one coefficient selection per species)
template <int NSP>
__device__ __forceinline__ void tt_ev_cv_mix(const ChemDev& cd, const TTDev& td, const double* massf, double Tve,
                                             double& e, double& cv) {
    const CeaT pw(Tve, log(Tve));
    e = 0.0;
    cv = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        double h, Cp;
        cea_h_Cp(cd.cea[s], pw, h, Cp);
        e += massf[s] * (h - td.Cp_tr_rot[s] * (Tve - TT_T_REF) - td.del_hf[s]);
        cv += massf[s] * (Cp - td.Cp_tr_rot[s]);
    }
}

// :698-740 vibElecTemperature: Newton from the state's Tv; false where the reference throws.
template <int NSP>
__device__ __forceinline__ bool tt_vib_temperature(const ChemDev& cd, const TTDev& td, const TTCell<NSP>& Q,
                                                   double& Tv_out) {
    const int MAX_ITERATIONS = 20;
    const double TOL = 1.0e-6, E_TOL = 0.01;
    double T_guess = Q.Tv;
    double e, Cv;
    tt_ev_cv_mix<NSP>(cd, td, Q.massf, T_guess, e, Cv);
    double f_guess = e - Q.uv;
    if (fabs(f_guess) < E_TOL) {
        Tv_out = Q.Tv;
        return true;
    }
    int count = 0;
    for (int iter = 0; iter < MAX_ITERATIONS; ++iter) {
        const double dT = -f_guess / Cv;
        T_guess += dT;
        if (fabs(dT) < TOL) break;
        tt_ev_cv_mix<NSP>(cd, td, Q.massf, T_guess, e, Cv);
        f_guess = e - Q.uv;
        ++count;
    }
    Tv_out = T_guess;
    return count != MAX_ITERATIONS;
}

// :275-291 update_thermo_from_rhou: T by direct inversion, Tv by Newton, p (perf_gas_mix_eos.d:43-49)
template <int NSP>
__device__ __forceinline__ bool tt_update_from_rhou(const ChemDev& cd, const TTDev& td, TTCell<NSP>& Q) {
    double sumA = 0.0, sumB = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        sumA += Q.massf[s] * td.sA[s];
        sumB += Q.massf[s] * td.sB[s];
    }
    Q.T = (Q.u + sumA) / sumB;
    double Tv;
    if (!tt_vib_temperature<NSP>(cd, td, Q, Tv)) return false;
    Q.Tv = Tv;
    double Rmix = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) Rmix += Q.massf[s] * cd.R[s];
    Q.p = Q.rho * Rmix * Q.T;
    return true;
}

// reaction_mechanism.d:50-63 with rate_constant.d:72-75 (Arrhenius at T) and :391-394
// (Park at T^s Tv^(1-s)); T clamped to the temperature limits, Tv not.
template <int NRM>
__device__ __forceinline__ void tt_rate_constants(const ChemDev& cd, double Tin, double Tv, double* kf, double* kb) {
    double T = Tin;
    if (T < cd.T_lower) T = cd.T_lower;
    if (T > cd.T_upper) T = cd.T_upper;
#pragma unroll
    for (int ir = 0; ir < NRM; ++ir) {
        if (ir >= cd.nr) {
            kf[ir] = 0.0;
            kb[ir] = 0.0;
            continue;
        }
        const ReacDev& r = cd.r[ir];
        double Tf = T, Tb = T;
        if (r.f_park) Tf = pow(T, r.f_s) * pow(Tv, 1.0 - r.f_s);
        if (r.b_park) Tb = pow(T, r.b_s) * pow(Tv, 1.0 - r.b_s);
        kf[ir] = r.fA * pow(Tf, r.fn) * exp(-r.fC / Tf);
        kb[ir] = r.bA * pow(Tb, r.bn) * exp(-r.bC / Tb);
    }
}

// two_temperature_air_kinetics.d:398-441 (relaxation time of molecular species isp, from the
// mole fractions and number densities of evalRate) and :444-460 + :521-543 (the exchange
// rate of u_v: Landau-Teller per molecular species plus the chemical source, whose reaction
// rates wf, wb are those of the chemistry step's last evaluation).
template <int NSP, int NRM>
__device__ __forceinline__ double tt_eval_rate(const ChemDev& cd, const TTDev& td, const TTCell<NSP>& Q,
                                               const double* wf, const double* wb) {
    double M_inv = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) M_inv += Q.massf[s] / cd.M[s];
    const double mixMolMass = 1.0 / M_inv;
    double molef[NSP], numden[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        molef[s] = Q.massf[s] * mixMolMass / cd.M[s];
        numden[s] = TT_AVOGADRO * Q.massf[s] * Q.rho / cd.M[s];
    }
    const double lgT = log(Q.T), lgTv = log(Q.Tv);
    const double Tm13 = pow(Q.T, -1. / 3);
    double evTv[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) evTv[s] = tt_ev(cd, td, s, Q.Tv, lgTv);
    double rate = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        if (!td.molecular[s]) continue;
        double totalND = 0.0, sum = 0.0;
#pragma unroll
        for (int c = 0; c < NSP; ++c) {
            if (molef[c] >= TT_SMALL_MOLE_FRACTION) {
                const double nd = numden[c];
                sum += nd * exp(td.A_MW[s] * (Tm13 - 0.015 * td.mu_q[s][c]) - 18.42);
                totalND += nd;
            }
        }
        double tauMW = sum / totalND;
        tauMW *= CH_P_ATM / Q.p;
        const double cBar_s = sqrt(8 * TT_BOLTZMANN * Q.T / (TT_PI * td.mpart[s]));
        const double sigma_s = 1e-20;
        const double tauP = 1.0 / (sigma_s * cBar_s * numden[s]);
        const double tau = tauMW + tauP;
        const double evStar = tt_ev(cd, td, s, Q.T, lgT);
        rate += Q.massf[s] * (evStar - evTv[s]) / tau;
    }
    // EnergyReactiveSourceTerm: per species, the net molar rate over the reactions it takes
    // part in (reaction order), times M and e_ve(Tv); then per unit mass of mixture.
    double src = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        double chemRate = 0.0;
#pragma unroll
        for (int ir = 0; ir < NRM; ++ir) {
            if (ir >= cd.nr) break;
            const RateRow& rr = cd.rr[ir];
            if ((rr.code[s] & 15) == 0) continue;  // not a participant
            const int nu = dense_row_nu(rr, s);
            const double prod = (nu > 0) ? nu * wf[ir] : (nu < 0) ? -nu * wb[ir] : 0.0;
            const double loss = (nu > 0) ? nu * wb[ir] : (nu < 0) ? -nu * wf[ir] : 0.0;
            chemRate += prod - loss;
        }
        chemRate *= cd.M[s];
        src += chemRate * evTv[s];
    }
    src /= Q.rho;
    rate += src;
    return rate;
}

// :546-706 EnergyStep: Cash-Karp on u_v at fixed total energy uTotal; every stage ends in
// update_thermo_from_rhou.  Returns 1 success, 0 failure (the state keeps the last stage,
// as in the reference), -1 where the Tv iteration fails.
template <int NSP, int NRM>
__device__ __forceinline__ int tt_energy_step(const ChemDev& cd, const TTDev& td, TTCell<NSP>& Q, double uTotal,
                                              double h, double& hSuggest, const double* wf, const double* wb) {
    const double a21 = 1. / 5., a31 = 3. / 40., a32 = 9. / 40., a41 = 3. / 10., a42 = -9. / 10., a43 = 6. / 5.,
                 a51 = -11. / 54., a52 = 5. / 2., a53 = -70. / 27., a54 = 35. / 27., a61 = 1631. / 55296.,
                 a62 = 175. / 512., a63 = 575. / 13824., a64 = 44275. / 110592., a65 = 253. / 4096.;
    const double b51 = 37. / 378., b53 = 250. / 621., b54 = 125. / 594., b56 = 512. / 1771., b41 = 2825. / 27648.,
                 b43 = 18575. / 48384., b44 = 13525. / 55296., b45 = 277. / 14336., b46 = 1.0 / 4.0;
    const double uv0 = Q.uv;
    const double k1 = tt_eval_rate<NSP, NRM>(cd, td, Q, wf, wb);
    Q.uv = uv0 + h * (a21 * k1);
    Q.u = uTotal - Q.uv;
    if (!tt_update_from_rhou<NSP>(cd, td, Q)) return -1;
    const double k2 = tt_eval_rate<NSP, NRM>(cd, td, Q, wf, wb);
    Q.uv = uv0 + h * (a31 * k1 + a32 * k2);
    Q.u = uTotal - Q.uv;
    if (!tt_update_from_rhou<NSP>(cd, td, Q)) return -1;
    const double k3 = tt_eval_rate<NSP, NRM>(cd, td, Q, wf, wb);
    Q.uv = uv0 + h * (a41 * k1 + a42 * k2 + a43 * k3);
    Q.u = uTotal - Q.uv;
    if (!tt_update_from_rhou<NSP>(cd, td, Q)) return -1;
    const double k4 = tt_eval_rate<NSP, NRM>(cd, td, Q, wf, wb);
    Q.uv = uv0 + h * (a51 * k1 + a52 * k2 + a53 * k3 + a54 * k4);
    Q.u = uTotal - Q.uv;
    if (!tt_update_from_rhou<NSP>(cd, td, Q)) return -1;
    const double k5 = tt_eval_rate<NSP, NRM>(cd, td, Q, wf, wb);
    Q.uv = uv0 + h * (a61 * k1 + a62 * k2 + a63 * k3 + a64 * k4 + a65 * k5);
    Q.u = uTotal - Q.uv;
    if (!tt_update_from_rhou<NSP>(cd, td, Q)) return -1;
    const double k6 = tt_eval_rate<NSP, NRM>(cd, td, Q, wf, wb);
    Q.uv = uv0 + h * (b51 * k1 + b53 * k3 + b54 * k4 + b56 * k6);
    Q.u = uTotal - Q.uv;
    if (!tt_update_from_rhou<NSP>(cd, td, Q)) return -1;
    const double errEst = Q.uv - (uv0 + h * (b41 * k1 + b43 * k3 + b44 * k4 + b45 * k5 + b46 * k6));
    const double sk = td.abs_tol + td.rel_tol * fmax(fabs(uv0), fabs(Q.uv));
    const double err = fabs(errEst / sk);
    const double maxscale = 10.0, minscale = 0.2, safe = 0.9, alpha = 0.2;
    double scale;
    if (err <= 1.0) {
        if (err == 0.0) {
            scale = maxscale;
        } else {
            scale = safe * pow(err, -alpha);
            if (scale < minscale) scale = minscale;
            if (scale > maxscale) scale = maxscale;
        }
        hSuggest = scale * h;
        return 1;
    }
    scale = fmax(safe * pow(err, -alpha), minscale);
    hSuggest = scale * h;
    return 0;
}

// gas_model.d:255-271 conc2massf into the cell state
template <int NSP>
__device__ __forceinline__ void tt_conc_to_massf(const ChemDev& cd, const double* conc, TTCell<NSP>& Q) {
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        Q.massf[s] = conc[s] * cd.M[s] / Q.rho;
        if (Q.massf[s] < CH_MIN_MASSF) Q.massf[s] = 0.0;
    }
}

struct TTCounts {
    int substeps, attempts;
};

// two_temperature_air_kinetics.d:162-358 opCall: 1 success, 0 where the reference throws
// ThermochemicalReactorUpdateException (the caller restores the input state), -1 where a
// GasModelException escapes (a failed Tv iteration, a bad density).
template <int NSP, int NRM>
__device__ __forceinline__ int tt_opcall(const ChemDev& cd, const TTDev& td, TTCell<NSP>& Q, double tInterval,
                                         double& dtSuggest, TTCounts& cnt) {
    const double DT_INCREASE_PERCENT = 10.0, DT_REDUCTION_FACTOR = 10.0, H_MIN = 1.0e-15,
                 ALLOWABLE_MASSF_ERROR = 1.0e-3;
    if (isnan(Q.rho) || !(Q.rho > 0.0)) return -1;  // massf2conc (gas_model.d:246-248)
    double conc0[NSP], concOut[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        conc0[s] = Q.massf[s] * Q.rho / cd.M[s];
        if (conc0[s] < CH_MIN_MOLES) conc0[s] = 0.0;
        concOut[s] = 0.0;
    }
    const double uTotal = Q.u + Q.uv;
    double kf[NRM], kb[NRM];
    tt_rate_constants<NRM>(cd, Q.T, Q.Tv, kf, kb);
    double t = 0.0, h, dtSave = 0.0;
    if (dtSuggest > tInterval) h = tInterval;
    else if (dtSuggest <= 0.0) h = estimate_step_size<NSP, NRM, true>(cd, kf, kb, conc0);
    else h = dtSuggest;
    int cycle = 0;
    for (; cycle < cd.maxSubcycles; ++cycle) {
        dtSave = dtSuggest;
        h = ((tInterval - t) < h) ? (tInterval - t) : h;
        int attempt = 0;
        for (; attempt < cd.maxAttempts; ++attempt) {
            ++cnt.attempts;
            double yLast[NSP];
            const bool r1 = rkf_step<NSP, NRM, true>(cd, kf, kb, conc0, h, concOut, dtSuggest, yLast);
            // the reaction rates of the chemistry step's last evaluation
            double wf[NRM], wb[NRM], c2[NSP];
#pragma unroll
            for (int s = 0; s < NSP; ++s) c2[s] = yLast[s] * yLast[s];
#pragma unroll
            for (int ir = 0; ir < NRM; ++ir) {
                wf[ir] = 0.0;
                wb[ir] = 0.0;
                if (ir < cd.nr) dense_row_rates<NSP>(cd.rr[ir], kf[ir], kb[ir], yLast, c2, wf[ir], wb[ir]);
            }
            const int r2 = tt_energy_step<NSP, NRM>(cd, td, Q, uTotal, h, dtSuggest, wf, wb);
            if (r2 < 0) return -1;
            const bool ok = r1 && (r2 == 1);
            bool passes = true;
            if (ok) {
                tt_conc_to_massf<NSP>(cd, concOut, Q);
                const double massfTotal = phobos_sum<NSP>(Q.massf);
                if (fabs(massfTotal - 1.0) > ALLOWABLE_MASSF_ERROR) passes = false;
            }
            if (ok && passes) {
                t += h;
#pragma unroll
                for (int s = 0; s < NSP; ++s) conc0[s] = concOut[s];
                const double hMax = h * (1.0 + DT_INCREASE_PERCENT / 100.0);
                if (dtSuggest > h) h = (dtSuggest < hMax) ? dtSuggest : hMax;
                break;
            }
            h /= DT_REDUCTION_FACTOR;
            if (h < H_MIN) return 0;
        }
        if (attempt == cd.maxAttempts) return 0;
        ++cnt.substeps;
        if (cd.tight) {
            tt_conc_to_massf<NSP>(cd, conc0, Q);
            Q.u = uTotal - Q.uv;
            if (!tt_update_from_rhou<NSP>(cd, td, Q)) return -1;
            tt_rate_constants<NRM>(cd, Q.T, Q.Tv, kf, kb);
        }
        if (t >= tInterval) {
            if (cycle == 0) dtSave = dtSuggest;
            break;
        }
    }
    if (cycle == cd.maxSubcycles) return 0;
    tt_conc_to_massf<NSP>(cd, concOut, Q);
    const double massfTotal = phobos_sum<NSP>(Q.massf);
#pragma unroll
    for (int s = 0; s < NSP; ++s) Q.massf[s] /= massfTotal;
    dtSuggest = dtSave;
    return 1;
}

}  // namespace gdtk
