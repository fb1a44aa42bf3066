// chem_device.hpp -- device-side thermally perfect gas and finite-rate chemistry (gfx950).
//
// One lane advances one cell.  The gas and mechanism tables (ChemDev) live in a read-only
// device buffer; every lane of a wave reads the same entries, so those loads are uniform
// (scalar).  Per-cell state is held in registers: arrays are sized by the template
// parameters NSP (species, exact) and NRM (an upper bound on the reactions, the actual
// count cd->nr is uniform at run time).  FP64 throughout, built without FMA contraction;
// the operation order of every expression follows the reference function cited beside it.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace gdtk {

constexpr int CH_MAX_SP = 8;
constexpr int CH_MAX_R = 32;

// src/gas/thermo/cea_thermo_curves.d:24-176; T_low..s_high are the constructor's
// end-point values (:45-54).
struct CEADev {
    int nseg;
    double T_breaks[4], T_blends[2], coeffs[3][9];
    double R, T_low, T_high, Cp_low, Cp_high, h_low, h_high, s_low, s_high;
};

// src/kinetics/reaction.d:58-457.  The reactant, product and efficiency lists keep the
// reference's order (reaction.d:213, 338, 483-491); nu[s] is the net stoichiometric
// coefficient of species s and part[s] marks the participants (reaction.d:100-143).
// Device code never indexes a register array with a run-time species index (that would
// put the per-cell arrays in scratch memory): it walks the species in compile-time order,
// or picks an entry through a select chain for lists that are not in ascending order.
struct ReacDev {
    int collider;
    int nreac, reac_sp[3], reac_nu[3];
    int nprod, prod_sp[3], prod_nu[3];
    int neff, eff_sp[CH_MAX_SP];
    double eff[CH_MAX_SP];
    // Dense form, used when every list is in ascending species order (then a walk over the
    // species in index order visits the entries in list order): coefficient per species
    // (0 = absent) and efficiency per species with a presence mask.
    int sorted;
    int reac_d[CH_MAX_SP], prod_d[CH_MAX_SP];
    unsigned eff_set;
    double eff_d[CH_MAX_SP];
    int f_from_eq, b_from_eq;
    double fA, fn, fC, bA, bn, bC;
    int f_park, b_park;  // Park2TRateConstant: the rate at T^s Tv^(1-s) (two-temperature reactor)
    double f_s, b_s;
    int nu[CH_MAX_SP];
    int part[CH_MAX_SP];
};

// One reaction of the branch-free rate evaluation, for a mechanism whose lists are all in
// ascending species order: per species the third-body efficiency (0 where the species is
// not in the list, or the reaction has no third body) and a code word holding the reactant
// exponent (bits 0-1, 0 = absent), the product exponent (bits 2-3), the net coefficient
// as a 4-bit two's complement number (bits 4-7) and the third-body list membership (bit
// 8).  flags bit 0: third-body reaction.  Rows
// past the reaction count are all zero: every species' net coefficient is 0 there, so they
// add nothing.
struct RateRow {
    double eff[CH_MAX_SP];
    int code[CH_MAX_SP];
    int flags;
    int pad_[3];
};

struct ChemDev {
    int nsp, nr;
    double M[CH_MAX_SP], R[CH_MAX_SP];
    CEADev cea[CH_MAX_SP];
    ReacDev r[CH_MAX_R];
    int all_sorted;  // every reaction in dense form: rr[] holds the mechanism
    RateRow rr[CH_MAX_R];
    double T_lower, T_upper;
    int method;  // 0 alpha-QSS, 1 RKF
    double eps1, eps2, delta;
    int maxIters;
    double errTol;
    int maxSubcycles, maxAttempts;
    double T_frozen;
    int tight;  // tightTempCoupling: T(u) and the rate constants after every subcycle
};

constexpr double CH_R_UNIVERSAL = 8.31451;               // gas/physical_constants.d
constexpr double CH_P_ATM = 101.325e3;
constexpr double CH_T_MIN = 10.0, CH_T_MAX = 100000.0;   // gas/gas_model.d:45-46
constexpr double CH_MIN_MOLES = 1.0e-30;                 // gas/gas_model.d:44
constexpr double CH_MIN_MASSF = 1.0e-30;                 // gas/gas_model.d:43

// Element k of a register array for a wave-uniform index k: a select chain over
// compile-time indices, so the array stays in registers.
template <int N>
__device__ __forceinline__ double pick(const double* a, int k) {
    double r = a[0];
#pragma unroll
    for (int s = 1; s < N; ++s) r = (k == s) ? a[s] : r;
    return r;
}

// cea_thermo_curves.d:137-176.  The reference throws where no segment matches (only a NaN
// temperature gets there with contiguous breaks); NaN coefficients stand in for that here.
__host__ __device__ inline void cea_coeffs(const CEADev& c, double T, double a[9]) {
    const int nb = c.nseg + 1;
    if (T < (c.T_breaks[1] - 0.5 * c.T_blends[0])) {
        for (int j = 0; j < 9; ++j) a[j] = c.coeffs[0][j];
        return;
    }
    if (T > (c.T_breaks[nb - 2] + 0.5 * c.T_blends[c.nseg - 2])) {
        for (int j = 0; j < 9; ++j) a[j] = c.coeffs[c.nseg - 1][j];
        return;
    }
    for (int i = 1; i < nb - 1; ++i) {
        double T_blend_low = c.T_breaks[i] - 0.5 * c.T_blends[i - 1];
        double T_blend_high = c.T_breaks[i] + 0.5 * c.T_blends[i - 1];
        if (T >= T_blend_low && T <= T_blend_high) {
            double wB = (1. / c.T_blends[i - 1]) * (T - T_blend_low);
            double wA = 1.0 - wB;
            for (int j = 0; j < 9; ++j) a[j] = wA * c.coeffs[i - 1][j] + wB * c.coeffs[i][j];
            return;
        }
        if (i < c.nseg - 1 && T > T_blend_high && T < (c.T_breaks[i + 1] - 0.5 * c.T_blends[i])) {
            for (int j = 0; j < 9; ++j) a[j] = c.coeffs[i][j];
            return;
        }
    }
    for (int j = 0; j < 9; ++j) a[j] = __builtin_nan("");
}

// cea_thermo_curves.d:56-70
__host__ __device__ inline double cea_Cp(const CEADev& c, double T) {
    if (T < c.T_low) return c.Cp_low;
    if (T > c.T_high) return c.Cp_high;
    double a[9];
    cea_coeffs(c, T, a);
    double Cp_on_R = a[0] / (T * T) + a[1] / T + a[2] + a[3] * T;
    Cp_on_R += a[4] * T * T + a[5] * T * T * T + a[6] * T * T * T * T;
    return c.R * Cp_on_R;
}

// cea_thermo_curves.d:71-102, with log(T) given (it is the same for every species)
__host__ __device__ inline double cea_h(const CEADev& c, double T, double logT) {
    if (T < c.T_low) return c.h_low - c.Cp_low * (c.T_low - T);
    if (T > c.T_high) return c.h_high + c.Cp_high * (T - c.T_high);
    double a[9];
    cea_coeffs(c, T, a);
    double h_on_RT = -a[0] / T + a[1] * logT + a[2] * T + a[3] * T * T / 2.0;
    h_on_RT += a[4] * T * T * T / 3.0 + a[5] * T * T * T * T / 4.0 + a[6] * T * T * T * T * T / 5.0 + a[7];
    return c.R * h_on_RT;
}
__host__ __device__ inline double cea_h(const CEADev& c, double T) { return cea_h(c, T, log(T)); }

// cea_h and cea_Cp at one temperature from one coefficient selection (the same values as
// the two separate calls).
__host__ __device__ inline void cea_h_Cp(const CEADev& c, double T, double logT, double& h, double& Cp) {
    if (T < c.T_low) {
        h = c.h_low - c.Cp_low * (c.T_low - T);
        Cp = c.Cp_low;
        return;
    }
    if (T > c.T_high) {
        h = c.h_high + c.Cp_high * (T - c.T_high);
        Cp = c.Cp_high;
        return;
    }
    double a[9];
    cea_coeffs(c, T, a);
    double Cp_on_R = a[0] / (T * T) + a[1] / T + a[2] + a[3] * T;
    Cp_on_R += a[4] * T * T + a[5] * T * T * T + a[6] * T * T * T * T;
    Cp = c.R * Cp_on_R;
    double h_on_RT = -a[0] / T + a[1] * logT + a[2] * T + a[3] * T * T / 2.0;
    h_on_RT += a[4] * T * T * T / 3.0 + a[5] * T * T * T * T / 4.0 + a[6] * T * T * T * T * T / 5.0 + a[7];
    h = c.R * h_on_RT;
}

// a / b correctly rounded from y = RN(1 / b): q = RN(a y) lies within an ulp of a / b, the
// remainder a - b q is exact in one fma, and RN(q + r y) is the correctly rounded quotient
// (Markstein's theorem; P. Markstein, IA-64 and Elementary Functions, 2000).  Bit for bit
// a / b for the normal operands met here; a true division costs ~10 instructions, this 3,
// and a divisor shared by several quotients needs its reciprocal once.
__host__ __device__ inline double div_by(double a, double b, double y) {
    const double q = a * y;
    const double r = fma(-b, q, a);
    return fma(r, y, q);
}

// The powers of T that every species' curve divides by, with their reciprocals, once per
// temperature.
struct CeaT {
    double T, TT, rT, rTT, logT;
    __host__ __device__ explicit CeaT(double t, double lg) : T(t), TT(t * t), rT(1.0 / t), rTT(1.0 / (t * t)), logT(lg) {}
};

// cea_h_Cp with the divisions of cea_thermo_curves.d:56-102 taken through div_by: the same
// values bit for bit.
__host__ __device__ inline void cea_h_Cp(const CEADev& c, const CeaT& p, double& h, double& Cp) {
    const double T = p.T;
    if (T < c.T_low) {
        h = c.h_low - c.Cp_low * (c.T_low - T);
        Cp = c.Cp_low;
        return;
    }
    if (T > c.T_high) {
        h = c.h_high + c.Cp_high * (T - c.T_high);
        Cp = c.Cp_high;
        return;
    }
    double a[9];
    cea_coeffs(c, T, a);
    double Cp_on_R = div_by(a[0], p.TT, p.rTT) + div_by(a[1], T, p.rT) + a[2] + a[3] * T;
    Cp_on_R += a[4] * T * T + a[5] * T * T * T + a[6] * T * T * T * T;
    Cp = c.R * Cp_on_R;
    double h_on_RT = div_by(-a[0], T, p.rT) + a[1] * p.logT + a[2] * T + a[3] * T * T / 2.0;
    h_on_RT += div_by(a[4] * T * T * T, 3.0, 1.0 / 3.0) + a[5] * T * T * T * T / 4.0 +
               div_by(a[6] * T * T * T * T * T, 5.0, 1.0 / 5.0) + a[7];
    h = c.R * h_on_RT;
}
__host__ __device__ inline double cea_h(const CEADev& c, const CeaT& p) {
    const double T = p.T;
    if (T < c.T_low) return c.h_low - c.Cp_low * (c.T_low - T);
    if (T > c.T_high) return c.h_high + c.Cp_high * (T - c.T_high);
    double a[9];
    cea_coeffs(c, T, a);
    double h_on_RT = div_by(-a[0], T, p.rT) + a[1] * p.logT + a[2] * T + a[3] * T * T / 2.0;
    h_on_RT += div_by(a[4] * T * T * T, 3.0, 1.0 / 3.0) + a[5] * T * T * T * T / 4.0 +
               div_by(a[6] * T * T * T * T * T, 5.0, 1.0 / 5.0) + a[7];
    return c.R * h_on_RT;
}
__host__ __device__ inline double cea_Cp(const CEADev& c, const CeaT& p) {
    const double T = p.T;
    if (T < c.T_low) return c.Cp_low;
    if (T > c.T_high) return c.Cp_high;
    double a[9];
    cea_coeffs(c, T, a);
    double Cp_on_R = div_by(a[0], p.TT, p.rTT) + div_by(a[1], T, p.rT) + a[2] + a[3] * T;
    Cp_on_R += a[4] * T * T + a[5] * T * T * T + a[6] * T * T * T * T;
    return c.R * Cp_on_R;
}

// cea_h and cea_s (cea_thermo_curves.d:71-130) at one temperature from one coefficient
// selection, log(T) and the reciprocals of CeaT: the same values bit for bit.  The
// divisor 2 T T of the entropy is exactly 2 TT, whose reciprocal is exactly rTT / 2.
__host__ __device__ inline void cea_h_s(const CEADev& c, const CeaT& p, double& h, double& sv) {
    const double T = p.T;
    if (T < c.T_low) {
        h = c.h_low - c.Cp_low * (c.T_low - T);
        sv = c.s_low - c.Cp_low * log(c.T_low / T);
        return;
    }
    if (T > c.T_high) {
        h = c.h_high + c.Cp_high * (T - c.T_high);
        sv = c.s_high + c.Cp_high * log(T / c.T_high);
        return;
    }
    double a[9];
    cea_coeffs(c, T, a);
    double h_on_RT = div_by(-a[0], T, p.rT) + a[1] * p.logT + a[2] * T + a[3] * T * T / 2.0;
    h_on_RT += div_by(a[4] * T * T * T, 3.0, 1.0 / 3.0) + a[5] * T * T * T * T / 4.0 +
               div_by(a[6] * T * T * T * T * T, 5.0, 1.0 / 5.0) + a[7];
    h = c.R * h_on_RT;
    double s_on_R = div_by(-a[0], 2.0 * p.TT, 0.5 * p.rTT) - div_by(a[1], T, p.rT) + a[2] * p.logT + a[3] * T;
    s_on_R += a[4] * T * T / 2.0 + div_by(a[5] * T * T * T, 3.0, 1.0 / 3.0) + a[6] * T * T * T * T / 4.0 + a[8];
    sv = c.R * s_on_R;
}

// cea_thermo_curves.d:103-130
__host__ __device__ inline double cea_s(const CEADev& c, double T) {
    if (T < c.T_low) return c.s_low - c.Cp_low * log(c.T_low / T);
    if (T > c.T_high) return c.s_high + c.Cp_high * log(T / c.T_high);
    double a[9];
    cea_coeffs(c, T, a);
    double logT = log(T);
    double s_on_R = -a[0] / (2.0 * T * T) - a[1] / T + a[2] * logT + a[3] * T;
    s_on_R += a[4] * T * T / 2.0 + a[5] * T * T * T / 3.0 + a[6] * T * T * T * T / 4.0 + a[8];
    return c.R * s_on_R;
}

// therm_perf_gas_mix_eos.d:61-67 with gas_model.d:410-419 (mass_average)
// SERIAL: a scheduling barrier after each species, so that the species' evaluations do
// not interleave (fewer registers live at once, for kernels whose occupancy they limit).
template <int NSP, bool SERIAL = false>
__device__ __forceinline__ double tpg_energy(const ChemDev& cd, const double* massf, double T) {
    const CeaT pw(T, log(T));
    double u = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        u += massf[i] * (cea_h(cd.cea[i], pw) - cd.R[i] * T);
        if (SERIAL) __builtin_amdgcn_sched_barrier(0);
    }
    return u;
}
// tpg_energy and tpg_Cv at one temperature (the Newton iteration of T(u) evaluates both at
// each iterate), sharing each species' coefficient selection and log(T).
template <int NSP, bool SERIAL = false>
__device__ __forceinline__ void tpg_energy_Cv(const ChemDev& cd, const double* massf, double T, double& u,
                                              double& Cv) {
    const CeaT pw(T, log(T));
    double uu = 0.0, v = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        double h, Cp;
        cea_h_Cp(cd.cea[i], pw, h, Cp);
        uu += massf[i] * (h - cd.R[i] * T);
        v += massf[i] * (Cp - cd.R[i]);
        if (SERIAL) __builtin_amdgcn_sched_barrier(0);
    }
    u = uu;
    Cv = v;
}
template <int NSP>
__device__ __forceinline__ double tpg_Cv(const ChemDev& cd, const double* massf, double T) {
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) v += massf[i] * (cea_Cp(cd.cea[i], T) - cd.R[i]);
    return v;
}
template <int NSP>
__device__ __forceinline__ double tpg_Cp(const ChemDev& cd, const double* massf, double T) {
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) v += massf[i] * cea_Cp(cd.cea[i], T);
    return v;
}
// perf_gas_mix_eos.d:43-76
template <int NSP>
__device__ __forceinline__ double tpg_Rmix(const ChemDev& cd, const double* massf) {
    double R = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) R += massf[i] * cd.R[i];
    return R;
}

// therm_perf_gas_mix_eos.d:68-151 with the safeguarded Newton of nm/newton.d:71-140.
// T and u are in/out as in the reference: every evaluation of the zero function writes T
// and u, and a converged root then overwrites T, so u ends as the energy of the last
// evaluated temperature.  Returns false where the reference throws; T is then restored
// and u = energy(T), which is the state the reference leaves behind.
//
// The reference evaluates u at both ends of the bracket [T1, T2] = [T - 500, T + 500] only
// for the signs of f = e_tgt - u there (and exact zeros), then starts the Newton iteration
// with an evaluation at T.  Here the evaluation at T comes first, and with it the signs are
// often fixed without the two end evaluations: u(T) rises at Cv >= 1.5 R_mix for any gas
// (translational energy alone; the CEA curves, their blends and their constant-Cp
// extensions all keep Cp >= 2.5 R per species), so with g = e_tgt - u(T),
//   f(T1) >= g + 1.4 R_mix (T - T1)   and   f(T2) <= g - 1.4 R_mix (T2 - T).
// When the first bound is above and the second below a rounding margin (1e-6 of the
// energies + 1 J/kg, against ~1e5 J/kg of slope over 500 K), f(T1) > 0 > f(T2) as the
// reference would compute them: its pass 0 takes the bracket (xH, xL) = (T1, T2) and the
// same iteration from T, so T and u come out bit for bit the same.  Otherwise (a root far
// from T, a temperature near the 10 K floor, a negative mass fraction) the ends are
// evaluated and the reference's passes run as written.  SERIAL as for tpg_energy.
template <int NSP, bool SERIAL = false>
__device__ __forceinline__ bool tpg_update_temperature(const ChemDev& cd, const double* massf, double& T, double& u) {
    const double Tsave = T;
    const double TOL = 1.0e-6;
    const double e_tgt = u;
    const double delT = 1000.0;
    const double T1 = fmax(T - 0.5 * delT, CH_T_MIN);
    const double T2 = T1 + delT;
    double u0, Cv0;
    tpg_energy_Cv<NSP, SERIAL>(cd, massf, Tsave, u0, Cv0);
    bool signs_known;
    {
        double Rm = 0.0;
        bool nonneg = true;
#pragma unroll
        for (int i = 0; i < NSP; ++i) {
            Rm += massf[i] * cd.R[i];
            nonneg = nonneg && (massf[i] >= 0.0);
        }
        const double g = e_tgt - u0;
        const double slope = 1.4 * Rm;
        const double margin = 1.0e-6 * (fabs(e_tgt) + fabs(u0)) + 1.0;
        signs_known = nonneg && (Tsave - T1 >= 1.0) && (T2 - Tsave >= 1.0) &&
                      (g + slope * (Tsave - T1) > margin) && (g - slope * (T2 - Tsave) < -margin);
    }
    for (int pass = 0; pass < 2; ++pass) {
        const double xMin = (pass == 0) ? T1 : CH_T_MIN;
        const double xMax = (pass == 0) ? T2 : CH_T_MAX;
        double xL, xH;
        if (pass == 0 && signs_known) {
            xH = xMin;
            xL = xMax;
        } else {
            xL = xMin;
            xH = xMax;
            T = xL;
            u = tpg_energy<NSP, SERIAL>(cd, massf, T);
            const double fL = e_tgt - u;
            T = xH;
            u = tpg_energy<NSP, SERIAL>(cd, massf, T);
            const double fH = e_tgt - u;
            if ((fL > 0.0 && fH > 0.0) || (fL < 0.0 && fH < 0.0)) continue;
            if (fL == 0.0) { T = xMin; return true; }
            if (fH == 0.0) { T = xMax; return true; }
            if (fL < 0.0) { xL = xMin; xH = xMax; }
            else { xH = xMin; xL = xMax; }
        }
        double rts = Tsave;
        double dxold = (xMax - xMin);
        double dx = dxold;
        T = rts;
        double Cv;
        if (pass == 0 && signs_known) {
            Cv = Cv0;
            u = u0;
        } else {
            // the same evaluation again (fewer registers live across the end evaluations)
            tpg_energy_Cv<NSP, SERIAL>(cd, massf, T, u, Cv);
        }
        double f0 = e_tgt - u;
        double df0 = -1.0 * Cv;
        for (int j = 0; j < 30; ++j) {
            if ((((rts - xH) * df0 - f0) * ((rts - xL) * df0 - f0) > 0.0) || (fabs(2.0 * f0) > fabs(dxold * df0))) {
                dxold = dx;
                dx = 0.5 * (xH - xL);
                rts = xL + dx;
                if (xL == rts) { T = rts; return true; }
            } else {
                dxold = dx;
                dx = f0 / df0;
                double tmp = rts;
                rts -= dx;
                if (tmp == rts) { T = rts; return true; }
            }
            if (fabs(dx) < TOL) { T = rts; return true; }
            T = rts;
            tpg_energy_Cv<NSP, SERIAL>(cd, massf, T, u, Cv);
            f0 = e_tgt - u;
            df0 = -1.0 * Cv;
            if (f0 < 0.0) xL = rts;
            else xH = rts;
        }
    }
    T = Tsave;
    u = tpg_energy<NSP, SERIAL>(cd, massf, T);
    return false;
}

// therm_perf_gas.d:394-404: frozen sound speed sqrt(gamma * (R T)), gamma = Cp/Cv, with
// tpg_Cp and tpg_Cv summed from one curve evaluation per species (the same sums).
template <int NSP>
__device__ __forceinline__ double tpg_sound_speed(const ChemDev& cd, const double* massf, double T) {
    const CeaT pw(T, 0.0);
    double Cp = 0.0, Cv = 0.0;
#pragma unroll
    for (int i = 0; i < NSP; ++i) {
        const double c = cea_Cp(cd.cea[i], pw);
        Cp += massf[i] * c;
        Cv += massf[i] * (c - cd.R[i]);
    }
    double gamma = Cp / Cv;
    double R = tpg_Rmix<NSP>(cd, massf);
    return sqrt(gamma * (R * T));
}

// c^n for the stoichiometric coefficients 1..3 of reaction.d's rate products (the host
// rejects other coefficients).
__device__ __forceinline__ double ch_ipow(double c, int n) {
    return (n == 1) ? c : (n == 2) ? c * c : c * (c * c);
}
// pow(P_atm/(R_u T), nuSum) of reaction.d:43-44 for whole nuSum.
__device__ __forceinline__ double ch_pow_nusum(double x, double nuSum) {
    return (nuSum == 0.0) ? 1.0 : (nuSum == 1.0) ? x : (nuSum == -1.0) ? 1.0 / x : (nuSum == 2.0) ? x * x : pow(x, nuSum);
}

// reaction_mechanism.d:50-63 with reaction.d:27-52 (equilibrium constant), 94-143 and
// rate_constant.d (Arrhenius A T^n exp(-C/T)); T clamped to the mechanism's tempLimits.
// Participants are summed in ascending species order, the order of the reference's
// sorted participant list.
template <int NSP, int NRM>
__device__ __forceinline__ void eval_rate_constants(const ChemDev& cd, double Tin, double* kf, double* kb) {
    double T = Tin;
    if (T < cd.T_lower) T = cd.T_lower;
    if (T > cd.T_upper) T = cd.T_upper;
    double g[NSP];
    const CeaT pw(T, log(T));
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        double h, sv;
        cea_h_s(cd.cea[s], pw, h, sv);
        g[s] = h - T * sv;
    }
#pragma unroll
    for (int ir = 0; ir < NRM; ++ir) {
        if (ir >= cd.nr) {
            kf[ir] = 0.0;
            kb[ir] = 0.0;
            continue;
        }
        const ReacDev& r = cd.r[ir];
        double Keq = 0.0;
        if (r.f_from_eq || r.b_from_eq) {
            double dG = 0.0, nuSum = 0.0;
#pragma unroll
            for (int s = 0; s < NSP; ++s) {
                if (!r.part[s]) continue;
                dG += r.nu[s] * g[s] * cd.M[s];
                nuSum += r.nu[s];
            }
            double K_p = exp(-dG / (CH_R_UNIVERSAL * T));
            Keq = K_p * ch_pow_nusum(CH_P_ATM / (CH_R_UNIVERSAL * T), nuSum);
        }
        double kfv, kbv;
        if (!r.f_from_eq) {
            kfv = r.fA * pow(T, r.fn) * exp(div_by(-r.fC, T, pw.rT));
            kbv = r.b_from_eq ? kfv / Keq : r.bA * pow(T, r.bn) * exp(div_by(-r.bC, T, pw.rT));
        } else {
            kbv = r.bA * pow(T, r.bn) * exp(div_by(-r.bC, T, pw.rT));
            kfv = kbv * Keq;
        }
        kf[ir] = kfv;
        kb[ir] = kbv;
    }
}

// The forward and backward rates of one reaction row (dense form) at conc, with c2 = conc^2:
// reaction.d:289-330 (eval_forward_rate / eval_backward_rate) for lists in ascending order.
template <int NSP>
__device__ __forceinline__ void dense_row_rates(const RateRow& rr, double kf, double kb, const double* conc,
                                                const double* c2, double& wf_out, double& wb_out) {
    int code[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) code[s] = rr.code[s];
    double collider = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) collider += (code[s] & 256) ? rr.eff[s] * conc[s] : 0.0;
    double wf = (rr.flags & 1) ? kf * collider : kf;
    double wb = (rr.flags & 1) ? kb * collider : kb;
    // c^e for e in 0..3 as (c or 1) * (c^2 or 1): c * (c * c) for e = 3, as ch_ipow
#pragma unroll
    for (int s = 0; s < NSP; ++s) wf *= ((code[s] & 1) ? conc[s] : 1.0) * ((code[s] & 2) ? c2[s] : 1.0);
#pragma unroll
    for (int s = 0; s < NSP; ++s) wb *= ((code[s] & 4) ? conc[s] : 1.0) * ((code[s] & 8) ? c2[s] : 1.0);
    wf_out = wf;
    wb_out = wb;
}

// Net coefficient of species s in a dense row (4-bit two's complement in bits 4-7).
__device__ __forceinline__ int dense_row_nu(const RateRow& rr, int s) { return (int)((unsigned)rr.code[s] << 24) >> 28; }

// reaction_mechanism.d:83-96 (eval_split_rates) with reaction.d's forward/backward rates
// (collider sum over the efficiency list, then the concentration products in list order)
// and production/loss per participant in reaction order.
//
// DENSE (every list of the mechanism in ascending species order, rr[] filled; the host
// picks the instantiation) takes a branch-free walk over the rows: an absent reactant or
// product multiplies by exactly 1, a species outside the third-body list or the reaction
// adds exactly 0 (a selected zero, not a product with zero), and a reaction without third
// body takes kf as it is, so every result equals the list walk bit for bit while the table
// words load in bulk instead of one load and branch per entry.
template <int NSP, int NRM, bool DENSE>
__device__ __forceinline__ void eval_split_rates(const ChemDev& cd, const double* kf, const double* kb,
                                                 const double* conc, double* q, double* L) {
#pragma unroll
    for (int s = 0; s < NSP; ++s) { q[s] = 0.0; L[s] = 0.0; }
    if constexpr (DENSE) {
        double c2[NSP];
#pragma unroll
        for (int s = 0; s < NSP; ++s) c2[s] = conc[s] * conc[s];
        const int nr = cd.nr;
#pragma unroll
        for (int ir = 0; ir < NRM; ++ir) {
            if (ir >= nr) break;
            const RateRow& rr = cd.rr[ir];
            double wf, wb;
            dense_row_rates<NSP>(rr, kf[ir], kb[ir], conc, c2, wf, wb);
#pragma unroll
            for (int s = 0; s < NSP; ++s) {
                const int nu = dense_row_nu(rr, s);
                const double m = (double)(nu < 0 ? -nu : nu);
                const bool pos = nu > 0;
                q[s] += (nu != 0) ? m * (pos ? wf : wb) : 0.0;
                L[s] += (nu != 0) ? m * (pos ? wb : wf) : 0.0;
            }
        }
    } else {
#pragma unroll
        for (int ir = 0; ir < NRM; ++ir) {
            if (ir >= cd.nr) continue;
            const ReacDev& r = cd.r[ir];
            double collider = 0.0, wf, wb;
            if (r.sorted) {
                if (r.collider) {
#pragma unroll
                    for (int s = 0; s < NSP; ++s)
                        if ((r.eff_set >> s) & 1u) collider += r.eff_d[s] * conc[s];
                }
                wf = r.collider ? kf[ir] * collider : kf[ir];
#pragma unroll
                for (int s = 0; s < NSP; ++s)
                    if (r.reac_d[s]) wf *= ch_ipow(conc[s], r.reac_d[s]);
                wb = r.collider ? kb[ir] * collider : kb[ir];
#pragma unroll
                for (int s = 0; s < NSP; ++s)
                    if (r.prod_d[s]) wb *= ch_ipow(conc[s], r.prod_d[s]);
            } else {
                if (r.collider) {
#pragma unroll
                    for (int e = 0; e < NSP; ++e)
                        if (e < r.neff) collider += r.eff[e] * pick<NSP>(conc, r.eff_sp[e]);
                }
                wf = r.collider ? kf[ir] * collider : kf[ir];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if (k < r.nreac) wf *= ch_ipow(pick<NSP>(conc, r.reac_sp[k]), r.reac_nu[k]);
                wb = r.collider ? kb[ir] * collider : kb[ir];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if (k < r.nprod) wb *= ch_ipow(pick<NSP>(conc, r.prod_sp[k]), r.prod_nu[k]);
            }
#pragma unroll
            for (int s = 0; s < NSP; ++s) {
                const int nu = r.nu[s];
                if (nu > 0) {
                    q[s] += nu * wf;
                    L[s] += nu * wb;
                } else if (nu < 0) {
                    q[s] += -nu * wb;
                    L[s] += -nu * wf;
                }
            }
        }
    }
}

// reaction_mechanism.d:162-203
template <int NSP, int NRM, bool DENSE>
__device__ __forceinline__ double estimate_step_size(const ChemDev& cd, const double* kf, const double* kb, const double* conc) {
    const double MUCH_LARGER_FACTOR = 10000.0, ZERO_EPS = 1.0e-50, ZERO_TOL = 1.0e-30, EPS1 = 1.0e-3;
    double q[NSP], L[NSP];
    eval_split_rates<NSP, NRM, DENSE>(cd, kf, kb, conc, q, L);
    double min_dt = 1.0e6, dt;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        double c = conc[s];
        if (c > 0.0 && (fabs(q[s] - L[s]) > ZERO_TOL)) {
            if (q[s] > MUCH_LARGER_FACTOR * L[s]) dt = c / (L[s] + ZERO_EPS);
            else dt = fabs(c / (q[s] - L[s]));
            min_dt = (dt < min_dt) ? dt : min_dt;
        }
    }
    double dt_chem = EPS1 * min_dt;
    dt_chem = (1.0e-10 < dt_chem) ? 1.0e-10 : dt_chem;
    dt_chem = (dt_chem < 1.0e-15) ? 1.0e-15 : dt_chem;
    return dt_chem;
}

// alpha(r) of chemistry_update.d:599-604 for r = 1/(p h + eps)
__device__ __forceinline__ double qss_alpha(double p, double h) {
    const double ZERO_EPS = 1.0e-50;
    double r = 1.0 / (p * h + ZERO_EPS);
    return (180.0 * r * r * r + 60.0 * r * r + 11.0 * r + 1.0) / (360.0 * r * r * r + 60.0 * r * r + 12.0 * r + 1.0);
}

// AlphaQssStep.opCall, chemistry_update.d:533-655: predictor, up to maxIters correctors
// with alpha-bar from p_p, convergence against the first predictor, and the step
// suggestion from three Newton iterations for sqrt(sigma).  yOut is written on success.
// Trip 0 of the loop is the predictor (rates at y0), trips 1..maxIters the correctors
// (rates at yp), so the rate code is inlined once; each trip evaluates exactly the
// reference's expressions for its kind:
//   predictor  yp = y0 + h (q0 - p0 y0) / (1 + alpha(p0) h p0)
//   corrector  yc = y0 + h (q~ - p- y0) / (1 + alpha(pp) h p-),
//              q~ = alpha(pp) qp + (1 - alpha(pp)) q0,  p- = (p0 + pp) / 2
template <int NSP, int NRM, bool DENSE>
__device__ __forceinline__ bool alpha_qss_step(const ChemDev& cd, const double* kf, const double* kb, const double* y0,
                                               double h, double* yOut, double& hSuggest, int& iters) {
    const double ZERO_EPS = 1.0e-50;
    double q0[NSP], p0[NSP], yp[NSP], yp0[NSP], yc[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        q0[s] = 0.0;
        p0[s] = 0.0;
        yp[s] = y0[s];
    }
    bool converged = false;
    for (int trip = 0; trip <= cd.maxIters; ++trip) {
        const bool pred = (trip == 0);
        double q[NSP], L[NSP];
        eval_split_rates<NSP, NRM, DENSE>(cd, kf, kb, yp, q, L);  // yp = y0 on the predictor trip
        bool passes = true;
#pragma unroll
        for (int s = 0; s < NSP; ++s) {
            const double p = L[s] / (yp[s] + ZERO_EPS);
            const double a = qss_alpha(p, h);
            const double qx = pred ? q[s] : a * q[s] + (1 - a) * q0[s];
            const double px = pred ? p : 0.5 * (p0[s] + p);
            const double ynew = y0[s] + (h * (qx - px * y0[s])) / (1.0 + a * h * px);
            if (pred) {
                q0[s] = q[s];
                p0[s] = p;
                yp0[s] = ynew;
            } else if (ynew >= ZERO_EPS) {
                if (fabs(ynew - yp0[s]) >= (cd.eps1 * (ynew + cd.delta))) passes = false;
            }
            yc[s] = ynew;
        }
        if (!pred) {
            ++iters;
            if (passes) {
                converged = true;
                break;
            }
        }
#pragma unroll
        for (int s = 0; s < NSP; ++s) yp[s] = yc[s];
    }
    double sigma = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        if (yc[s] < ZERO_EPS) continue;
        double test = fabs(yc[s] - yp0[s]) / (cd.eps2 * (yc[s] + cd.delta));
        if (test > sigma) sigma = test;
    }
    if (sigma <= 0.0) {
        hSuggest = 10.0 * h;
    } else {
        double x = sigma;
        x = x - 0.5 * (x * x - sigma) / x;
        x = x - 0.5 * (x * x - sigma) / x;
        x = x - 0.5 * (x * x - sigma) / x;
        hSuggest = h * ((1.0 / x) + 0.005);
    }
    if (converged) {
#pragma unroll
        for (int s = 0; s < NSP; ++s) yOut[s] = yc[s];
    }
    return converged;
}

// RKFStep.opCall, chemistry_update.d:385-484 (Cash-Karp tableau); yOut is always written.
template <int NSP, int NRM, bool DENSE>
__device__ __forceinline__ bool rkf_step(const ChemDev& cd, const double* kf, const double* kb, const double* y0, double h,
                         double* yOut, double& hSuggest, double* yLast = nullptr) {
    const double a21 = 1. / 5., a31 = 3. / 40., a32 = 9. / 40., a41 = 3. / 10., a42 = -9. / 10., a43 = 6. / 5.,
                 a51 = -11. / 54., a52 = 5. / 2., a53 = -70. / 27., a54 = 35. / 27., a61 = 1631. / 55296.,
                 a62 = 175. / 512., a63 = 575. / 13824., a64 = 44275. / 110592., a65 = 253. / 4096.;
    const double b51 = 37. / 378., b53 = 250. / 621., b54 = 125. / 594., b56 = 512. / 1771., b41 = 2825. / 27648.,
                 b43 = 18575. / 48384., b44 = 13525. / 55296., b45 = 277. / 14336., b46 = 1.0 / 4.0;
    double k1[NSP], k2[NSP], k3[NSP], k4[NSP], k5[NSP], k6[NSP], yT[NSP], q[NSP], L[NSP];
    auto rates = [&](const double* y, double* k) {
        eval_split_rates<NSP, NRM, DENSE>(cd, kf, kb, y, q, L);
#pragma unroll
        for (int s = 0; s < NSP; ++s) k[s] = q[s] - L[s];
    };
    rates(y0, k1);
#pragma unroll
    for (int s = 0; s < NSP; ++s) yT[s] = y0[s] + h * (a21 * k1[s]);
    rates(yT, k2);
#pragma unroll
    for (int s = 0; s < NSP; ++s) yT[s] = y0[s] + h * (a31 * k1[s] + a32 * k2[s]);
    rates(yT, k3);
#pragma unroll
    for (int s = 0; s < NSP; ++s) yT[s] = y0[s] + h * (a41 * k1[s] + a42 * k2[s] + a43 * k3[s]);
    rates(yT, k4);
#pragma unroll
    for (int s = 0; s < NSP; ++s) yT[s] = y0[s] + h * (a51 * k1[s] + a52 * k2[s] + a53 * k3[s] + a54 * k4[s]);
    rates(yT, k5);
#pragma unroll
    for (int s = 0; s < NSP; ++s)
        yT[s] = y0[s] + h * (a61 * k1[s] + a62 * k2[s] + a63 * k3[s] + a64 * k4[s] + a65 * k5[s]);
    rates(yT, k6);
    // the concentrations of the last rate evaluation (whose reaction rates the
    // two-temperature reactor's energy source reads)
    if (yLast) {
#pragma unroll
        for (int s = 0; s < NSP; ++s) yLast[s] = yT[s];
    }
    double err = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        yOut[s] = y0[s] + h * (b51 * k1[s] + b53 * k3[s] + b54 * k4[s] + b56 * k6[s]);
        double yE = yOut[s] - (y0[s] + h * (b41 * k1[s] + b43 * k3[s] + b44 * k4[s] + b45 * k5[s] + b46 * k6[s]));
        double ay0 = fabs(y0[s]), ay1 = fabs(yOut[s]);
        double sk = cd.errTol + cd.errTol * ((ay0 < ay1) ? ay1 : ay0);
        err += (yE / sk) * (yE / sk);
    }
    err = sqrt(err / NSP);
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
        return true;
    }
    scale = safe * pow(err, -alpha);
    scale = (scale < minscale) ? minscale : scale;
    hSuggest = scale * h;
    return false;
}

// Phobos std.algorithm.sum over NSP doubles: 0.0 + the pairwise sum of the D runtime
// (pairs summed into a store collapsed at powers of two, the store then summed from its
// last entry down).
template <int NSP>
__device__ __forceinline__ double phobos_sum(const double* x) {
    double store[NSP];
    int idx = 0;
#pragma unroll
    for (int k = 0; k < NSP / 2; ++k) {
        store[idx] = x[2 * k] + x[2 * k + 1];
        int lastToKeep = idx - __builtin_ctz((unsigned)(k + 1));
        while (idx > lastToKeep) {
            store[idx - 1] += store[idx];
            --idx;
        }
        ++idx;
    }
    if (NSP % 2 == 1) {
        store[idx] = x[NSP - 1];
        ++idx;
    }
    double s = store[idx - 1];
    for (int j = idx - 2; j >= 0; --j) s += store[j];
    return 0.0 + s;
}

// gas_model.d:255-271 (conc2massf) followed by the mass-fraction sum of
// chemistry_update.d:180-185 and 292-294.
template <int NSP>
__device__ __forceinline__ double conc_to_massf(const ChemDev& cd, const double* conc, double rho, double* massf) {
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        massf[s] = conc[s] * cd.M[s] / rho;
        if (massf[s] < CH_MIN_MASSF) massf[s] = 0.0;
    }
    return phobos_sum<NSP>(massf);
}

}  // namespace gdtk
