// gdtk_device.hpp -- device-side physics of the structured-block explicit update (gfx950).
//
// Everything here is FP64 and is compiled with -ffp-contract=off so that the operation
// order of each expression is the one written, which is the order of the D reference
// (cited per function).  The functions are written against plain values so that the
// same code serves the face-per-thread kernels and the LDS-tiled fused stage kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace gdtk {

// Flow-state field slots in the structure-of-arrays cell storage.
enum FsField : int { FS_RHO = 0, FS_VX, FS_VY, FS_VZ, FS_U, FS_P, FS_T, FS_A, FS_NBASE };

struct Vec3d {
    double x, y, z;
};

// A face's frame and area (src/eilmer/fvinterface.d:49-78).
struct FaceGeom {
    Vec3d n, t1, t2;
    double area;
};

// src/geom/elements/vector3.d:403-411
__device__ __forceinline__ Vec3d to_local(const Vec3d v, const FaceGeom& g) {
    Vec3d r;
    r.x = v.x * g.n.x + v.y * g.n.y + v.z * g.n.z;
    r.y = v.x * g.t1.x + v.y * g.t1.y + v.z * g.t1.z;
    r.z = v.x * g.t2.x + v.y * g.t2.y + v.z * g.t2.z;
    return r;
}
// src/geom/elements/vector3.d:416-423 and 683-692
__device__ __forceinline__ Vec3d to_global(const Vec3d v, const FaceGeom& g) {
    Vec3d r;
    r.x = v.x * g.n.x + v.y * g.t1.x + v.z * g.t2.x;
    r.y = v.x * g.n.y + v.y * g.t1.y + v.z * g.t2.y;
    r.z = v.x * g.n.z + v.y * g.t1.z + v.z * g.t2.z;
    return r;
}

// Ideal gas (src/gas/ideal_gas.d:66-143).  The sound speed is evaluated as
// sqrt((gamma*R)*T) and the pressure as (rho*R)*T, the reference's order.
struct IdealGasDev {
    double R, Cv, Cvinv, gamma;
    // ideal_gas.d:98-106; returns false where the reference throws GasModelException.
    __device__ __forceinline__ bool update_rhou(double rho, double u, double& T, double& p) const {
        if (u <= 0.0 || rho <= 0.0) return false;
        T = u * Cvinv;
        p = rho * R * T;
        return true;
    }
    // ideal_gas.d:107-115
    __device__ __forceinline__ void update_rhoT(double rho, double T, double& p, double& u) const {
        p = rho * R * T;
        u = Cv * T;
    }
    __device__ __forceinline__ double sound_speed(double T) const { return sqrt(gamma * R * T); }
};

// src/nm/limiters.d:41-49
__device__ __forceinline__ double clip_to_limits(double q, double A, double B) {
    const double lower_limit = (A <= B) ? A : B;
    const double upper_limit = (A > B) ? A : B;
    const double qclipped = (q > lower_limit) ? q : lower_limit;
    return (qclipped <= upper_limit) ? qclipped : upper_limit;
}

// Geometry-only coefficients of the l2r2 reconstruction (onedinterp.d:338-354).
struct L2R2 {
    double aL0, aR0, lenL0, lenR0, tL01, tR0L0, tR1R0, twoL0L1, twoR0R1;
    __device__ __forceinline__ void prepare(double lenL1, double lL0, double lR0, double lenR1) {
        lenL0 = lL0;
        lenR0 = lR0;
        aL0 = 0.5 * lL0 / (lenL1 + 2.0 * lL0 + lR0);
        aR0 = 0.5 * lR0 / (lL0 + 2.0 * lR0 + lenR1);
        tL01 = 2.0 / (lL0 + lenL1);
        tR0L0 = 2.0 / (lR0 + lL0);
        tR1R0 = 2.0 / (lenR1 + lR0);
        twoL0L1 = (2.0 * lL0 + lenL1);
        twoR0R1 = (2.0 * lR0 + lenR1);
    }
    // onedinterp.d:357-384 with the defaults apply_limiter = extrema_clipping = true, beta = 1.
    // LMR: Eilmer 5's limiter smoothing, epsilon_van_albada scaled by max(1e-12, |qL0|, |qR0|)
    // and by 2 / (lenR0 + lenL0) (lmr/onedinterp.d:137-169).
    template <bool LMR = false>
    __device__ __forceinline__ void scalar(double qL1, double qL0, double qR0, double qR1, double eps_va, double& qL,
                                           double& qR) const {
        const double beta = 1.0;
        double delLminus = (qL0 - qL1) * tL01;
        double del = (qR0 - qL0) * tR0L0;
        double delRplus = (qR1 - qR0) * tR1R0;
        double eps = eps_va;
        if (LMR) {
            const double qq = fmax(fmax(1e-12, fabs(qL0)), fmax(1e-12, fabs(qR0)));
            eps = qq * eps_va * tR0L0;
        }
        double sL = (delLminus * del + fabs(delLminus * del) + eps) / (delLminus * delLminus + del * del + eps);
        double sR = (del * delRplus + fabs(del * delRplus) + eps) / (del * del + delRplus * delRplus + eps);
        qL = qL0 + beta * sL * aL0 * (del * twoL0L1 + delLminus * lenR0);
        qR = qR0 - beta * sR * aR0 * (delRplus * lenL0 + del * twoR0R1);
        qL = clip_to_limits(qL, qL0, qR0);
        qR = clip_to_limits(qR, qL0, qR0);
    }
};

// One side of the l2r2 reconstruction: the coefficients and operations of L2R2::prepare and
// L2R2::scalar that the left value (from L1, L0, R0) or the right value (from L0, R0, R1)
// needs, in the same order, so that each side equals scalar()'s bit for bit.
template <bool RIGHT>
struct L2R2Side {
    double a, len_other, tO, tD, two;  // left: aL0, lenR0, tL01, tR0L0, twoL0L1
    __device__ __forceinline__ void prepare(double lenL1, double lL0, double lR0, double lenR1) {
        if (!RIGHT) {
            len_other = lR0;
            a = 0.5 * lL0 / (lenL1 + 2.0 * lL0 + lR0);
            tO = 2.0 / (lL0 + lenL1);
            tD = 2.0 / (lR0 + lL0);
            two = (2.0 * lL0 + lenL1);
        } else {
            len_other = lL0;
            a = 0.5 * lR0 / (lL0 + 2.0 * lR0 + lenR1);
            tD = 2.0 / (lR0 + lL0);
            tO = 2.0 / (lenR1 + lR0);
            two = (2.0 * lR0 + lenR1);
        }
    }
    // left: value(qL1, qL0, qR0); right: value(qL0, qR0, qR1).  LMR: Eilmer 5's epsilon,
    // eps_va scaled by max(1e-12, |qL0|, |qR0|) and by 2 / (lenR0 + lenL0), as L2R2::scalar.
    template <bool LMR = false>
    __device__ __forceinline__ double value(double q0, double q1, double q2, double eps_va) const {
        const double beta = 1.0;
        double eps = eps_va;
        if (LMR) {
            const double qL0 = RIGHT ? q0 : q1, qR0 = RIGHT ? q1 : q2;
            const double qq = fmax(fmax(1e-12, fabs(qL0)), fmax(1e-12, fabs(qR0)));
            eps = qq * eps_va * tD;
        }
        if (!RIGHT) {
            const double delLminus = (q1 - q0) * tO;
            const double del = (q2 - q1) * tD;
            const double sL = (delLminus * del + fabs(delLminus * del) + eps) / (delLminus * delLminus + del * del + eps);
            const double qL = q1 + beta * sL * a * (del * two + delLminus * len_other);
            return clip_to_limits(qL, q1, q2);
        } else {
            const double del = (q1 - q0) * tD;
            const double delRplus = (q2 - q1) * tO;
            const double sR = (del * delRplus + fabs(del * delRplus) + eps) / (del * del + delRplus * delRplus + eps);
            const double qR = q1 - beta * sR * a * (delRplus * len_other + del * two);
            return clip_to_limits(qR, q0, q1);
        }
    }
};

// A reconstructed face state; massf is used for NSP > 1 (multi-species gas).
template <int NSP>
struct FaceStateT {
    double rho, p, T, u, a;
    Vec3d vel;
    double massf[NSP];
};
using FaceState = FaceStateT<1>;

// Conserved-quantity layout for the transient solver (conservedquantities.d:147-193):
// mass, momentum, total energy, then the species densities of a multi-species gas.
template <int DIM, int NSP = 1>
struct CQ {
    static constexpr int mass = 0, xMom = 1, yMom = 2;
    static constexpr int zMom = (DIM == 3) ? 3 : -1;
    static constexpr int totEnergy = (DIM == 3) ? 4 : 3;
    static constexpr int species = (DIM == 3) ? 5 : 4;
    static constexpr int n = ((DIM == 3) ? 5 : 4) + ((NSP > 1) ? NSP : 0);
};

// fluxcalc.d:474-647, factor == 1, accumulating into F (which starts at zero).  LMR: Eilmer
// 5's common sound speed, a smooth maximum of aL and aR (lmr/fluxcalc.d ausmdv).
template <int DIM, int NSP = 1, bool LMR = false>
__device__ __forceinline__ void ausmdv(const FaceStateT<NSP>& L, const FaceStateT<NSP>& R, double* F) {
    using C = CQ<DIM, NSP>;
    const double factor = 1.0;
    double rL = L.rho, pL = L.p;
    double pLrL = pL / rL;
    double uL = L.vel.x, vL = L.vel.y, wL = L.vel.z;
    double eL = L.u;
    double aL = L.a;
    double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    double HL = eL + pLrL + keL;
    double rR = R.rho, pR = R.p;
    double pRrR = pR / rR;
    double uR = R.vel.x, vR = R.vel.y, wR = R.vel.z;
    double eR = R.u;
    double aR = R.a;
    double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    double HR = eR + pRrR + keR;
    double alphaL = 2.0 * pLrL / (pLrL + pRrR);
    double alphaR = 2.0 * pRrR / (pLrL + pRrR);
    double am;
    if (LMR) {
        const double da = aL - aR;
        const double scale = 0.5 * (aL + aR);
        const double eps = 1e-6 * scale + 1e-12;
        am = 0.5 * (aL + aR) + 0.5 * sqrt(da * da + eps * eps);
    } else {
        am = fmax(aL, aR);
    }
    double ML = uL / am;
    double MR = uR / am;
    double pLplus, uLplus;
    double duL = 0.5 * (uL + fabs(uL));
    if (fabs(ML) <= 1.0) {
        pLplus = pL * (ML + 1.0) * (ML + 1.0) * (2.0 - ML) * 0.25;
        uLplus = alphaL * ((uL + am) * (uL + am) / (4.0 * am) - duL) + duL;
    } else {
        pLplus = pL * duL / uL;
        uLplus = duL;
    }
    double pRminus, uRminus;
    double duR = 0.5 * (uR - fabs(uR));
    if (fabs(MR) <= 1.0) {
        pRminus = pR * (MR - 1.0) * (MR - 1.0) * (2.0 + MR) * 0.25;
        uRminus = alphaR * (-(uR - am) * (uR - am) / (4.0 * am) - duR) + duR;
    } else {
        pRminus = pR * duR / uR;
        uRminus = duR;
    }
    double ru_half = uLplus * rL + uRminus * rR;
    double p_half = pLplus + pRminus;
    double dp = pL - pR;
    const double K_SWITCH = 10.0;
    dp = K_SWITCH * fabs(dp) / fmin(pL, pR);
    double s = 0.5 * fmin(1.0, dp);
    double ru2_AUSMV = uLplus * rL * uL + uRminus * rR * uR;
    double ru2_AUSMD = 0.5 * (ru_half * (uL + uR) - fabs(ru_half) * (uR - uL));
    double ru2_half = (0.5 + s) * ru2_AUSMV + (0.5 - s) * ru2_AUSMD;
    F[C::mass] += factor * ru_half;
    const bool fromL = (ru_half >= 0.0);
    const double vU = fromL ? vL : vR, wU = fromL ? wL : wR, HU = fromL ? HL : HR;
    F[C::xMom] += (ru2_half + p_half) * factor;
    F[C::yMom] += (ru_half * vU) * factor;
    if (DIM == 3) F[C::zMom] += (ru_half * wU) * factor;
    F[C::totEnergy] += factor * ru_half * HU;
    if constexpr (NSP > 1) {
#pragma unroll
        for (int i = 0; i < NSP; ++i) F[C::species + i] += factor * ru_half * (fromL ? L.massf[i] : R.massf[i]);
    }
    // Entropy fix (apply_entropy_fix defaults to true, globalconfig.d:1088).
    const double C_EFIX = 0.125;
    bool caseA = ((uL - aL) < 0.0) && ((uR - aR) > 0.0);
    bool caseB = ((uL + aL) < 0.0) && ((uR + aR) > 0.0);
    double d_ua = 0.0;
    if (caseA && !caseB) d_ua = C_EFIX * ((uR - aR) - (uL - aL));
    if (caseB && !caseA) d_ua = C_EFIX * ((uR + aR) - (uL + aL));
    if (d_ua != 0.0) {
        F[C::mass] -= factor * d_ua * (rR - rL);
        F[C::xMom] -= factor * d_ua * (rR * uR - rL * uL);
        F[C::yMom] -= factor * d_ua * (rR * vR - rL * vL);
        if (DIM == 3) F[C::zMom] -= factor * d_ua * (rR * wR - rL * wL);
        F[C::totEnergy] -= factor * d_ua * (rR * HR - rL * HL);
        if constexpr (NSP > 1) {
#pragma unroll
            for (int i = 0; i < NSP; ++i) F[C::species + i] -= factor * d_ua * (rR * R.massf[i] - rL * L.massf[i]);
        }
    }
}

// fluxcalc.d:1280-1312 (Abramowitz & Stegun 7.1.26 as coded in the reference)
__device__ __forceinline__ void exxef(double sn, double& exx, double& ef) {
    const double P = 0.327591100;
    const double A1 = 0.254829592;
    const double A2 = -0.284496736;
    const double A3 = 1.421413741;
    const double A4 = -1.453152027;
    const double A5 = 1.061405429;
    const double LIMIT = 5.0;
    const double EXLIM = 0.138879e-10;
    const double EFLIM = 1.0;
    double ef1;
    if (fabs(sn) > LIMIT) {
        exx = EXLIM;
        ef1 = EFLIM;
    } else {
        double snsq = sn * sn;
        exx = exp(-snsq);
        double y = 1.0 / (1.0 + P * fabs(sn));
        ef1 = 1.0 - y * (A1 + y * (A2 + y * (A3 + y * (A4 + A5 * y)))) * exx;
    }
    ef = copysign(ef1, sn);
}

// fluxcalc.d:1131-1277, factor == 1.  cvL, cvR: the gas model's Cv of each side.
template <int DIM, int NSP = 1>
__device__ __forceinline__ void efmflx(const FaceStateT<NSP>& L, const FaceStateT<NSP>& R, double cvL, double cvR,
                                       double* F) {
    using C = CQ<DIM, NSP>;
    const double factor = 1.0;
    const double PHI = 1.0;
    const double dtwspi = 0.282094792;
    double rhoL = L.rho, presL = L.p;
    double eL = L.u;
    double hL = eL + presL / rhoL;
    double tL = L.T;
    double vnL = L.vel.x, vpL = L.vel.y, vqL = L.vel.z;
    double rhoR = R.rho, presR = R.p;
    double eR = R.u;
    double hR = eR + presR / rhoR;
    double tR = R.T;
    double vnR = R.vel.x, vpR = R.vel.y, vqR = R.vel.z;
    double RgasL = presL / (rhoL * tL);
    double RgasR = presR / (rhoR * tR);
    double rLsqrt = sqrt(rhoL);
    double rRsqrt = sqrt(rhoR);
    double alpha = rLsqrt / (rLsqrt + rRsqrt);
    double cv = alpha * cvL + (1.0 - alpha) * cvR;
    double Rgas = alpha * RgasL + (1.0 - alpha) * RgasR;
    double cp = cv + Rgas;
    double gam = cp / cv;
    double con = 0.5 * (gam + 1.0) / (gam - 1.0);
    double rtL = Rgas * tL;
    double cmpL = sqrt(2.0 * rtL);
    double hvsqL = 0.5 * (vnL * vnL + vpL * vpL + vqL * vqL);
    double snL = vnL / (PHI * cmpL);
    double exL, efL;
    exxef(snL, exL, efL);
    double wL = 0.5 * (1.0 + efL);
    double dL = exL * dtwspi;
    double rtR = presR / rhoR;  // as coded in the reference
    double cmpR = sqrt(2.0 * rtR);
    double hvsqR = 0.5 * (vnR * vnR + vpR * vpR + vqR * vqR);
    double snR = vnR / (PHI * cmpR);
    double exR, efR;
    exxef(snR, exR, efR);
    double wR = 0.5 * (1.0 - efR);
    double dR = -exR * dtwspi;
    double fmsL = (wL * rhoL * vnL) + (dL * cmpL * rhoL);
    double fmsR = (wR * rhoR * vnR) + (dR * cmpR * rhoR);
    double mass_flux = factor * (fmsL + fmsR);
    F[C::mass] += mass_flux;
    F[C::xMom] += factor * (fmsL * vnL + fmsR * vnR + wL * presL + wR * presR);
    F[C::yMom] += factor * (fmsL * vpL + fmsR * vpR);
    if (DIM == 3) F[C::zMom] += factor * (fmsL * vqL + fmsR * vqR);
    F[C::totEnergy] += factor * ((wL * rhoL * vnL) * (hvsqL + hL) + (wR * rhoR * vnR) * (hvsqR + hR) +
                                 (dL * cmpL * rhoL) * (hvsqL + con * rtL) + (dR * cmpR * rhoR) * (hvsqR + con * rtR));
    if constexpr (NSP > 1) {
        const bool fromL = mass_flux > 0.0;
#pragma unroll
        for (int i = 0; i < NSP; ++i) F[C::species + i] += mass_flux * (fromL ? L.massf[i] : R.massf[i]);
    }
}

// One stencil cell as read by the reconstruction.
struct StencilCell {
    double rho, u;
    Vec3d vel;
};

// Flux of one interior face from its 4-cell stencil, single-species gas:
// onedinterp.d:117-273,751-988 (l2r2, rhou, local frame) followed by
// fluxcalc.d:54-184 (frame rotations, calculator, gvel == 0 terms, rotation back).
// cL0/cR0 carry the extra fields the reference copies into Lft/Rght before
// reconstruction: a (never updated by the rhou EOS call), and p, T for the
// EOS-failure fallback.
// Where the face EOS call fails, the reference copies the whole cell state (onedinterp.d:45-68);
// p and T of the cells L0/R0 are then read from the flow-state planes.
// Flux-calculator codes: 0 AUSMDV, 1 EFM, 2 adaptive EFM/AUSMDV, 3 Hanel, 4 adaptive
// Hanel/AUSMDV.  The adaptive ones read the faces' shock indicators.
__host__ __device__ constexpr bool flux_is_adaptive(int fc) { return fc == 2 || fc == 4; }

struct FallbackFS {
    const double* fs;
    int64_t np, cL0, cR0;
};

template <int DIM, int FC, bool LMR = false>
__device__ __forceinline__ void face_flux_l2r2(const IdealGasDev& gas, const FaceGeom& g, const StencilCell& L1,
                                               const StencilCell& L0, const StencilCell& R0, const StencilCell& R1,
                                               double aL0, double aR0, const FallbackFS& fb, double lenL1,
                                               double lenL0, double lenR0, double lenR1, double eps, double* F,
                                               bool shocked = false) {
    using C = CQ<DIM>;
    Vec3d vL1 = to_local(L1.vel, g), vL0 = to_local(L0.vel, g), vR0 = to_local(R0.vel, g), vR1 = to_local(R1.vel, g);
    L2R2 c;
    c.prepare(lenL1, lenL0, lenR0, lenR1);
    FaceState Ls, Rs;
    c.scalar<LMR>(vL1.x, vL0.x, vR0.x, vR1.x, eps, Ls.vel.x, Rs.vel.x);
    c.scalar<LMR>(vL1.y, vL0.y, vR0.y, vR1.y, eps, Ls.vel.y, Rs.vel.y);
    c.scalar<LMR>(vL1.z, vL0.z, vR0.z, vR1.z, eps, Ls.vel.z, Rs.vel.z);
    c.scalar<LMR>(L1.rho, L0.rho, R0.rho, R1.rho, eps, Ls.rho, Rs.rho);
    c.scalar<LMR>(L1.u, L0.u, R0.u, R1.u, eps, Ls.u, Rs.u);
    Ls.a = aL0;
    Rs.a = aR0;
    if (!gas.update_rhou(Ls.rho, Ls.u, Ls.T, Ls.p)) {
        // onedinterp.d:45-55: Lft.copy_values_from(cL0.fs), velocity already local.
        Ls.rho = L0.rho; Ls.u = L0.u; Ls.vel = vL0;
        Ls.p = fb.fs[fb.cL0 + FS_P * fb.np];
        Ls.T = fb.fs[fb.cL0 + FS_T * fb.np];
    }
    if (!gas.update_rhou(Rs.rho, Rs.u, Rs.T, Rs.p)) {
        Rs.rho = R0.rho; Rs.u = R0.u; Rs.vel = vR0;
        Rs.p = fb.fs[fb.cR0 + FS_P * fb.np];
        Rs.T = fb.fs[fb.cR0 + FS_T * fb.np];
    }
    // Back to the global frame (onedinterp.d:979-987), then into the face frame
    // again for the flux calculator (fluxcalc.d:65-67).
    Ls.vel = to_local(to_global(Ls.vel, g), g);
    Rs.vel = to_local(to_global(Rs.vel, g), g);
    for (int j = 0; j < C::n; ++j) F[j] = 0.0;
    // FC 2, 4: adaptive_efm_ausmdv, adaptive_hanel_ausmdv (fluxcalc.d:1315-1348) with the
    // face's indicator 0 or 1, so one calculator runs with factor 1.
    if (FC == 1 || (FC == 2 && shocked)) efmflx<DIM>(Ls, Rs, gas.Cv, gas.Cv, F);
    else if (FC == 3 || (FC == 4 && shocked)) hanel<DIM>(Ls, Rs, F);
    else ausmdv<DIM, 1, LMR>(Ls, Rs, F);
    // fluxcalc.d:159-171 add grid-velocity terms, here with gvel = 0: each adds a signed
    // zero to a flux component.  Every component starts at +0 and is only ever changed by
    // x + y or x - y, so it is never -0 (a sum is -0 only if both terms are -0), and
    // adding a signed zero to a value that is not -0 returns that value bit for bit.  The
    // terms are therefore omitted without changing any result.
    double fz = (DIM == 3) ? F[C::zMom] : 0.0;
    Vec3d m = to_global(Vec3d{F[C::xMom], F[C::yMom], fz}, g);
    F[C::xMom] = m.x;
    F[C::yMom] = m.y;
    if (DIM == 3) F[C::zMom] = m.z;
}

// boundary_flux_effect.d:586-645, the simple-outflow flux from the interior cell (massf:
// its mass fractions, read for NSP > 1).
// Hanel, Schwane and Seider's flux-vector splitting, fluxcalc.d:1028-1128 (factor 1, so
// each component is the expression itself): van Leer velocity and pressure splitting with
// each side's own sound speed.
template <int DIM, int NSP = 1>
__device__ __forceinline__ void hanel(const FaceStateT<NSP>& L, const FaceStateT<NSP>& R, double* F) {
    using C = CQ<DIM, NSP>;
    const double rL = L.rho, pL = L.p;
    const double pLrL = pL / rL;
    const double uL = L.vel.x, vL = L.vel.y, wL = L.vel.z;
    const double eL = L.u, aL = L.a;
    const double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    const double HL = eL + pLrL + keL;
    const double rR = R.rho, pR = R.p;
    const double pRrR = pR / rR;
    const double uR = R.vel.x, vR = R.vel.y, wR = R.vel.z;
    const double eR = R.u, aR = R.a;
    const double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    const double HR = eR + pRrR + keR;
    double pLplus, uLplus;
    if (fabs(uL) <= aL) {
        uLplus = 1.0 / (4.0 * aL) * (uL + aL) * (uL + aL);
        pLplus = pL * uLplus * (1.0 / aL * (2.0 - uL / aL));
    } else {
        uLplus = 0.5 * (uL + fabs(uL));
        pLplus = pL * uLplus * (1.0 / uL);
    }
    double pRminus, uRminus;
    if (fabs(uR) <= aR) {
        uRminus = -1.0 / (4.0 * aR) * (uR - aR) * (uR - aR);
        pRminus = pR * uRminus * (1.0 / aR * (-2.0 - uR / aR));
    } else {
        uRminus = 0.5 * (uR - fabs(uR));
        pRminus = pR * uRminus * (1.0 / uR);
    }
    const double p_half = pLplus + pRminus;
    F[C::mass] += uLplus * rL + uRminus * rR;
    F[C::xMom] += uLplus * rL * uL + uRminus * rR * uR + p_half;
    F[C::yMom] += uLplus * rL * vL + uRminus * rR * vR;
    if (DIM == 3) F[C::zMom] += uLplus * rL * wL + uRminus * rR * wR;
    F[C::totEnergy] += uLplus * rL * HL + uRminus * rR * HR;
    if (NSP > 1)
#pragma unroll
        for (int i = 0; i < NSP; ++i) F[C::species + i] += uLplus * rL * L.massf[i] + uRminus * rR * R.massf[i];
}

template <int DIM, int NSP = 1>
__device__ __forceinline__ void outflow_flux(const FaceGeom& g, int outsign, double rho, double u, double p, Vec3d vel,
                                             double* F, const double* massf = nullptr) {
    using C = CQ<DIM, NSP>;
    double vn = vel.x * g.n.x + vel.y * g.n.y + vel.z * g.n.z;
    double mass_flux = rho * vn;
    if ((outsign * mass_flux) > 0.0) {
        F[C::mass] = mass_flux;
        F[C::xMom] = p * g.n.x + vel.x * mass_flux;
        F[C::yMom] = p * g.n.y + vel.y * mass_flux;
        if (DIM == 3) F[C::zMom] = p * g.n.z + vel.z * mass_flux;
        double utot = u + 0.5 * (vel.x * vel.x + vel.y * vel.y + vel.z * vel.z);
        F[C::totEnergy] = mass_flux * utot + p * vn;
        if constexpr (NSP > 1) {
#pragma unroll
            for (int i = 0; i < NSP; ++i) F[C::species + i] = mass_flux * massf[i];
        }
    } else {
        F[C::mass] = 0.0;
        F[C::xMom] = g.n.x * p;
        F[C::yMom] = g.n.y * p;
        if (DIM == 3) F[C::zMom] = g.n.z * p;
        double gdotn = 0.0 * g.n.x + 0.0 * g.n.y + 0.0 * g.n.z;
        F[C::totEnergy] = p * gdotn;
        if constexpr (NSP > 1) {
#pragma unroll
            for (int i = 0; i < NSP; ++i) F[C::species + i] = 0.0;
        }
    }
}

// Decode of one cell's conserved quantities, single species (fvcell.d:586-821):
// returns 1 where the reference throws (non-positive density), else 0.  With
// ignore_low_T_thermo_update_failure (default true) an EOS failure resets T to
// suggested_low_T_value and re-encodes U (reencoded = true).
template <int DIM>
__device__ __forceinline__ int decode_ideal(const IdealGasDev& gas, double* U, double low_T, double& rho, Vec3d& vel,
                                            double& u, double& p, double& T, double& a, bool& reencoded) {
    using C = CQ<DIM>;
    reencoded = false;
    rho = U[C::mass];
    if (!(rho > 0.0)) return 1;
    double dinv = 1.0 / rho;
    double zMom = (DIM == 3) ? U[C::zMom] : 0.0;
    vel.x = U[C::xMom] * dinv;
    vel.y = U[C::yMom] * dinv;
    vel.z = zMom * dinv;
    double rE = U[C::totEnergy];
    double uu = rE * dinv;
    double ke = 0.5 * (vel.x * vel.x + vel.y * vel.y + vel.z * vel.z);
    uu -= ke;
    u = uu;
    if (!gas.update_rhou(rho, u, T, p)) {
        T = low_T;
        gas.update_rhoT(rho, T, p, u);
        // encode_conserved (fvcell.d:511-583)
        U[C::mass] = rho;
        U[C::xMom] = rho * vel.x;
        U[C::yMom] = rho * vel.y;
        if (DIM == 3) U[C::zMom] = rho * vel.z;
        double ke2 = 0.5 * (vel.x * vel.x + vel.y * vel.y + vel.z * vel.z);
        U[C::totEnergy] = rho * (u + ke2);
        reencoded = true;
    }
    if (T <= 0.0) return 1;
    a = gas.sound_speed(T);
    return 0;
}

// FlowState.check_data (flowstate.d:363-390) with GasState.check_values
// (gas_state.d:226-268) for a single-species gas; true if valid.
__device__ __forceinline__ bool check_data(double rho, double p, double T, double a, Vec3d vel, double max_vel,
                                           double min_T, double max_T) {
    bool ok = true;
    if (!isfinite(rho) || rho < 1.01 * 0.0) ok = false;
    if (!isfinite(T) || T < 1.01 * 0.0) ok = false;
    if (!isfinite(p)) ok = false;
    if (!isfinite(a)) ok = false;
    if (fabs(vel.x) > max_vel || fabs(vel.y) > max_vel || fabs(vel.z) > max_vel) ok = false;
    if (T < min_T) ok = false;
    if (T > max_T) ok = false;
    return ok;
}

}  // namespace gdtk
