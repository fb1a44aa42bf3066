// eilmer_oracle.cpp -- CPU restatement of Eilmer 4's structured-block explicit update.
//
// TEST INFRASTRUCTURE ONLY (the oracle).  See eilmer_oracle.hpp.
//
// Build: g++ -O2 -std=c++17 -ffp-contract=off -fPIC (no -ffast-math), see oracle/Makefile.
//
// Parity pins (what checks each function; the tests live in tests/test_oracle_*.py):
//   ideal gas EOS              closed form (p = rho R T, a = sqrt(gamma R T)), R from
//                              src/gas/sample-data/ideal-air-gas-model.lua:5-7
//   l2r2 + van Albada          exactness for linear data on uniform cells, first order at
//                              extrema, clipping bounds (onedinterp.d:338-384)
//   ausmdv / efmflx            consistency F(U,U) = Euler flux, supersonic upwinding,
//                              mirror symmetry (fluxcalc.d:474-647, 1131-1312)
//   exxef                      Abramowitz & Stegun 7.1.26 error bound |erf - ef| < 1.5e-7
//   whole explicit update      examples/eilmer/3D/sod-shock-tube/sg/sod-test.rb:32,36,68
//                              (75 +/- 3 steps to 0.6 ms, probe values within 1 %) and
//                              the analytic Sod solution; conservation of mass/energy in a
//                              closed slip-wall box; free-stream preservation on skewed grids
//   dt controller              step counts of the Sod test above (steps depend on it)
// Bit-level per-step field values of the D code cannot be pinned without a D compiler
// (none in this image); see DESIGN.md.
#include "eilmer_oracle.hpp"
#include "chem_oracle.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <limits>

namespace orc {

//---------------------------------------------------------------------------------
// Ideal gas, src/gas/ideal_gas.d:30-143
IdealGas::IdealGas(double mMass_, double gamma_) {
    n_species = 1;
    mol_masses = {mMass_};
    mMass = mMass_;
    gamma = gamma_;
    // ideal_gas.d:66-70
    Rgas = R_universal / mMass;
    Cv_ = Rgas / (gamma - 1.0);
    Cvinv = 1.0 / Cv_;
    Cp_ = Rgas * gamma / (gamma - 1.0);
}
void IdealGas::update_thermo_from_pT(GasState& Q) const {
    // ideal_gas.d:89-97
    if (Q.T <= 0.0 || Q.p <= 0.0) throw GasModelException("negative T or p in update_thermo_from_pT");
    Q.rho = Q.p / (Q.T * Rgas);
    Q.u = Cv_ * Q.T;
}
void IdealGas::update_thermo_from_rhou(GasState& Q) const {
    // ideal_gas.d:98-106
    if (Q.u <= 0.0 || Q.rho <= 0.0) throw GasModelException("negative u or rho in update_thermo_from_rhou");
    Q.T = Q.u * Cvinv;
    Q.p = Q.rho * Rgas * Q.T;
}
void IdealGas::update_thermo_from_rhoT(GasState& Q) const {
    // ideal_gas.d:107-115
    if (Q.T <= 0.0 || Q.rho <= 0.0) throw GasModelException("negative T or rho in update_thermo_from_rhoT");
    Q.p = Q.rho * Rgas * Q.T;
    Q.u = Cv_ * Q.T;
}
void IdealGas::update_sound_speed(GasState& Q) const {
    // ideal_gas.d:136-143
    if (Q.T <= 0.0) throw GasModelException("negative T in update_sound_speed");
    Q.a = std::sqrt(gamma * Rgas * Q.T);
}

//---------------------------------------------------------------------------------
// src/nm/limiters.d:41-49
double clip_to_limits(double q, double A, double B) {
    const double lower_limit = (A <= B) ? A : B;
    const double upper_limit = (A > B) ? A : B;
    const double qclipped = (q > lower_limit) ? q : lower_limit;
    return (qclipped <= upper_limit) ? qclipped : upper_limit;
}

// src/gas/gas_model.d:373-407
void scale_mass_fractions(std::vector<double>& massf, double tolerance, double assert_error_tolerance) {
    size_t nsp = massf.size();
    if (nsp == 1) {
        if (std::fabs(massf[0] - 1.0) > assert_error_tolerance)
            throw GasModelException("Single species mass fraction far from 1.0.");
        massf[0] = 1.0;
    } else {
        double s = 0.0;
        for (size_t i = 0; i < nsp; ++i) {
            massf[i] = massf[i] >= 0.0 ? massf[i] : 0.0;
            s += massf[i];
        }
        if (std::fabs(s - 1.0) > assert_error_tolerance)
            throw GasModelException("Sum of species mass fractions far from 1.0");
        if (std::fabs(s - 1.0) > tolerance)
            for (size_t i = 0; i < nsp; ++i) massf[i] /= s;
    }
}

//---------------------------------------------------------------------------------
// Flux calculators, src/eilmer/fluxcalc.d

// fluxcalc.d:474-647
void ausmdv(const FlowState& Lft, const FlowState& Rght, std::vector<double>& F, const CQI& cqi,
            const GasModel& gm, bool entropy_fix, double factor, bool lmr) {
    double rL = Lft.gas.rho;
    double pL = Lft.gas.p;
    double pLrL = pL / rL;
    double uL = Lft.vel.x;
    double vL = Lft.vel.y;
    double wL = Lft.vel.z;
    double eL = gm.internal_energy(Lft.gas);
    double aL = Lft.gas.a;
    double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    double HL = eL + pLrL + keL;
    double rR = Rght.gas.rho;
    double pR = Rght.gas.p;
    double pRrR = pR / rR;
    double uR = Rght.vel.x;
    double vR = Rght.vel.y;
    double wR = Rght.vel.z;
    double eR = gm.internal_energy(Rght.gas);
    double aR = Rght.gas.a;
    double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    double HR = eR + pRrR + keR;
    double alphaL = 2.0 * pLrL / (pLrL + pRrR);
    double alphaR = 2.0 * pRrR / (pLrL + pRrR);
    double am = std::fmax(aL, aR);
    if (lmr) {
        // lmr/fluxcalc.d, ausmdv: the smooth maximum of aL and aR (Biswas et al., CVPR 2021).
        double da = aL - aR;
        double scale = 0.5 * (aL + aR);
        double eps = 1e-6 * scale + 1e-12;
        am = 0.5 * (aL + aR) + 0.5 * std::sqrt(da * da + eps * eps);
    }
    double ML = uL / am;
    double MR = uR / am;
    double pLplus, uLplus;
    double duL = 0.5 * (uL + std::fabs(uL));
    if (std::fabs(ML) <= 1.0) {
        pLplus = pL * (ML + 1.0) * (ML + 1.0) * (2.0 - ML) * 0.25;
        uLplus = alphaL * ((uL + am) * (uL + am) / (4.0 * am) - duL) + duL;
    } else {
        pLplus = pL * duL / uL;
        uLplus = duL;
    }
    double pRminus, uRminus;
    double duR = 0.5 * (uR - std::fabs(uR));
    if (std::fabs(MR) <= 1.0) {
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
    dp = K_SWITCH * std::fabs(dp) / std::fmin(pL, pR);
    double s = 0.5 * std::fmin(1.0, dp);
    double ru2_AUSMV = uLplus * rL * uL + uRminus * rR * uR;
    double ru2_AUSMD = 0.5 * (ru_half * (uL + uR) - std::fabs(ru_half) * (uR - uL));
    double ru2_half = (0.5 + s) * ru2_AUSMV + (0.5 - s) * ru2_AUSMD;
    F[cqi.mass] += factor * ru_half;
    if (ru_half >= 0.0) {
        F[cqi.xMom] += (ru2_half + p_half) * factor;
        F[cqi.yMom] += (ru_half * vL) * factor;
        if (cqi.threeD) F[cqi.zMom] += (ru_half * wL) * factor;
        F[cqi.totEnergy] += factor * ru_half * HL;
        if (cqi.n_species > 1)
            for (int i = 0; i < cqi.n_species; ++i) F[cqi.species + i] += factor * ru_half * Lft.gas.massf[i];
    } else {
        F[cqi.xMom] += (ru2_half + p_half) * factor;
        F[cqi.yMom] += (ru_half * vR) * factor;
        if (cqi.threeD) F[cqi.zMom] += (ru_half * wR) * factor;
        F[cqi.totEnergy] += factor * ru_half * HR;
        if (cqi.n_species > 1)
            for (int i = 0; i < cqi.n_species; ++i) F[cqi.species + i] += factor * ru_half * Rght.gas.massf[i];
    }
    if (entropy_fix) {
        const double C_EFIX = 0.125;
        bool caseA = ((uL - aL) < 0.0) && ((uR - aR) > 0.0);
        bool caseB = ((uL + aL) < 0.0) && ((uR + aR) > 0.0);
        double d_ua = 0.0;
        if (caseA && !caseB) d_ua = C_EFIX * ((uR - aR) - (uL - aL));
        if (caseB && !caseA) d_ua = C_EFIX * ((uR + aR) - (uL + aL));
        if (d_ua != 0.0) {
            F[cqi.mass] -= factor * d_ua * (rR - rL);
            F[cqi.xMom] -= factor * d_ua * (rR * uR - rL * uL);
            F[cqi.yMom] -= factor * d_ua * (rR * vR - rL * vL);
            if (cqi.threeD) F[cqi.zMom] -= factor * d_ua * (rR * wR - rL * wL);
            F[cqi.totEnergy] -= factor * d_ua * (rR * HR - rL * HL);
            if (cqi.n_species > 1)
                for (int i = 0; i < cqi.n_species; ++i)
                    F[cqi.species + i] -= factor * d_ua * (rR * Rght.gas.massf[i] - rL * Lft.gas.massf[i]);
        }
    }
}

// fluxcalc.d:1280-1312
void exxef(double sn, double& exx, double& ef) {
    double snsq, ef1, y;
    const double P = 0.327591100;
    const double A1 = 0.254829592;
    const double A2 = -0.284496736;
    const double A3 = 1.421413741;
    const double A4 = -1.453152027;
    const double A5 = 1.061405429;
    const double LIMIT = 5.0;
    const double EXLIM = 0.138879e-10;
    const double EFLIM = 1.0;
    if (std::fabs(sn) > LIMIT) {
        exx = EXLIM;
        ef1 = EFLIM;
    } else {
        snsq = sn * sn;
        exx = std::exp(-snsq);
        y = 1.0 / (1.0 + P * std::fabs(sn));
        ef1 = 1.0 - y * (A1 + y * (A2 + y * (A3 + y * (A4 + A5 * y)))) * exx;
    }
    ef = std::copysign(ef1, sn);
}


// fluxcalc.d:1028-1128 (Hanel, Schwane and Seider's flux-vector splitting)
void hanel(const FlowState& Lft, const FlowState& Rght, std::vector<double>& F, const CQI& cqi, const GasModel& gm,
           double factor) {
    double rL = Lft.gas.rho;
    double pL = Lft.gas.p;
    double pLrL = pL / rL;
    double uL = Lft.vel.x;
    double vL = Lft.vel.y;
    double wL = Lft.vel.z;
    double eL = gm.internal_energy(Lft.gas);
    double aL = Lft.gas.a;
    double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    double HL = eL + pLrL + keL;
    double rR = Rght.gas.rho;
    double pR = Rght.gas.p;
    double pRrR = pR / rR;
    double uR = Rght.vel.x;
    double vR = Rght.vel.y;
    double wR = Rght.vel.z;
    double eR = gm.internal_energy(Rght.gas);
    double aR = Rght.gas.a;
    double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    double HR = eR + pRrR + keR;
    double pLplus, uLplus;
    if (std::fabs(uL) <= aL) {
        uLplus = 1.0 / (4.0 * aL) * (uL + aL) * (uL + aL);
        pLplus = pL * uLplus * (1.0 / aL * (2.0 - uL / aL));
    } else {
        uLplus = 0.5 * (uL + std::fabs(uL));
        pLplus = pL * uLplus * (1.0 / uL);
    }
    double pRminus, uRminus;
    if (std::fabs(uR) <= aR) {
        uRminus = -1.0 / (4.0 * aR) * (uR - aR) * (uR - aR);
        pRminus = pR * uRminus * (1.0 / aR * (-2.0 - uR / aR));
    } else {
        uRminus = 0.5 * (uR - std::fabs(uR));
        pRminus = pR * uRminus * (1.0 / uR);
    }
    double p_half = pLplus + pRminus;
    F[cqi.mass] += factor * (uLplus * rL + uRminus * rR);
    F[cqi.xMom] += factor * (uLplus * rL * uL + uRminus * rR * uR + p_half);
    F[cqi.yMom] += factor * (uLplus * rL * vL + uRminus * rR * vR);
    if (cqi.threeD) F[cqi.zMom] += factor * (uLplus * rL * wL + uRminus * rR * wR);
    F[cqi.totEnergy] += factor * (uLplus * rL * HL + uRminus * rR * HR);
    if (cqi.n_species > 1)
        for (int i = 0; i < cqi.n_species; ++i)
            F[cqi.species + i] += factor * (uLplus * rL * Lft.gas.massf[i] + uRminus * rR * Rght.gas.massf[i]);
}

// fluxcalc.d:1131-1277
void efmflx(const FlowState& Lft, const FlowState& Rght, std::vector<double>& F, const CQI& cqi,
            const GasModel& gm, double factor) {
    const double PHI = 1.0;
    const double dtwspi = 0.282094792;
    double rhoL = Lft.gas.rho;
    double presL = Lft.gas.p;
    double eL = gm.internal_energy(Lft.gas);
    double hL = eL + presL / rhoL;
    double tL = Lft.gas.T;
    double vnL = Lft.vel.x, vpL = Lft.vel.y, vqL = Lft.vel.z;
    double rhoR = Rght.gas.rho;
    double presR = Rght.gas.p;
    double eR = gm.internal_energy(Rght.gas);
    double hR = eR + presR / rhoR;
    double tR = Rght.gas.T;
    double vnR = Rght.vel.x, vpR = Rght.vel.y, vqR = Rght.vel.z;
    double cvL = gm.Cv(Lft.gas);
    double RgasL = presL / (rhoL * tL);
    double cvR = gm.Cv(Rght.gas);
    double RgasR = presR / (rhoR * tR);
    double rLsqrt = std::sqrt(rhoL);
    double rRsqrt = std::sqrt(rhoR);
    double alpha = rLsqrt / (rLsqrt + rRsqrt);
    double cv = alpha * cvL + (1.0 - alpha) * cvR;
    double Rgas = alpha * RgasL + (1.0 - alpha) * RgasR;
    double cp = cv + Rgas;
    double gam = cp / cv;
    double con = 0.5 * (gam + 1.0) / (gam - 1.0);
    double rtL = Rgas * tL;
    double cmpL = std::sqrt(2.0 * rtL);
    double hvsqL = 0.5 * (vnL * vnL + vpL * vpL + vqL * vqL);
    double snL = vnL / (PHI * cmpL);
    double exL, efL;
    exxef(snL, exL, efL);
    double wL = 0.5 * (1.0 + efL);
    double dL = exL * dtwspi;
    double rtR = presR / rhoR;  // as written in the reference (not Rgas * tR)
    double cmpR = std::sqrt(2.0 * rtR);
    double hvsqR = 0.5 * (vnR * vnR + vpR * vpR + vqR * vqR);
    double snR = vnR / (PHI * cmpR);
    double exR, efR;
    exxef(snR, exR, efR);
    double wR = 0.5 * (1.0 - efR);
    double dR = -exR * dtwspi;
    double fmsL = (wL * rhoL * vnL) + (dL * cmpL * rhoL);
    double fmsR = (wR * rhoR * vnR) + (dR * cmpR * rhoR);
    double mass_flux = factor * (fmsL + fmsR);
    F[cqi.mass] += mass_flux;
    F[cqi.xMom] += factor * (fmsL * vnL + fmsR * vnR + wL * presL + wR * presR);
    F[cqi.yMom] += factor * (fmsL * vpL + fmsR * vpR);
    if (cqi.threeD) F[cqi.zMom] += factor * (fmsL * vqL + fmsR * vqR);
    F[cqi.totEnergy] += factor * ((wL * rhoL * vnL) * (hvsqL + hL) + (wR * rhoR * vnR) * (hvsqR + hR) +
                                  (dL * cmpL * rhoL) * (hvsqL + con * rtL) +
                                  (dR * cmpR * rhoR) * (hvsqR + con * rtR));
    if (mass_flux > 0.0) {
        if (cqi.n_species > 1)
            for (int i = 0; i < cqi.n_species; ++i) F[cqi.species + i] += mass_flux * Lft.gas.massf[i];
    } else {
        if (cqi.n_species > 1)
            for (int i = 0; i < cqi.n_species; ++i) F[cqi.species + i] += mass_flux * Rght.gas.massf[i];
    }
}

//---------------------------------------------------------------------------------
// Geometry, src/geom/elements/properties.d

static void cross_product(double ax, double ay, double az, double bx, double by, double bz,
                          double& cx, double& cy, double& cz) {
    // vector3.d:646-653
    cx = ay * bz - az * by;
    cy = az * bx - ax * bz;
    cz = ax * by - ay * bx;
}

// properties.d:97-141
static void quad_properties(const Vec3& p0, const Vec3& p1, const Vec3& p2, const Vec3& p3, Vec3& centroid,
                            Vec3& n, Vec3& t1, Vec3& t2, double& area) {
    const double area_tol = 1.0e-20;
    centroid = Vec3(0.25 * (p0.x + p1.x + p2.x + p3.x), 0.25 * (p0.y + p1.y + p2.y + p3.y),
                    0.25 * (p0.z + p1.z + p2.z + p3.z));
    double p01x = p1.x - p0.x + p2.x - p3.x;
    double p01y = p1.y - p0.y + p2.y - p3.y;
    double p01z = p1.z - p0.z + p2.z - p3.z;
    double p03x = p3.x - p0.x + p2.x - p1.x;
    double p03y = p3.y - p0.y + p2.y - p1.y;
    double p03z = p3.z - p0.z + p2.z - p1.z;
    double vax, vay, vaz;
    cross_product(p01x, p01y, p01z, p03x, p03y, p03z, vax, vay, vaz);
    vax *= 0.25; vay *= 0.25; vaz *= 0.25;
    area = std::sqrt(vax * vax + vay * vay + vaz * vaz);
    if (area > area_tol) {
        n = Vec3(vax / area, vay / area, vaz / area);
        t1 = Vec3(p01x, p01y, p01z);
        t1.normalize();
        cross(t2, n, t1);
        t2.normalize();
    } else {
        n = Vec3(1.0, 0.0, 0.0); t1 = Vec3(0.0, 1.0, 0.0); t2 = Vec3(0.0, 0.0, 1.0);
    }
}

// properties.d:177-237 (the outputs used here: area, centroid, iLen, jLen)
static void xyplane_quad_cell_properties(const Vec3& p0, const Vec3& p1, const Vec3& p2, const Vec3& p3,
                                         Vec3& centroid, double& xyplane_area, double& iLen, double& jLen) {
    double xA = p1.x, yA = p1.y, xB = p2.x, yB = p2.y, xC = p3.x, yC = p3.y, xD = p0.x, yD = p0.y;
    xyplane_area = 0.5 * ((xB + xA) * (yB - yA) + (xC + xB) * (yC - yB) + (xD + xC) * (yD - yC) +
                          (xA + xD) * (yA - yD));
    double x = 1.0 / (xyplane_area * 6.0) *
               ((yB - yA) * (xA * xA + xA * xB + xB * xB) + (yC - yB) * (xB * xB + xB * xC + xC * xC) +
                (yD - yC) * (xC * xC + xC * xD + xD * xD) + (yA - yD) * (xD * xD + xD * xA + xA * xA));
    double y = -1.0 / (xyplane_area * 6.0) *
               ((xB - xA) * (yA * yA + yA * yB + yB * yB) + (xC - xB) * (yB * yB + yB * yC + yC * yC) +
                (xD - xC) * (yC * yC + yC * yD + yD * yD) + (xA - xD) * (yD * yD + yD * yA + yA * yA));
    centroid = Vec3(x, y, 0.0);
    double xN = 0.5 * (xC + xB), yN = 0.5 * (yC + yB);
    double xS = 0.5 * (xD + xA), yS = 0.5 * (yD + yA);
    double xE = 0.5 * (xA + xB), yE = 0.5 * (yA + yB);
    double xW = 0.5 * (xD + xC), yW = 0.5 * (yD + yC);
    double dx = xN - xS, dy = yN - yS;
    jLen = std::sqrt(dx * dx + dy * dy);
    dx = xE - xW; dy = yE - yW;
    iLen = std::sqrt(dx * dx + dy * dy);
}

// properties.d:293-304
static double tetrahedron_volume(const Vec3& p0, const Vec3& p1, const Vec3& p2, const Vec3& p3) {
    double dx01 = p1.x - p0.x, dy01 = p1.y - p0.y, dz01 = p1.z - p0.z;
    double dx02 = p2.x - p0.x, dy02 = p2.y - p0.y, dz02 = p2.z - p0.z;
    double cx, cy, cz;
    cross_product(dx01, dy01, dz01, dx02, dy02, dz02, cx, cy, cz);
    double dx03 = p3.x - p0.x, dy03 = p3.y - p0.y, dz03 = p3.z - p0.z;
    // vector3.d:640 dot_product
    return (dx03 * cx + dy03 * cy + dz03 * cz) / 6.0;
}

// properties.d:390-429 (volume only; true_centroid is false by default, globalconfig.d:928)
static double pyramid_volume(const Vec3& p0, const Vec3& p1, const Vec3& p2, const Vec3& p3, const Vec3& p4) {
    Vec3 pmB(0.25 * (p0.x + p1.x + p2.x + p3.x), 0.25 * (p0.y + p1.y + p2.y + p3.y),
             0.25 * (p0.z + p1.z + p2.z + p3.z));
    double volume = 0.0;
    volume += tetrahedron_volume(p0, p1, pmB, p4);
    volume += tetrahedron_volume(p1, p2, pmB, p4);
    volume += tetrahedron_volume(p2, p3, pmB, p4);
    volume += tetrahedron_volume(p3, p0, pmB, p4);
    return volume;
}

// properties.d:590-680
static void hex_cell_properties(const Vec3 p[8], Vec3& centroid, double& volume, double& iLen, double& jLen,
                                double& kLen) {
    const Vec3 &p0 = p[0], &p1 = p[1], &p2 = p[2], &p3 = p[3], &p4 = p[4], &p5 = p[5], &p6 = p[6], &p7 = p[7];
    centroid = Vec3(0.125 * (p0.x + p1.x + p2.x + p3.x + p4.x + p5.x + p6.x + p7.x),
                    0.125 * (p0.y + p1.y + p2.y + p3.y + p4.y + p5.y + p6.y + p7.y),
                    0.125 * (p0.z + p1.z + p2.z + p3.z + p4.z + p5.z + p6.z + p7.z));
    Vec3 pmN(0.25 * (p3.x + p2.x + p6.x + p7.x), 0.25 * (p3.y + p2.y + p6.y + p7.y), 0.25 * (p3.z + p2.z + p6.z + p7.z));
    Vec3 pmE(0.25 * (p1.x + p2.x + p6.x + p5.x), 0.25 * (p1.y + p2.y + p6.y + p5.y), 0.25 * (p1.z + p2.z + p6.z + p5.z));
    Vec3 pmS(0.25 * (p0.x + p1.x + p5.x + p4.x), 0.25 * (p0.y + p1.y + p5.y + p4.y), 0.25 * (p0.z + p1.z + p5.z + p4.z));
    Vec3 pmW(0.25 * (p0.x + p3.x + p7.x + p4.x), 0.25 * (p0.y + p3.y + p7.y + p4.y), 0.25 * (p0.z + p3.z + p7.z + p4.z));
    Vec3 pmT(0.25 * (p4.x + p5.x + p6.x + p7.x), 0.25 * (p4.y + p5.y + p6.y + p7.y), 0.25 * (p4.z + p5.z + p6.z + p7.z));
    Vec3 pmB(0.25 * (p0.x + p1.x + p2.x + p3.x), 0.25 * (p0.y + p1.y + p2.y + p3.y), 0.25 * (p0.z + p1.z + p2.z + p3.z));
    double dx, dy, dz;
    dx = pmE.x - pmW.x; dy = pmE.y - pmW.y; dz = pmE.z - pmW.z;
    iLen = std::sqrt(dx * dx + dy * dy + dz * dz);
    dx = pmN.x - pmS.x; dy = pmN.y - pmS.y; dz = pmN.z - pmS.z;
    jLen = std::sqrt(dx * dx + dy * dy + dz * dz);
    dx = pmT.x - pmB.x; dy = pmT.y - pmB.y; dz = pmT.z - pmB.z;
    kLen = std::sqrt(dx * dx + dy * dy + dz * dz);
    volume = 0.0;
    volume += pyramid_volume(p6, p7, p3, p2, centroid);
    volume += pyramid_volume(p5, p6, p2, p1, centroid);
    volume += pyramid_volume(p4, p5, p1, p0, centroid);
    volume += pyramid_volume(p7, p4, p0, p3, centroid);
    volume += pyramid_volume(p7, p6, p5, p4, centroid);
    volume += pyramid_volume(p0, p1, p2, p3, centroid);
    const double smallButSignificantVolume = 1.0e-12;
    const double verySmallVolume = 1.0e-20;
    if ((volume < 0.0 && std::fabs(volume) < smallButSignificantVolume) ||
        (volume >= 0.0 && volume < verySmallVolume))
        volume = 0.0;
}

//---------------------------------------------------------------------------------
// SFluidBlock

FVCell* SFluidBlock::new_ghost() {
    FVCell* c = new FVCell();
    c->is_ghost = true;
    c->id = -1;
    c->fs.gas.massf.assign(gmodel->n_species, 0.0);
    c->fs.gas.rho_s.assign(gmodel->n_species, 0.0);
    ghost_cells.push_back(c);
    return c;
}

// sfluidblock.d:68-110, 443-700 (construction of cells, faces, ghost cells and stencils)
SFluidBlock::SFluidBlock(int id_, int nic_, int njc_, int nkc_, const Config& cfg_, const GasModel* gm,
                         const std::vector<Vec3>& vtx, const BoundaryCondition bcs[6])
    : id(id_), nic(nic_), njc(njc_), nkc(nkc_), cfg(cfg_), gmodel(gm) {
    niv = nic + 1;
    njv = njc + 1;
    if (cfg.dimensions == 2) {
        if (nkc != 1) throw FlowSolverException("Inconsistent dimensions and value for nkc.");
        nkv = 1;
    } else {
        nkv = nkc + 1;
    }
    cqi.init(cfg.dimensions, gm->n_species);
    vertices = vtx;
    for (int b = 0; b < 6; ++b) bc[b] = bcs[b];
    int nsp = gm->n_species;
    auto new_cell = [&]() {
        FVCell* c = new FVCell();
        c->fs.gas.massf.assign(nsp, 0.0);
        c->fs.gas.rho_s.assign(nsp, 0.0);
        for (auto& u : c->U) u.assign(cqi.n, 0.0);
        for (auto& d : c->dUdt) d.assign(cqi.n, 0.0);
        c->Q.assign(cqi.n, 0.0);
        return c;
    };
    for (int n = 0; n < nic * njc * nkc; ++n) { cells.push_back(new_cell()); cells.back()->id = n; }
    for (int n = 0; n < niv * njc * nkc; ++n) { auto f = new FVInterface(); f->idir = IndexDirection::i; faces.push_back(f); }
    for (int n = 0; n < nic * njv * nkc; ++n) { auto f = new FVInterface(); f->idir = IndexDirection::j; faces.push_back(f); }
    if (cfg.dimensions == 3)
        for (int n = 0; n < nic * njc * nkv; ++n) { auto f = new FVInterface(); f->idir = IndexDirection::k; faces.push_back(f); }
    for (size_t n = 0; n < faces.size(); ++n) { faces[n]->id = (int)n; faces[n]->F.assign(cqi.n, 0.0); }
    // Ghost cells, attached to boundary faces (sfluidblock.d:467-531); order N,S,E,W,T,B.
    const int ng = n_ghost_cell_layers;
    if (bc[3].ghost_cell_data_available)
        for (int k = 0; k < nkc; ++k) for (int i = 0; i < nic; ++i) { auto f = get_ifj(i, njc, k); for (int n = 0; n < ng; ++n) f->right_cells.push_back(new_ghost()); }
    if (bc[2].ghost_cell_data_available)
        for (int k = 0; k < nkc; ++k) for (int i = 0; i < nic; ++i) { auto f = get_ifj(i, 0, k); for (int n = 0; n < ng; ++n) f->left_cells.push_back(new_ghost()); }
    if (bc[1].ghost_cell_data_available)
        for (int k = 0; k < nkc; ++k) for (int j = 0; j < njc; ++j) { auto f = get_ifi(nic, j, k); for (int n = 0; n < ng; ++n) f->right_cells.push_back(new_ghost()); }
    if (bc[0].ghost_cell_data_available)
        for (int k = 0; k < nkc; ++k) for (int j = 0; j < njc; ++j) { auto f = get_ifi(0, j, k); for (int n = 0; n < ng; ++n) f->left_cells.push_back(new_ghost()); }
    if (cfg.dimensions == 3) {
        if (bc[5].ghost_cell_data_available)
            for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) { auto f = get_ifk(i, j, nkc); for (int n = 0; n < ng; ++n) f->right_cells.push_back(new_ghost()); }
        if (bc[4].ghost_cell_data_available)
            for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) { auto f = get_ifk(i, j, 0); for (int n = 0; n < ng; ++n) f->left_cells.push_back(new_ghost()); }
    }
    for (auto g : ghost_cells) {
        for (auto& u : g->U) u.assign(cqi.n, 0.0);
        for (auto& d : g->dUdt) d.assign(cqi.n, 0.0);
        g->Q.assign(cqi.n, 0.0);
    }
    // Stencils (sfluidblock.d:548-612).
    for (int k = 0; k < nkc; ++k) for (int j = 0; j < njc; ++j) for (int i = 1; i < niv; ++i) {
        auto fL = get_ifi(i, j, k);
        fL->left_cells.push_back(get_cell(i - 1, j, k));
        auto fL1 = get_ifi(i - 1, j, k);
        for (int n = 0; n < ng - 1; ++n) if ((int)fL1->left_cells.size() > n) fL->left_cells.push_back(fL1->left_cells[n]);
        int iR = niv - 1 - i;
        auto fR = get_ifi(iR, j, k);
        fR->right_cells.push_back(get_cell(iR, j, k));
        auto fR1 = get_ifi(iR + 1, j, k);
        for (int n = 0; n < ng - 1; ++n) if ((int)fR1->right_cells.size() > n) fR->right_cells.push_back(fR1->right_cells[n]);
    }
    for (int k = 0; k < nkc; ++k) for (int i = 0; i < nic; ++i) for (int j = 1; j < njv; ++j) {
        auto fL = get_ifj(i, j, k);
        fL->left_cells.push_back(get_cell(i, j - 1, k));
        auto fL1 = get_ifj(i, j - 1, k);
        for (int n = 0; n < ng - 1; ++n) if ((int)fL1->left_cells.size() > n) fL->left_cells.push_back(fL1->left_cells[n]);
        int jR = njv - 1 - j;
        auto fR = get_ifj(i, jR, k);
        fR->right_cells.push_back(get_cell(i, jR, k));
        auto fR1 = get_ifj(i, jR + 1, k);
        for (int n = 0; n < ng - 1; ++n) if ((int)fR1->right_cells.size() > n) fR->right_cells.push_back(fR1->right_cells[n]);
    }
    if (cfg.dimensions == 3)
        for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) for (int k = 1; k < nkv; ++k) {
            auto fL = get_ifk(i, j, k);
            fL->left_cells.push_back(get_cell(i, j, k - 1));
            auto fL1 = get_ifk(i, j, k - 1);
            for (int n = 0; n < ng - 1; ++n) if ((int)fL1->left_cells.size() > n) fL->left_cells.push_back(fL1->left_cells[n]);
            int kR = nkv - 1 - k;
            auto fR = get_ifk(i, j, kR);
            fR->right_cells.push_back(get_cell(i, j, kR));
            auto fR1 = get_ifk(i, j, kR + 1);
            for (int n = 0; n < ng - 1; ++n) if ((int)fR1->right_cells.size() > n) fR->right_cells.push_back(fR1->right_cells[n]);
        }
    // Boundary face lists and outsigns (sfluidblock.d:620-674).
    for (int k = 0; k < nkc; ++k) for (int i = 0; i < nic; ++i) {
        auto f = get_ifj(i, njc, k); bc[3].faces.push_back(f); bc[3].outsigns.push_back(1); f->i_bndry = (int)bc[3].outsigns.size() - 1;
        f->is_on_boundary = true; f->bc_id = 3;
    }
    for (int k = 0; k < nkc; ++k) for (int j = 0; j < njc; ++j) {
        auto f = get_ifi(nic, j, k); bc[1].faces.push_back(f); bc[1].outsigns.push_back(1); f->i_bndry = (int)bc[1].outsigns.size() - 1;
        f->is_on_boundary = true; f->bc_id = 1;
    }
    for (int k = 0; k < nkc; ++k) for (int i = 0; i < nic; ++i) {
        auto f = get_ifj(i, 0, k); bc[2].faces.push_back(f); bc[2].outsigns.push_back(-1); f->i_bndry = (int)bc[2].outsigns.size() - 1;
        f->is_on_boundary = true; f->bc_id = 2;
    }
    for (int k = 0; k < nkc; ++k) for (int j = 0; j < njc; ++j) {
        auto f = get_ifi(0, j, k); bc[0].faces.push_back(f); bc[0].outsigns.push_back(-1); f->i_bndry = (int)bc[0].outsigns.size() - 1;
        f->is_on_boundary = true; f->bc_id = 0;
    }
    if (cfg.dimensions == 3) {
        for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) {
            auto f = get_ifk(i, j, nkc); bc[5].faces.push_back(f); bc[5].outsigns.push_back(1); f->i_bndry = (int)bc[5].outsigns.size() - 1;
            f->is_on_boundary = true; f->bc_id = 5;
        }
        for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) {
            auto f = get_ifk(i, j, 0); bc[4].faces.push_back(f); bc[4].outsigns.push_back(-1); f->i_bndry = (int)bc[4].outsigns.size() - 1;
            f->is_on_boundary = true; f->bc_id = 4;
        }
    }
    // Cell-face binding (sfluidblock.d:678-703).
    for (int k = 0; k < nkc; ++k) for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) {
        auto c = get_cell(i, j, k);
        int nf = (cfg.dimensions == 3) ? 6 : 4;
        c->iface.assign(nf, nullptr);
        c->outsign.assign(nf, 0);
        c->iface[0] = get_ifi(i, j, k); c->outsign[0] = -1;
        c->iface[1] = get_ifi(i + 1, j, k); c->outsign[1] = 1;
        c->iface[2] = get_ifj(i, j, k); c->outsign[2] = -1;
        c->iface[3] = get_ifj(i, j + 1, k); c->outsign[3] = 1;
        if (cfg.dimensions == 3) {
            c->iface[4] = get_ifk(i, j, k); c->outsign[4] = -1;
            c->iface[5] = get_ifk(i, j, k + 1); c->outsign[5] = 1;
        }
    }
    compute_geometry();
}

SFluidBlock::~SFluidBlock() {
    for (auto c : cells) delete c;
    for (auto c : ghost_cells) delete c;
    for (auto f : faces) delete f;
}

// sfluidblock.d:865-1000 (compute_primary_cell_geometric_data), fvcell.d:396-480,
// fvinterface.d:301-350.  Ghost cells receive mirrored cell lengths (sfluidblock.d:884-).
void SFluidBlock::compute_geometry() {
    auto V = [&](int i, int j, int k) -> const Vec3& { return vertices[vertex_index(i, j, k)]; };
    if (cfg.dimensions == 2) {
        for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) {
            auto c = get_cell(i, j);
            double xyarea, iL, jL;
            xyplane_quad_cell_properties(V(i, j, 0), V(i + 1, j, 0), V(i + 1, j + 1, 0), V(i, j + 1, 0), c->pos, xyarea,
                                         iL, jL);
            if (xyarea < 0.0) throw FlowSolverException("Negative cell volume");
            c->volume = xyarea;
            c->iLength = iL; c->jLength = jL; c->kLength = 0.0;
        }
        // faces: fvinterface.d:301-332, vertex order from sfluidblock.d:737-796
        auto face2d = [&](FVInterface* f, const Vec3& A, const Vec3& B) {
            double xA = A.x, yA = A.y, xB = B.x, yB = B.y;
            double LAB = std::sqrt((xB - xA) * (xB - xA) + (yB - yA) * (yB - yA));
            if (LAB > 1.0e-12) {
                f->n = Vec3((yB - yA) / LAB, -(xB - xA) / LAB, 0.0);
                f->t2 = Vec3(0.0, 0.0, 1.0);
                cross(f->t1, f->n, f->t2);
                f->length = LAB;
            } else {
                f->n = Vec3(1.0, 0.0, 0.0); f->t2 = Vec3(0.0, 0.0, 1.0); f->t1 = Vec3(0.0, 1.0, 0.0);
                f->length = 0.0;
            }
            f->area = f->length;
            f->pos = Vec3(0.5 * (xA + xB), 0.5 * (yA + yB), 0.0);
        };
        for (int j = 0; j < njc; ++j) for (int i = 0; i < niv; ++i) face2d(get_ifi(i, j), V(i, j, 0), V(i, j + 1, 0));
        for (int j = 0; j < njv; ++j) for (int i = 0; i < nic; ++i) face2d(get_ifj(i, j), V(i + 1, j, 0), V(i, j, 0));
    } else {
        for (int k = 0; k < nkc; ++k) for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) {
            auto c = get_cell(i, j, k);
            Vec3 p[8] = {V(i, j, k), V(i + 1, j, k), V(i + 1, j + 1, k), V(i, j + 1, k),
                         V(i, j, k + 1), V(i + 1, j, k + 1), V(i + 1, j + 1, k + 1), V(i, j + 1, k + 1)};
            double vol, iL, jL, kL;
            hex_cell_properties(p, c->pos, vol, iL, jL, kL);
            if (vol <= 0.0) throw FlowSolverException("Invalid volume");
            c->volume = vol; c->iLength = iL; c->jLength = jL; c->kLength = kL;
        }
        for (int k = 0; k < nkc; ++k) for (int j = 0; j < njc; ++j) for (int i = 0; i < niv; ++i) {
            auto f = get_ifi(i, j, k);
            quad_properties(V(i, j, k), V(i, j + 1, k), V(i, j + 1, k + 1), V(i, j, k + 1), f->pos, f->n, f->t1, f->t2, f->area);
            f->length = std::sqrt(f->area);
        }
        for (int k = 0; k < nkc; ++k) for (int j = 0; j < njv; ++j) for (int i = 0; i < nic; ++i) {
            auto f = get_ifj(i, j, k);
            quad_properties(V(i, j, k), V(i, j, k + 1), V(i + 1, j, k + 1), V(i + 1, j, k), f->pos, f->n, f->t1, f->t2, f->area);
            f->length = std::sqrt(f->area);
        }
        for (int k = 0; k < nkv; ++k) for (int j = 0; j < njc; ++j) for (int i = 0; i < nic; ++i) {
            auto f = get_ifk(i, j, k);
            quad_properties(V(i, j, k), V(i + 1, j, k), V(i + 1, j + 1, k), V(i, j + 1, k), f->pos, f->n, f->t1, f->t2, f->area);
            f->length = std::sqrt(f->area);
        }
    }
    // Ghost cell lengths: copy_values_from(cell, cell_lengths_only) of the mirror-image
    // interior cell (sfluidblock.d:884-1000; fvcell.d:300-330).
    auto copy_len = [](FVCell* g, FVCell* c) { g->iLength = c->iLength; g->jLength = c->jLength; g->kLength = c->kLength; };
    const int ng = n_ghost_cell_layers;
    if (bc[1].ghost_cell_data_available)
        for (int j = 0; j < njc; ++j) for (int k = 0; k < nkc; ++k) { auto f = get_ifi(nic, j, k); for (int n = 0; n < ng; ++n) copy_len(f->right_cells[n], get_cell(nic - 1 - n, j, k)); }
    if (bc[0].ghost_cell_data_available)
        for (int j = 0; j < njc; ++j) for (int k = 0; k < nkc; ++k) { auto f = get_ifi(0, j, k); for (int n = 0; n < ng; ++n) copy_len(f->left_cells[n], get_cell(n, j, k)); }
    if (bc[3].ghost_cell_data_available)
        for (int i = 0; i < nic; ++i) for (int k = 0; k < nkc; ++k) { auto f = get_ifj(i, njc, k); for (int n = 0; n < ng; ++n) copy_len(f->right_cells[n], get_cell(i, njc - 1 - n, k)); }
    if (bc[2].ghost_cell_data_available)
        for (int i = 0; i < nic; ++i) for (int k = 0; k < nkc; ++k) { auto f = get_ifj(i, 0, k); for (int n = 0; n < ng; ++n) copy_len(f->left_cells[n], get_cell(i, n, k)); }
    if (cfg.dimensions == 3) {
        if (bc[5].ghost_cell_data_available)
            for (int i = 0; i < nic; ++i) for (int j = 0; j < njc; ++j) { auto f = get_ifk(i, j, nkc); for (int n = 0; n < ng; ++n) copy_len(f->right_cells[n], get_cell(i, j, nkc - 1 - n)); }
        if (bc[4].ghost_cell_data_available)
            for (int i = 0; i < nic; ++i) for (int j = 0; j < njc; ++j) { auto f = get_ifk(i, j, 0); for (int n = 0; n < ng; ++n) copy_len(f->left_cells[n], get_cell(i, j, n)); }
    }
}

// fvcell.d:511-583
void SFluidBlock::encode_conserved(FVCell* c, int ftl) {
    auto& U = c->U[ftl];
    const FlowState& fs = c->fs;
    U[cqi.mass] = fs.gas.rho;
    U[cqi.xMom] = fs.gas.rho * fs.vel.x;
    U[cqi.yMom] = fs.gas.rho * fs.vel.y;
    if (cqi.threeD) U[cqi.zMom] = fs.gas.rho * fs.vel.z;
    double u = gmodel->internal_energy(fs.gas);
    double ke = 0.5 * (fs.vel.x * fs.vel.x + fs.vel.y * fs.vel.y + fs.vel.z * fs.vel.z);
    U[cqi.totEnergy] = fs.gas.rho * (u + ke);
    if (cqi.n_species > 1)
        for (int isp = 0; isp < cqi.n_species; ++isp) U[cqi.species + isp] = fs.gas.rho * fs.gas.massf[isp];
}

// fvcell.d:586-821 (transient ordering: cqi.mass == 0; default options)
int SFluidBlock::decode_conserved(FVCell* c, int ftl) {
    auto& U = c->U[ftl];
    FlowState& fs = c->fs;
    double rho = U[cqi.mass];
    if (!(rho > 0.0)) throw FlowSolverException("Bad cell with negative mass.");
    fs.gas.rho = rho;
    double dinv = 1.0 / rho;
    double zMom = cqi.threeD ? U[cqi.zMom] : 0.0;
    fs.vel = Vec3(U[cqi.xMom] * dinv, U[cqi.yMom] * dinv, zMom * dinv);
    double rE = U[cqi.totEnergy];
    double u = rE * dinv;
    double ke = 0.5 * (fs.vel.x * fs.vel.x + fs.vel.y * fs.vel.y + fs.vel.z * fs.vel.z);
    u -= ke;
    fs.gas.u = u;
    if (cqi.n_species > 1) {
        // fvcell.d:716-738 (n_species > 1 && cqi.mass == 0), tolerance=0, assert tol 0.1
        double rhos_sum = 0.0;
        for (int isp = 0; isp < cqi.n_species; ++isp) {
            if (U[cqi.species + isp] < 0.0) U[cqi.species + isp] = 0.0;
            rhos_sum += U[cqi.species + isp];
        }
        if (std::fabs(rhos_sum - rho) > 0.1) throw FlowSolverException("Sum of species densities far from total density");
        if (std::fabs(rhos_sum - rho) > 0.0) {
            double scale_factor = rho / rhos_sum;
            for (int isp = 0; isp < cqi.n_species; ++isp) U[cqi.species + isp] *= scale_factor;
        }
        for (int isp = 0; isp < cqi.n_species; ++isp) {
            fs.gas.massf[isp] = U[cqi.species + isp] * dinv;
            fs.gas.rho_s[isp] = U[cqi.species + isp];
        }
    } else {
        fs.gas.massf[0] = 1.0;
    }
    try {
        try {
            gmodel->update_thermo_from_rhou(fs.gas);
        } catch (GasModelException&) {
            if (cfg.ignore_low_T_thermo_update_failure && (rho > 0.0)) {
                fs.gas.T = cfg.suggested_low_T_value;
                gmodel->update_thermo_from_rhoT(fs.gas);
                encode_conserved(c, ftl);
            } else {
                throw;
            }
        }
        if (fs.gas.T <= 0.0) throw FlowSolverException("update_thermo returned negative temperature.");
        gmodel->update_sound_speed(fs.gas);
    } catch (GasModelException&) {
        throw FlowSolverException("Bad cell with failed thermodynamic update.");
    }
    return 0;
}

// fvcell.d:824-854
void SFluidBlock::time_derivatives(FVCell* c, int ftl) {
    auto& d = c->dUdt[ftl];
    double vol_inv = 1.0 / c->volume;
    for (int j = 0; j < cqi.n; ++j) {
        double surface_integral = 0.0;
        for (size_t i = 0; i < c->iface.size(); ++i) {
            double area = c->outsign[i] * c->iface[i]->area;
            surface_integral -= c->iface[i]->F[j] * area;
        }
        d[j] = vol_inv * surface_integral + c->Q[j];
    }
}

// fvcell.d:975-1058 (inviscid, not stringent, structured grid)
double SFluidBlock::signal_frequency(FVCell* c) {
    double signal = 0.0;
    size_t first_face = 3, second_face = 1, third_face = c->iface.size() - 1;
    double un_N = std::fabs(c->fs.vel.dot(c->iface[first_face]->n));
    double un_E = std::fabs(c->fs.vel.dot(c->iface[second_face]->n));
    double un_T = (cfg.dimensions == 3) ? std::fabs(c->fs.vel.dot(c->iface[third_face]->n)) : 0.0;
    double signalN = (un_N + c->fs.gas.a) / c->jLength;
    signal = std::fmax(signal, signalN);
    double signalE = (un_E + c->fs.gas.a) / c->iLength;
    signal = std::fmax(signal, signalE);
    if (cfg.dimensions == 3) {
        double signalT = (un_T + c->fs.gas.a) / c->kLength;
        signal = std::fmax(signal, signalT);
    }
    return signal;
}

// fluidblock.d:987-1084 (no STS, no LTS)
void SFluidBlock::determine_time_step_size(double dt_current, double cfl_value, bool check_cfl, double out[3]) {
    double cfl_allow = 1.2;  // two-stage explicit (fluidblock.d:1011)
    double cfl_adjust = 0.5;
    bool first = true;
    double signal = 0.0, cfl_local, dt_local, cfl_min = 0.0, cfl_max = 0.0, dt_allow = 0.0;
    for (auto c : cells) {
        signal = signal_frequency(c);
        cfl_local = dt_current * signal;
        dt_local = cfl_value / signal;
        if (first) {
            cfl_min = cfl_local; cfl_max = cfl_local; dt_allow = dt_local; first = false;
        } else {
            cfl_min = std::fmin(cfl_min, cfl_local);
            cfl_max = std::fmax(cfl_max, cfl_local);
            dt_allow = std::fmin(dt_allow, dt_local);
        }
    }
    if (check_cfl && (cfl_max < 0.0 || cfl_max > cfl_allow)) {
        cfl_max = cfl_adjust * cfl_allow;
        dt_allow = cfl_max / signal;  // 'signal' of the last cell, as in the reference
    }
    out[0] = dt_allow; out[1] = cfl_max; out[2] = 0.0;
}

// fluidblock.d:772-776
void SFluidBlock::clear_fluxes() {
    for (auto f : faces) std::fill(f->F.begin(), f->F.end(), 0.0);
}

// bc/ghost_cell_effect/ghost_cell.d:33-43
static void reflect_normal_velocity(FlowState& fs, const FVInterface* f) {
    fs.vel.to_local(f->n, f->t1, f->t2);
    fs.vel.x = -(fs.vel.x);
    fs.vel.to_global(f->n, f->t1, f->t2);
}

// fluidblock.d:1086-1096 -> ghost-cell effects applied per boundary in bc[] order.
void SFluidBlock::apply_pre_recon(std::vector<SFluidBlock*>& all_blocks) {
    const int ng = n_ghost_cell_layers;
    for (int b = 0; b < ((cfg.dimensions == 3) ? 6 : 4); ++b) {
        BoundaryCondition& B = bc[b];
        switch (B.kind) {
        case BCKind::slip_wall:
            // internal_copy_then_reflect.d:111-134
            for (size_t i = 0; i < B.faces.size(); ++i) {
                FVInterface* f = B.faces[i];
                for (int n = 0; n < ng; ++n) {
                    FVCell *src, *dest;
                    if (B.outsigns[i] == 1) { src = f->left_cells[n]; dest = f->right_cells[n]; }
                    else { src = f->right_cells[n]; dest = f->left_cells[n]; }
                    dest->fs = src->fs;
                    reflect_normal_velocity(dest->fs, f);
                }
            }
            break;
        case BCKind::inflow_supersonic:
            // flow_state_copy.d:92-107
            for (size_t i = 0; i < B.faces.size(); ++i) {
                FVInterface* f = B.faces[i];
                for (int n = 0; n < ng; ++n) {
                    FVCell* ghost = (B.outsigns[i] == 1) ? f->right_cells[n] : f->left_cells[n];
                    ghost->fs = B.inflow;
                }
            }
            break;
        case BCKind::outflow_simple_flux:
        case BCKind::outflow_simple_extrapolate:
            // extrapolate_copy.d:124-140 (xOrder == 0)
            for (size_t i = 0; i < B.faces.size(); ++i) {
                FVInterface* f = B.faces[i];
                for (int n = 0; n < ng; ++n) {
                    FVCell *src, *dest;
                    if (B.outsigns[i] == 1) { src = f->left_cells[0]; dest = f->right_cells[n]; }
                    else { src = f->right_cells[0]; dest = f->left_cells[n]; }
                    dest->fs = src->fs;
                }
            }
            break;
        case BCKind::exchange:
            break;  // filled by Simulation::exchange_ghost_cell_boundary_data()
        }
    }
    (void)all_blocks;
}

// onedinterp.d:338-354
void L2R2::prepare(double lenL1, double lenL0, double lenR0, double lenR1) {
    lenL0_ = lenL0;
    lenR0_ = lenR0;
    aL0 = 0.5 * lenL0 / (lenL1 + 2.0 * lenL0 + lenR0);
    aR0 = 0.5 * lenR0 / (lenL0 + 2.0 * lenR0 + lenR1);
    two_over_lenL0_plus_lenL1 = 2.0 / (lenL0 + lenL1);
    two_over_lenR0_plus_lenL0 = 2.0 / (lenR0 + lenL0);
    two_over_lenR1_plus_lenR0 = 2.0 / (lenR1 + lenR0);
    two_lenL0_plus_lenL1 = (2.0 * lenL0 + lenL1);
    two_lenR0_plus_lenR1 = (2.0 * lenR0 + lenR1);
}

// onedinterp.d:357-384 (beta == 1: apply_heuristic_pressure_based_limiting is off by default)
void L2R2::scalar(double qL1, double qL0, double qR0, double qR1, double& qL, double& qR, bool apply_limiter,
                  bool extrema_clipping, double epsilon_van_albada, bool lmr) const {
    const double beta = 1.0;
    double delLminus = (qL0 - qL1) * two_over_lenL0_plus_lenL1;
    double del = (qR0 - qL0) * two_over_lenR0_plus_lenL0;
    double delRplus = (qR1 - qR0) * two_over_lenR1_plus_lenR0;
    if (lmr) {
        // lmr/onedinterp.d:147-151: epsilon dimensionalised by the larger cell value.
        double qqL = std::fmax(1e-12, std::fabs(qL0));
        double qqR = std::fmax(1e-12, std::fabs(qR0));
        double qq = std::fmax(qqL, qqR);
        epsilon_van_albada = qq * epsilon_van_albada * two_over_lenR0_plus_lenL0;
    }
    double sL = 1.0, sR = 1.0;
    if (apply_limiter) {
        sL = (delLminus * del + std::fabs(delLminus * del) + epsilon_van_albada) /
             (delLminus * delLminus + del * del + epsilon_van_albada);
        sR = (del * delRplus + std::fabs(del * delRplus) + epsilon_van_albada) /
             (del * del + delRplus * delRplus + epsilon_van_albada);
    }
    qL = qL0 + beta * sL * aL0 * (del * two_lenL0_plus_lenL1 + delLminus * lenR0_);
    qR = qR0 - beta * sR * aR0 * (delRplus * lenL0_ + del * two_lenR0_plus_lenR1);
    if (extrema_clipping) {
        qL = clip_to_limits(qL, qL0, qR0);
        qR = clip_to_limits(qR, qL0, qR0);
    }
}

// onedinterp.d:751-988 (thermo_interpolator == rhou, the default)
void SFluidBlock::interp_l2r2(FVInterface* f, FVCell* cL1, FVCell* cL0, FVCell* cR0, FVCell* cR1, double lL1,
                              double lL0, double lR0, double lR1, FlowState& L, FlowState& R) {
    int nsp = gmodel->n_species;
    // The reference rotates the stencil cells' velocities in place and back again
    // (onedinterp.d:761-770, 983-986).  With mutate_cell_velocities == false the
    // rotation acts on copies, leaving the cells untouched.
    Vec3 vL1 = cL1->fs.vel, vL0 = cL0->fs.vel, vR0 = cR0->fs.vel, vR1 = cR1->fs.vel;
    Vec3 *pL1 = &vL1, *pL0 = &vL0, *pR0 = &vR0, *pR1 = &vR1;
    if (cfg.mutate_cell_velocities) { pL1 = &cL1->fs.vel; pL0 = &cL0->fs.vel; pR0 = &cR0->fs.vel; pR1 = &cR1->fs.vel; }
    if (cfg.interpolate_in_local_frame) {
        pL1->to_local(f->n, f->t1, f->t2);
        pL0->to_local(f->n, f->t1, f->t2);
        pR0->to_local(f->n, f->t1, f->t2);
        pR1->to_local(f->n, f->t1, f->t2);
    }
    l2r2.prepare(lL1, lL0, lR0, lR1);
    interp_l2r2_scalar(pL1->x, pL0->x, pR0->x, pR1->x, L.vel.x, R.vel.x);
    interp_l2r2_scalar(pL1->y, pL0->y, pR0->y, pR1->y, L.vel.y, R.vel.y);
    interp_l2r2_scalar(pL1->z, pL0->z, pR0->z, pR1->z, L.vel.z, R.vel.z);
    const GasState &gL1 = cL1->fs.gas, &gL0 = cL0->fs.gas, &gR0 = cR0->fs.gas, &gR1 = cR1->fs.gas;
    if (nsp > 1 && cfg.lmr) {
        // lmr/onedinterp.d:268-286, 314-315: Eilmer 5 reconstructs the mass fractions
        // themselves, scales them, then the density (thermo interpolation rhou).  The face
        // states' rho_s keep the copied cell values, which the flux does not read.
        if (cfg.allow_reconstruction_for_species) {
            for (int isp = 0; isp < nsp; ++isp)
                interp_l2r2_scalar(gL1.massf[isp], gL0.massf[isp], gR0.massf[isp], gR1.massf[isp], L.gas.massf[isp],
                                   R.gas.massf[isp]);
            if (cfg.scale_species_after_reconstruction) {
                scale_mass_fractions(L.gas.massf);
                scale_mass_fractions(R.gas.massf);
            }
        } else {
            L.gas.massf = gL0.massf;
            R.gas.massf = gR0.massf;
        }
        interp_l2r2_scalar(gL1.rho, gL0.rho, gR0.rho, gR1.rho, L.gas.rho, R.gas.rho);
    } else if (nsp > 1) {
        if (cfg.allow_reconstruction_for_species) {
            for (int isp = 0; isp < nsp; ++isp)
                interp_l2r2_scalar(gL1.rho_s[isp], gL0.rho_s[isp], gR0.rho_s[isp], gR1.rho_s[isp], L.gas.rho_s[isp],
                                   R.gas.rho_s[isp]);
        } else {
            L.gas.rho_s = gL0.rho_s;
            R.gas.rho_s = gR0.rho_s;
        }
        double rho_L = 0.0, rho_R = 0.0;
        for (int isp = 0; isp < nsp; ++isp) { rho_L += L.gas.rho_s[isp]; rho_R += R.gas.rho_s[isp]; }
        L.gas.rho = rho_L;
        R.gas.rho = rho_R;
        for (int isp = 0; isp < nsp; ++isp) {
            L.gas.massf[isp] = L.gas.rho_s[isp] / L.gas.rho;
            R.gas.massf[isp] = R.gas.rho_s[isp] / R.gas.rho;
        }
        if (cfg.scale_species_after_reconstruction) {
            scale_mass_fractions(L.gas.massf);
            scale_mass_fractions(R.gas.massf);
        }
    } else {
        interp_l2r2_scalar(gL1.rho, gL0.rho, gR0.rho, gR1.rho, L.gas.rho, R.gas.rho);
    }
    interp_l2r2_scalar(gL1.u, gL0.u, gR0.u, gR1.u, L.gas.u, R.gas.u);
    // codeForThermoUpdateBoth("rhou"), onedinterp.d:45-68: on failure, copy the cell state.
    // Note that the sound speed is NOT updated here: Lft.gas.a keeps cL0's value.
    try { gmodel->update_thermo_from_rhou(L.gas); }
    catch (std::exception&) { L = cL0->fs; if (!cfg.mutate_cell_velocities) L.vel = *pL0; }
    try { gmodel->update_thermo_from_rhou(R.gas); }
    catch (std::exception&) { R = cR0->fs; if (!cfg.mutate_cell_velocities) R.vel = *pR0; }
    if (cfg.interpolate_in_local_frame) {
        L.vel.to_global(f->n, f->t1, f->t2);
        R.vel.to_global(f->n, f->t1, f->t2);
        pL1->to_global(f->n, f->t1, f->t2);
        pL0->to_global(f->n, f->t1, f->t2);
        pR0->to_global(f->n, f->t1, f->t2);
        pR1->to_global(f->n, f->t1, f->t2);
    }
}

// onedinterp.d:117-273 (interpolation_order 2, both stencils of length >= 2)
void SFluidBlock::interp(FVInterface* f, FlowState& L, FlowState& R) {
    FVCell* cL0 = (f->left_cells.size() > 0) ? f->left_cells[0] : f->right_cells[0];
    FVCell* cR0 = (f->right_cells.size() > 0) ? f->right_cells[0] : f->left_cells[0];
    L = cL0->fs;
    R = cR0->fs;
    if (f->left_cells.size() >= 2 && f->right_cells.size() >= 2) {
        FVCell* cL1 = f->left_cells[1];
        FVCell* cR1 = f->right_cells[1];
        double lL0, lL1, lR0, lR1;
        switch (f->idir) {
        case IndexDirection::i: lL0 = cL0->iLength; lL1 = cL1->iLength; lR0 = cR0->iLength; lR1 = cR1->iLength; break;
        case IndexDirection::j: lL0 = cL0->jLength; lL1 = cL1->jLength; lR0 = cR0->jLength; lR1 = cR1->jLength; break;
        default: lL0 = cL0->kLength; lL1 = cL1->kLength; lR0 = cR0->kLength; lR1 = cR1->kLength; break;
        }
        interp_l2r2(f, cL1, cL0, cR0, cR1, lL1, lL0, lR0, lR1, L, R);
    } else {
        throw FlowSolverException("Stencils not suitable for the restated interpolation (ghost cells required).");
    }
}

// fluxcalc.d:54-184 (fixed grid: gvel == 0, omegaz == 0)
void SFluidBlock::compute_interface_flux(FlowState& L, FlowState& R, FVInterface* f) {
    // Subtracting a zero grid velocity and the gvel energy terms leave values unchanged.
    L.vel.to_local(f->n, f->t1, f->t2);
    R.vel.to_local(f->n, f->t1, f->t2);
    switch (cfg.flux_calculator) {
    case FluxCalc::ausmdv: ausmdv(L, R, f->F, cqi, *gmodel, cfg.apply_entropy_fix, 1.0, cfg.lmr); break;
    case FluxCalc::efm: efmflx(L, R, f->F, cqi, *gmodel); break;
    case FluxCalc::adaptive_efm_ausmdv: {
        // fluxcalc.d:1315-1332: EFM weighted by the face's shock indicator, AUSMDV by the rest.
        double alpha = f->S;
        if (alpha > 0.0) efmflx(L, R, f->F, cqi, *gmodel, alpha);
        if (alpha < 1.0) ausmdv(L, R, f->F, cqi, *gmodel, cfg.apply_entropy_fix, 1.0 - alpha, cfg.lmr);
        break;
    }
    case FluxCalc::hanel: hanel(L, R, f->F, cqi, *gmodel); break;
    case FluxCalc::adaptive_hanel_ausmdv: {
        // fluxcalc.d:1335-1348
        double alpha = f->S;
        if (alpha > 0.0) hanel(L, R, f->F, cqi, *gmodel, alpha);
        if (alpha < 1.0) ausmdv(L, R, f->F, cqi, *gmodel, cfg.apply_entropy_fix, 1.0 - alpha, cfg.lmr);
        break;
    }
    }
    auto& F = f->F;
    // Energy flux: F[E] += 0.5*F[mass]*v_sqr + (F[x]*gvx + F[y]*gvy + F[z]*gvz) with gvel = 0.
    double gx = 0.0, gy = 0.0, gz = 0.0;
    double v_sqr = gx * gx + gy * gy + gz * gz;
    F[cqi.totEnergy] += 0.5 * F[cqi.mass] * v_sqr +
                        (F[cqi.xMom] * gx + F[cqi.yMom] * gy + (cqi.threeD ? F[cqi.zMom] * gz : 0.0));
    F[cqi.xMom] += gx * F[cqi.mass];
    F[cqi.yMom] += gy * F[cqi.mass];
    double zDummy = 0.0;
    double& Fz = cqi.threeD ? F[cqi.zMom] : zDummy;
    if (cqi.threeD) F[cqi.zMom] += gz * F[cqi.mass];
    // vector3.d:683-692
    double ax = F[cqi.xMom], ay = F[cqi.yMom], az = Fz;
    double vx = ax * f->n.x + ay * f->t1.x + az * f->t2.x;
    double vy = ax * f->n.y + ay * f->t1.y + az * f->t2.y;
    double vz = ax * f->n.z + ay * f->t1.z + az * f->t2.z;
    F[cqi.xMom] = vx; F[cqi.yMom] = vy; Fz = vz;
}

// shockdetectors.d:21-113: compression and shear between the face's two cells, in the
// face frame and relative to the sound speeds; 1 where the compression is stronger than
// comp_tol and the shear weaker than shear_tol.
double PJ_shock_detector(const FVInterface& f, double comp_tol, double shear_tol) {
    const FVCell* cL = f.left_cell();
    const FVCell* cR = f.right_cell();
    double comp, shear;
    if (cL && cR) {
        double uL = cL->fs.vel.dot(f.n);
        double uR = cR->fs.vel.dot(f.n);
        double aL = cL->fs.gas.a;
        double aR = cR->fs.gas.a;
        double a_min = (aL < aR) ? aL : aR;
        comp = ((uR - uL) / a_min);
        double vL = cL->fs.vel.dot(f.t1);
        double vR = cR->fs.vel.dot(f.t1);
        double wL = cL->fs.vel.dot(f.t2);
        double wR = cR->fs.vel.dot(f.t2);
        double sound_speed = 0.5 * (aL + aR);
        double shear_y = std::fabs(vL - vR) / sound_speed;
        double shear_z = std::fabs(wL - wR) / sound_speed;
        shear = std::fmax(shear_y, shear_z);
    } else if (cL) {
        // A wall on the right; the grid velocity is zero here.
        double uL = cL->fs.vel.dot(f.n);
        double aL = cL->fs.gas.a;
        comp = ((-uL) / aL);
        double vL = cL->fs.vel.dot(f.t1);
        double wL = cL->fs.vel.dot(f.t2);
        shear = std::fmax(std::fabs(vL) / aL, std::fabs(wL) / aL);
    } else if (cR) {
        double uR = cR->fs.vel.dot(f.n);
        double aR = cR->fs.gas.a;
        comp = (uR / aR);
        double vR = cR->fs.vel.dot(f.t1);
        double wR = cR->fs.vel.dot(f.t2);
        shear = std::fmax(std::fabs(vR) / aR, std::fabs(wR) / aR);
    } else {
        return 0.0;
    }
    if ((shear < shear_tol) && (comp < comp_tol)) return 1.0;
    return 0.0;
}

// fluidblock.d:478-520 (the PJ detector; damped_outflow marks the faces of
// outflow_simple_extrapolate boundaries)
void SFluidBlock::detect_shock_points() {
    double comp_tol = cfg.compression_tolerance;
    double shear_tol = cfg.shear_tolerance;
    if (comp_tol > 0.0) throw std::runtime_error("compression_tolerance should be negative!");
    for (auto f : faces) f->S = PJ_shock_detector(*f, comp_tol, shear_tol);
    if (cfg.damped_outflow) {
        for (int b = 0; b < ((cfg.dimensions == 3) ? 6 : 4); ++b)
            if (bc[b].kind == BCKind::outflow_simple_extrapolate)
                for (auto f : bc[b].faces) f->S = 1;
    }
}

// fluidblock.d:568-580: a cell is shocked if any of its faces is.
void SFluidBlock::shock_faces_to_cells() {
    for (auto c : cells) {
        c->fs.S = 0.0;
        for (auto f : c->iface) c->fs.S = std::fmax(c->fs.S, f->S);
    }
}

// fluidblock.d:582-605: a face is shocked if it or either of its cells is.  Ghost cells keep
// the indicator they were given by the last ghost-cell fill.
void SFluidBlock::enforce_strict_shock_detector() {
    for (auto f : faces) {
        if (f->S > 0.0) {
            f->S = 1.0;
            continue;
        }
        if (f->left_cell() && f->left_cell()->fs.S > 0.0) {
            f->S = 1.0;
            continue;
        }
        if (f->right_cell() && f->right_cell()->fs.S > 0.0) {
            f->S = 1.0;
            continue;
        }
    }
}

// sfluidblock.d:2188-2245
void SFluidBlock::convective_flux_phase0(bool allow_high_order) {
    bool do_reconstruction = allow_high_order && (cfg.interpolation_order > 1);
    for (auto f : faces) {
        if (do_reconstruction) {
            interp(f, Lft, Rght);
        } else {
            FVCell* cL0 = (f->left_cells.size() > 0) ? f->left_cells[0] : f->right_cells[0];
            FVCell* cR0 = (f->right_cells.size() > 0) ? f->right_cells[0] : f->left_cells[0];
            Lft = cL0->fs;
            Rght = cR0->fs;
        }
        if (f->is_on_boundary && bc[f->bc_id].convective_flux_computed_in_bc) continue;
        compute_interface_flux(Lft, Rght, f);
    }
}

// boundary_flux_effect.d:586-645
void SFluidBlock::outflow_flux(const FlowState& fs, int outsign, FVInterface* face) {
    auto& F = face->F;
    Vec3 v_rel = fs.vel;  // minus gvel == 0
    double mass_flux = fs.gas.rho * v_rel.dot(face->n);
    if ((outsign * mass_flux) > 0.0) {
        F[cqi.mass] = mass_flux;
        F[cqi.xMom] = fs.gas.p * face->n.x + fs.vel.x * mass_flux;
        F[cqi.yMom] = fs.gas.p * face->n.y + fs.vel.y * mass_flux;
        if (cqi.threeD) F[cqi.zMom] = fs.gas.p * face->n.z + fs.vel.z * mass_flux;
        double utot = fs.gas.u + 0.5 * fs.vel.dot(fs.vel);
        F[cqi.totEnergy] = mass_flux * utot + fs.gas.p * fs.vel.dot(face->n);
        if (cqi.n_species > 1)
            for (int i = 0; i < cqi.n_species; ++i) F[cqi.species + i] = mass_flux * fs.gas.massf[i];
    } else {
        F[cqi.mass] = 0.0;
        F[cqi.xMom] = face->n.x * fs.gas.p;
        F[cqi.yMom] = face->n.y * fs.gas.p;
        if (cqi.threeD) F[cqi.zMom] = face->n.z * fs.gas.p;
        Vec3 gvel(0.0, 0.0, 0.0);
        F[cqi.totEnergy] = fs.gas.p * gvel.dot(face->n);
        if (cqi.n_species > 1)
            for (int i = 0; i < cqi.n_species; ++i) F[cqi.species + i] = 0.0;
    }
}

// fluidblock.d:1098 -> boundary_flux_effect.d:557-584 (structured grid)
void SFluidBlock::apply_post_conv_flux() {
    for (int b = 0; b < ((cfg.dimensions == 3) ? 6 : 4); ++b) {
        BoundaryCondition& B = bc[b];
        if (B.kind != BCKind::outflow_simple_flux) continue;
        for (size_t i = 0; i < B.faces.size(); ++i) {
            int outsign = B.outsigns[i];
            FVInterface* f = B.faces[i];
            FVCell* interior = (outsign == 1) ? f->left_cells[0] : f->right_cells[0];
            outflow_flux(interior->fs, outsign, f);
        }
    }
}

// gas/gas_state.d:226-268 and flowstate.d:363-390
bool SFluidBlock::check_data(const FlowState& fs) const {
    bool ok = true;
    const GasState& g = fs.gas;
    if (!std::isfinite(g.rho) || g.rho < 1.01 * 0.0) ok = false;
    if (!std::isfinite(g.T) || g.T < 1.01 * 0.0) ok = false;
    if (!std::isfinite(g.p)) ok = false;
    if (!std::isfinite(g.p_e)) ok = false;
    if (!std::isfinite(g.a)) ok = false;
    double f_sum = 0.0;
    for (double m : g.massf) f_sum += m;
    if (f_sum < 0.99 || f_sum > 1.01 || !std::isfinite(f_sum)) ok = false;
    if (std::fabs(fs.vel.x) > cfg.max_velocity || std::fabs(fs.vel.y) > cfg.max_velocity ||
        std::fabs(fs.vel.z) > cfg.max_velocity)
        ok = false;
    if (g.T < cfg.min_temp) ok = false;
    if (g.T > cfg.max_temp) ok = false;
    return ok;
}

// fluidblock.d:607-676 (adjust_invalid_cell_data == false)
int SFluidBlock::count_invalid_cells(int ftl) {
    (void)ftl;
    int count = 0;
    for (auto c : cells)
        if (c->data_is_bad || !check_data(c->fs)) ++count;
    return count;
}

// Initial state: FlowState(p, T, vel) -> update_thermo_from_pT + update_sound_speed,
// then encode_conserved + decode_conserved as at simcore.d:326-330.
void SFluidBlock::set_flowstate(int ci, const FlowState& fs) {
    FVCell* c = cells[ci];
    c->fs = fs;
    encode_conserved(c, 0);
    decode_conserved(c, 0);
}

//---------------------------------------------------------------------------------
// Simulation driver

Simulation::~Simulation() {
    for (auto b : blocks) delete b;
}

// simcore_gasdynamic_step.d:60-159 (global time stepping, no STS, no dt_override)
void Simulation::determine_time_step_size() {
    bool do_dt_check_now = (st.step % cfg.cfl_count) == 0;
    if (!do_dt_check_now) return;
    double cfl_value = cfg.cfl_value;
    double dt_allow = std::numeric_limits<double>::max();
    double cfl_max = 0.0;
    for (auto b : blocks) {
        double r[3];
        b->determine_time_step_size(st.dt_global, cfl_value, st.step > 0, r);
        dt_allow = std::min(dt_allow, r[0]);
        cfl_max = std::max(cfl_max, r[1]);
    }
    st.dt_allow = dt_allow;
    st.cfl_max = cfl_max;
    if (st.step == 0) st.dt_allow = std::fmin(cfg.dt_init, st.dt_allow);
    if (st.dt_allow <= st.dt_global) {
        st.dt_global = st.dt_allow;
    } else {
        st.dt_global = std::min(st.dt_global * 1.5, st.dt_allow);
        st.dt_global = std::min(st.dt_global, cfg.dt_max);
    }
}

// Source cell of a full-face copy: restates the case tables of
// set_up_cell_mapping_phase0 (full_face_copy.d:696-869 in 2-D, 870-1480 in 3-D).
// Each face has in-face coordinates (a, b): west/east (j, k), south/north (i, k),
// bottom/top (i, j); in 2-D the single coordinate a is j on west/east and i on
// south/north.  The tables map the destination face's (a, b) to the source face's (p, q):
// orientation 0 keeps (a, b), orientation 2 reverses both, and orientations 1 and 3 are
// the quarter turns (b, B-1-a) and (A-1-b, a), in that order when the destination
// face's coordinates and its outward normal form a right-handed set (east, south, top)
// and in the opposite order otherwise; then p is reversed, p -> A-1-p, when both faces
// have the same handedness.  (A, B) are the source face's extents.  Source layer n lies
// n cells inside the source face.
namespace {
const int kFaceHand[6] = {-1, +1, +1, -1, -1, +1};  // W, E, S, N, B, T

void face_extents(const SFluidBlock* b, int face, int& A, int& B) {
    if (face <= 1) { A = b->njc; B = b->nkc; }
    else if (face <= 3) { A = b->nic; B = b->nkc; }
    else { A = b->nic; B = b->njc; }
}
}  // namespace

void ffc_source_index(int dims, int dface, int sface, int orient, int a, int bb, int n, int snic, int snjc,
                      int snkc, int& i, int& j, int& k) {
    int A, B;
    if (sface <= 1) { A = snjc; B = snkc; }
    else if (sface <= 3) { A = snic; B = snkc; }
    else { A = snic; B = snjc; }
    int p, q;
    if (dims == 2) {
        p = a; q = 0;
    } else if (orient == 0) {
        p = a; q = bb;
    } else if (orient == 2) {
        p = A - 1 - a; q = B - 1 - bb;
    } else if ((orient == 1) == (kFaceHand[dface] > 0)) {
        p = bb; q = B - 1 - a;
    } else {
        p = A - 1 - bb; q = a;
    }
    if (kFaceHand[dface] == kFaceHand[sface]) p = A - 1 - p;
    switch (sface) {
    case 0: i = n; j = p; k = q; break;
    case 1: i = snic - 1 - n; j = p; k = q; break;
    case 2: j = n; i = p; k = q; break;
    case 3: j = snjc - 1 - n; i = p; k = q; break;
    case 4: k = n; i = p; j = q; break;
    default: k = snkc - 1 - n; i = p; j = q; break;
    }
}

namespace {
FVCell* ffc_source(SFluidBlock* dst, int dface, SFluidBlock* src, int sface, int orient, int a, int bb, int n) {
    int i, j, k;
    ffc_source_index(dst->cfg.dimensions, dface, sface, orient, a, bb, n, src->nic, src->njc, src->nkc, i, j, k);
    if (i < 0 || i >= src->nic || j < 0 || j >= src->njc || k < 0 || k >= src->nkc)
        throw FlowSolverException("Full-face copy maps outside the source block.");
    return src->get_cell(i, j, k);
}
}  // namespace

// check_cell_mapping + set_up_cell_mapping_phase0..2 (full_face_copy.d:140-1480) and
// exchange_geometry_phase2 (:1576-1625, same process): each ghost cell of an exchange
// face is paired with its source cell and takes that cell's lengths (copy_values_from
// with CopyDataOption.grid, fvcell.d; lengths are copied as they are, not permuted).
void Simulation::set_up_exchange() {
    const int nfaces = (cfg.dimensions == 3) ? 6 : 4;
    for (auto b : blocks) {
        for (int face = 0; face < nfaces; ++face) {
            BoundaryCondition& B = b->bc[face];
            if (B.kind != BCKind::exchange) continue;
            if (B.other_block < 0 || B.other_block >= (int)blocks.size() || B.other_face < 0 ||
                B.other_face >= nfaces || B.other_orientation < 0 || B.other_orientation > 3)
                throw FlowSolverException("Incorrect boundary connection.");
            SFluidBlock* o = blocks[B.other_block];
            int A, Bx, oA, oB;
            face_extents(b, face, A, Bx);
            face_extents(o, B.other_face, oA, oB);
            bool turned = (cfg.dimensions == 3) && (B.other_orientation % 2 == 1);
            if (cfg.dimensions == 2 ? (A != oA) : (turned ? (A != oB || Bx != oA) : (A != oA || Bx != oB)))
                throw FlowSolverException("Exchange faces have inconsistent numbers of cells.");
            B.ghost_cells.clear();
            B.mapped_cells.clear();
            const int ng = b->n_ghost_cell_layers;
            for (int bb = 0; bb < (cfg.dimensions == 3 ? Bx : 1); ++bb)
                for (int a = 0; a < A; ++a) {
                    FVInterface* f;
                    switch (face) {
                    case 0: f = b->get_ifi(0, a, bb); break;
                    case 1: f = b->get_ifi(b->nic, a, bb); break;
                    case 2: f = b->get_ifj(a, 0, bb); break;
                    case 3: f = b->get_ifj(a, b->njc, bb); break;
                    case 4: f = b->get_ifk(a, bb, 0); break;
                    default: f = b->get_ifk(a, bb, b->nkc); break;
                    }
                    for (int n = 0; n < ng; ++n) {
                        FVCell* g = (face % 2 == 1) ? f->right_cells[n] : f->left_cells[n];
                        FVCell* m = ffc_source(b, face, o, B.other_face, B.other_orientation, a, bb, n);
                        B.ghost_cells.push_back(g);
                        B.mapped_cells.push_back(m);
                        g->iLength = m->iLength;
                        g->jLength = m->jLength;
                        g->kLength = m->kLength;
                        g->volume = m->volume;
                        g->pos = m->pos;
                    }
                }
        }
    }
    exchange_ready = true;
}

// simcore_gasdynamic_step.d:3195-3224 without smoothing (shock_detector_smoothing 0).
void Simulation::detect_shocks() {
    for (auto b : blocks) {
        b->detect_shock_points();
        b->shock_faces_to_cells();
    }
    if (cfg.strict_shock_detector)
        for (auto b : blocks) b->enforce_strict_shock_detector();
}

// simcore_exchange.d:96-135 -> full_face_copy.d:1888-1900 (same process): every ghost
// cell of an exchange face takes its source cell's flow state.
void Simulation::exchange_ghost_cell_boundary_data() {
    if (!exchange_ready) set_up_exchange();
    for (auto b : blocks)
        for (int face = 0; face < 6; ++face) {
            BoundaryCondition& B = b->bc[face];
            if (B.kind != BCKind::exchange) continue;
            for (size_t i = 0; i < B.ghost_cells.size(); ++i) B.ghost_cells[i]->fs = B.mapped_cells[i]->fs;
        }
}

// One stage of simcore_gasdynamic_step.d:946-1545 for the pc scheme, without retries.
void Simulation::stage(int stage, double dt) {
    int ftl = stage - 1;
    for (auto b : blocks) {
        b->clear_fluxes();
        for (auto c : b->cells) std::fill(c->Q.begin(), c->Q.end(), 0.0);
    }
    exchange_ghost_cell_boundary_data();
    for (auto b : blocks) b->apply_pre_recon(blocks);
    // simcore_gasdynamic_step.d:1030-1037: detection in stage 1, after the ghost-cell fill.
    if (cfg.do_shock_detect && stage == 1 &&
        ((!cfg.frozen_shock_detector) || (cfg.shock_detector_freeze_step > st.step)))
        detect_shocks();
    for (auto b : blocks) b->convective_flux_phase0(true);
    for (auto b : blocks) b->apply_post_conv_flux();
    for (auto b : blocks) {
        const CQI& cqi = b->cqi;
        for (auto c : b->cells) b->time_derivatives(c, ftl);
        if (stage == 1) {
            double gamma_1 = 1.0;
            for (auto c : b->cells) {
                for (int j = 0; j < cqi.n; ++j) c->U[1][j] = c->U[0][j] + dt * gamma_1 * c->dUdt[0][j];
                b->decode_conserved(c, ftl + 1);
            }
        } else {
            double gamma_1 = 0.5, gamma_2 = 0.5;
            for (auto c : b->cells) {
                for (int j = 0; j < cqi.n; ++j)
                    c->U[2][j] = c->U[0][j] + dt * (gamma_1 * c->dUdt[0][j] + gamma_2 * c->dUdt[1][j]);
                b->decode_conserved(c, ftl + 1);
            }
        }
    }
    for (auto b : blocks) {
        int nbad = b->count_invalid_cells(ftl + 1);
        if (nbad > cfg.max_invalid_cells) throw FlowSolverException("Too many bad cells during explicit gasdynamic update.");
    }
}

// simcore_gasdynamic_step.d:906-1575 (pc scheme, fixed grid)
void Simulation::gasdynamic_step() {
    double t0 = st.time;
    int attempt_number = 0;
    bool step_failed = false;
    for (auto b : blocks) for (auto c : b->cells) c->data_is_bad = false;
    do {
        ++attempt_number;
        step_failed = false;
        for (int s = 1; s <= 2; ++s) {
            try {
                stage(s, st.dt_global);
            } catch (FlowSolverException& e) {
                if (std::string(e.what()).find("Too many bad cells") != std::string::npos) throw;
                step_failed = true;
            } catch (GasModelException&) {
                step_failed = true;
            }
            if (step_failed) {
                st.dt_global = st.dt_global * 0.2;
                break;
            }
        }
    } while (step_failed && (attempt_number < cfg.max_attempts_for_step));
    st.attempts_last_step = attempt_number;
    if (step_failed) throw FlowSolverException("Explicit update failed after max attempts; giving up.");
    for (auto b : blocks) for (auto c : b->cells) std::swap(c->U[0], c->U[2]);
    st.time = t0 + st.dt_global;
}

// simcore.d:1619-1638 -> fvcell.d:858-972 for every cell: the reactor over dt with the
// cell's dt_chem (one retry with dt_chem = -1), the EOS from rho, u, then the species
// densities of U[0] from the new mass fractions.  The sound speed and the cell's species
// densities rho_s are not updated (the reference leaves them until the next decode).
void Simulation::chemistry_step(double dt) {
    for (auto b : blocks) {
        const CQI& cqi = b->cqi;
        for (auto c : b->cells) {
            if (thermochemical_increment(*chem, *chem_gas, c->fs.gas, dt, c->dt_chem, T_frozen) != 0)
                throw FlowSolverException("The thermochemical_increment() failed.");
            if (cqi.n_species > 1)
                for (int isp = 0; isp < cqi.n_species; ++isp)
                    c->U[0][cqi.species + isp] = c->fs.gas.rho * c->fs.gas.massf[isp];
        }
    }
}

// simcore.d:1014-1377 (the parts of the loop on this path)
int Simulation::run(int max_steps, double max_time) {
    int taken = 0;
    while (true) {
        determine_time_step_size();
        gasdynamic_step();
        if (chem) chemistry_step(st.dt_global);
        st.step += 1;
        ++taken;
        if (st.time >= max_time) break;
        if (taken >= max_steps) break;
    }
    return taken;
}

}  // namespace orc
