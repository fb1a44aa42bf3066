// chem_oracle.cpp -- CPU restatement of the thermally perfect gas and the alpha-QSS
// chemistry update (TEST INFRASTRUCTURE ONLY: the oracle).
//
// Parity pins (tests/test_oracle_chem.py):
//   CEA curves        O Cp(500), h(3700), s(10000): cea_thermo_curves.d:186-217
//   TPG mixture EOS   O2/N2/H2 u(1000 K) and T back from 1500 K: therm_perf_gas_mix_eos.d:177-200
//   5-species air     pT, rhou, rhoT, rhop known answers: therm_perf_gas.d:565-609
//   rate constants    H2+I2 at 700 K, N2 dissociation at 4000 K: rate_constant.d:599-618
//   alpha-QSS step    H2-I2-HI analytic value and error ratio 4.001 on halving h:
//                     chemistry_update.d:725-737
//   ChemistryUpdate   full update of H2-I2-HI to 60 000 s ([HI] = 7.14201983840, 1e-9)
//                     with the file's RKF step: chemistry_update.d:739-750
//   equilibrium constant  Kc against a closed form from the same Gibbs energies and
//                     detailed balance of the 5-species air mechanism at equilibrium.
// The summation order of Phobos's std.algorithm.sum (pairwise) and the 80-bit
// intermediates of its pow/exp are third-party (LDC runtime) details not under the
// reference tree: parity at the last-ulp level there is unpinned.
#include "chem_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace orc {

static constexpr double T_MIN = 10.0;       // gas_model.d:45
static constexpr double T_MAX = 100000.0;   // gas_model.d:46
static constexpr double MIN_MOLES = 1.0e-30;          // gas_model.d:44
static constexpr double MIN_MASS_FRACTION = 1.0e-30;  // gas_model.d:43

//---------------------------------------------------------------------------------
// Phobos std.algorithm.sum for floating point ranges with length and slicing:
// seed + sumPairwise(r) with a 64-entry store collapsed at powers of two.
double phobos_sum(const double* x, int n) {
    if (n <= 0) return 0.0;
    double store[64];
    int idx = 0;
    int k = 0;
    for (; k < n / 2; ++k) {
        store[idx] = x[2 * k] + x[2 * k + 1];
        int lastToKeep = idx - __builtin_ctz((unsigned)(k + 1));
        while (idx > lastToKeep) {
            store[idx - 1] += store[idx];
            --idx;
        }
        ++idx;
    }
    if (n % 2 == 1) {
        store[idx] = x[n - 1];
        ++idx;
    }
    double s = store[idx - 1];
    for (int j = idx - 2; j >= 0; --j) s += store[j];
    return 0.0 + s;
}

//---------------------------------------------------------------------------------
// cea_thermo_curves.d:24-176
CEAThermoCurve::CEAThermoCurve(double R_, const std::vector<double>& T_breaks, const std::vector<double>& T_blends,
                               const std::vector<std::array<double, 9>>& coeffs)
    : R(R_), T_breaks_(T_breaks), T_blends_(T_blends), coeffs_(coeffs) {
    T_low = T_breaks_.front();
    T_high = T_breaks_.back();
    Cp_low = eval_Cp(T_low);
    Cp_high = eval_Cp(T_high);
    h_low = eval_h(T_low);
    h_high = eval_h(T_high);
    s_low = eval_s(T_low);
    s_high = eval_s(T_high);
}

void CEAThermoCurve::determineCoefficients(double T, double a[9]) const {
    const size_t nb = T_breaks_.size();
    if (T < (T_breaks_[1] - 0.5 * T_blends_[0])) {
        for (int j = 0; j < 9; ++j) a[j] = coeffs_[0][j];
        return;
    }
    if (T > (T_breaks_[nb - 2] + 0.5 * T_blends_.back())) {
        for (int j = 0; j < 9; ++j) a[j] = coeffs_.back()[j];
        return;
    }
    for (size_t i = 1; i < nb - 1; ++i) {
        double T_blend_low = T_breaks_[i] - 0.5 * T_blends_[i - 1];
        double T_blend_high = T_breaks_[i] + 0.5 * T_blends_[i - 1];
        if (T >= T_blend_low && T <= T_blend_high) {
            double wB = (1. / T_blends_[i - 1]) * (T - T_blend_low);
            double wA = 1.0 - wB;
            for (int j = 0; j < 9; ++j) a[j] = wA * coeffs_[i - 1][j] + wB * coeffs_[i][j];
            return;
        }
        if (i < T_blends_.size() && T > T_blend_high && T < (T_breaks_[i + 1] - 0.5 * T_blends_[i])) {
            for (int j = 0; j < 9; ++j) a[j] = coeffs_[i][j];
            return;
        }
    }
    throw GasModelException("Coefficients for CEA curve could not be determined.");
}

double CEAThermoCurve::eval_Cp(double T) const {
    if (T < T_low) return Cp_low;
    if (T > T_high) return Cp_high;
    double a[9];
    determineCoefficients(T, a);
    double Cp_on_R = a[0] / (T * T) + a[1] / T + a[2] + a[3] * T;
    Cp_on_R += a[4] * T * T + a[5] * T * T * T + a[6] * T * T * T * T;
    return R * Cp_on_R;
}

double CEAThermoCurve::eval_h(double T) const { return eval_h(T, std::log(T)); }

double CEAThermoCurve::eval_h(double T, double logT) const {
    if (T < T_low) return h_low - Cp_low * (T_low - T);
    if (T > T_high) return h_high + Cp_high * (T - T_high);
    double a[9];
    determineCoefficients(T, a);
    double h_on_RT = -a[0] / T + a[1] * logT + a[2] * T + a[3] * T * T / 2.0;
    h_on_RT += a[4] * T * T * T / 3.0 + a[5] * T * T * T * T / 4.0 + a[6] * T * T * T * T * T / 5.0 + a[7];
    return R * h_on_RT;
}

double CEAThermoCurve::eval_s(double T) const { return eval_s(T, std::log(T)); }

double CEAThermoCurve::eval_s(double T, double logT) const {
    if (T < T_low) return s_low - Cp_low * std::log(T_low / T);
    if (T > T_high) return s_high + Cp_high * std::log(T / T_high);
    double a[9];
    determineCoefficients(T, a);
    double s_on_R = -a[0] / (2.0 * T * T) - a[1] / T + a[2] * logT + a[3] * T;
    s_on_R += a[4] * T * T / 2.0 + a[5] * T * T * T / 3.0 + a[6] * T * T * T * T / 4.0 + a[8];
    return R * s_on_R;
}

//---------------------------------------------------------------------------------
// therm_perf_gas.d:40-470, therm_perf_gas_mix_eos.d, perf_gas_mix_eos.d
ThermallyPerfectGas::ThermallyPerfectGas(const std::vector<double>& M, const std::vector<CEAThermoCurve>& c)
    : curves(c) {
    n_species = (int)M.size();
    mol_masses = M;
    Rsp.resize(M.size());
    for (size_t i = 0; i < M.size(); ++i) Rsp[i] = R_universal / M[i];
}

// gas_model.d:410-419 mass_average
static double mass_average(const GasState& Q, const double* phi, int n) {
    double result = 0.0;
    for (int i = 0; i < n; ++i) result += Q.massf[i] * phi[i];
    return result;
}

double ThermallyPerfectGas::R(const GasState& Q) const { return mass_average(Q, Rsp.data(), n_species); }

double ThermallyPerfectGas::Cp(const GasState& Q) const {
    double v[16];
    for (int i = 0; i < n_species; ++i) v[i] = curves[i].eval_Cp(Q.T);
    return mass_average(Q, v, n_species);
}

double ThermallyPerfectGas::Cv(const GasState& Q) const {
    double v[16];
    for (int i = 0; i < n_species; ++i) v[i] = curves[i].eval_Cp(Q.T) - Rsp[i];
    return mass_average(Q, v, n_species);
}

void ThermallyPerfectGas::update_energy(GasState& Q) const {
    double v[16];
    for (int i = 0; i < n_species; ++i) v[i] = curves[i].eval_h(Q.T) - Rsp[i] * Q.T;
    Q.u = mass_average(Q, v, n_species);
}

// nm/newton.d:71-140 applied as in therm_perf_gas_mix_eos.d:68-151.  The zero function
// writes Q.T and Q.u (update_energy) on every evaluation, as the reference's does.
void ThermallyPerfectGas::update_temperature(GasState& Q) const {
    const double Tsave = Q.T;
    const double TOL = 1.0e-6;
    const double e_tgt = Q.u;
    const double delT = 1000.0;
    const double T1 = std::fmax(Q.T - 0.5 * delT, T_MIN);
    const double T2 = T1 + delT;
    newton_evals = 0;
    auto zeroFn = [&](double T) {
        ++newton_evals;
        Q.T = T;
        update_energy(Q);
        return e_tgt - Q.u;
    };
    auto dzdT = [&](double T) {
        double v[16];
        for (int i = 0; i < n_species; ++i) v[i] = curves[i].eval_Cp(T) - Rsp[i];
        return -1.0 * mass_average(Q, v, n_species);
    };
    auto solve = [&](double x0, double xMin, double xMax, double tol, double& out) -> bool {
        const int MAXIT = 30;
        double xL = xMin, xH = xMax;
        double fL = zeroFn(xL);
        double fH = zeroFn(xH);
        if ((fL > 0.0 && fH > 0.0) || (fL < 0.0 && fH < 0.0)) return false;
        if (fL == 0.0) { out = xMin; return true; }
        if (fH == 0.0) { out = xMax; return true; }
        if (fL < 0.0) { xL = xMin; xH = xMax; }
        else { xH = xMin; xL = xMax; }
        double rts = x0;
        double dxold = (xMax - xMin);
        double dx = dxold;
        double f0 = zeroFn(rts);
        double df0 = dzdT(rts);
        for (int j = 0; j < MAXIT; ++j) {
            if ((((rts - xH) * df0 - f0) * ((rts - xL) * df0 - f0) > 0.0) || (std::fabs(2.0 * f0) > std::fabs(dxold * df0))) {
                dxold = dx;
                dx = 0.5 * (xH - xL);
                rts = xL + dx;
                if (xL == rts) { out = rts; return true; }
            } else {
                dxold = dx;
                dx = f0 / df0;
                double tmp = rts;
                rts -= dx;
                if (tmp == rts) { out = rts; return true; }
            }
            if (std::fabs(dx) < tol) { out = rts; return true; }
            f0 = zeroFn(rts);
            df0 = dzdT(rts);
            if (f0 < 0.0) xL = rts;
            else xH = rts;
        }
        return false;
    };
    double T;
    if (solve(Tsave, T1, T2, TOL, T)) { Q.T = T; return; }
    if (solve(Tsave, T_MIN, T_MAX, TOL, T)) { Q.T = T; return; }
    Q.T = Tsave;
    update_energy(Q);
    throw GasModelException("There was a problem iterating to find temperature");
}

// perf_gas_mix_eos.d:43-76 (no separate electron temperature)
static double Rmix(const ThermallyPerfectGas& g, const GasState& Q) {
    double R = 0.0;
    for (int i = 0; i < g.n_species; ++i) R += Q.massf[i] * g.Rsp[i];
    return R;
}

void ThermallyPerfectGas::update_thermo_from_pT(GasState& Q) const {
    double Rm = Rmix(*this, Q);
    double denom = Rm * Q.T;
    Q.rho = Q.p / denom;
    update_energy(Q);
}

void ThermallyPerfectGas::update_thermo_from_rhou(GasState& Q) const {
    update_temperature(Q);
    double Rm = Rmix(*this, Q);
    Q.p = Q.rho * Rm * Q.T;
}

void ThermallyPerfectGas::update_thermo_from_rhoT(GasState& Q) const {
    update_energy(Q);
    double Rm = Rmix(*this, Q);
    Q.p = Q.rho * Rm * Q.T;
}

void ThermallyPerfectGas::update_thermo_from_rhop(GasState& Q) const {
    double Rm = Rmix(*this, Q);
    Q.T = Q.p / (Rm * Q.rho);
    update_energy(Q);
}

// therm_perf_gas.d:394-404: frozen sound speed sqrt(gamma * R T), gamma = Cp/Cv.
void ThermallyPerfectGas::update_sound_speed(GasState& Q) const {
    double gamma = Cp(Q) / Cv(Q);
    double R = mass_average(Q, Rsp.data(), n_species);
    Q.a = std::sqrt(gamma * (R * Q.T));
}

double ThermallyPerfectGas::gibbs_free_energy(double T, int isp) const {
    double h = curves[isp].eval_h(T);
    double s = curves[isp].eval_s(T);
    return h - T * s;
}

void ThermallyPerfectGas::massf2conc(const GasState& Q, double* conc) const {
    if (std::isnan(Q.rho) || (Q.rho <= 0.0)) throw GasModelException("Invalid density.");
    for (int i = 0; i < n_species; ++i) {
        conc[i] = Q.massf[i] * Q.rho / mol_masses[i];
        if (conc[i] < MIN_MOLES) conc[i] = 0.0;
    }
}

void ThermallyPerfectGas::conc2massf(const double* conc, GasState& Q) const {
    if (std::isnan(Q.rho) || (Q.rho <= 0.0)) throw GasModelException("Invalid density.");
    for (int i = 0; i < n_species; ++i) {
        Q.massf[i] = conc[i] * mol_masses[i] / Q.rho;
        if (Q.massf[i] < MIN_MASS_FRACTION) Q.massf[i] = 0.0;
    }
}

//---------------------------------------------------------------------------------
// Integer powers of concentrations (reaction.d eval_forward_rate/eval_backward_rate).
static double ipow(double c, int n) {
    switch (n) {
    case 0: return 1.0;
    case 1: return c;
    case 2: return c * c;
    case 3: return c * (c * c);
    default: return std::pow(c, (double)n);
    }
}

// pow(P_atm/(R_u T), nuSum) with nuSum a small whole number (reaction.d:43-44).
static double pow_nusum(double x, double nuSum) {
    if (nuSum == 0.0) return 1.0;
    if (nuSum == 1.0) return x;
    if (nuSum == -1.0) return 1.0 / x;
    if (nuSum == 2.0) return x * x;
    return std::pow(x, nuSum);
}

// reaction_mechanism.d:23-63, reaction.d:58-457
ReactionMechanism::ReactionMechanism(const ThermallyPerfectGas* gm, const std::vector<ReactionDef>& defs,
                                     double T_lower, double T_upper)
    : nsp(gm->n_species), nr((int)defs.size()), gm_(gm), defs_(defs), T_lower_(T_lower), T_upper_(T_upper) {
    k_f.assign(nr, 0.0);
    k_b.assign(nr, 0.0);
    w_f_.assign(nr, 0.0);
    w_b_.assign(nr, 0.0);
    gibbs_.assign(nsp, 0.0);
    qw_.assign(nsp, 0.0);
    Lw_.assign(nsp, 0.0);
    for (auto& d : defs_) {
        std::vector<int> part;
        for (auto& r : d.reactants) part.push_back(r.first);
        for (auto& p : d.products) part.push_back(p.first);
        std::sort(part.begin(), part.end());
        part.erase(std::unique(part.begin(), part.end()), part.end());
        participants_.push_back(part);
        std::vector<int> nu(nsp, 0);
        for (int isp = 0; isp < nsp; ++isp) {
            int nu1 = 0, nu2 = 0;
            for (auto& r : d.reactants) if (r.first == isp) { nu1 = r.second; break; }
            for (auto& p : d.products) if (p.first == isp) { nu2 = p.second; break; }
            nu[isp] = nu2 - nu1;
        }
        nu_.push_back(nu);
    }
}

void ReactionMechanism::eval_rate_constants(const GasState& Qin) {
    double T = Qin.T;
    if (T < T_lower_) T = T_lower_;
    if (T > T_upper_) T = T_upper_;
    // eval_gibbs_free_energies at p = P_atm: the TPG Gibbs energy per species needs T only.
    for (int isp = 0; isp < nsp; ++isp) gibbs_[isp] = gm_->gibbs_free_energy(T, isp);
    for (int ir = 0; ir < nr; ++ir) {
        const ReactionDef& d = defs_[ir];
        auto arr = [&](const RateConstantDef& rc) { return rc.A * std::pow(T, rc.n) * std::exp(-rc.C / T); };
        auto Keq = [&]() {
            // reaction.d:27-52 compute_equilibrium_constant
            double dG = 0.0, nuSum = 0.0;
            for (int isp : participants_[ir]) {
                double g = gibbs_[isp];
                dG += nu_[ir][isp] * g * gm_->mol_masses[isp];
                nuSum += nu_[ir][isp];
            }
            double K_p = std::exp(-dG / (R_universal * T));
            return K_p * pow_nusum(P_atm / (R_universal * T), nuSum);
        };
        if (!d.frc.fromEqConst) {
            k_f[ir] = arr(d.frc);
            if (d.brc.fromEqConst) k_b[ir] = k_f[ir] / Keq();
            else k_b[ir] = arr(d.brc);
        } else {
            k_b[ir] = arr(d.brc);
            k_f[ir] = k_b[ir] * Keq();
        }
    }
}

void ReactionMechanism::eval_rate_constants_2T(double Tin, double Tv) {
    double T = Tin;
    if (T < T_lower_) T = T_lower_;
    if (T > T_upper_) T = T_upper_;
    for (int ir = 0; ir < nr; ++ir) {
        const ReactionDef& d = defs_[ir];
        if (d.frc.fromEqConst || d.brc.fromEqConst)
            throw std::runtime_error("equilibrium-constant rates are not supported with two temperatures");
        // rate_constant.d:72-75 (Arrhenius at T) and :391-394 (Park at T^s Tv^(1-s))
        auto rc = [&](const RateConstantDef& k) {
            double Te = T;
            if (k.park) Te = std::pow(T, k.s) * std::pow(Tv, 1.0 - k.s);
            return k.A * std::pow(Te, k.n) * std::exp(-k.C / Te);
        };
        k_f[ir] = rc(d.frc);
        k_b[ir] = rc(d.brc);
    }
}

double ReactionMechanism::rate(int ir, int isp) const {
    // reaction.d:268-288 production(isp) - loss(isp)
    const int nu = nu_[ir][isp];
    const double prod = (nu > 0) ? nu * w_f_[ir] : (nu < 0) ? -nu * w_b_[ir] : 0.0;
    const double loss = (nu > 0) ? nu * w_b_[ir] : (nu < 0) ? -nu * w_f_[ir] : 0.0;
    return prod - loss;
}

bool ReactionMechanism::has_participant(int ir, int isp) const {
    for (int p : participants_[ir])
        if (p == isp) return true;
    return false;
}

void ReactionMechanism::eval_split_rates(const double* conc, double* q, double* L) {
    for (int ir = 0; ir < nr; ++ir) {
        const ReactionDef& d = defs_[ir];
        double collider = 0.0;
        if (d.collider)
            for (auto& e : d.efficiencies) collider += e.second * conc[e.first];
        double vf = d.collider ? k_f[ir] * collider : k_f[ir];
        for (auto& r : d.reactants) vf *= ipow(conc[r.first], r.second);
        double vb = d.collider ? k_b[ir] * collider : k_b[ir];
        for (auto& p : d.products) vb *= ipow(conc[p.first], p.second);
        w_f_[ir] = vf;
        w_b_[ir] = vb;
    }
    for (int isp = 0; isp < nsp; ++isp) { q[isp] = 0.0; L[isp] = 0.0; }
    for (int ir = 0; ir < nr; ++ir) {
        for (int isp : participants_[ir]) {
            int nu = nu_[ir][isp];
            double prod = (nu > 0) ? nu * w_f_[ir] : (nu < 0) ? -nu * w_b_[ir] : 0.0;
            double loss = (nu > 0) ? nu * w_b_[ir] : (nu < 0) ? -nu * w_f_[ir] : 0.0;
            q[isp] += prod;
            L[isp] += loss;
        }
    }
}

void ReactionMechanism::eval_rates(const double* conc, double* rates) {
    eval_split_rates(conc, qw_.data(), Lw_.data());
    for (int isp = 0; isp < nsp; ++isp) rates[isp] = qw_[isp] - Lw_[isp];
}

double ReactionMechanism::estimateStepSize(const double* conc) {
    const double MUCH_LARGER_FACTOR = 10000.0, ZERO_EPS = 1.0e-50, ZERO_TOL = 1.0e-30, EPS1 = 1.0e-3;
    const double UPPER = 1.0e-10, LOWER = 1.0e-15;
    double min_dt = 1.0e6, dt;
    eval_split_rates(conc, qw_.data(), Lw_.data());
    for (int isp = 0; isp < nsp; ++isp) {
        double c = conc[isp];
        if (c > 0.0 && (std::fabs(qw_[isp] - Lw_[isp]) > ZERO_TOL)) {
            if (qw_[isp] > MUCH_LARGER_FACTOR * Lw_[isp]) dt = c / (Lw_[isp] + ZERO_EPS);
            else dt = std::fabs(c / (qw_[isp] - Lw_[isp]));
            min_dt = std::min(min_dt, dt);
        }
    }
    double dt_chem = EPS1 * min_dt;
    dt_chem = std::min(dt_chem, UPPER);
    dt_chem = std::max(dt_chem, LOWER);
    return dt_chem;
}

//---------------------------------------------------------------------------------
ChemistryUpdate::ChemistryUpdate(const ThermallyPerfectGas* gm, const std::vector<ReactionDef>& defs,
                                 const ChemConfig& c)
    : rmech(gm, defs, c.T_lower, c.T_upper), cfg(c), gm_(gm), n_(gm->n_species) {
    for (auto* v : {&conc0_, &concOut_, &yp_, &yp0_, &yc_, &q0_, &L0_, &p0_, &qtilda_, &pbar_, &qp_, &Lp_, &pp_,
                    &alpha_, &alphabar_, &k1_, &k2_, &k3_, &k4_, &k5_, &k6_, &yTmp_, &yErr_})
        v->assign(n_, 0.0);
}

// chemistry_update.d:533-655
bool ChemistryUpdate::alpha_qss_step(const double* y0, double h, double* yOut, double& hSuggest, int* iters) {
    const double ZERO_EPS = 1.0e-50;
    auto p_on_y = [&](const std::vector<double>& L, const double* y, std::vector<double>& p) {
        for (int i = 0; i < n_; ++i) p[i] = L[i] / (y[i] + ZERO_EPS);
    };
    auto alpha_compute = [&](const std::vector<double>& p, std::vector<double>& alpha) {
        for (int i = 0; i < n_; ++i) {
            double r = 1.0 / (p[i] * h + ZERO_EPS);
            alpha[i] = (180.0 * r * r * r + 60.0 * r * r + 11.0 * r + 1.0) /
                       (360.0 * r * r * r + 60.0 * r * r + 12.0 * r + 1.0);
        }
    };
    auto update_conc = [&](std::vector<double>& yT, const std::vector<double>& q, const std::vector<double>& p,
                           const std::vector<double>& alpha) {
        for (int i = 0; i < n_; ++i) yT[i] = y0[i] + (h * (q[i] - p[i] * y0[i])) / (1.0 + alpha[i] * h * p[i]);
    };
    auto test_converged = [&]() {
        for (int i = 0; i < n_; ++i) {
            if (yc_[i] < ZERO_EPS) continue;
            double test = std::fabs(yc_[i] - yp0_[i]);
            if (test >= (cfg.eps1 * (yc_[i] + cfg.delta))) return false;
        }
        return true;
    };
    auto step_suggest = [&]() {
        double sigma = 0.0;
        for (int i = 0; i < n_; ++i) {
            if (yc_[i] < ZERO_EPS) continue;
            double test = std::fabs(yc_[i] - yp0_[i]) / (cfg.eps2 * (yc_[i] + cfg.delta));
            if (test > sigma) sigma = test;
        }
        double h_new;
        if (sigma <= 0.0) {
            h_new = 10.0 * h;
        } else {
            double x = sigma;
            x = x - 0.5 * (x * x - sigma) / x;
            x = x - 0.5 * (x * x - sigma) / x;
            x = x - 0.5 * (x * x - sigma) / x;
            h_new = h * ((1.0 / x) + 0.005);
        }
        return h_new;
    };
    rmech.eval_split_rates(y0, q0_.data(), L0_.data());
    p_on_y(L0_, y0, p0_);
    alpha_compute(p0_, alpha_);
    update_conc(yp_, q0_, p0_, alpha_);
    yp0_ = yp_;
    for (int corr = 0; corr < cfg.maxIters; ++corr) {
        if (iters) ++*iters;
        rmech.eval_split_rates(yp_.data(), qp_.data(), Lp_.data());
        p_on_y(Lp_, yp_.data(), pp_);
        alpha_compute(pp_, alphabar_);
        for (int i = 0; i < n_; ++i) {
            qtilda_[i] = alphabar_[i] * qp_[i] + (1 - alphabar_[i]) * q0_[i];
            pbar_[i] = 0.5 * (p0_[i] + pp_[i]);
        }
        update_conc(yc_, qtilda_, pbar_, alphabar_);
        if (test_converged()) {
            for (int i = 0; i < n_; ++i) yOut[i] = yc_[i];
            hSuggest = step_suggest();
            return true;
        }
        yp_ = yc_;
    }
    hSuggest = step_suggest();
    return false;
}

// chemistry_update.d:385-484
bool ChemistryUpdate::rkf_step(const double* y0, double h, double* yOut, double& hSuggest) {
    const double a21 = 1. / 5., a31 = 3. / 40., a32 = 9. / 40., a41 = 3. / 10., a42 = -9. / 10., a43 = 6. / 5.,
                 a51 = -11. / 54., a52 = 5. / 2., a53 = -70. / 27., a54 = 35. / 27., a61 = 1631. / 55296.,
                 a62 = 175. / 512., a63 = 575. / 13824., a64 = 44275. / 110592., a65 = 253. / 4096.;
    const double b51 = 37. / 378., b53 = 250. / 621., b54 = 125. / 594., b56 = 512. / 1771.,
                 b41 = 2825. / 27648., b43 = 18575. / 48384., b44 = 13525. / 55296., b45 = 277. / 14336.,
                 b46 = 1.0 / 4.0;
    const int n = n_;
    rmech.eval_rates(y0, k1_.data());
    for (int i = 0; i < n; ++i) yTmp_[i] = y0[i] + h * (a21 * k1_[i]);
    rmech.eval_rates(yTmp_.data(), k2_.data());
    for (int i = 0; i < n; ++i) yTmp_[i] = y0[i] + h * (a31 * k1_[i] + a32 * k2_[i]);
    rmech.eval_rates(yTmp_.data(), k3_.data());
    for (int i = 0; i < n; ++i) yTmp_[i] = y0[i] + h * (a41 * k1_[i] + a42 * k2_[i] + a43 * k3_[i]);
    rmech.eval_rates(yTmp_.data(), k4_.data());
    for (int i = 0; i < n; ++i) yTmp_[i] = y0[i] + h * (a51 * k1_[i] + a52 * k2_[i] + a53 * k3_[i] + a54 * k4_[i]);
    rmech.eval_rates(yTmp_.data(), k5_.data());
    for (int i = 0; i < n; ++i)
        yTmp_[i] = y0[i] + h * (a61 * k1_[i] + a62 * k2_[i] + a63 * k3_[i] + a64 * k4_[i] + a65 * k5_[i]);
    rmech.eval_rates(yTmp_.data(), k6_.data());
    for (int i = 0; i < n; ++i) yOut[i] = y0[i] + h * (b51 * k1_[i] + b53 * k3_[i] + b54 * k4_[i] + b56 * k6_[i]);
    for (int i = 0; i < n; ++i)
        yErr_[i] = yOut[i] - (y0[i] + h * (b41 * k1_[i] + b43 * k3_[i] + b44 * k4_[i] + b45 * k5_[i] + b46 * k6_[i]));
    double err = 0.0, sk = 0.0;
    const double atol = cfg.errTol, rtol = cfg.errTol;
    for (int i = 0; i < n; ++i) {
        sk = atol + rtol * std::max(std::fabs(y0[i]), std::fabs(yOut[i]));
        err += (yErr_[i] / sk) * (yErr_[i] / sk);
    }
    err = std::sqrt(err / n);
    double scale = 0.0;
    const double maxscale = 10.0, minscale = 0.2, safe = 0.9, alpha = 0.2;
    if (err <= 1.0) {
        if (err == 0.0) scale = maxscale;
        else {
            scale = safe * std::pow(err, -alpha);
            if (scale < minscale) scale = minscale;
            if (scale > maxscale) scale = maxscale;
        }
        hSuggest = scale * h;
        return true;
    }
    scale = std::max(safe * std::pow(err, -alpha), minscale);
    hSuggest = scale * h;
    return false;
}

// chemistry_update.d:140-326
bool ChemistryUpdate::update(GasState& Q, double tInterval, double& dtSuggest, ChemStats* st) {
    const double DT_INCREASE_PERCENT = 10.0, DT_REDUCTION_FACTOR = 10.0, H_MIN = 1.0e-15,
                 ALLOWABLE_MASSF_ERROR = 1.0e-3;
    GasState Qinit = Q;
    gm_->massf2conc(Q, conc0_.data());
    rmech.eval_rate_constants(Q);
    double t = 0.0, h, dtSave = 0.0;
    if (dtSuggest > tInterval) h = tInterval;
    else if (dtSuggest <= 0.0) h = rmech.estimateStepSize(conc0_.data());
    else h = dtSuggest;
    int cycle = 0, attempt = 0;
    for (; cycle < cfg.maxSubcycles; ++cycle) {
        dtSave = dtSuggest;
        h = std::min(h, tInterval - t);
        attempt = 0;
        for (; attempt < cfg.maxAttempts; ++attempt) {
            if (st) ++st->attempts;
            int it = 0;
            bool ok = (cfg.method == OdeMethod::alpha_qss)
                          ? alpha_qss_step(conc0_.data(), h, concOut_.data(), dtSuggest, &it)
                          : rkf_step(conc0_.data(), h, concOut_.data(), dtSuggest);
            if (st) st->corrector_iters += it;
            bool passes = true;
            if (ok) {
                gm_->conc2massf(concOut_.data(), Q);
                double massfTotal = phobos_sum(Q.massf.data(), n_);
                if (std::fabs(massfTotal - 1.0) > ALLOWABLE_MASSF_ERROR) passes = false;
            }
            if (ok && passes) {
                t += h;
                conc0_ = concOut_;
                double hMax = h * (1.0 + DT_INCREASE_PERCENT / 100.0);
                if (dtSuggest > h) h = std::min(dtSuggest, hMax);
                break;
            } else {
                if (st) ++st->failures;
                h /= DT_REDUCTION_FACTOR;
                if (h < H_MIN) {
                    Q = Qinit;
                    return false;
                }
            }
        }
        if (attempt == cfg.maxAttempts) {
            Q = Qinit;
            return false;
        }
        if (st) ++st->subcycles;
        if (cfg.tightTempCoupling) {
            gm_->conc2massf(conc0_.data(), Q);
            gm_->update_thermo_from_rhou(Q);
            rmech.eval_rate_constants(Q);
        }
        if (t >= tInterval) {
            if (cycle == 0) dtSave = dtSuggest;
            break;
        }
    }
    if (cycle == cfg.maxSubcycles) {
        Q = Qinit;
        return false;
    }
    gm_->conc2massf(concOut_.data(), Q);
    double massfTotal = phobos_sum(Q.massf.data(), n_);
    for (auto& mf : Q.massf) mf /= massfTotal;
    dtSuggest = dtSave;
    return true;
}

// fvcell.d:858-972 (single temperature, no ignition zones, species written back by the caller).
// The reference retries only on ThermochemicalReactorUpdateException (here: update() returning
// false).  A GasModelException raised inside the update (a bad density in massf2conc, a NaN
// temperature in the CEA curves, a failed T(u) iteration of the tight temperature coupling)
// is not caught there and ends the run; here the cell fails, without the retry.
int thermochemical_increment(ChemistryUpdate& chem, const ThermallyPerfectGas& gm, GasState& Q, double dt,
                             double& dt_chem, double T_frozen, ChemStats* stats) {
    if (Q.T <= T_frozen) return 0;
    bool ok = false;
    try {
        ok = chem.update(Q, dt, dt_chem, stats);
        if (!ok) {
            dt_chem = -1.0;
            ok = chem.update(Q, dt, dt_chem, stats);
            if (!ok) return -1;
        }
    } catch (GasModelException&) {
        dt_chem = -1.0;
        return -1;
    }
    try {
        gm.update_thermo_from_rhou(Q);
    } catch (GasModelException&) {
        return -1;
    }
    return 0;
}

}  // namespace orc
