// tt_oracle.cpp -- CPU restatement of the two-temperature air gas model and its reactor
// (TEST INFRASTRUCTURE ONLY: the oracle).
//
// Parity pins (tests/test_oracle_tt.py):
//   vibro-electronic energy  zero at 298.15 K for every species (the heat of formation is
//                     h(298.15)); Cv_ve the T-derivative of e_ve (central difference)
//   rhou <-> rhoT     round trip of T (direct inversion) and Tv (Newton) over random states
//   relaxation time   Millikan-White from the published correlation with A from the
//                     characteristic temperature and reduced mass (p tau = exp(A (T^-1/3 -
//                     0.015 mu^1/4) - 18.42) atm s), and the Park limit 1/(sigma cbar n) with n
//                     from the ideal-gas law, not from the code's mass fractions
//   Landau-Teller     a trace of N2 relaxing in a bath of atoms (T nearly constant) follows
//                     e_v(t) = e_v* - (e_v* - e_v0) exp(-t/tau)
//   Park rate         k(T, Tv) at Tv = T equals the Arrhenius value; at Tv != T it equals
//                     A Te^n exp(-C/Te) with Te = exp(s ln T + (1-s) ln Tv)
//   equilibrium       Tv = T and no reactions: the state is left exactly unchanged
// The reactor's full sequence (stage order, the energy step's effect on the state when it
// fails, the chemistry step's last rates feeding the energy source) has no golden vector in
// the reference tree: it is restated line by line, and parity of that sequence is
// unpinned beyond the checks above.
#include "tt_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace orc {

static constexpr double T_REF = 298.15;                 // two_temperature_air.d:24
static constexpr double Avogadro_number = 6.02214e23;   // physical_constants.d:18
static constexpr double Boltzmann_constant = 1.380658e-23;  // physical_constants.d:17
static constexpr double SMALL_MOLE_FRACTION = 1.0e-15;  // gas_model.d:42
static constexpr double MIN_MASS_FRACTION = 1.0e-30;    // gas_model.d:43

//---------------------------------------------------------------------------------
// two_temperature_air.d:158-220 (constructor: R, heat of formation, Cp_tr_rot)
TwoTemperatureAir::TwoTemperatureAir(const std::vector<double>& M_, const std::vector<CEAThermoCurve>& curves_,
                                     const std::vector<int>& molecular_)
    : nsp((int)M_.size()), M(M_), molecular(molecular_), curves(curves_) {
    R.resize(nsp);
    del_hf.resize(nsp);
    Cp_tr_rot.resize(nsp);
    for (int i = 0; i < nsp; ++i) {
        R[i] = R_universal / M[i];
        del_hf[i] = curves[i].eval_h(298.15);
        Cp_tr_rot[i] = (5. / 2.) * R[i];
        if (molecular[i]) Cp_tr_rot[i] += R[i];
    }
}

// :651-660
double TwoTemperatureAir::transRotEnergy(const TTState& Q) const {
    double e = 0.0;
    for (int i = 0; i < nsp; ++i) {
        double h_tr_rot = Cp_tr_rot[i] * (Q.T - T_REF) + del_hf[i];
        e += Q.massf[i] * (h_tr_rot - R[i] * Q.T);
    }
    return e;
}

// :573-583
double TwoTemperatureAir::vibElecEnergy(double Tve, int isp) const {
    double h_at_Tve = curves[isp].eval_h(Tve);
    return h_at_Tve - Cp_tr_rot[isp] * (Tve - T_REF) - del_hf[isp];
}

// :584-591
double TwoTemperatureAir::vibElecEnergy(const TTState& Q, double Tve) const {
    double e = 0.0;
    for (int i = 0; i < nsp; ++i) e += Q.massf[i] * vibElecEnergy(Tve, i);
    return e;
}

// :662-671
double TwoTemperatureAir::vibElecCv(double Tve, int isp) const { return curves[isp].eval_Cp(Tve) - Cp_tr_rot[isp]; }

// :673-680
double TwoTemperatureAir::vibElecCv(const TTState& Q, double Tve) const {
    double c = 0.0;
    for (int i = 0; i < nsp; ++i) c += Q.massf[i] * vibElecCv(Tve, i);
    return c;
}

// :698-740
double TwoTemperatureAir::vibElecTemperature(const TTState& Q) const {
    const int MAX_ITERATIONS = 20;
    const double TOL = 1.0e-6, E_TOL = 0.01;
    double T_guess = Q.Tv;
    double f_guess = vibElecEnergy(Q, T_guess) - Q.uv;
    newton_iters = 0;
    if (std::fabs(f_guess) < E_TOL) return Q.Tv;
    int count = 0;
    for (int iter = 0; iter < MAX_ITERATIONS; ++iter) {
        double Cv = vibElecCv(Q, T_guess);
        double dT = -f_guess / Cv;
        T_guess += dT;
        ++newton_iters;
        if (std::fabs(dT) < TOL) break;
        f_guess = vibElecEnergy(Q, T_guess) - Q.uv;
        ++count;
    }
    if (count == MAX_ITERATIONS) throw GasModelException("The 'vibTemperature' function failed to converge.");
    return T_guess;
}

// perf_gas_mix_eos.d:43-49 with heavyParticleGasConstant (:86-94), no free electrons
void TwoTemperatureAir::update_pressure(TTState& Q) const {
    double Rmix = 0.0;
    for (int i = 0; i < nsp; ++i) Rmix += Q.massf[i] * R[i];
    Q.p = Q.rho * Rmix * Q.T;
}

// :275-291
void TwoTemperatureAir::update_thermo_from_rhou(TTState& Q) const {
    double sumA = 0.0, sumB = 0.0;
    for (int i = 0; i < nsp; ++i) {
        sumA += Q.massf[i] * (Cp_tr_rot[i] * T_REF - del_hf[i]);
        sumB += Q.massf[i] * (Cp_tr_rot[i] - R[i]);
    }
    Q.T = (Q.u + sumA) / sumB;
    Q.Tv = vibElecTemperature(Q);
    update_pressure(Q);
}

// :293-299
void TwoTemperatureAir::update_thermo_from_rhoT(TTState& Q) const {
    update_pressure(Q);
    Q.u = transRotEnergy(Q);
    Q.uv = vibElecEnergy(Q, Q.Tv);
}

// :264-270 with perf_gas_mix_eos.d:55-62
void TwoTemperatureAir::update_thermo_from_pT(TTState& Q) const {
    double Rmix = 0.0;
    for (int i = 0; i < nsp; ++i) Rmix += Q.massf[i] * R[i];
    double denom = Rmix * Q.T;
    Q.rho = Q.p / denom;
    Q.u = transRotEnergy(Q);
    Q.uv = vibElecEnergy(Q, Q.Tv);
}

//---------------------------------------------------------------------------------
// two_temperature_air_kinetics.d:55-75 (constructor) and :387-439 (initModel)
TwoTemperatureAirKinetics::TwoTemperatureAirKinetics(const TwoTemperatureAir* gas, const ThermallyPerfectGas* tpg,
                                                     const std::vector<ReactionDef>& defs, const ChemConfig& cfg,
                                                     const std::vector<double>& A_MW, double abs_tol, double rel_tol)
    : chem(tpg, defs, cfg), gas_(gas), tpg_(tpg), n_(gas->nsp), atol_(abs_tol), rtol_(rel_tol) {
    A_.assign(n_, 0.0);
    for (int i = 0; i < n_; ++i)
        if (gas->molecular[i]) A_[i] = A_MW[i];
    mu_.assign(n_, std::vector<double>(n_, 0.0));
    for (int i = 0; i < n_; ++i)
        for (int j = 0; j < n_; ++j) {
            double Mi = gas->M[i], Mj = gas->M[j];
            mu_[i][j] = (Mi * Mj) / (Mi + Mj);
            mu_[i][j] *= 1000.0;  // kg/mol -> g/mol
        }
    particleMass_.resize(n_);
    for (int i = 0; i < n_; ++i) particleMass_[i] = gas->M[i] / Avogadro_number;
    reactionsBySpecies_.resize(n_);
    for (int i = 0; i < n_; ++i)
        for (int ir = 0; ir < chem.rmech.nr; ++ir)
            if (chem.rmech.has_participant(ir, i)) reactionsBySpecies_[i].push_back(ir);
    molef_.assign(n_, 0.0);
    numden_.assign(n_, 0.0);
    conc0_.assign(n_, 0.0);
    concOut_.assign(n_, 0.0);
}

void TwoTemperatureAirKinetics::to_gas(const TTState& Q, GasState& G) const {
    G.rho = Q.rho;
    G.p = Q.p;
    G.T = Q.T;
    G.u = Q.u;
    G.massf = Q.massf;
}

// :398-441
double TwoTemperatureAirKinetics::evalRelaxationTime(const TTState& Q, int isp) {
    double totalND = 0.0, sum = 0.0;
    for (int csp = 0; csp < n_; ++csp) {
        if (molef_[csp] >= SMALL_MOLE_FRACTION) {
            double nd = numden_[csp];
            sum += nd * std::exp(A_[isp] * (std::pow(Q.T, -1. / 3) - 0.015 * std::pow(mu_[isp][csp], 0.25)) - 18.42);
            totalND += nd;
        }
    }
    double tauMW = sum / totalND;  // p*tauMW in atm.s
    tauMW *= P_atm / Q.p;
    double nd = numden_[isp];
    double kB = Boltzmann_constant;
    double cBar_s = std::sqrt(8 * kB * Q.T / (M_PI * particleMass_[isp]));
    double sigma_s = 1e-20;
    double tauP = 1.0 / (sigma_s * cBar_s * nd);
    return tauMW + tauP;
}

// :444-460 (no free electrons: no electron-translational exchange)
double TwoTemperatureAirKinetics::evalRate(const TTState& Q) {
    // gas_model.d:464-472 massf2molef (mixture_molecular_mass :454-462), :282-291 massf2numden
    double M_inv = 0.0;
    for (int i = 0; i < n_; ++i) M_inv += Q.massf[i] / gas_->M[i];
    double mixMolMass = 1.0 / M_inv;
    for (int i = 0; i < n_; ++i) molef_[i] = Q.massf[i] * mixMolMass / gas_->M[i];
    for (int i = 0; i < n_; ++i) numden_[i] = Avogadro_number * Q.massf[i] * Q.rho / gas_->M[i];
    double rate = 0.0;
    for (int isp = 0; isp < n_; ++isp) {
        if (!gas_->molecular[isp]) continue;
        double tau = evalRelaxationTime(Q, isp);
        double evStar = gas_->vibElecEnergy(Q.T, isp);
        double ev = gas_->vibElecEnergy(Q.Tv, isp);
        rate += Q.massf[isp] * (evStar - ev) / tau;
    }
    rate += energyReactiveSourceTerm(Q);
    return rate;
}

// :521-543
double TwoTemperatureAirKinetics::energyReactiveSourceTerm(const TTState& Q) {
    double rate = 0.0;
    for (int isp = 0; isp < n_; ++isp) {
        double ev = gas_->vibElecEnergy(Q.Tv, isp);
        double chemRate = 0.0;
        for (int ir : reactionsBySpecies_[isp]) chemRate += chem.rmech.rate(ir, isp);
        chemRate *= gas_->M[isp];
        rate += chemRate * ev;
    }
    rate /= Q.rho;
    return rate;
}

// :546-706 (Cash-Karp step of u_ve at fixed total energy)
bool TwoTemperatureAirKinetics::energyStep(TTState& Q, double h, double& hSuggest) {
    const double a21 = 1. / 5., a31 = 3. / 40., a32 = 9. / 40., a41 = 3. / 10., a42 = -9. / 10., a43 = 6. / 5.,
                 a51 = -11. / 54., a52 = 5. / 2., a53 = -70. / 27., a54 = 35. / 27., a61 = 1631. / 55296.,
                 a62 = 175. / 512., a63 = 575. / 13824., a64 = 44275. / 110592., a65 = 253. / 4096.;
    const double b51 = 37. / 378., b53 = 250. / 621., b54 = 125. / 594., b56 = 512. / 1771.,
                 b41 = 2825. / 27648., b43 = 18575. / 48384., b44 = 13525. / 55296., b45 = 277. / 14336.,
                 b46 = 1.0 / 4.0;
    const double uv0 = Q.uv;
    double k1 = evalRate(Q);
    Q.uv = uv0 + h * (a21 * k1);
    Q.u = uTotal_ - Q.uv;
    gas_->update_thermo_from_rhou(Q);
    double k2 = evalRate(Q);
    Q.uv = uv0 + h * (a31 * k1 + a32 * k2);
    Q.u = uTotal_ - Q.uv;
    gas_->update_thermo_from_rhou(Q);
    double k3 = evalRate(Q);
    Q.uv = uv0 + h * (a41 * k1 + a42 * k2 + a43 * k3);
    Q.u = uTotal_ - Q.uv;
    gas_->update_thermo_from_rhou(Q);
    double k4 = evalRate(Q);
    Q.uv = uv0 + h * (a51 * k1 + a52 * k2 + a53 * k3 + a54 * k4);
    Q.u = uTotal_ - Q.uv;
    gas_->update_thermo_from_rhou(Q);
    double k5 = evalRate(Q);
    Q.uv = uv0 + h * (a61 * k1 + a62 * k2 + a63 * k3 + a64 * k4 + a65 * k5);
    Q.u = uTotal_ - Q.uv;
    gas_->update_thermo_from_rhou(Q);
    double k6 = evalRate(Q);
    Q.uv = uv0 + h * (b51 * k1 + b53 * k3 + b54 * k4 + b56 * k6);
    Q.u = uTotal_ - Q.uv;
    gas_->update_thermo_from_rhou(Q);
    double errEst = Q.uv - (uv0 + h * (b41 * k1 + b43 * k3 + b44 * k4 + b45 * k5 + b46 * k6));
    double sk = atol_ + rtol_ * std::fmax(std::fabs(uv0), std::fabs(Q.uv));
    double err = std::fabs(errEst / sk);
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
    scale = std::fmax(safe * std::pow(err, -alpha), minscale);
    hSuggest = scale * h;
    return false;
}

// :162-358 (opCall)
bool TwoTemperatureAirKinetics::opCall(TTState& Q, double tInterval, double& dtSuggest, ChemStats* st) {
    const double DT_INCREASE_PERCENT = 10.0, DT_REDUCTION_FACTOR = 10.0, H_MIN = 1.0e-15,
                 ALLOWABLE_MASSF_ERROR = 1.0e-3;
    const ChemConfig& cfg = chem.cfg;
    TTState Qinit = Q;
    GasState G;
    to_gas(Q, G);
    tpg_->massf2conc(G, conc0_.data());
    uTotal_ = Q.u + Q.uv;
    chem.rmech.eval_rate_constants_2T(Q.T, Q.Tv);
    double t = 0.0, h, dtSave = 0.0;
    if (dtSuggest > tInterval) h = tInterval;
    else if (dtSuggest <= 0.0) h = chem.rmech.estimateStepSize(conc0_.data());
    else h = dtSuggest;
    int cycle = 0, attempt = 0;
    for (; cycle < cfg.maxSubcycles; ++cycle) {
        dtSave = dtSuggest;
        h = std::min(h, tInterval - t);
        attempt = 0;
        for (; attempt < cfg.maxAttempts; ++attempt) {
            if (st) ++st->attempts;
            int it = 0;
            bool r1 = (cfg.method == OdeMethod::alpha_qss)
                          ? chem.alpha_qss_step(conc0_.data(), h, concOut_.data(), dtSuggest, &it)
                          : chem.rkf_step(conc0_.data(), h, concOut_.data(), dtSuggest);
            if (st) st->corrector_iters += it;
            bool r2 = energyStep(Q, h, dtSuggest);
            bool ok = r1 && r2;
            bool passes = true;
            if (ok) {
                // gas_model.d:255-271 conc2massf into Q
                for (int i = 0; i < n_; ++i) {
                    Q.massf[i] = concOut_[i] * gas_->M[i] / Q.rho;
                    if (Q.massf[i] < MIN_MASS_FRACTION) Q.massf[i] = 0.0;
                }
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
            for (int i = 0; i < n_; ++i) {
                Q.massf[i] = conc0_[i] * gas_->M[i] / Q.rho;
                if (Q.massf[i] < MIN_MASS_FRACTION) Q.massf[i] = 0.0;
            }
            Q.u = uTotal_ - Q.uv;
            gas_->update_thermo_from_rhou(Q);
            chem.rmech.eval_rate_constants_2T(Q.T, Q.Tv);
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
    for (int i = 0; i < n_; ++i) {
        Q.massf[i] = concOut_[i] * gas_->M[i] / Q.rho;
        if (Q.massf[i] < MIN_MASS_FRACTION) Q.massf[i] = 0.0;
    }
    double massfTotal = phobos_sum(Q.massf.data(), n_);
    for (auto& mf : Q.massf) mf /= massfTotal;
    dtSuggest = dtSave;
    return true;
}

// fvcell.d:858-972.  A ThermochemicalReactorUpdateException (opCall false) leaves the input
// state and is retried once with dt_chem = -1; a GasModelException (the Tv iteration, a bad
// density) is not caught there and ends the run: here the cell fails at once.
int tt_thermochemical_increment(TwoTemperatureAirKinetics& kin, const TwoTemperatureAir& gas, TTState& Q, double dt,
                                double& dt_chem, double T_frozen, ChemStats* stats) {
    if (Q.T <= T_frozen) return 0;
    try {
        bool ok = kin.opCall(Q, dt, dt_chem, stats);
        if (!ok) {
            dt_chem = -1.0;
            ok = kin.opCall(Q, dt, dt_chem, stats);
            if (!ok) return -1;
        }
    } catch (GasModelException&) {
        return -1;
    }
    try {
        gas.update_thermo_from_rhou(Q);
    } catch (GasModelException&) {
        return -1;
    }
    return 0;
}

}  // namespace orc
