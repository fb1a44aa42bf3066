// chem_oracle.hpp -- CPU restatement of the thermally perfect gas and the finite-rate
// chemistry step (TEST INFRASTRUCTURE ONLY: the oracle).
//
// Follows src/gas/thermo/cea_thermo_curves.d, src/gas/thermo/therm_perf_gas_mix_eos.d,
// src/gas/thermo/perf_gas_mix_eos.d, src/gas/therm_perf_gas.d, src/nm/newton.d,
// src/gas/gas_model.d, src/kinetics/rate_constant.d, src/kinetics/reaction.d,
// src/kinetics/reaction_mechanism.d, src/kinetics/chemistry_update.d and
// src/eilmer/fvcell.d:858-972 (cited per function in chem_oracle.cpp).
#pragma once
#include <array>
#include <vector>

#include "eilmer_oracle.hpp"

namespace orc {

// src/gas/thermo/cea_thermo_curves.d:24-176
class CEAThermoCurve {
public:
    CEAThermoCurve(double R, const std::vector<double>& T_breaks, const std::vector<double>& T_blends,
                   const std::vector<std::array<double, 9>>& coeffs);
    double eval_Cp(double T) const;
    double eval_h(double T) const;
    double eval_h(double T, double logT) const;
    double eval_s(double T) const;
    double eval_s(double T, double logT) const;
    double R;

private:
    std::vector<double> T_breaks_, T_blends_;
    std::vector<std::array<double, 9>> coeffs_;
    double T_low, T_high, Cp_low, Cp_high, h_low, h_high, s_low, s_high;
    void determineCoefficients(double T, double a[9]) const;
};

// src/gas/therm_perf_gas.d:40-470 (thermodynamics only)
class ThermallyPerfectGas : public GasModel {
public:
    ThermallyPerfectGas(const std::vector<double>& M, const std::vector<CEAThermoCurve>& curves);
    std::vector<double> Rsp;
    std::vector<CEAThermoCurve> curves;
    void update_thermo_from_pT(GasState& Q) const override;
    void update_thermo_from_rhou(GasState& Q) const override;
    void update_thermo_from_rhoT(GasState& Q) const override;
    void update_thermo_from_rhop(GasState& Q) const;
    void update_sound_speed(GasState& Q) const override;
    double internal_energy(const GasState& Q) const override { return Q.u; }
    double Cv(const GasState& Q) const override;
    double Cp(const GasState& Q) const override;
    double R(const GasState& Q) const override;
    void update_energy(GasState& Q) const;            // therm_perf_gas_mix_eos.d:61-67
    void update_temperature(GasState& Q) const;       // therm_perf_gas_mix_eos.d:68-151
    double gibbs_free_energy(double T, int isp) const;  // gas_model.d:187-194 with TPG h, s
    void massf2conc(const GasState& Q, double* conc) const;   // gas_model.d:238-253
    void conc2massf(const double* conc, GasState& Q) const;   // gas_model.d:255-271
    // Diagnostic counters of the last update_temperature call.
    mutable int newton_evals = 0;
};

struct RateConstantDef {
    bool fromEqConst = false;
    double A = 0.0, n = 0.0, C = 0.0;
    // Park2TRateConstant (rate_constant.d:365-395): A T^n exp(-C/T) at T = T^s Tv^(1-s)
    bool park = false;
    double s = 1.0;
};

// src/kinetics/reaction.d:58-457 (ElementaryReaction and AnonymousColliderReaction)
struct ReactionDef {
    bool collider = false;
    std::vector<std::pair<int, int>> reactants, products;
    std::vector<std::pair<int, double>> efficiencies;
    RateConstantDef frc, brc;
};

class ReactionMechanism {
public:
    ReactionMechanism(const ThermallyPerfectGas* gm, const std::vector<ReactionDef>& defs, double T_lower,
                      double T_upper);
    void eval_rate_constants(const GasState& Q);                       // reaction_mechanism.d:50-63
    // The same with a vibro-electronic temperature for Park rate constants (T is clamped to
    // the temperature limits, Tv is not); equilibrium-constant rates are not supported.
    void eval_rate_constants_2T(double T, double Tv);
    // reaction_mechanism.d:118-121: production - loss of species isp by reaction ir at the
    // rates of the last eval_split_rates call.
    double rate(int ir, int isp) const;
    bool has_participant(int ir, int isp) const;       // reaction_mechanism.d:133-142
    void eval_split_rates(const double* conc, double* q, double* L);   // reaction_mechanism.d:83-96
    void eval_rates(const double* conc, double* rates);                // reaction_mechanism.d:75-82
    double estimateStepSize(const double* conc);                       // reaction_mechanism.d:162-203
    int nsp, nr;
    std::vector<double> k_f, k_b;

private:
    const ThermallyPerfectGas* gm_;
    std::vector<ReactionDef> defs_;
    std::vector<std::vector<int>> participants_, nu_;
    std::vector<double> w_f_, w_b_, gibbs_, qw_, Lw_;
    double T_lower_, T_upper_;
};

enum class OdeMethod { alpha_qss = 0, rkf = 1 };

struct ChemConfig {
    OdeMethod method = OdeMethod::alpha_qss;
    double eps1 = 1.0e-3, eps2 = 5.0e-4, delta = 1.0e-10;
    int maxIters = 10;
    double errTol = 1.0e-3;
    bool tightTempCoupling = false;
    int maxSubcycles = 10000, maxAttempts = 4;
    double T_lower = 300.0, T_upper = 50000.0;
};

struct ChemStats {
    int subcycles = 0, attempts = 0, corrector_iters = 0, failures = 0;
};

// src/kinetics/chemistry_update.d:36-326 with AlphaQssStep (:506-656) and RKFStep (:385-484)
class ChemistryUpdate {
public:
    ChemistryUpdate(const ThermallyPerfectGas* gm, const std::vector<ReactionDef>& defs, const ChemConfig& cfg);
    // Returns false where the reference throws ThermochemicalReactorUpdateException.
    bool update(GasState& Q, double tInterval, double& dtSuggest, ChemStats* stats = nullptr);
    // The single ODE steps (for the reference's step-order tests).
    bool alpha_qss_step(const double* y0, double h, double* yOut, double& hSuggest, int* iters = nullptr);
    bool rkf_step(const double* y0, double h, double* yOut, double& hSuggest);
    ReactionMechanism rmech;
    ChemConfig cfg;

private:
    const ThermallyPerfectGas* gm_;
    int n_;
    std::vector<double> conc0_, concOut_;
    std::vector<double> yp_, yp0_, yc_, q0_, L0_, p0_, qtilda_, pbar_, qp_, Lp_, pp_, alpha_, alphabar_;
    std::vector<double> k1_, k2_, k3_, k4_, k5_, k6_, yTmp_, yErr_;
};

// fvcell.d:858-972 for a single cell: returns 0 on success, -1 where the reference
// throws FlowSolverException.  T_frozen from globalconfig.d:1247.
int thermochemical_increment(ChemistryUpdate& chem, const ThermallyPerfectGas& gm, GasState& Q, double dt,
                             double& dt_chem, double T_frozen = 300.0, ChemStats* stats = nullptr);

// Phobos std.algorithm.sum over a random-access range of doubles: seed 0.0 plus the
// pairwise sum (the summation order of the D runtime the reference links).
double phobos_sum(const double* x, int n);

}  // namespace orc
