// tt_oracle.hpp -- CPU restatement of the two-temperature air gas model and its kinetics
// (TEST INFRASTRUCTURE ONLY: the oracle).
//
// Follows src/gas/two_temperature_air.d (5-species neutral air: thermodynamics with a
// trans-rotational temperature T and a vibro-electronic temperature Tv) and
// src/kinetics/two_temperature_air_kinetics.d (the reactor: the chemistry ODE step and a
// Cash-Karp step of the vibro-electronic energy with Landau-Teller relaxation, Millikan-White
// and Park relaxation times and the chemical source of vibro-electronic energy), with the
// per-cell caller of src/eilmer/fvcell.d:858-972.  Cited per function in tt_oracle.cpp.
#pragma once
#include <vector>

#include "chem_oracle.hpp"

namespace orc {

// A two-temperature gas state: GasState with T_modes[0] = Tv and u_modes[0] = uv.
struct TTState {
    double rho = 0.0, p = 0.0, T = 0.0, Tv = 0.0, u = 0.0, uv = 0.0;
    std::vector<double> massf;
};

// src/gas/two_temperature_air.d:50-291, 502-740 (neutral species only)
class TwoTemperatureAir {
public:
    TwoTemperatureAir(const std::vector<double>& M, const std::vector<CEAThermoCurve>& curves,
                      const std::vector<int>& molecular);
    int nsp;
    std::vector<double> M, R, Cp_tr_rot, del_hf;
    std::vector<int> molecular;
    std::vector<CEAThermoCurve> curves;

    double transRotEnergy(const TTState& Q) const;
    double vibElecEnergy(double Tve, int isp) const;
    double vibElecEnergy(const TTState& Q, double Tve) const;
    double vibElecCv(double Tve, int isp) const;
    double vibElecCv(const TTState& Q, double Tve) const;
    double vibElecTemperature(const TTState& Q) const;   // throws GasModelException
    void update_pressure(TTState& Q) const;
    void update_thermo_from_rhou(TTState& Q) const;      // throws GasModelException
    void update_thermo_from_rhoT(TTState& Q) const;
    void update_thermo_from_pT(TTState& Q) const;
    // Diagnostic: Newton iterations of the last vibElecTemperature call.
    mutable int newton_iters = 0;
};

// src/kinetics/two_temperature_air_kinetics.d:44-706 for the neutral species
class TwoTemperatureAirKinetics {
public:
    // tpg: the same species' masses and curves (massf <-> conc); A_MW per species (0 for
    // atoms); tol: the energy step's absolute and relative tolerances.
    TwoTemperatureAirKinetics(const TwoTemperatureAir* gas, const ThermallyPerfectGas* tpg,
                              const std::vector<ReactionDef>& defs, const ChemConfig& cfg,
                              const std::vector<double>& A_MW, double abs_tol = 1.0e-9, double rel_tol = 1.0e-9);
    // opCall: false where the reference throws ThermochemicalReactorUpdateException (Q is
    // then the input state); a failed Tv iteration throws GasModelException.
    bool opCall(TTState& Q, double tInterval, double& dtSuggest, ChemStats* st = nullptr);
    double evalRelaxationTime(const TTState& Q, int isp);
    double evalRate(const TTState& Q);
    double energyReactiveSourceTerm(const TTState& Q);
    bool energyStep(TTState& Q, double h, double& hSuggest);
    ChemistryUpdate chem;

private:
    const TwoTemperatureAir* gas_;
    const ThermallyPerfectGas* tpg_;
    int n_;
    std::vector<double> A_, particleMass_, molef_, numden_, conc0_, concOut_;
    std::vector<std::vector<double>> mu_;
    std::vector<std::vector<int>> reactionsBySpecies_;
    double atol_, rtol_, uTotal_ = 0.0;
    void to_gas(const TTState& Q, GasState& G) const;
};

// fvcell.d:858-972 for one cell of a two-temperature gas: 0 on success, -1 where the
// reference ends the run (both reactor attempts failed, or a gas-model exception).
int tt_thermochemical_increment(TwoTemperatureAirKinetics& kin, const TwoTemperatureAir& gas, TTState& Q, double dt,
                                double& dt_chem, double T_frozen = 300.0, ChemStats* stats = nullptr);

}  // namespace orc
