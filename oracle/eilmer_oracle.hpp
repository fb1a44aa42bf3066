// eilmer_oracle.hpp -- CPU restatement of Eilmer 4's structured-block explicit update.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP implementation is
// checked against.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may load it, and only as the checker or the timed CPU baseline,
// never as part of the product path.
//
// It restates, in serial C++ with IEEE double arithmetic and no FMA contraction
// (built with -ffp-contract=off), the D reference code paths listed below.  Each
// function cites the reference file:line it follows.  Object structure mirrors
// the reference on purpose (faces carrying left_cells/right_cells stencils, cells
// carrying iface/outsign lists, the face loop order of convective_flux_phase0,
// and the in-place local-frame rotation of *cell* velocities inside interp_l2r2),
// because the goal is fidelity of operation order, not speed.
//
// Parity status of each part is recorded in the header of eilmer_oracle.cpp.
#pragma once
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// src/gas/physical_constants.d:12
constexpr double R_universal = 8.31451;
// src/gas/physical_constants.d (P_atm used by the kinetics)
constexpr double P_atm = 101.325e3;

struct Vec3 {
    double x = 0.0, y = 0.0, z = 0.0;
    Vec3() = default;
    Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
    // src/geom/elements/vector3.d:384-387
    double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    // src/geom/elements/vector3.d:403-411
    void to_local(const Vec3& n, const Vec3& t1, const Vec3& t2) {
        double vx = dot(n), vy = dot(t1), vz = dot(t2);
        x = vx; y = vy; z = vz;
    }
    // src/geom/elements/vector3.d:416-423
    void to_global(const Vec3& n, const Vec3& t1, const Vec3& t2) {
        double vx = x * n.x + y * t1.x + z * t2.x;
        double vy = x * n.y + y * t1.y + z * t2.y;
        double vz = x * n.z + y * t1.z + z * t2.z;
        x = vx; y = vy; z = vz;
    }
    // src/geom/elements/vector3.d:362-381
    void normalize() {
        double m = std::sqrt(dot(*this));
        if (m > 0.0) { x /= m; y /= m; z /= m; } else { x = y = z = 0.0; }
        const double small = 1.0e-30;
        if (std::fabs(x) < small) x = 0.0;
        if (std::fabs(y) < small) y = 0.0;
        if (std::fabs(z) < small) z = 0.0;
    }
};

// src/geom/elements/vector3.d:616-621
inline void cross(Vec3& v3, const Vec3& v1, const Vec3& v2) {
    v3.x = v1.y * v2.z - v2.y * v1.z;
    v3.y = v2.x * v1.z - v1.x * v2.z;
    v3.z = v1.x * v2.y - v2.x * v1.y;
}

struct GasModelException : std::runtime_error {
    explicit GasModelException(const std::string& m) : std::runtime_error(m) {}
};
struct FlowSolverException : std::runtime_error {
    explicit FlowSolverException(const std::string& m) : std::runtime_error(m) {}
};

// src/gas/gas_state.d:18-60 (the fields used on this path)
struct GasState {
    double rho = 0.0, p = 0.0, T = 0.0, u = 0.0, a = 0.0, p_e = 0.0;
    std::vector<double> massf, rho_s;
};

// src/eilmer/flowstate.d:49-390 (gas + velocity)
struct FlowState {
    GasState gas;
    Vec3 vel;
    double S = 0.0;  // shock indicator, flowstate.d:63 (copied with the flow state)
};

class GasModel {
public:
    virtual ~GasModel() = default;
    int n_species = 1;
    std::vector<double> mol_masses;
    virtual void update_thermo_from_pT(GasState& Q) const = 0;
    virtual void update_thermo_from_rhou(GasState& Q) const = 0;
    virtual void update_thermo_from_rhoT(GasState& Q) const = 0;
    virtual void update_sound_speed(GasState& Q) const = 0;
    virtual double internal_energy(const GasState& Q) const = 0;
    virtual double Cv(const GasState& Q) const = 0;
    virtual double Cp(const GasState& Q) const = 0;
    virtual double R(const GasState& Q) const = 0;
};

// src/gas/ideal_gas.d:30-143
class IdealGas : public GasModel {
public:
    double mMass, gamma, Rgas, Cv_, Cvinv, Cp_;
    IdealGas(double mMass_, double gamma_);
    void update_thermo_from_pT(GasState& Q) const override;
    void update_thermo_from_rhou(GasState& Q) const override;
    void update_thermo_from_rhoT(GasState& Q) const override;
    void update_sound_speed(GasState& Q) const override;
    double internal_energy(const GasState& Q) const override { return Q.u; }
    double Cv(const GasState&) const override { return Cv_; }
    double Cp(const GasState&) const override { return Cp_; }
    double R(const GasState&) const override { return Rgas; }
};

enum class FluxCalc { ausmdv = 0, efm = 1, adaptive_efm_ausmdv = 2, hanel = 3, adaptive_hanel_ausmdv = 4 };
enum class BCKind { slip_wall = 0, inflow_supersonic = 1, outflow_simple_flux = 2,
                    outflow_simple_extrapolate = 3, exchange = 4 };

// The option subset of src/eilmer/globalconfig.d:936-1294 that this path reads.
struct Config {
    int dimensions = 2;
    FluxCalc flux_calculator = FluxCalc::ausmdv;
    int interpolation_order = 2;
    bool apply_limiter = true;               // globalconfig.d:1077
    bool extrema_clipping = true;            // globalconfig.d:1078
    double epsilon_van_albada = 1.0e-12;     // globalconfig.d:1079
    // Eilmer 5 (src/lmr) numerics on this path: the dimensionalised van Albada epsilon
    // (lmr/onedinterp.d:147-151) and AUSMDV's smooth-maximum sound speed (lmr/fluxcalc.d).
    bool lmr = false;
    bool interpolate_in_local_frame = true;  // globalconfig.d:1081
    bool apply_entropy_fix = true;           // globalconfig.d:1088
    bool allow_reconstruction_for_species = true;   // globalconfig.d:1073
    bool scale_species_after_reconstruction = true; // globalconfig.d:1093
    bool ignore_low_T_thermo_update_failure = true; // globalconfig.d:1005
    double suggested_low_T_value = 200.0;           // globalconfig.d:1006
    int max_invalid_cells = 0;                      // globalconfig.d:1014
    int max_attempts_for_step = 3;                  // globalconfig.d:947
    double cfl_value = 0.5;                         // globalconfig.d:1287
    int cfl_count = 10;                             // globalconfig.d:1294
    double dt_init = 1.0e-3;                        // globalconfig.d:1281
    double dt_max = 1.0e-3;                         // globalconfig.d:1282
    // src/eilmer/globalconfig.d:79-86
    double max_velocity = 30000.0, max_temp = 50000.0, min_temp = 0.0;
    // Reference-fidelity switch: rotate the stencil cells' velocities in place
    // (onedinterp.d:761-770, 983-986), as the reference does.
    bool mutate_cell_velocities = true;
    // Shock detection (globalconfig.d:1061-1131): on for the adaptive flux calculators
    // (configCheckPoint1, globalconfig.d:2682-2690); the PJ detector.
    bool do_shock_detect = false;
    double compression_tolerance = -0.30;    // globalconfig.d:1123
    double shear_tolerance = 0.20;           // globalconfig.d:1108
    bool strict_shock_detector = true;       // globalconfig.d:1064
    bool damped_outflow = false;             // globalconfig.d:1061
    bool frozen_shock_detector = false;      // globalconfig.d:1130
    int shock_detector_freeze_step = 1000000000;  // globalconfig.d:1131
};

struct FVInterface;

// l2r2 prepared coefficients and scalar reconstruction, onedinterp.d:82-93, 338-384
struct L2R2 {
    double aL0, aR0, lenL0_, lenR0_, two_over_lenL0_plus_lenL1, two_over_lenR0_plus_lenL0,
        two_over_lenR1_plus_lenR0, two_lenL0_plus_lenL1, two_lenR0_plus_lenR1;
    void prepare(double lenL1, double lenL0, double lenR0, double lenR1);
    void scalar(double qL1, double qL0, double qR0, double qR1, double& qL, double& qR, bool apply_limiter,
                bool extrema_clipping, double epsilon_van_albada, bool lmr = false) const;
};

// src/eilmer/fvcell.d (fields used on this path)
struct FVCell {
    int id = -1;
    bool is_ghost = false;
    FlowState fs;
    std::vector<double> U[3];     // flow time levels 0..2 (pc uses 0,1,2)
    std::vector<double> dUdt[2];
    std::vector<double> Q;
    double volume = 0.0, iLength = 0.0, jLength = 0.0, kLength = 0.0;
    Vec3 pos;
    std::vector<FVInterface*> iface;
    std::vector<int> outsign;
    bool data_is_bad = false;
    double signal = 0.0;
    double dt_chem = -1.0;        // simcore.d:335 (set_cell_dt_chem(-1.0))
};

enum class IndexDirection { i, j, k };

// src/eilmer/fvinterface.d:49-78
struct FVInterface {
    int id = -1;
    IndexDirection idir = IndexDirection::i;
    Vec3 n, t1, t2, pos;
    double area = 0.0, length = 0.0;
    std::vector<double> F;
    std::vector<FVCell*> left_cells, right_cells;
    bool is_on_boundary = false;
    int bc_id = -1;
    int i_bndry = -1;
    double S = 0.0;  // the face flow state's shock indicator (fvinterface.d fs.S)
    // left_cell/right_cell of sfluidblock.d:757-773: the adjacent cells, ghost cells included
    // (every boundary kind here has ghost-cell data).
    FVCell* left_cell() const { return left_cells.empty() ? nullptr : left_cells[0]; }
    FVCell* right_cell() const { return right_cells.empty() ? nullptr : right_cells[0]; }
};

struct BoundaryCondition {
    BCKind kind = BCKind::slip_wall;
    bool ghost_cell_data_available = true;
    bool convective_flux_computed_in_bc = false;
    FlowState inflow;                 // for inflow_supersonic
    std::vector<FVInterface*> faces;
    std::vector<int> outsigns;
    // exchange: other block, face and orientation (same-process full-face copy), and the
    // cell mapping built once all blocks exist (full_face_copy.d:67-125, 696-1480)
    int other_block = -1, other_face = -1, other_orientation = 0;
    std::vector<FVCell*> ghost_cells, mapped_cells;
};

// src/eilmer/conservedquantities.d:147-193 (transient ordering)
struct CQI {
    int n = 4, mass = 0, xMom = 1, yMom = 2, zMom = -1, totEnergy = 3, species = -1, n_species = 1;
    bool threeD = false;
    void init(int dimensions, int nsp) {
        mass = 0; xMom = 1; yMom = 2;
        if (dimensions == 3) { threeD = true; zMom = 3; totEnergy = 4; n = 5; }
        else { threeD = false; zMom = -1; totEnergy = 3; n = 4; }
        n_species = nsp;
        if (nsp > 1) { species = n; n += nsp; } else { species = -1; }
    }
};

// src/eilmer/sfluidblock.d:68 + src/eilmer/fluidblock.d
class SFluidBlock {
public:
    int id = 0;
    int nic, njc, nkc, niv, njv, nkv;
    int n_ghost_cell_layers = 2;
    Config cfg;
    const GasModel* gmodel;
    CQI cqi;
    std::vector<FVCell*> cells;          // interior, ordered (k*njc + j)*nic + i
    std::vector<FVCell*> ghost_cells;
    std::vector<FVInterface*> faces;     // ifi, then ifj, then ifk
    std::vector<Vec3> vertices;
    BoundaryCondition bc[6];

    SFluidBlock(int id, int nic, int njc, int nkc, const Config& cfg, const GasModel* gm,
                const std::vector<Vec3>& vtx, const BoundaryCondition bcs[6]);
    ~SFluidBlock();

    int cell_index(int i, int j, int k = 0) const { return (k * njc + j) * nic + i; }
    int vertex_index(int i, int j, int k = 0) const { return (k * njv + j) * niv + i; }
    int ifi_index(int i, int j, int k = 0) const { return (k * njc + j) * niv + i; }
    int ifj_index(int i, int j, int k = 0) const { return (k * njv + j) * nic + i + niv * njc * nkc; }
    int ifk_index(int i, int j, int k = 0) const {
        return (k * njc + j) * nic + i + niv * njc * nkc + nic * njv * nkc;
    }
    FVCell* get_cell(int i, int j, int k = 0) { return cells[cell_index(i, j, k)]; }
    FVInterface* get_ifi(int i, int j, int k = 0) { return faces[ifi_index(i, j, k)]; }
    FVInterface* get_ifj(int i, int j, int k = 0) { return faces[ifj_index(i, j, k)]; }
    FVInterface* get_ifk(int i, int j, int k = 0) { return faces[ifk_index(i, j, k)]; }

    void compute_geometry();
    void set_flowstate(int cell, const FlowState& fs);
    void encode_conserved(FVCell* c, int ftl);
    int decode_conserved(FVCell* c, int ftl);
    void time_derivatives(FVCell* c, int ftl);
    double signal_frequency(FVCell* c);
    void clear_fluxes();
    void apply_pre_recon(std::vector<SFluidBlock*>& all_blocks);
    void convective_flux_phase0(bool allow_high_order);
    void apply_post_conv_flux();
    int count_invalid_cells(int ftl);
    void determine_time_step_size(double dt_current, double cfl_value, bool check_cfl, double out[3]);
    bool check_data(const FlowState& fs) const;
    void detect_shock_points();              // fluidblock.d:478-520
    void shock_faces_to_cells();             // fluidblock.d:568-580
    void enforce_strict_shock_detector();    // fluidblock.d:582-605

    // Work states (fluidblock.d:93, 208-209)
    FlowState Lft, Rght;
    void interp(FVInterface* f, FlowState& L, FlowState& R);
    void interp_l2r2(FVInterface* f, FVCell* cL1, FVCell* cL0, FVCell* cR0, FVCell* cR1,
                     double lL1, double lL0, double lR0, double lR1, FlowState& L, FlowState& R);
    void compute_interface_flux(FlowState& L, FlowState& R, FVInterface* f);
    void outflow_flux(const FlowState& fs, int outsign, FVInterface* f);
private:
    L2R2 l2r2;
    void interp_l2r2_scalar(double qL1, double qL0, double qR0, double qR1, double& qL, double& qR) {
        l2r2.scalar(qL1, qL0, qR0, qR1, qL, qR, cfg.apply_limiter, cfg.extrema_clipping, cfg.epsilon_van_albada,
                    cfg.lmr);
    }
    FVCell* new_ghost();
};

// Source cell (i, j, k) of ghost layer n at in-face position (a, b) of face dface,
// exchanged from face sface of a block of snic x snjc x snkc cells (full_face_copy.d).
void ffc_source_index(int dims, int dface, int sface, int orient, int a, int b, int n, int snic, int snjc,
                      int snkc, int& i, int& j, int& k);

// Pure functions exposed for unit tests (same code the block uses).
void ausmdv(const FlowState& L, const FlowState& R, std::vector<double>& F, const CQI& cqi,
            const GasModel& gm, bool entropy_fix, double factor = 1.0, bool lmr = false);
void efmflx(const FlowState& L, const FlowState& R, std::vector<double>& F, const CQI& cqi,
            const GasModel& gm, double factor = 1.0);
void hanel(const FlowState& L, const FlowState& R, std::vector<double>& F, const CQI& cqi, const GasModel& gm,
           double factor = 1.0);
void exxef(double sn, double& exx, double& ef);
// shockdetectors.d:21-113
double PJ_shock_detector(const FVInterface& f, double comp_tol, double shear_tol);
double clip_to_limits(double q, double A, double B);
void scale_mass_fractions(std::vector<double>& massf, double tolerance = 0.0,
                          double assert_error_tolerance = 0.1);

// The simulation driver: src/eilmer/simcore.d:929-1377 (explicit, fixed grid, pc)
struct SimState {
    double time = 0.0, dt_global = 0.0, dt_allow = 0.0, cfl_max = 0.0;
    int step = 0;
    int attempts_last_step = 0;
};

class ChemistryUpdate;
class ThermallyPerfectGas;

class Simulation {
public:
    Config cfg;
    std::vector<SFluidBlock*> blocks;
    SimState st;
    bool exchange_ready = false;
    // Operator-split finite-rate chemistry (config.reacting, full_T_full_R splitting).
    ChemistryUpdate* chem = nullptr;
    const ThermallyPerfectGas* chem_gas = nullptr;
    double T_frozen = 300.0;
    void chemistry_step(double dt);  // simcore.d:1619-1638
    explicit Simulation(const Config& c) : cfg(c) {
        st.dt_global = cfg.dt_init;
        // configCheckPoint1, globalconfig.d:2682-2690
        if (cfg.flux_calculator == FluxCalc::adaptive_efm_ausmdv ||
            cfg.flux_calculator == FluxCalc::adaptive_hanel_ausmdv)
            cfg.do_shock_detect = true;
    }
    ~Simulation();
    void determine_time_step_size();
    void detect_shocks();     // simcore_gasdynamic_step.d:3195-3224
    void set_up_exchange();   // cell mapping + geometry exchange, once all blocks exist
    void exchange_ghost_cell_boundary_data();
    void gasdynamic_step();   // simcore_gasdynamic_step.d:906-1575 (pc, 2 stages)
    void stage(int stage, double dt);  // one stage, no retry logic (for unit tests)
    int run(int max_steps, double max_time);
};

}  // namespace orc
