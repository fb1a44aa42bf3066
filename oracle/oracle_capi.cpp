// oracle_capi.cpp -- C entry points over the oracle, for the Python test harness (ctypes).
//
// TEST INFRASTRUCTURE ONLY.  Loaded by tests/, smoke() and the CPU-baseline leg of
// bench.py as the checker; never by the product path.
#include <cstring>
#include <string>

#include "eilmer_oracle.hpp"

using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_gas_create_ideal(double mMass, double gamma) { return new IdealGas(mMass, gamma); }
void orc_gas_destroy(void* g) { delete static_cast<GasModel*>(g); }
int orc_gas_nsp(void* g) { return static_cast<GasModel*>(g)->n_species; }

// Simple gas-state functions for unit tests.  st = [rho, p, T, u, a], massf[nsp].
static void to_gs(const GasModel* gm, const double* st, const double* massf, GasState& Q) {
    Q.rho = st[0]; Q.p = st[1]; Q.T = st[2]; Q.u = st[3]; Q.a = st[4];
    Q.massf.assign(gm->n_species, 1.0);
    Q.rho_s.assign(gm->n_species, 0.0);
    if (massf) for (int i = 0; i < gm->n_species; ++i) Q.massf[i] = massf[i];
}
static void from_gs(const GasState& Q, double* st) {
    st[0] = Q.rho; st[1] = Q.p; st[2] = Q.T; st[3] = Q.u; st[4] = Q.a;
}
int orc_gas_update(void* g, int mode, double* st, const double* massf) {
    GasModel* gm = static_cast<GasModel*>(g);
    GasState Q;
    to_gs(gm, st, massf, Q);
    try {
        switch (mode) {
        case 0: gm->update_thermo_from_pT(Q); break;
        case 1: gm->update_thermo_from_rhou(Q); break;
        case 2: gm->update_thermo_from_rhoT(Q); break;
        default: g_err = "bad mode"; return -2;
        }
        gm->update_sound_speed(Q);
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
    from_gs(Q, st);
    return 0;
}

struct Sim {
    Simulation* sim;
    const GasModel* gm;
};

void* orc_sim_create(void* gas, int dims, int flux_calc, double cfl, double dt_init, double dt_max, int cfl_count,
                     int mutate_cell_velocities) {
    Config c;
    c.dimensions = dims;
    c.flux_calculator = static_cast<FluxCalc>((flux_calc >= 0 && flux_calc <= 4) ? flux_calc : 0);
    c.cfl_value = cfl;
    c.dt_init = dt_init;
    c.dt_max = dt_max;
    c.cfl_count = cfl_count;
    c.mutate_cell_velocities = mutate_cell_velocities != 0;
    Sim* s = new Sim{new Simulation(c), static_cast<GasModel*>(gas)};
    return s;
}
// Shock-detector options (globalconfig.d:1061-1131); call before adding blocks.
void orc_sim_set_shock_detector(void* h, double comp_tol, double shear_tol, int strict, int damped_outflow,
                                int frozen, int freeze_step) {
    Config& c = static_cast<Sim*>(h)->sim->cfg;
    c.compression_tolerance = comp_tol;
    c.shear_tolerance = shear_tol;
    c.strict_shock_detector = strict != 0;
    c.damped_outflow = damped_outflow != 0;
    c.frozen_shock_detector = frozen != 0;
    c.shock_detector_freeze_step = freeze_step;
}
// Eilmer 5 (lmr) numerics on the flux path (Config::lmr).
void orc_sim_set_numerics(void* h, int lmr) { static_cast<Sim*>(h)->sim->cfg.lmr = lmr != 0; }
void orc_sim_destroy(void* h) {
    Sim* s = static_cast<Sim*>(h);
    delete s->sim;
    delete s;
}

// Inflow state given as [p, T, vx, vy, vz] + massf[nsp]; set through pT + sound speed.
static FlowState make_fs(const GasModel* gm, const double* pTv, const double* massf) {
    FlowState fs;
    fs.gas.p = pTv[0];
    fs.gas.T = pTv[1];
    fs.vel = Vec3(pTv[2], pTv[3], pTv[4]);
    fs.gas.massf.assign(gm->n_species, 1.0);
    if (massf) for (int i = 0; i < gm->n_species; ++i) fs.gas.massf[i] = massf[i];
    gm->update_thermo_from_pT(fs.gas);
    gm->update_sound_speed(fs.gas);
    fs.gas.rho_s.assign(gm->n_species, 0.0);
    for (int i = 0; i < gm->n_species; ++i) fs.gas.rho_s[i] = fs.gas.rho * fs.gas.massf[i];
    return fs;
}

// bc_kinds[6]: BCKind values; other_block/other_face/other_orientation used by exchange
// faces (other_orientation may be null: orientation 0).
// inflow: 6 x (5 + nsp) doubles, one row per face (only read for inflow faces).
int orc_sim_add_block(void* h, int nic, int njc, int nkc, const double* vtx, const int* bc_kinds,
                      const double* inflow, const int* other_block, const int* other_face,
                      const int* other_orientation) {
    Sim* s = static_cast<Sim*>(h);
    try {
        const Config& c = s->sim->cfg;
        int niv = nic + 1, njv = njc + 1, nkv = (c.dimensions == 3) ? nkc + 1 : 1;
        std::vector<Vec3> v((size_t)niv * njv * nkv);
        for (size_t n = 0; n < v.size(); ++n) v[n] = Vec3(vtx[3 * n], vtx[3 * n + 1], vtx[3 * n + 2]);
        BoundaryCondition bcs[6];
        int nsp = s->gm->n_species;
        for (int b = 0; b < 6; ++b) {
            bcs[b].kind = static_cast<BCKind>(bc_kinds[b]);
            bcs[b].ghost_cell_data_available = true;
            bcs[b].convective_flux_computed_in_bc = (bcs[b].kind == BCKind::outflow_simple_flux);
            if (bcs[b].kind == BCKind::inflow_supersonic)
                bcs[b].inflow = make_fs(s->gm, inflow + b * (5 + nsp), inflow + b * (5 + nsp) + 5);
            if (bcs[b].kind == BCKind::exchange) {
                bcs[b].other_block = other_block[b];
                bcs[b].other_face = other_face[b];
                bcs[b].other_orientation = other_orientation ? other_orientation[b] : 0;
            }
        }
        int id = (int)s->sim->blocks.size();
        s->sim->blocks.push_back(new SFluidBlock(id, nic, njc, nkc, c, s->gm, v, bcs));
        return id;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void orc_ffc_source(int dims, int dface, int sface, int orient, int a, int b, int n, int snic, int snjc, int snkc,
                    int* ijk) {
    ffc_source_index(dims, dface, sface, orient, a, b, n, snic, snjc, snkc, ijk[0], ijk[1], ijk[2]);
}

// Per interior cell: p, T, vx, vy, vz, and massf (ncell x nsp, may be null for nsp == 1).
int orc_block_set_initial(void* h, int b, const double* p, const double* T, const double* vx, const double* vy,
                          const double* vz, const double* massf) {
    Sim* s = static_cast<Sim*>(h);
    try {
        SFluidBlock* B = s->sim->blocks.at(b);
        int nsp = s->gm->n_species;
        for (size_t ci = 0; ci < B->cells.size(); ++ci) {
            double row[5] = {p[ci], T[ci], vx[ci], vy[ci], vz[ci]};
            FlowState fs = make_fs(s->gm, row, massf ? massf + ci * nsp : nullptr);
            B->set_flowstate((int)ci, fs);
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// The cold start from a flow file's cell values (fluidblockio_old.d:1693-1800 then
// simcore.d:318-336): rho, vel, p, a, u, T, massf (cell-major, may be null for nsp == 1) and
// S set directly, then encode, decode and dt_chem = -1.
int orc_block_set_fs(void* h, int b, const double* rho, const double* vx, const double* vy, const double* vz,
                     const double* p, const double* a, const double* u, const double* T, const double* massf,
                     const double* S) {
    Sim* s = static_cast<Sim*>(h);
    try {
        SFluidBlock* B = s->sim->blocks.at(b);
        const int nsp = s->gm->n_species;
        for (size_t ci = 0; ci < B->cells.size(); ++ci) {
            FlowState fs = B->cells[ci]->fs;
            fs.gas.rho = rho[ci];
            fs.vel = Vec3(vx[ci], vy[ci], vz[ci]);
            fs.gas.p = p[ci];
            fs.gas.a = a[ci];
            fs.gas.u = u[ci];
            fs.gas.T = T[ci];
            fs.gas.massf.assign(nsp, 1.0);
            if (massf)
                for (int i = 0; i < nsp; ++i) fs.gas.massf[i] = massf[ci * nsp + i];
            fs.S = S ? S[ci] : 0.0;
            B->set_flowstate((int)ci, fs);
            B->cells[ci]->dt_chem = -1.0;
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// field: 0 rho, 1 vx, 2 vy, 3 vz, 4 u, 5 p, 6 T, 7 a, 8 volume, 9 iLength, 10 jLength,
// 11 kLength, 12 S (shock indicator), 13-15 pos (centroid), 16 dt_chem, 100+j: U[0][j], 200+j: dUdt[0][j], 300+isp: massf[isp]
int orc_block_get(void* h, int b, int field, double* out) {
    Sim* s = static_cast<Sim*>(h);
    try {
        SFluidBlock* B = s->sim->blocks.at(b);
        for (size_t ci = 0; ci < B->cells.size(); ++ci) {
            const FVCell* c = B->cells[ci];
            double v;
            if (field >= 300) v = c->fs.gas.massf.at(field - 300);
            else if (field >= 200) v = c->dUdt[0].at(field - 200);
            else if (field >= 100) v = c->U[0].at(field - 100);
            else switch (field) {
                case 0: v = c->fs.gas.rho; break;
                case 1: v = c->fs.vel.x; break;
                case 2: v = c->fs.vel.y; break;
                case 3: v = c->fs.vel.z; break;
                case 4: v = c->fs.gas.u; break;
                case 5: v = c->fs.gas.p; break;
                case 6: v = c->fs.gas.T; break;
                case 7: v = c->fs.gas.a; break;
                case 8: v = c->volume; break;
                case 9: v = c->iLength; break;
                case 10: v = c->jLength; break;
                case 11: v = c->kLength; break;
                case 12: v = c->fs.S; break;
                case 13: v = c->pos.x; break;
                case 14: v = c->pos.y; break;
                case 15: v = c->pos.z; break;
                case 16: v = c->dt_chem; break;
                default: g_err = "bad field"; return -2;
            }
            out[ci] = v;
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// PJ detector on one face between two cells: vel*, n, t1, t2 are 3-vectors.
double orc_pj_detector(const double* velL, double aL, const double* velR, double aR, const double* n,
                       const double* t1, const double* t2, double comp_tol, double shear_tol) {
    FVCell L, R;
    L.fs.vel = Vec3(velL[0], velL[1], velL[2]);
    L.fs.gas.a = aL;
    R.fs.vel = Vec3(velR[0], velR[1], velR[2]);
    R.fs.gas.a = aR;
    FVInterface f;
    f.n = Vec3(n[0], n[1], n[2]);
    f.t1 = Vec3(t1[0], t1[1], t1[2]);
    f.t2 = Vec3(t2[0], t2[1], t2[2]);
    f.left_cells.push_back(&L);
    f.right_cells.push_back(&R);
    return PJ_shock_detector(f, comp_tol, shear_tol);
}

// Shock indicator of every face (fs.S), in face order.
int orc_block_face_S(void* h, int b, double* out) {
    Sim* s = static_cast<Sim*>(h);
    SFluidBlock* B = s->sim->blocks.at(b);
    for (size_t n = 0; n < B->faces.size(); ++n) out[n] = B->faces[n]->S;
    return (int)B->faces.size();
}

// Face geometry of face index range: out rows of [nx,ny,nz,t1x,t1y,t1z,t2x,t2y,t2z,area]
int orc_block_faces(void* h, int b, double* out) {
    Sim* s = static_cast<Sim*>(h);
    SFluidBlock* B = s->sim->blocks.at(b);
    for (size_t n = 0; n < B->faces.size(); ++n) {
        const FVInterface* f = B->faces[n];
        double* r = out + 10 * n;
        r[0] = f->n.x; r[1] = f->n.y; r[2] = f->n.z;
        r[3] = f->t1.x; r[4] = f->t1.y; r[5] = f->t1.z;
        r[6] = f->t2.x; r[7] = f->t2.y; r[8] = f->t2.z;
        r[9] = f->area;
    }
    return (int)B->faces.size();
}

// Face fluxes after the last stage, rows of F[cqi.n] in face order.
int orc_block_face_flux(void* h, int b, double* out) {
    Sim* s = static_cast<Sim*>(h);
    SFluidBlock* B = s->sim->blocks.at(b);
    int n = B->cqi.n;
    for (size_t k = 0; k < B->faces.size(); ++k)
        for (int j = 0; j < n; ++j) out[k * n + j] = B->faces[k]->F[j];
    return (int)B->faces.size();
}

// st_out = [time, dt_global, step, cfl_max, dt_allow, attempts_last_step]
static void sim_state(Sim* s, double* st_out) {
    if (!st_out) return;
    st_out[0] = s->sim->st.time;
    st_out[1] = s->sim->st.dt_global;
    st_out[2] = s->sim->st.step;
    st_out[3] = s->sim->st.cfl_max;
    st_out[4] = s->sim->st.dt_allow;
    st_out[5] = s->sim->st.attempts_last_step;
}

int orc_sim_run(void* h, int max_steps, double max_time, double* st_out) {
    Sim* s = static_cast<Sim*>(h);
    try {
        int n = s->sim->run(max_steps, max_time);
        sim_state(s, st_out);
        return n;
    } catch (std::exception& e) {
        g_err = e.what();
        sim_state(s, st_out);
        return -1;
    }
}

// One full gasdynamic step (with retries) at the given dt (no dt control).
int orc_sim_step_fixed_dt(void* h, double dt, double* st_out) {
    Sim* s = static_cast<Sim*>(h);
    try {
        s->sim->st.dt_global = dt;
        s->sim->gasdynamic_step();
        s->sim->st.step += 1;
        sim_state(s, st_out);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// One stage at dt; stage 1 leaves fs = decode(U1), stage 2 fs = decode(U2) (U not swapped).
int orc_sim_stage(void* h, int stage, double dt) {
    Sim* s = static_cast<Sim*>(h);
    try {
        s->sim->stage(stage, dt);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Block-level dt estimate: out = [dt_allow, cfl_max, 0]
int orc_block_dt(void* h, int b, double dt_current, double cfl, int check_cfl, double* out) {
    Sim* s = static_cast<Sim*>(h);
    s->sim->blocks.at(b)->determine_time_step_size(dt_current, cfl, check_cfl != 0, out);
    return 0;
}

void orc_sim_set_dt(void* h, double dt_global, int step) {
    Sim* s = static_cast<Sim*>(h);
    s->sim->st.dt_global = dt_global;
    s->sim->st.step = step;
}

void orc_sim_set_time(void* h, double time) { static_cast<Sim*>(h)->sim->st.time = time; }

// Unit-level flux calculators: L, R = [rho, p, T, u, a, velx, vely, velz] in the local frame.
int orc_flux(void* gas, int flux_calc, int dims, const double* L, const double* R, const double* massfL,
             const double* massfR, double* F, int lmr) {
    GasModel* gm = static_cast<GasModel*>(gas);
    CQI cqi;
    cqi.init(dims, gm->n_species);
    FlowState l, r;
    l.gas.rho = L[0]; l.gas.p = L[1]; l.gas.T = L[2]; l.gas.u = L[3]; l.gas.a = L[4]; l.vel = Vec3(L[5], L[6], L[7]);
    r.gas.rho = R[0]; r.gas.p = R[1]; r.gas.T = R[2]; r.gas.u = R[3]; r.gas.a = R[4]; r.vel = Vec3(R[5], R[6], R[7]);
    l.gas.massf.assign(gm->n_species, 1.0);
    r.gas.massf.assign(gm->n_species, 1.0);
    if (massfL) for (int i = 0; i < gm->n_species; ++i) l.gas.massf[i] = massfL[i];
    if (massfR) for (int i = 0; i < gm->n_species; ++i) r.gas.massf[i] = massfR[i];
    std::vector<double> Fv(cqi.n, 0.0);
    if (flux_calc == 1) efmflx(l, r, Fv, cqi, *gm);
    else if (flux_calc == 3) hanel(l, r, Fv, cqi, *gm);
    else ausmdv(l, r, Fv, cqi, *gm, true, 1.0, lmr != 0);
    for (int j = 0; j < cqi.n; ++j) F[j] = Fv[j];
    return cqi.n;
}

void orc_exxef(double sn, double* exx, double* ef) { exxef(sn, *exx, *ef); }

// l2r2 scalar reconstruction through the oracle's own L2R2 (same code the block uses).
void orc_l2r2_scalar(const double* len4, const double* q4, int apply_limiter, int extrema_clipping, double eps,
                     double* qLqR, int lmr) {
    L2R2 c;
    c.prepare(len4[0], len4[1], len4[2], len4[3]);
    c.scalar(q4[0], q4[1], q4[2], q4[3], qLqR[0], qLqR[1], apply_limiter != 0, extrema_clipping != 0, eps, lmr != 0);
}

}  // extern "C"

//---------------------------------------------------------------------------------
// Thermochemistry (chem_oracle.hpp).  The structures below have the same layout as the
// gdtk_tpg_gas / gdtk_mechanism records of the C ABI so that one ctypes record can be
// handed to both the oracle and the library.
#include "chem_oracle.hpp"

namespace {
constexpr int MS = 8, MR = 32;
struct OGas {
    int nsp;
    double M[MS];
    int nseg[MS];
    double T_breaks[MS][4];
    double T_blends[MS][2];
    double coeffs[MS][3][9];
};
struct OReaction {
    int collider;
    int nreac, reac_sp[3], reac_nu[3];
    int nprod, prod_sp[3], prod_nu[3];
    int neff, eff_sp[MS];
    double eff[MS];
    int f_from_eq; double fA, fn, fC;
    int b_from_eq; double bA, bn, bC;
    int f_park; double f_s;
    int b_park; double b_s;
};
struct OMech {
    int nr;
    OReaction r[MR];
    double T_lower, T_upper;
    int method;
    double eps1, eps2, delta;
    int maxIters;
    double errTol;
    int tightTempCoupling, maxSubcycles, maxAttempts;
    double T_frozen;
};
struct OChem {
    ThermallyPerfectGas* gas;
    ChemistryUpdate* chem;
    double T_frozen;
};
ThermallyPerfectGas* make_tpg(const OGas* g) {
    std::vector<double> M(g->M, g->M + g->nsp);
    std::vector<CEAThermoCurve> curves;
    for (int i = 0; i < g->nsp; ++i) {
        double R = R_universal / g->M[i];
        const int ns = g->nseg[i];
        std::vector<double> tb(g->T_breaks[i], g->T_breaks[i] + ns + 1), bl(g->T_blends[i], g->T_blends[i] + ns - 1);
        std::vector<std::array<double, 9>> co(ns);
        for (int s = 0; s < ns; ++s) for (int j = 0; j < 9; ++j) co[s][j] = g->coeffs[i][s][j];
        curves.emplace_back(R, tb, bl, co);
    }
    return new ThermallyPerfectGas(M, curves);
}
}  // namespace

extern "C" {

int orc_tpg_sizes(int* sizes) {
    sizes[0] = (int)sizeof(OGas);
    sizes[1] = (int)sizeof(OMech);
    sizes[2] = (int)sizeof(OReaction);
    return 0;
}

void* orc_tpg_create(const void* gas) { return make_tpg(static_cast<const OGas*>(gas)); }

// A single CEA curve (species isp of gas): out = {Cp, h, s} at T.
void orc_cea_eval(void* gas, int isp, double T, double* out) {
    auto* g = static_cast<ThermallyPerfectGas*>(gas);
    out[0] = g->curves[isp].eval_Cp(T);
    out[1] = g->curves[isp].eval_h(T);
    out[2] = g->curves[isp].eval_s(T);
}

// st = [rho, p, T, u, a]; mode 0 pT, 1 rhou, 2 rhoT, 3 rhop; then sound speed.
int orc_tpg_update(void* gas, int mode, double* st, const double* massf, int* newton_evals) {
    auto* g = static_cast<ThermallyPerfectGas*>(gas);
    GasState Q;
    Q.rho = st[0]; Q.p = st[1]; Q.T = st[2]; Q.u = st[3]; Q.a = st[4];
    Q.massf.assign(massf, massf + g->n_species);
    try {
        switch (mode) {
        case 0: g->update_thermo_from_pT(Q); break;
        case 1: g->update_thermo_from_rhou(Q); break;
        case 2: g->update_thermo_from_rhoT(Q); break;
        case 3: g->update_thermo_from_rhop(Q); break;
        default: g_err = "bad mode"; return -2;
        }
        g->update_sound_speed(Q);
    } catch (std::exception& e) {
        g_err = e.what();
        st[0] = Q.rho; st[1] = Q.p; st[2] = Q.T; st[3] = Q.u;
        return -1;
    }
    if (newton_evals) *newton_evals = g->newton_evals;
    st[0] = Q.rho; st[1] = Q.p; st[2] = Q.T; st[3] = Q.u; st[4] = Q.a;
    return 0;
}

static std::vector<ReactionDef> make_defs(const OMech* md) {
    std::vector<ReactionDef> defs;
    for (int ir = 0; ir < md->nr; ++ir) {
        const OReaction& r = md->r[ir];
        ReactionDef d;
        d.collider = r.collider != 0;
        for (int k = 0; k < r.nreac; ++k) d.reactants.push_back({r.reac_sp[k], r.reac_nu[k]});
        for (int k = 0; k < r.nprod; ++k) d.products.push_back({r.prod_sp[k], r.prod_nu[k]});
        for (int k = 0; k < r.neff; ++k) d.efficiencies.push_back({r.eff_sp[k], r.eff[k]});
        d.frc.fromEqConst = r.f_from_eq != 0; d.frc.A = r.fA; d.frc.n = r.fn; d.frc.C = r.fC;
        d.brc.fromEqConst = r.b_from_eq != 0; d.brc.A = r.bA; d.brc.n = r.bn; d.brc.C = r.bC;
        d.frc.park = r.f_park != 0; d.frc.s = r.f_s;
        d.brc.park = r.b_park != 0; d.brc.s = r.b_s;
        defs.push_back(d);
    }
    return defs;
}

static ChemConfig make_cfg(const OMech* md) {
    ChemConfig cfg;
    cfg.method = md->method == 1 ? OdeMethod::rkf : OdeMethod::alpha_qss;
    cfg.eps1 = md->eps1; cfg.eps2 = md->eps2; cfg.delta = md->delta; cfg.maxIters = md->maxIters;
    cfg.errTol = md->errTol; cfg.tightTempCoupling = md->tightTempCoupling != 0;
    cfg.maxSubcycles = md->maxSubcycles; cfg.maxAttempts = md->maxAttempts;
    cfg.T_lower = md->T_lower; cfg.T_upper = md->T_upper;
    return cfg;
}

void* orc_chem_create(const void* gasv, const void* mechv) {
    const OGas* gd = static_cast<const OGas*>(gasv);
    const OMech* md = static_cast<const OMech*>(mechv);
    ThermallyPerfectGas* gas = make_tpg(gd);
    std::vector<ReactionDef> defs = make_defs(md);
    ChemConfig cfg;
    cfg.method = md->method == 1 ? OdeMethod::rkf : OdeMethod::alpha_qss;
    cfg.eps1 = md->eps1; cfg.eps2 = md->eps2; cfg.delta = md->delta; cfg.maxIters = md->maxIters;
    cfg.errTol = md->errTol; cfg.tightTempCoupling = md->tightTempCoupling != 0;
    cfg.maxSubcycles = md->maxSubcycles; cfg.maxAttempts = md->maxAttempts;
    cfg.T_lower = md->T_lower; cfg.T_upper = md->T_upper;
    return new OChem{gas, new ChemistryUpdate(gas, defs, cfg), md->T_frozen};
}

// Operator-split chemistry of a simulation (its gas must be the chemistry's gas).
int orc_sim_set_chemistry(void* h, void* chem) {
    Sim* s = static_cast<Sim*>(h);
    auto* c = static_cast<OChem*>(chem);
    s->sim->chem = c->chem;
    s->sim->chem_gas = c->gas;
    s->sim->T_frozen = c->T_frozen;
    return 0;
}

void orc_chem_destroy(void* h) {
    auto* c = static_cast<OChem*>(h);
    delete c->chem;
    delete c->gas;
    delete c;
}

// Rate constants at the state (rho, T, massf): kf[nr], kb[nr].
int orc_chem_rate_constants(void* h, double T, double* kf, double* kb) {
    auto* c = static_cast<OChem*>(h);
    GasState Q;
    Q.T = T;
    Q.massf.assign(c->gas->n_species, 0.0);
    c->chem->rmech.eval_rate_constants(Q);
    for (int i = 0; i < c->chem->rmech.nr; ++i) { kf[i] = c->chem->rmech.k_f[i]; kb[i] = c->chem->rmech.k_b[i]; }
    return 0;
}

// Split rates q, L at concentrations conc (rate constants from the last evaluation).
int orc_chem_split_rates(void* h, const double* conc, double* q, double* L) {
    auto* c = static_cast<OChem*>(h);
    c->chem->rmech.eval_split_rates(conc, q, L);
    return 0;
}

// One ODE step from y0 over h: returns 1 success, 0 failure; hSuggest out.
int orc_chem_ode_step(void* h, int method, const double* y0, double hh, double* yOut, double* hSuggest) {
    auto* c = static_cast<OChem*>(h);
    if (method == 1) return c->chem->rkf_step(y0, hh, yOut, *hSuggest) ? 1 : 0;
    return c->chem->alpha_qss_step(y0, hh, yOut, *hSuggest) ? 1 : 0;
}

// ChemistryUpdate.opCall on one state: st = [rho, p, T, u], massf in/out; returns 0 ok, -1 failed.
int orc_chem_update(void* h, double* st, double* massf, double tInterval, double* dtSuggest, double* stats) {
    auto* c = static_cast<OChem*>(h);
    GasState Q;
    Q.rho = st[0]; Q.p = st[1]; Q.T = st[2]; Q.u = st[3];
    Q.massf.assign(massf, massf + c->gas->n_species);
    ChemStats cs;
    bool ok = false;
    try {
        ok = c->chem->update(Q, tInterval, *dtSuggest, &cs);
    } catch (std::exception& e) {
        g_err = e.what();
        ok = false;
    }
    for (int i = 0; i < c->gas->n_species; ++i) massf[i] = Q.massf[i];
    st[0] = Q.rho; st[1] = Q.p; st[2] = Q.T; st[3] = Q.u;
    if (stats) { stats[0] = cs.subcycles; stats[1] = cs.corrector_iters; stats[2] = cs.attempts; stats[3] = cs.failures; }
    return ok ? 0 : -1;
}

// Batched thermochemical_increment (fvcell.d:858-972) over n cells, species-major massf,
// with nthreads OpenMP threads (one oracle object per thread).  sub[n] = substeps per cell,
// att[n] = ODE step attempts per cell (either may be NULL).
int orc_chem_step_batch(const void* gasv, const void* mechv, int n, const double* rho, double* T, double* u,
                        double* p, double* massf, double* dt_chem, double dt, int* status, int* sub, int* att,
                        int nthreads) {
    int failed = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : failed)
    {
        OChem* c = static_cast<OChem*>(orc_chem_create(gasv, mechv));
        const int nsp = c->gas->n_species;
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            GasState Q;
            Q.rho = rho[i]; Q.T = T[i]; Q.u = u[i]; Q.p = p[i];
            Q.massf.resize(nsp);
            for (int s = 0; s < nsp; ++s) Q.massf[s] = massf[(size_t)s * n + i];
            ChemStats cs;
            int rc = thermochemical_increment(*c->chem, *c->gas, Q, dt, dt_chem[i], c->T_frozen, &cs);
            status[i] = rc;
            if (sub) sub[i] = cs.subcycles;
            if (att) att[i] = cs.attempts;
            if (rc != 0) { ++failed; continue; }
            T[i] = Q.T; u[i] = Q.u; p[i] = Q.p;
            for (int s = 0; s < nsp; ++s) massf[(size_t)s * n + i] = Q.massf[s];
        }
        orc_chem_destroy(c);
    }
    return failed;
}

void orc_gas_destroy_tpg(void* g) { delete static_cast<ThermallyPerfectGas*>(g); }

}  // extern "C"

//---------------------------------------------------------------------------------
// Two-temperature air (tt_oracle.hpp).  st = [rho, p, T, Tv, u, uv].
#include "tt_oracle.hpp"

namespace {
struct OTT {
    ThermallyPerfectGas* tpg;
    TwoTemperatureAir* gas;
    TwoTemperatureAirKinetics* kin;
    double T_frozen;
};
void to_tt(const double* st, const double* massf, int nsp, TTState& Q) {
    Q.rho = st[0]; Q.p = st[1]; Q.T = st[2]; Q.Tv = st[3]; Q.u = st[4]; Q.uv = st[5];
    Q.massf.assign(massf, massf + nsp);
}
void from_tt(const TTState& Q, double* st, double* massf) {
    st[0] = Q.rho; st[1] = Q.p; st[2] = Q.T; st[3] = Q.Tv; st[4] = Q.u; st[5] = Q.uv;
    if (massf) for (size_t i = 0; i < Q.massf.size(); ++i) massf[i] = Q.massf[i];
}
}  // namespace

extern "C" {

// molecular[nsp]: 1 for species with a vibrational mode; A_MW[nsp]: Millikan-White A.
void* orc_tt_create(const void* gasv, const void* mechv, const int* molecular, const double* A_MW) {
    const OGas* gd = static_cast<const OGas*>(gasv);
    const OMech* md = static_cast<const OMech*>(mechv);
    ThermallyPerfectGas* tpg = make_tpg(gd);
    std::vector<int> mol(molecular, molecular + gd->nsp);
    auto* gas = new TwoTemperatureAir(tpg->mol_masses, tpg->curves, mol);
    std::vector<double> A(A_MW, A_MW + gd->nsp);
    auto* kin = new TwoTemperatureAirKinetics(gas, tpg, make_defs(md), make_cfg(md), A);
    return new OTT{tpg, gas, kin, md->T_frozen};
}

void orc_tt_destroy(void* h) {
    auto* t = static_cast<OTT*>(h);
    delete t->kin;
    delete t->gas;
    delete t->tpg;
    delete t;
}

// mode 0 pT (rho, u, uv out), 1 rhou (T, Tv, p out), 2 rhoT (p, u, uv out); returns 0 or -1.
int orc_tt_update(void* h, int mode, double* st, const double* massf, int* newton_iters) {
    auto* t = static_cast<OTT*>(h);
    TTState Q;
    to_tt(st, massf, t->gas->nsp, Q);
    try {
        switch (mode) {
        case 0: t->gas->update_thermo_from_pT(Q); break;
        case 1: t->gas->update_thermo_from_rhou(Q); break;
        case 2: t->gas->update_thermo_from_rhoT(Q); break;
        default: g_err = "bad mode"; return -2;
        }
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
    if (newton_iters) *newton_iters = t->gas->newton_iters;
    from_tt(Q, st, nullptr);
    return 0;
}

// out = {e_ve(T, isp), Cv_ve(T, isp)} of one species.
void orc_tt_species_energy(void* h, int isp, double T, double* out) {
    auto* t = static_cast<OTT*>(h);
    out[0] = t->gas->vibElecEnergy(T, isp);
    out[1] = t->gas->vibElecCv(T, isp);
}

// Relaxation time of each species (0 for atoms) and the energy-exchange rate at a state.
void orc_tt_relaxation(void* h, const double* st, const double* massf, double* tau, double* rate) {
    auto* t = static_cast<OTT*>(h);
    TTState Q;
    to_tt(st, massf, t->gas->nsp, Q);
    *rate = t->kin->evalRate(Q);
    for (int i = 0; i < t->gas->nsp; ++i) tau[i] = t->gas->molecular[i] ? t->kin->evalRelaxationTime(Q, i) : 0.0;
}

int orc_tt_rate_constants(void* h, double T, double Tv, double* kf, double* kb) {
    auto* t = static_cast<OTT*>(h);
    try {
        t->kin->chem.rmech.eval_rate_constants_2T(T, Tv);
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
    for (int i = 0; i < t->kin->chem.rmech.nr; ++i) {
        kf[i] = t->kin->chem.rmech.k_f[i];
        kb[i] = t->kin->chem.rmech.k_b[i];
    }
    return 0;
}

// Batched two-temperature thermochemical_increment over n cells (species-major massf), with
// nthreads OpenMP threads, one oracle object per thread.  status[n] 0 / -1; sub, att per cell.
int orc_tt_step_batch(const void* gasv, const void* mechv, const int* molecular, const double* A_MW, int n,
                      const double* rho, double* T, double* Tv, double* u, double* uv, double* p, double* massf,
                      double* dt_chem, double dt, int* status, int* sub, int* att, int nthreads) {
    int failed = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : failed)
    {
        OTT* t = static_cast<OTT*>(orc_tt_create(gasv, mechv, molecular, A_MW));
        const int nsp = t->gas->nsp;
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; ++i) {
            TTState Q;
            Q.rho = rho[i]; Q.T = T[i]; Q.Tv = Tv[i]; Q.u = u[i]; Q.uv = uv[i]; Q.p = p[i];
            Q.massf.resize(nsp);
            for (int s = 0; s < nsp; ++s) Q.massf[s] = massf[(size_t)s * n + i];
            ChemStats cs;
            int rc = tt_thermochemical_increment(*t->kin, *t->gas, Q, dt, dt_chem[i], t->T_frozen, &cs);
            status[i] = rc;
            if (sub) sub[i] = cs.subcycles;
            if (att) att[i] = cs.attempts;
            if (rc != 0) { ++failed; continue; }
            T[i] = Q.T; Tv[i] = Q.Tv; u[i] = Q.u; uv[i] = Q.uv; p[i] = Q.p;
            for (int s = 0; s < nsp; ++s) massf[(size_t)s * n + i] = Q.massf[s];
        }
        orc_tt_destroy(t);
    }
    return failed;
}

}  // extern "C"
