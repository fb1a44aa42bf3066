/* gdtk_hip.h -- C ABI of the MI355X (gfx950) structured-block explicit update.
 *
 * This is the drop-in boundary for Eilmer 4's fixed-grid explicit transient update.
 * The entry points are the per-block and per-step operations that the reference's
 * time-marching core performs on its blocks, so that a D host (or any FFI) can keep
 * its own prep, I/O and control flow and hand the hot loop to the GPU.
 *
 * Conventions follow the reference's own C ABI, src/gas/gas_cwrap.d:33-80, 892-921:
 * integer handles into library-owned tables, int status returns (0 success, negative
 * failure), no exceptions across the ABI, and a retrievable message for the last error.
 * Arrays are plain host pointers in the reference's cell order
 * (k*njc + j)*nic + i (src/eilmer/sfluidblock.d:299-312) and vertex order
 * (k*njv + j)*niv + i (sfluidblock.d:315-328).  All arithmetic is IEEE FP64.
 */
#ifndef GDTK_HIP_H
#define GDTK_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define GDTK_OK 0
#define GDTK_ERR_ARG (-1)             /* bad handle or argument */
#define GDTK_ERR_HIP (-2)             /* HIP runtime error */
#define GDTK_ERR_BAD_CELL (-3)        /* decode failed in some cell: the reference's step_failed */
#define GDTK_ERR_TOO_MANY_BAD_CELLS (-4) /* count_invalid_cells > max_invalid_cells (fatal) */
#define GDTK_ERR_STEP_FAILED (-5)     /* step failed after max_attempts_for_step */
#define GDTK_ERR_UNSUPPORTED (-6)

/* Largest species count of a gas or mechanism record. */
#define GDTK_MAX_SPECIES 8

/* Flux calculators, a subset of src/eilmer/globalconfig.d:270-290. */
#define GDTK_FLUX_AUSMDV 0
#define GDTK_FLUX_EFM 1
/* adaptive_efm_ausmdv (fluxcalc.d:1315-1332, "adaptive"): EFM on faces the shock detector
 * marks, AUSMDV elsewhere; turns shock detection on (globalconfig.d:2682-2690). */
#define GDTK_FLUX_ADAPTIVE_EFM_AUSMDV 2
/* hanel (fluxcalc.d:1028-1128), and adaptive_hanel_ausmdv (fluxcalc.d:1335-1348, the
 * reference's default calculator, globalconfig.d:1031): Hanel on marked faces, AUSMDV
 * elsewhere, with shock detection on. */
#define GDTK_FLUX_HANEL 3
#define GDTK_FLUX_ADAPTIVE_HANEL_AUSMDV 4

/* Boundary conditions (src/eilmer/bc.lua:757, 1320, 1580, 1601, 1650). */
#define GDTK_BC_SLIP_WALL 0
#define GDTK_BC_INFLOW_SUPERSONIC 1
#define GDTK_BC_OUTFLOW_SIMPLE_FLUX 2
#define GDTK_BC_OUTFLOW_SIMPLE_EXTRAPOLATE 3
#define GDTK_BC_EXCHANGE 4

/* Flow-state fields for gdtk_hip_block_get_field: the FlowState/GasState members. */
#define GDTK_FIELD_RHO 0
#define GDTK_FIELD_VX 1
#define GDTK_FIELD_VY 2
#define GDTK_FIELD_VZ 3
#define GDTK_FIELD_U 4
#define GDTK_FIELD_P 5
#define GDTK_FIELD_T 6
#define GDTK_FIELD_A 7
#define GDTK_FIELD_VOLUME 8
#define GDTK_FIELD_ILENGTH 9
#define GDTK_FIELD_JLENGTH 10
#define GDTK_FIELD_KLENGTH 11
#define GDTK_FIELD_SHOCK 12 /* the cell's shock indicator fs.S (0 or 1; adaptive flux calculator) */
#define GDTK_FIELD_DT_CHEM 13 /* the cell's dt_chem (fvcell.d:78; after a chemistry step) */
#define GDTK_FIELD_U0 100   /* + conserved-quantity index: U[0][j] */
#define GDTK_FIELD_DUDT0 200 /* + conserved-quantity index: dUdt[0][j] */
#define GDTK_FIELD_MASSF 300 /* + species index: mass fraction (multi-species gas) */
#define GDTK_FIELD_RHO_S 400 /* + species index: species density (multi-species gas) */

/* Simulation-wide options: the subset of GlobalConfig (src/eilmer/globalconfig.d:936-1294)
 * read on this path.  Gas: ideal gas from mMass and gamma (src/gas/ideal_gas.d:30-72). */
typedef struct gdtk_sim_config {
    int dimensions;            /* 2 or 3 */
    int flux_calculator;       /* GDTK_FLUX_* */
    int interpolation_order;   /* 2 (l2r2); the only order implemented */
    double cfl_value;          /* 0.5 */
    double dt_init;            /* 1e-3 */
    double dt_max;             /* 1e-3 */
    int cfl_count;             /* 10 */
    int max_attempts_for_step; /* 3 */
    int max_invalid_cells;     /* 0 */
    double mMass;              /* kg/mol */
    double gamma;
    int use_fused;             /* 1: fused LDS-tiled stage kernel where the block shape allows */
} gdtk_sim_config;

/* One structured block (src/eilmer/sfluidblock.d:68-110).  Faces are numbered as the
 * reference's Face enum (src/geom/elements/nomenclature.d:12-20): west 0, east 1,
 * south 2, north 3, bottom 4, top 5. */
typedef struct gdtk_block_desc {
    int nic, njc, nkc;         /* nkc = 1 in 2-D */
    const double* vertices;    /* niv*njv*nkv x 3 coordinates */
    int bc_kind[6];            /* per face west, east, south, north, bottom, top */
    double inflow[6][5];       /* p, T, vx, vy, vz of an inflow face (FlowState(p,T,vel)) */
    /* Exchange faces (GDTK_BC_EXCHANGE, the same-process full-face copy of
     * src/eilmer/bc/ghost_cell_effect/full_face_copy.d:67-125): the neighbour block id,
     * its face and the orientation 0..3 of the connection (src/eilmer/blk_conn.lua conventions; 0 in
     * 2-D).  Connections are checked when the first stage runs, once every block exists. */
    int other_block[6];
    int other_face[6];
    int other_orientation[6];
    /* Mass fractions of an inflow face's state (multi-species gas only). */
    double inflow_massf[6][GDTK_MAX_SPECIES];
} gdtk_block_desc;

/* Replaces the GlobalConfig/SimState set-up of init_simulation (src/eilmer/simcore.d:83-720). */
int gdtk_hip_sim_create(const gdtk_sim_config* cfg, int* sim);
int gdtk_hip_sim_destroy(int sim);

/* Replaces SFluidBlock construction and compute_primary_cell_geometric_data
 * (sfluidblock.d:443-1000): builds the block on the device from its vertices. */
int gdtk_hip_sim_add_block(int sim, const gdtk_block_desc* desc, int* block_id);

/* Replaces the initial FlowState fill plus encode/decode of init_simulation
 * (simcore.d:326-330): FlowState(p, T, vel) per cell, U[0] = encode(fs), fs = decode(U[0]). */
int gdtk_hip_block_set_initial(int sim, int block, const double* p, const double* T, const double* vx,
                               const double* vy, const double* vz);

/* Multi-species thermally perfect gas for the flux path (ThermallyPerfectGas,
 * src/gas/therm_perf_gas.d, with its mixture EOS): replaces the ideal gas of the
 * configuration; call before adding blocks.  The conserved quantities then carry the
 * species densities (conservedquantities.d:147-193), the reconstruction interpolates them
 * (onedinterp.d:855-895) and decode/encode go through the mixture's Newton T(u).  2 or 5
 * species; the face-per-thread stage kernels serve this gas. */
int gdtk_hip_sim_set_gas_tpg(int sim, const struct gdtk_tpg_gas* gas);

/* gdtk_hip_block_set_initial with mass fractions, species-major massf[isp*ncell + cell]:
 * FlowState(p, T, vel, massf) per cell (multi-species gas). */
int gdtk_hip_block_set_initial_massf(int sim, int block, const double* p, const double* T, const double* vx,
                                     const double* vy, const double* vz, const double* massf);

/* Operator-split finite-rate chemistry (config.reacting with the default full_T_full_R
 * splitting, simcore.d:1089-1096): after every gasdynamic step, chemistry_step
 * (simcore.d:1619-1638) runs FVCell.thermochemical_increment (fvcell.d:858-972) on every
 * cell with dt_global and the cell's dt_chem (initially -1, simcore.d:335) through the
 * chemistry handle `chem` (gdtk_hip_chem_create, same species as the gas), or through the
 * learned surrogate `surrogate` (gdtk_hip_surrogate_create; -1: none).  A failed cell is the
 * reference's fatal FlowSolverException (fvcell.d:923): gdtk_hip_sim_run and
 * gdtk_hip_sim_step_fixed_dt return GDTK_ERR_STEP_FAILED.  The chemistry step does not wait
 * for the device: its failure count is checked when the next step's flux attempt has waited
 * for the stream, or before the call returns, so a failure can surface one flux step later. */
int gdtk_hip_sim_set_chemistry(int sim, int chem, int surrogate);
/* {chemistry steps, total ms, network cells the surrogate sent back to the integrator in the
 * last step}; returns the number of values written (<= 3). */
int gdtk_hip_sim_chem_stats(int sim, double* stats, int nstats);

/* Copies one field of all active cells to host memory (GDTK_FIELD_*). */
int gdtk_hip_block_get_field(int sim, int block, int field, double* out);

/* ---------------------------------------------------------------------------------
 * Flow-solution files (the legacy structured-block format, config.flow_format "gziptext"
 * or "rawbinary").  A cell record holds, in this order: pos.x pos.y pos.z volume rho
 * vel.x vel.y vel.z p a mu k mu_t k_t S, massf[i]-<species> for each species (1.0 for a
 * single-species gas), dt_chem (multi-species gas only), u, T -- the variable list of
 * build_flow_variable_list (src/eilmer/fluidblockio_old.d:107-151) for an inviscid,
 * laminar, single-temperature gas without MHD, radiation or local time stepping. */
#define GDTK_FLOW_GZIPTEXT 0
#define GDTK_FLOW_RAWBINARY 1

/* Replaces write_solution for an SFluidBlock (fluidblockio_old.d:388-447; called per block by
 * write_solution_files, simcore_io.d:54-136): writes the block's current flow state with
 * the given label and sim_time.  species_names: whitespace-separated species names for the
 * mass-fraction variables (NULL: "air" for the ideal gas). */
int gdtk_hip_block_write_flow(int sim, int block, const char* filename, int format, const char* label,
                              double sim_time, const char* species_names);

/* Replaces read_legacy_solution for an SFluidBlock (fluidblockio_old.d:156-273) followed by the
 * cold start of init_simulation (simcore.d:299-336): the file's rho, vel, p, a, u, T, mass
 * fractions, S, mu and k become the cell states (the file may order its variables freely;
 * position and volume are not taken), then U[0] = encode(fs), fs = decode(U[0]) and the
 * cells' dt_chem = -1.  Writes the file's sim_time.  A size mismatch or a missing variable
 * fails with GDTK_ERR_ARG, as the reference throws. */
int gdtk_hip_block_read_flow(int sim, int block, const char* filename, int format, const char* species_names,
                             double* sim_time);

/* The gas transport coefficients mu and k of a block's cells (an inviscid run carries them
 * unchanged from the initial FlowState, luaflowstate.d:490); written to the flow files
 * (0 when never set or read). */
int gdtk_hip_block_set_transport(int sim, int block, const double* mu, const double* k);

/* Sets SimState.time (a restart takes it from the flow files, simcore.d:344). */
int gdtk_hip_sim_set_time(int sim, double time);
/* The SimState as gdtk_hip_sim_run reports it: {time, dt_global, step, cfl_max, dt_allow,
 * attempts_of_last_step}. */
int gdtk_hip_sim_get_state(int sim, double* state);

/* Host-only halves of the flow files (no GPU): the number of doubles in a cell record, and
 * the writer and reader on host records (nic*njc*nkc records in the reference's cell order;
 * nkc = 1 in 2-D). */
int gdtk_hip_flow_nvars(int nsp);
int gdtk_hip_flow_write_records(const char* filename, int format, const char* label, double sim_time, int dims,
                                int nic, int njc, int nkc, int nsp, const char* species_names, const double* records);
int gdtk_hip_flow_read_records(const char* filename, int format, int dims, int nic, int njc, int nkc, int nsp,
                               const char* species_names, double* records, double* sim_time);

/* Shock-detector options of GlobalConfig (src/eilmer/globalconfig.d:1061-1131), read when the
 * flux calculator is adaptive (GDTK_FLUX_ADAPTIVE_EFM_AUSMDV, GDTK_FLUX_ADAPTIVE_HANEL_AUSMDV).  Detection (detect_shocks,
 * simcore_gasdynamic_step.d:3195-3224) runs in stage 1 of every step after the ghost-cell
 * fill, with the PJ detector (shockdetectors.d:21-113), no smoothing and the strict rule
 * (fluidblock.d:582-605). */
typedef struct gdtk_shock_detector {
    double compression_tolerance;   /* -0.30 */
    double shear_tolerance;         /* 0.20 */
    int strict_shock_detector;      /* 1 (0 is not supported) */
    int damped_outflow;             /* 0: 1 marks outflow_simple_extrapolate faces shocked */
    int frozen_shock_detector;      /* 0: 1 stops detection from shock_detector_freeze_step on */
    int shock_detector_freeze_step; /* 1000000000 */
} gdtk_shock_detector;

/* Replaces the shock-detector fields of the GlobalConfig read (globalconfig.d:1957-1990).
 * Call before the first stage; defaults as commented above. */
int gdtk_hip_sim_set_shock_detector(int sim, const gdtk_shock_detector* opts);

/* Numerics of the flux path.  GDTK_NUMERICS_EILMER4 (the default) follows src/eilmer.
 * GDTK_NUMERICS_LMR binds the same path for Eilmer 5 (src/lmr), whose structured-grid l2r2
 * (lmr/onedinterp.d:137-169, 212-380) scales epsilon_van_albada by
 * max(1e-12, |qL0|, |qR0|) * 2 / (lenR0 + lenL0), and whose AUSMDV (lmr/fluxcalc.d, ausmdv)
 * takes the common sound speed as the smooth maximum
 * 0.5 (aL + aR) + 0.5 sqrt((aL - aR)^2 + eps^2), eps = 1e-6 * 0.5 (aL + aR) + 1e-12.
 * For a multi-species gas lmr reconstructs the mass fractions themselves (scaled after
 * reconstruction) and the density (lmr/onedinterp.d:268-286, 314-315), where Eilmer 4
 * reconstructs the species densities.  The stage then runs the face-per-thread kernels. */
#define GDTK_NUMERICS_EILMER4 0
#define GDTK_NUMERICS_LMR 1
int gdtk_hip_sim_set_numerics(int sim, int numerics);

/* One stage of the explicit update over all blocks (simcore_gasdynamic_step.d:946-1426):
 * exchange, ghost-cell effects, convective_flux_phase0 (sfluidblock.d:2188), boundary flux
 * effects, time_derivatives, the pc stage update and decode_conserved.  Returns
 * GDTK_ERR_BAD_CELL where the reference would set step_failed.  All blocks are updated by
 * one launch of the ghost-cell kernel and one persistent launch of the stage kernel. */
int gdtk_hip_stage(int sim, int stage, double dt);

/* FluidBlock.determine_time_step_size (fluidblock.d:987-1084): out = {dt_allow, cfl_max, 0}. */
int gdtk_hip_block_dt(int sim, int block, double dt_current, double cfl_value, int check_cfl, double* out);

/* determine_time_step_size (simcore_gasdynamic_step.d:60-159) followed by
 * gasdynamic_explicit_increment_with_fixed_grid (simcore_gasdynamic_step.d:906-1575),
 * steps until max_steps or max_time (simcore.d:1014-1377).
 * state = {time, dt_global, step, cfl_max, dt_allow, attempts_of_last_step}. */
int gdtk_hip_sim_run(int sim, int max_steps, double max_time, double* state);

/* One gasdynamic step at a fixed dt (no dt control), with the reference's retries. */
int gdtk_hip_sim_step_fixed_dt(int sim, double dt, double* state);

/* Sets SimState.dt_global and SimState.step (for restarts and tests). */
int gdtk_hip_sim_set_dt(int sim, double dt_global, int step);

/* Kernel timing on the library's stream: enable collects HIP events around every kernel
 * launch of a stage.  stats holds 3 values {launches, total_ms, items_per_launch} for
 * each kernel kind in the order face flux (items = faces), cell update (cells), fused
 * stage (cells), ghost fill (cells), shock detection (cells): nstats up to 15. */
int gdtk_hip_sim_profile(int sim, int enable);
int gdtk_hip_sim_kernel_stats(int sim, double* stats, int nstats);

/* Project autotuner (no reference counterpart): times both stages of each launch-shape
 * candidate (face-per-thread kernels; fused kernel with tiles 64, 32 and 16 cells wide and
 * k chunks of 8..64 planes) on the current state, keeps the fastest for the rest of the run
 * and returns the number of values written to out: {best_variant, best_kchunk, best_ms,
 * best_tile_x, then (variant, kchunk, ms, tile_x) per candidate, ms per stage}.  The flow
 * state and U[0] are unchanged. */
int gdtk_hip_sim_autotune(int sim, double dt, int reps, double* out, int nout);

/* Source cell of an exchange ghost cell, as the exchange kernel maps it (the case tables
 * of set_up_cell_mapping_phase0, full_face_copy.d:696-1480): ghost layer n at in-face
 * position (a, b) of `face` -- west/east (j, k), south/north (i, k), bottom/top (i, j) --
 * connected to `other_face` of a block of nic x njc x nkc cells with `orientation`.
 * Writes the source cell's (i, j, k) to ijk[3].  Host-only; needs no GPU. */
int gdtk_hip_exchange_source(int dims, int face, int other_face, int orientation, int a, int b, int n, int nic,
                             int njc, int nkc, int* ijk);

/* Launch knobs (no reference counterpart; KCHUNK and VARIANT are chosen by the autotuner):
 * GDTK_TUNE_KCHUNK   planes in k marched by one workgroup of the fused stage kernel;
 * GDTK_TUNE_GRAPH    1 (default): each step's attempt runs as one captured HIP graph
 *                    (both stages of all blocks), 0: launched kernel by kernel;
 * GDTK_TUNE_WG_PER_CU workgroups per CU of the persistent stage launch (0: occupancy);
 * GDTK_TUNE_VARIANT  -1 (from use_fused), 0 face-per-thread kernels, 1 fused kernel;
 * GDTK_TUNE_TILE_X   tile width of the fused kernel: 64 (x 4), 32 (x 8) or 16 (x 16) cells. */
#define GDTK_TUNE_KCHUNK 1
#define GDTK_TUNE_GRAPH 2
#define GDTK_TUNE_WG_PER_CU 3
#define GDTK_TUNE_VARIANT 4
#define GDTK_TUNE_TILE_X 5
int gdtk_hip_sim_set_tuning(int sim, int key, int value);

/* ---------------------------------------------------------------------------------
 * Finite-rate chemistry behind the ThermochemicalReactor interface
 * (src/kinetics/thermochemical_reactor.d:45-47), batched over cells like the
 * reference's GPU plug-in boundary GPUChem.thermochemical_increment
 * (src/eilmer/cuda_gpu_chem.d:155-231).  Gas: thermally perfect mixture with CEA
 * NASA-9 curves (src/gas/therm_perf_gas.d); mechanism: elementary and
 * anonymous-collider reactions with Arrhenius or equilibrium-constant rates
 * (src/kinetics/reaction.d, rate_constant.d); ODE step alpha-QSS or RKF
 * (src/kinetics/chemistry_update.d). */
#define GDTK_MAX_REACTIONS 32
#define GDTK_ODE_ALPHA_QSS 0
#define GDTK_ODE_RKF 1

typedef struct gdtk_tpg_gas {
    int nsp;
    double M[GDTK_MAX_SPECIES];                 /* kg/mol */
    int nseg[GDTK_MAX_SPECIES];                 /* CEA segments per species, 2 or 3 */
    double T_breaks[GDTK_MAX_SPECIES][4];       /* nseg+1 break points */
    double T_blends[GDTK_MAX_SPECIES][2];
    double coeffs[GDTK_MAX_SPECIES][3][9];
} gdtk_tpg_gas;

typedef struct gdtk_reaction {
    int collider;                               /* anonymous collider "M" */
    int nreac, reac_sp[3], reac_nu[3];
    int nprod, prod_sp[3], prod_nu[3];
    int neff, eff_sp[GDTK_MAX_SPECIES];
    double eff[GDTK_MAX_SPECIES];
    int f_from_eq; double fA, fn, fC;           /* forward: Arrhenius or from K_c */
    int b_from_eq; double bA, bn, bC;           /* backward: Arrhenius or from K_c */
    /* Park2TRateConstant (rate_constant.d:365-395): A Te^n exp(-C/Te) at Te = T^s Tv^(1-s);
     * the two-temperature reactor only (gdtk_hip_chem_create refuses it) */
    int f_park; double f_s;
    int b_park; double b_s;
} gdtk_reaction;

typedef struct gdtk_mechanism {
    int nr;
    gdtk_reaction r[GDTK_MAX_REACTIONS];
    double T_lower, T_upper;                    /* config.tempLimits */
    int method;                                 /* GDTK_ODE_* */
    double eps1, eps2, delta;                   /* alpha-QSS */
    int maxIters;
    double errTol;                              /* RKF */
    /* tightTempCoupling (chemistry_update.d:265-269): T(u) and the rate constants after every
     * subcycle; with the alpha-QSS step only (GDTK_ERR_UNSUPPORTED with RKF) */
    int tightTempCoupling, maxSubcycles, maxAttempts;
    double T_frozen;                            /* globalconfig T_frozen (300 K) */
} gdtk_mechanism;

/* Replaces init_thermochemical_reactor (src/kinetics/init_thermochemical_reactor.d:54-138)
 * and the GPUChem constructor (cuda_gpu_chem.d:71-150). */
int gdtk_hip_chem_create(const gdtk_tpg_gas* gas, const gdtk_mechanism* mech, int* chem);
int gdtk_hip_chem_destroy(int chem);

/* FVCell.thermochemical_increment (src/eilmer/fvcell.d:858-972) for n cells at once:
 * ChemistryUpdate.opCall over dt with the cell's dt_chem (retried once with dt_chem = -1
 * on failure), then update_thermo_from_rhou.  Host arrays: rho[n]; T, u, p, dt_chem [n]
 * in/out; massf[nsp*n] in/out, species-major (massf[isp*n + i]); status[n] out
 * (0 ok, -1 failed cell).  stats (may be NULL, 4 values) = {total substeps, total
 * corrector iterations, max substeps of a cell, kernel ms}.  Returns the number of
 * failed cells (>= 0) or a negative GDTK_ERR_* code. */
int gdtk_hip_chem_step(int chem, int n, const double* rho, double* T, double* u, double* p, double* massf,
                       double* dt_chem, double dt, int* status, double* stats);

/* gdtk_hip_chem_step on device-resident arrays, out of place and asynchronous: the
 * work is queued on `stream` (a hipStream_t; NULL = the default stream) and nothing
 * is copied to or from the host.  Inputs rho, T, u, p, dt_chem [n] and massf [nsp*n];
 * outputs T_out, u_out, p_out, dt_chem_out, status [n] and massf_out [nsp*n], every entry
 * written (cells that fail or are frozen get their input values, as the reference leaves
 * them).  Outputs may be the input arrays.  Same per-cell semantics as gdtk_hip_chem_step;
 * the same chemistry handle must not run two steps at once. */
int gdtk_hip_chem_step_device(int chem, int n, const double* rho, const double* T, const double* u, const double* p,
                              const double* massf, const double* dt_chem, double dt, double* T_out, double* u_out,
                              double* p_out, double* massf_out, double* dt_chem_out, int* status, void* stream);

/* gdtk_hip_chem_step_device with the same arguments and results, but allocating nothing: the
 * step's work buffers (rate constants, concentrations, step sizes, flags, queue counters,
 * statistics) live in a fixed device array of GDTK_CHEM_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  n above gdtk_hip_chem_noalloc_max_cells (which depends on the
 * mechanism's species and reaction counts) is refused with GDTK_ERR_ARG.  All calls on one
 * GPU share its array: calls that overlap on one GPU race, and gdtk_hip_chem_stats and
 * gdtk_hip_chem_substeps read the array as the last call on that GPU left it. */
#define GDTK_CHEM_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
int gdtk_hip_chem_step_device_noalloc(int chem, int n, const double* rho, const double* T, const double* u,
                                      const double* p, const double* massf, const double* dt_chem, double dt,
                                      double* T_out, double* u_out, double* p_out, double* massf_out,
                                      double* dt_chem_out, int* status, void* stream);
/* The largest cell count of one gdtk_hip_chem_step_device_noalloc call for this handle. */
int gdtk_hip_chem_noalloc_max_cells(int chem);

/* Statistics of the last step (waits for it): {total substeps, total corrector
 * iterations, max substeps of a cell, total ODE step attempts, ms setup kernel, ms
 * integration kernel, ms finish kernel, cells}; returns the number of values written
 * (<= nstats, at most 8). */
int gdtk_hip_chem_stats(int chem, double* stats, int nstats);

/* Substeps taken by each cell of the last step (both attempts of a retried cell; 0 for
 * cells not integrated).  n must be that step's cell count. */
int gdtk_hip_chem_substeps(int chem, int n, int* substeps);

/* Launch shape of the integration kernel (set by gdtk_hip_chem_autotune):
 * GDTK_CHEM_TUNE_PERSISTENT  1 persistent lanes fed by a work queue, 0 one lane per cell,
 *                            -1 (default) persistent above 2048 cells per CU
 * GDTK_CHEM_TUNE_CHUNK       cells taken from the queue per wave grab
 * GDTK_CHEM_TUNE_REFILL      substep trips between refills of finished lanes
 * GDTK_CHEM_TUNE_BLOCKS_PER_CU  persistent workgroups per CU (0 = occupancy limit)
 * GDTK_CHEM_TUNE_HEAVY_RATIO cells whose first substep is shorter than dt / value are
 *                            queued first in the persistent variant (0 = no such list;
 *                            default 4), so the longest integrations start at once */
#define GDTK_CHEM_TUNE_PERSISTENT 1
#define GDTK_CHEM_TUNE_CHUNK 2
#define GDTK_CHEM_TUNE_REFILL 3
#define GDTK_CHEM_TUNE_BLOCKS_PER_CU 4
#define GDTK_CHEM_TUNE_HEAVY_RATIO 5
int gdtk_hip_chem_set_tuning(int chem, int key, int value);

/* Times the integration variants on a sample batch of device arrays (left unchanged) and
 * keeps the fastest.  out (may be NULL) receives 5 values per candidate {persistent,
 * chunk, refill, blocks_per_cu, ms}; returns the number of candidates. */
int gdtk_hip_chem_autotune(int chem, int n, const double* rho, const double* T, const double* u, const double* p,
                           const double* massf, const double* dt_chem, double dt, double* out, int nout);

/* Thermodynamic updates of n gas states on the device, for checking the gas model:
 * mode 0 pT (p, T -> rho, u), 1 rhou (rho, u -> T, p; T in is the Newton start),
 * 2 rhoT (rho, T -> p, u); then the sound speed a.  massf species-major as above. */
int gdtk_hip_tpg_update(int chem, int mode, int n, double* rho, double* p, double* T, double* u, double* a,
                        const double* massf);

/* ---------------------------------------------------------------------------------
 * Learned chemistry surrogate (no counterpart in the reference: the building blocks of a
 * drop-in for gdtk_hip_chem_step_device).  A small MLP maps standardised features of a cell
 * (log T, log rho, Y[nsp], log dt) to its mass-fraction change over dt; it is trained on
 * this library's own alpha-QSS results with the FP32 kernels below, and evaluated in BF16
 * on the matrix cores with FP32 accumulation. */
typedef struct gdtk_mlp_desc {
    int n_in;            /* 1..32 */
    int hidden;          /* 256 (3 hidden layers) or 64 (2 hidden layers) */
    int n_hidden_layers;
    int n_out;           /* 1..16 */
} gdtk_mlp_desc;

/* He-normal initial weights from seed, zero biases. */
int gdtk_hip_mlp_create(const gdtk_mlp_desc* desc, unsigned long long seed, int* net);
int gdtk_hip_mlp_destroy(int net);
/* Parameter count (>= 0) or a negative error.  Layout: per layer W [out][in] row-major, then b [out]. */
int gdtk_hip_mlp_nparams(int net);
int gdtk_hip_mlp_get_params(int net, float* params, int n);
int gdtk_hip_mlp_set_params(int net, const float* params, int n);
/* BF16-aware training (on != 0): the training forward pass rounds inputs, weights and hidden
 * activations to BF16 as the inference kernel does; gradients pass the roundings unchanged
 * and update the FP32 parameters.  Off by default. */
int gdtk_hip_mlp_set_bf16_training(int net, int on);
/* FP32 forward pass on host arrays x [batch][n_in] -> y [batch][n_out] (cell-major). */
int gdtk_hip_mlp_forward(int net, int batch, const float* x, float* y);
/* Mean-squared-error loss of a host batch and its gradient (grad [nparams], may be NULL). */
int gdtk_hip_mlp_loss_grad(int net, int batch, const float* x, const float* y, float* loss, float* grad);
/* Adam training on device arrays x [n_in][n], y [n_out][n] (species-major, FP32): shuffled
 * minibatches, `epochs` passes; loss_hist[epoch] (may be NULL) = mean batch loss.  Returns the
 * number of epochs run. */
int gdtk_hip_mlp_train(int net, int n, const float* x, const float* y, int batch, int epochs, float lr,
                       unsigned long long seed, float* loss_hist, int nhist);
/* BF16 matrix-core inference on device arrays x [n_in][n] -> y [n_out][n] on `stream`. */
int gdtk_hip_mlp_infer(int net, int n, const float* x, float* y, void* stream);
/* The matrix-core lane layout used by the inference kernel, on one 16x32 by 32x16 product
 * (host arrays, row-major; values exact in BF16): D = A B. */
int gdtk_hip_mfma_probe(const float* A, const float* B, float* D);

#define GDTK_SURR_MAX_FEAT 11
#define GDTK_SURR_MAX_ELEM 4
typedef struct gdtk_surrogate_cfg {
    double x_mean[GDTK_SURR_MAX_FEAT];     /* features: log T, log rho, Y[0..nsp-1], log dt */
    double x_inv_std[GDTK_SURR_MAX_FEAT];
    double y_scale[GDTK_MAX_SPECIES];      /* mass-fraction change = network output * y_scale */
    double x_abs_max;                      /* a standardised feature beyond this queues the cell */
    double disagree_tol;                   /* max |dY_a - dY_b| beyond this queues the cell */
    double elem[GDTK_SURR_MAX_ELEM][GDTK_MAX_SPECIES]; /* element mass per unit species mass */
    int n_elem;
    /* Routing: a cell above T_frozen goes to the networks if its integration would be long,
     * i.e. dt > route_ratio * its dt_chem (the last suggested substep), or, with dt_chem <= 0,
     * T > route_T; otherwise to the integrator.  route_ratio <= 0 sends every cell above
     * T_frozen to the networks. */
    double route_ratio, route_T;
} gdtk_surrogate_cfg;

/* Two networks (nsp + 3 inputs, nsp outputs) and the chemistry handle that serves the
 * queued cells. */
int gdtk_hip_surrogate_create(int chem, int net_a, int net_b, const gdtk_surrogate_cfg* cfg, int* surrogate);
int gdtk_hip_surrogate_destroy(int surrogate);
/* Same arguments and per-cell results as gdtk_hip_chem_step_device, plus path [n] (may be
 * NULL): 0 network, 1 alpha-QSS integrator, 2 frozen, 3 failed.  Network cells (routed by
 * the configuration's route_ratio and route_T): mean of the two networks' changes, projected
 * so that every element's mass is conserved and no mass fraction is negative, then T by
 * Newton from the unchanged u and p = rho R T; dt_chem is passed through.  A network cell
 * outside the trained range, whose networks disagree or whose projection is infeasible goes
 * to the integrator.  A network composition that admits no temperature is reported failed
 * (status -1, state unchanged, dt_chem -1), the reference's failure mode for such cells.
 * Asynchronous on `stream`: nothing is read back to the host. */
int gdtk_hip_surrogate_step_device(int surrogate, int n, const double* rho, const double* T, const double* u,
                                   const double* p, const double* massf, const double* dt_chem, double dt,
                                   double* T_out, double* u_out, double* p_out, double* massf_out,
                                   double* dt_chem_out, int* status, int* path, void* stream);
/* {cells, network cells sent back to the integrator, ms routing + features + networks, ms
 * projection + T(u), ms integrator, max network disagreement, network cells} of the last
 * step (waits for it). */
int gdtk_hip_surrogate_stats(int surrogate, double* stats, int nstats);
/* The element-conservation projection of the surrogate step on its own, for device arrays
 * Y0, d0, d [nsp][n] (species-major) and ok [n], on `stream`: d is the change nearest d0 (least
 * squares) with E d = 0 and Y0 + d >= 0, where E = elem [n_elem][nsp] (host, row-major; element
 * mass per unit species mass); ok[i] = 0 and d = 0 where the free species cannot carry the
 * element balance.  No counterpart in the reference. */
int gdtk_hip_surrogate_project(int nsp, int n_elem, const double* elem, int n, const double* Y0, const double* d0,
                               double* d, int* ok, void* stream);

/* Message of the last failing call on this thread. */
const char* gdtk_hip_last_error(void);

/* ---- Two-temperature air reactor ------------------------------------------------------
 * TwoTemperatureAirKinetics (src/kinetics/two_temperature_air_kinetics.d:44-706) on the
 * 5-species neutral TwoTemperatureAir gas (src/gas/two_temperature_air.d): per substep the
 * chemistry step (RKF; Park rates at T^s Tv^(1-s)) and a Cash-Karp step of the
 * vibro-electronic energy u_v at fixed total energy (Landau-Teller exchange with
 * Millikan-White + Park relaxation times and the chemical source of u_v), with tight
 * temperature coupling as configured; per cell as FVCell.thermochemical_increment does
 * (fvcell.d:858-972: retry with dt_chem = -1, then update_thermo_from_rhou). */
typedef struct gdtk_tt_params {
    int molecular[GDTK_MAX_SPECIES];   /* 1: species with a vibrational mode (N2, O2, NO;
                                          two_temperature_air.d:25) */
    double A_MW[GDTK_MAX_SPECIES];     /* Millikan-White A of those species (N2 220, O2 129,
                                          NO 168: two_temperature_air_kinetics.d:714-720) */
    double abs_tol, rel_tol;           /* energy step tolerances (1e-9: :375-376) */
} gdtk_tt_params;

/* Replaces init_thermochemical_reactor's TwoTemperatureAirKinetics branch
 * (init_thermochemical_reactor.d:126-128).  The mechanism: RKF step, explicit backward
 * rates (no equilibrium-constant rates), lists in ascending species order (as prep-chem
 * writes them); 1 to 8 species, no free electrons. */
int gdtk_hip_tt_create(const gdtk_tpg_gas* gas, const gdtk_mechanism* mech, const gdtk_tt_params* prm, int* tt);
int gdtk_hip_tt_destroy(int tt);

/* n cells on device arrays (species-major massf, outputs may alias inputs), asynchronous on
 * stream: rho, T, Tv (T_modes[0]), u, uv (u_modes[0]), p, massf, dt_chem in; the same out;
 * status 0 or -1 (both reactor attempts failed, or the Tv iteration did not converge). */
int gdtk_hip_tt_step_device(int tt, int n, const double* rho, const double* T, const double* Tv, const double* u,
                            const double* uv, const double* p, const double* massf, const double* dt_chem, double dt,
                            double* T_out, double* Tv_out, double* u_out, double* uv_out, double* p_out,
                            double* massf_out, double* dt_chem_out, int* status, void* stream);

/* The last step (waits for it): {total substeps, max substeps, total step attempts, ms, cells}. */
int gdtk_hip_tt_stats(int tt, double* stats, int nstats);

#ifdef __cplusplus
}
#endif
#endif /* GDTK_HIP_H */
