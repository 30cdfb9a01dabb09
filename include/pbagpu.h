/*
 * pbagpu.h — C ABI of the MI355X-native per-blastwave radiation hot path.
 *
 * The path has three stages, each one entry point over a whole batch:
 *
 *   1. pbagpu_dyn_evolve   blastwave dynamics of every run (one ODE system per
 *                          run, shared adaptive DOP853 step), plus the per-step
 *                          derived quantities and injection-spectrum limits.
 *      Replaces the step loop PyBlastAfterglow::run (src/main.cpp:157-195) ->
 *      EvolveODEsystem::advanceTimeSubStep (src/model_evolve.h:439-566) with the
 *      RHS callback RHS_comb (src/model_evolve.h:1587-1689, bound through
 *      IntegratorBase::SetRHS, src/utilitites/ode_solvers.h:2061-2086), and the
 *      per-step addOtherVars / computeInjectionElectronDistributionLimits{FS,RS}
 *      (src/blastwave/blastwave.h:525-756, src/blastwave/blastwave_radiation.h:976-1107).
 *
 *   2. pbagpu_ele_evolve   electron kinetics (Chang-Cooper) and comoving
 *                          synchrotron + SSC emissivity of every population
 *                          (forward / reverse shock of every blastwave).
 *      Replaces the per-step call EvolveODEsystem::computeMicrophysics
 *      (src/model_evolve.h:934-953) -> evolveElectronDistirubitonComputeComoving
 *      RadiationSpectra{FS,RS} (src/blastwave/blastwave_radiation.h:1109-1182) ->
 *      ElectronAndRadiation::evaluateElectronDistributionNumericMixed and
 *      computeRadiationNumericMixed (src/microphysics/microphysics.h:359-599).
 *
 *   3. pbagpu_eats_lightcurve  equal-time-arrival-surface quadrature of every
 *                          layer at every (t_obs, nu_obs) pair.
 *      Replaces EjectaBase::evalEjectaLightCurves (src/ejecta/ejecta_base.h:1164-1192)
 *      -> EATS::evalLightCurve / evalFluxDensA (src/eats.h:808-856) with the
 *      integrand callback fluxFuncA = fluxDensAdaptiveWithComov
 *      (src/blastwave/blastwave_radiation.h:75-141, bound at :867-874).
 *
 * All compute entry points take DEVICE pointers and a HIP stream (NULL: the
 * null stream).  Structures are plain data; the library never frees caller
 * memory.  Every entry point returns 0 on success or a PBAGPU_E* code; the
 * per-item status arrays carry a PBAGPU_S* bitmask so the host can reproduce
 * the reference's log-and-exit behaviour.
 */
#ifndef PBAGPU_H
#define PBAGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- codes */
enum {
    PBAGPU_OK = 0,
    PBAGPU_E_ARG = -1,        /* bad argument / unsupported size */
    PBAGPU_E_HIP = -2,        /* HIP runtime error */
    PBAGPU_E_UNSUPPORTED = -3 /* option combination outside the implemented set */
};
/* per-item status bits (OR-ed) */
enum {
    PBAGPU_S_NONFINITE = 1,       /* NaN/inf met where the reference exits */
    PBAGPU_S_TERMINATED = 2,      /* blastwave termination (allow_termination=no -> exit in the reference) */
    PBAGPU_S_NMAX = 4,            /* ODE / quadrature budget exceeded */
    PBAGPU_S_BAD_MPHYS = 8,       /* B<=0, gamma_min>=gamma_max, ... (microphysics.h:371-376) */
    PBAGPU_S_NU_OUT_OF_GRID = 16, /* comoving frequency outside [freq1, freq2] (microphysics.h:1191-1202) */
    PBAGPU_S_HEAP_FULL = 32       /* adaptive-quadrature interval store exhausted */
};

/* ------------------------------------------------------- configuration */
/* Microphysics of one shock; parfile keys *_fs / *_rs
 * (ElectronAndRaidaionBase::setBasePars, src/microphysics/microphysics_base.h:59-560;
 *  ElectronAndRadiation::setPars / allocateForNumericSpectra, microphysics.h:72-345). */
typedef struct {
    double eps_e, eps_b, eps_t, p;
    double gam2;          /* top of the Chang-Cooper grid (gam1 is not used by that grid) */
    double freq1, freq2;  /* comoving frequency grid */
    double beta_min;      /* no microphysics below this shock velocity */
    double epsilon_e_rad; /* < 0: radiative losses computed inside the RHS */
    int ngam, nfreq;
    int max_substeps;
    int ssc;              /* method_ssc = numeric */
    int adi_loss_ele;     /* num_ele_use_adi_loss */
    int ssa;              /* use_ssa: synchrotron self-absorption (computeSSA, numeric_model.h:190-238) */
    int syn;              /* method_synchrotron: 0 = CSYN (SynKernel::cycSynch, kernels.h:414-437),
                             1 = Dermer09 (SynKernel::synchDermer, kernels.h:451-473),
                             2 = Bessel (SynKernel::synchBessel, kernels.h:396-412),
                             3 = GSL (SynKernel::synchGSL, kernels.h:211-342; F(x) from the
                                 integral representation of K_5/3, pbagpu_synchrotron_f) */
    int pad_;
} pba_shock_params;

/* One afterglow model (all its blastwaves share these); parfile main + GRB sections. */
typedef struct {
    double n_ism, d_l, z, theta_obs;
    double rtol;                  /* ODE rtol = atol */
    double tb0;                   /* ODE start time as given in the parfile (the grid's first
                                     point, 10^log10(tb0), can differ from it by an ulp) */
    double tprompt, rs_shutoff_rho, min_gamma0_for_rs, rs_gamma0_frac_no_exceed;
    double mom_when_start_spread, theta_max;
    double ism_A0, ism_s, ism_r_ej, ism_r_ism; /* density profile when n_ism <= 0 (parfile A0, s, r_ej,
                                     r_ism; RhoISM::evaluateRhoDrhoDrDefault,
                                     src/blastwave/blastwave_components.h:846-893): n = A0 r_ej^-s
                                     inside r_ej, A0 R^-s up to r_ism, A0 r_ism^-s beyond */
    double rtol_theta, rtol_phi;  /* EATS tolerances */
    int nmax_theta, nmax_phi;
    int fsrs;                     /* bw_type = fsrs */
    int do_rs, do_rs_radiation;
    int adiab_loss, adiab_loss_rs;
    int exponential_rho4, init_deltaR4;
    int counter_jet;
    int nmax;                     /* ODE step budget per output interval (0: unlimited) */
    int integrator;               /* 0: DOP853E, adaptive (IntegratorEmbedded<dop853_integrator>);
                                     1: RK4, one classical step per output interval
                                     (IntegratorStatic<rk4_integrator>, ode_solvers.h:319-366, 2099-2118) */
    pba_shock_params fs, rs;
} pba_run_params;

/* Initial data of one blastwave (one layer; EjectaID2 m_data,
 * src/ejecta/ejecta_id.h:13-341, as BlastWaveBase::setAllParametersForOneLayer
 * reads it, src/blastwave/blastwave_base.h:295-312). */
typedef struct {
    double E0, M0, mom0;                  /* piece-wise: already divided by the layer's 2i+1 cells */
    double theta_c_l, theta_c_h, theta_c; /* layer angles */
    double ctheta0;
    double theta_wing, theta_core;
    double theta_b0;  /* initial half-opening angle: theta_c_h (adaptive) or theta_wing (piece-wise) */
    double ncells;    /* 1 (adaptive) or the total cell count n^2 of the jet (piece-wise); >= 1 */
    int ilayer, nlayers;
    int ngam_fs, ngam_rs; /* this blastwave's Chang-Cooper grid size for each shock, or 0 for the
                             run's (pba_shock_params ngam).  The reference shares one grid per run
                             (CommonTables, src/microphysics/common.h:100-110); a run of the
                             piece-wise workload with per-layer electron grids sets them per layer. */
} pba_layer_init;

/* ------------------------------------------------ per-step table layout */
/* The dynamics stage writes, for every blastwave, a [PBA_NV][nt] table of
 * doubles (variable-major, so that one variable's time series is contiguous).
 * Names follow BW::VARNAMES (src/blastwave/blastwave_pars.h:211-250). */
enum pba_var {
    PV_tburst = 0, PV_R, PV_Rsh, PV_tt, PV_tcomov, PV_Gamma, PV_beta, PV_mom, PV_theta, PV_ctheta,
    PV_M2, PV_Eint2, PV_Erad2, PV_Esh2, PV_Ead2, PV_rho, PV_drhodr, PV_adi, PV_rho2, PV_GammaFsh,
    PV_thickness, PV_U_p, PV_gm, PV_gc, PV_gM, PV_B, PV_Theta, PV_z_cool, PV_nprime, PV_accel_frac,
    PV_Eint3, PV_Erad3, PV_Ead3, PV_Esh3, PV_M3, PV_deltaR4, PV_Gamma43, PV_adi3, PV_rho4, PV_rho3,
    PV_GammaRsh, PV_Rrsh, PV_thickness_rs, PV_U_p3, PV_gm_rs, PV_gc_rs, PV_gM_rs, PV_B3, PV_Theta_rs,
    PV_z_cool_rs, PV_nprime_rs, PV_accel_frac_rs, PV_GammaRel,
    PV_mflags,  /* bit0: forward-shock electrons computed at this step; bit1: reverse shock
                   (isBlastWaveValidForElectronCalc / considerReverseShock,
                   blastwave_radiation.h:881-974) */
    PBA_NV
};

/* Per-blastwave scalar results of the dynamics stage. */
typedef struct {
    int status;      /* PBAGPU_S_* */
    int i_end_r;     /* last usable step + 1 for the observer integral */
    int fsrs;        /* 1 if evolved with the reverse shock */
    int n_steps;     /* accepted ODE steps (all intervals) */
    int n_rhs;       /* RHS evaluations */
    int pad_;
} pba_bw_result;

/* ----------------------------------------------------------- lifecycle */
int pbagpu_device_count(int *count);
int pbagpu_set_device(int device);
int pbagpu_version(void);
/* Text of the last failure in this thread (failing call and HIP error), "" if none.
 * The reference logs the cause before exit(1) (e.g. numeric_model.h:688-691); this
 * string lets the host reproduce that message. */
const char *pbagpu_last_error(void);

/* Streams and workspaces.  Every compute entry point runs on the stream it is given
 * and keeps its device workspace (interval stores, staged parameters, kernel tables)
 * per (device, stream): the stream is the context.  Calls on different streams may run
 * at the same time from different host threads (one thread per stream); calls on one
 * stream must come from one thread at a time, as the reference drives one model from
 * one thread (src/main.cpp:157-195).  No counterpart in the reference, which owns no
 * device memory: this frees the workspace of `stream` on the current device, after
 * waiting for the stream; the next call on it allocates again. */
int pbagpu_release_stream(void *stream);
/* A stream whose kernels run only on compute units [cu_first, cu_first + cu_count) of the
 * current device (cu_count <= 0: a plain non-blocking stream on every unit).  No
 * counterpart in the reference.  The dynamics of a run is a long chain of dependent ODE
 * steps on one wavefront, while the persistent observer kernel holds every unit it is
 * given until it ends; giving the dynamics a few units of their own lets batches of
 * runs be pipelined.  The persistent observer grid is sized to its stream's units. */
int pbagpu_stream_create(int cu_first, int cu_count, void **stream);
/* pbagpu_release_stream, then destroy a stream made by pbagpu_stream_create. */
int pbagpu_stream_destroy(void *stream);

/* ---------------------------------------------------- kernel timing */
/* No counterpart in the reference: instrumentation for benchmarks.  When
 * enabled, every stage records HIP events on the caller's stream immediately
 * before and after its main kernel and accumulates the elapsed device time
 * (guarded by a mutex: stages may be driven from several host threads, one
 * stream each). */
enum {
    PBAGPU_STAGE_DYN = 0,      /* dyn_kernel */
    PBAGPU_STAGE_ELE = 1,      /* ele_kernel */
    PBAGPU_STAGE_EATS = 2,     /* eats_kernel */
    PBAGPU_STAGE_SSC_TABLE = 3,/* Compton kernel table build */
    PBAGPU_STAGE_EATS_PW = 4,  /* eats_pw_kernel (piece-wise observer stage) */
    PBAGPU_STAGE_SKY_HIST = 5, /* sky_bin_kernel (sky-map histogram) */
    PBAGPU_STAGE_EATS_VEC = 6, /* eats_vec_kernel (frequency-vectorised observer stage) */
    PBAGPU_N_STAGES = 7
};
int pbagpu_timing_enable(int on);
int pbagpu_timing_read(int stage, double *ms_total, long long *launches, int reset);
/* Work units counted while timing is enabled: integrand samples of the observer
 * stage (PBAGPU_STAGE_EATS), cell fluxes of the piece-wise observer stage
 * (PBAGPU_STAGE_EATS_PW); other stages report 0 (their units follow from the
 * per-blastwave results and the step flags on the host). */
int pbagpu_timing_units(int stage, double *units, int reset);

/* Library options (process-wide; 0 by default).  No counterpart in the reference.
 *   PBAGPU_OPT_ELE_GRAPH  1: the electron stage captures its per-step launches (three or more
 *                         per output step) into one HIP graph and launches it once per call,
 *                         instead of launching them one by one (FP64 and BF16 contractions;
 *                         the surrogate mode always launches directly), on a created stream
 *                         (calls on the null stream, which cannot be captured, stay direct).
 *                         Results are identical.
 *   PBAGPU_OPT_DYN_RUNS_PER_WAVE  runs the dynamics stage packs into one 64-lane wavefront: 0 (the
 *                         default) only as many as the runs outnumber the SIMDs of the stream's
 *                         compute units (at most floor(64 / bw_per_run)); n >= 1 at most n.  A
 *                         run's results do not depend on the packing.
 * Returns PBAGPU_E_ARG for an unknown option or a value out of its range (graph 0 or 1, runs per
 * wavefront 0 .. 64). */
enum {
    PBAGPU_OPT_ELE_GRAPH = 0,
    PBAGPU_OPT_DYN_RUNS_PER_WAVE = 1,
    PBAGPU_N_OPTIONS = 2
};
int pbagpu_set_option(int which, long long value);
int pbagpu_get_option(int which, long long *value);

/* ------------------------------------------------------------- stage 1 */
/* Evolve n_runs independent models of bw_per_run blastwaves each.
 *   params   [n_runs]            host array
 *   init     [n_runs*bw_per_run] host array
 *   d_tgrid  [nt] device          burster-frame output grid (MakeLogspaceVec(tb0,tb1,ntb))
 *   d_table  [n_runs*bw_per_run][PBA_NV][nt] device, written
 *   res      [n_runs*bw_per_run] host array, written
 */
int pbagpu_dyn_evolve(const pba_run_params *params, int n_runs, int bw_per_run,
                      const pba_layer_init *init, int nt, const double *d_tgrid,
                      double *d_table, pba_bw_result *res, void *stream);

/* The same for runs whose time grids differ in length (and values), in one launch: run r has
 * its own parfile grid of nt_run[r] steps (MakeLogspaceVec(tb0, tb1, ntb) of that run).
 *   nt_run   [n_runs] host, 2 <= nt_run[r] <= nt_max
 *   d_tgrid  [n_runs][nt_max] device: row r holds run r's grid in its first nt_run[r] entries
 *   d_table  [n_runs*bw_per_run][PBA_NV][nt_max] device, written: run r's steps in columns
 *            [0, nt_run[r]), zeros after them (no step flags there, so the electron stage
 *            skips them; every blastwave's i_end_r <= nt_run[r] bounds the observer stage)
 * The electron and observer stages then take nt = nt_max. */
int pbagpu_dyn_evolve_ragged(const pba_run_params *params, int n_runs, int bw_per_run,
                             const pba_layer_init *init, int nt_max, const int *nt_run,
                             const double *d_tgrid, double *d_table, pba_bw_result *res, void *stream);

/* Kilonova ejecta: every blastwave of every run evolved as the reference evolves ejecta,
 * setStandardISM + BlastWave::rhs_fs_dense (src/model_evolve.h:1627-1651,
 * src/blastwave/blastwave.h:769-781, 2120-2497), with method_dgdr = our and no energy injection,
 * nuclear heating or shell collisions; replaces the same RHS callback and step loop as
 * pbagpu_dyn_evolve.  The blastwaves of a run are in the reference's state-vector order,
 * layer-major (for each layer, its velocity shells; src/model_evolve.h:225-234); init carries
 * the per-blastwave initial data of the 2-D [shells][layers] file (EjectaID2, use_1d_id = no).
 * spread: method_spread set (the spreading form of dt_obs/dR and of the initial observer time;
 * dtheta/dR stays 0).  Besides the jet's columns the table holds GammaRel, and rho / drhodr are
 * those of the last RHS evaluation, as the reference stores them.  Runs of at most 64
 * blastwaves (one wavefront per run; PBAGPU_E_UNSUPPORTED beyond); params[r].fsrs must be 0.
 * The electron and observer stages take the resulting tables as a jet's. */
int pbagpu_dyn_evolve_ejecta(const pba_run_params *params, int n_runs, int bw_per_run,
                             const pba_layer_init *init, int spread, int nt, const double *d_tgrid,
                             double *d_table, pba_bw_result *res, void *stream);

/* ------------------------------------------------------------- stage 2 */
/* One electron population: shock (0 forward, 1 reverse) of blastwave `bw`
 * (index into the dynamics table), with the microphysics of run `run`. */
typedef struct {
    int bw;
    int run;
    int shock;
    int ngam;   /* Chang-Cooper grid size of this population, or 0 for its shock's ngam (the
                   blastwave's pba_layer_init ngam_fs / ngam_rs) */
} pba_population;

/* Electron kinetics + emissivity of n_pop populations over all nt steps.
 *   d_table  dynamics table from stage 1
 *   d_jall   [n_pop][nt][nfreq] device, written: comoving emissivity per electron
 *            (total_rad.j_all, base_state.h:257-310)
 *   d_aall   [n_pop][nt][nfreq] device, written: comoving absorption coefficient
 *            (total_rad.a_all; 0 where the shock has ssa = 0); may be NULL when no
 *            population's shock has ssa = 1
 *   d_fall   [n_pop][nt][ngam]  device or NULL: electron spectrum per step (ele.f_all)
 *   d_status [n_pop] device int, written
 * nfreq comes from the shock parameters and must be equal within one call; ngam
 * (the population's, or its shock's) may differ between populations: every population
 * keeps its own grid, Compton table and state, and d_fall rows are laid out with the
 * largest ngam of the call as their stride ([n_pop][nt][ngam_max], a population's
 * spectrum in the first ngam entries of a row).  precision: 0 = FP64 everywhere; 1 = SSC
 * contraction in BF16 hi/lo split on MFMA with FP32 accumulation. */
int pbagpu_ele_evolve(const pba_run_params *params, int n_runs,
                      const pba_population *pops, int n_pop,
                      const double *d_table, int bw_total, int nt,
                      double *d_jall, double *d_aall, double *d_fall, int *d_status,
                      int precision, void *stream);

/* The comoving spectra that the reference keeps per step with save_spec = yes
 * (ElectronAndRadiation::storeDenseOutput, src/microphysics/microphysics.h:602-633, and
 * Electrons::save_to_all, src/microphysics/base_state.h:200-211) and writes into spec.h5
 * (Ejecta::computeSaveEjectaSpectrum, src/ejecta/ejecta.h:354-509).  Every pointer is a
 * device array or NULL (not written); a population's rows stay 0 at steps where it does not
 * radiate, as in the reference.
 *   syn_j, syn_i   [n_pop][nt][nfreq]  syn.j_all = syn.j accel_frac, syn.i_all (its intensity
 *                                     with the absorption depth a_all dr', computeIntensity's
 *                                     iAPPROX rule)
 *   ssc_j          [n_pop][nt][nfreq]  ssc.j_all = ssc.j accel_frac (ssc.i_all equals it: the
 *                                     SSC absorption is 0 without pair production)
 *   total_f        [n_pop][nt][nfreq]  total_rad.f_all, the photon density of the step
 *   total_i        [n_pop][nt][nfreq]  total_rad.i_all
 *   gam_dot_syn, gam_dot_adi, gam_dot_ssc  [n_pop][nt][ngam_max]  the cooling rates of the last
 *                                     Chang-Cooper step (0 in the last bin and before the first)
 * syn.f_all and ssc.f_all are never set by the reference (0 in its files); syn.a_all and
 * total_rad.a_all equal d_aall. */
typedef struct {
    double *syn_j, *syn_i;
    double *ssc_j;
    double *total_f, *total_i;
    double *gam_dot_syn, *gam_dot_adi, *gam_dot_ssc;
} pba_spectra;

/* pbagpu_ele_evolve plus the spectra above (the reference's save_spec = yes path). */
int pbagpu_ele_evolve_spectra(const pba_run_params *params, int n_runs,
                              const pba_population *pops, int n_pop,
                              const double *d_table, int bw_total, int nt,
                              double *d_jall, double *d_aall, double *d_fall, int *d_status,
                              int precision, const pba_spectra *spec, void *stream);

/* ------------------------------------------------------------- stage 3 */
/* Observer light curve of every layer at n_pairs (t_obs, nu_obs) pairs.
 *   layers     [n_layers] host: blastwave index in the dynamics table and the
 *              population indices of its forward / reverse shock emissivity (-1: none)
 *   d_out      [n_layers][n_pairs] device, written, mJy
 *   d_aall     absorption table from stage 2, or NULL (no absorption: the optical
 *              depth is 0, as with use_ssa = no)
 */
typedef struct {
    double theta_c_l, theta_c_h; /* angular extent of the layer (EjectaID2 adaptive grid) */
    int bw;
    int run;
    int pop_fs;
    int pop_rs;      /* -1 if the layer has no reverse-shock emission */
    int i_end_r;     /* from pba_bw_result */
} pba_eats_layer;

int pbagpu_eats_lightcurve(const pba_run_params *params, int n_runs,
                           const pba_eats_layer *layers, int n_layers,
                           const double *d_table, int nt, const double *d_jall, const double *d_aall, int nfreq,
                           int n_pairs, const double *d_tobs, const double *d_nuobs,
                           double *d_out, int *d_status, void *stream);

/* The same light curves on a tensor grid of n_t observer times x n_nu frequencies
 * (pairs t-major, as cast_times_freqs, src/utilitites/utils.h:1419-1448), by a quadrature
 * shared across frequencies: one wavefront per (layer, t_obs) integrates every nu_obs with
 * one adaptive Gauss-Kronrod 7-15 tree in phi and in 1 - cos(theta), refined until
 * |K15 - G7| <= rtol |I| on every interval share for every frequency.  The integrand, its
 * domain and its arithmetic (FP64) are those of pbagpu_eats_lightcurve; the CADRE trees of
 * the reference (one per pair, rtol_theta / rtol_phi) are replaced, so the result agrees
 * with it to the quadrature tolerance, not bit for bit.  The surface geometry of a sample
 * (mu, the arrival-time search, the state interpolation) is computed once for all
 * frequencies instead of once per pair.  No counterpart in the reference; it replaces the
 * same call site as pbagpu_eats_lightcurve (EjectaBase::evalEjectaLightCurves,
 * src/ejecta/ejecta_base.h:1164-1192).  n_nu <= 128; synchrotron self-absorption is not
 * implemented here (PBAGPU_E_UNSUPPORTED).
 *   d_tobs [n_t], d_nuobs [n_nu] device; d_out [n_layers][n_t * n_nu] device, written, mJy */
int pbagpu_eats_lightcurve_vec(const pba_run_params *params, int n_runs,
                               const pba_eats_layer *layers, int n_layers,
                               const double *d_table, int nt, const double *d_jall, int nfreq,
                               int n_t, const double *d_tobs, int n_nu, const double *d_nuobs,
                               double rtol, double *d_out, int *d_status, void *stream);

/* Piece-wise EATS: the layer's light curve as the sum over its 2i+1 phi cells
 * at the layer centre angle, primary jet and (counter_jet) counter-jet.
 * Replaces EATS::evalLightCurve for method_eats = piece-wise ->
 * EATS::evalSkyMapPW (src/eats.h:808-816, 859-883, 1064-1213) with
 * fluxDensPieceWiseWithComov (src/blastwave/blastwave_radiation.h:14-72).
 * As there, the reverse-shock term (rs = 1) takes the forward-shock
 * emissivity (pop_fs) with the reverse-shock downstream state.
 *   layers  [n_layers] host
 *   d_out   [n_layers][n_pairs] device, written, mJy
 *   d_status[n_layers] device, written: PBAGPU_S_* bits (a cell before the
 *           first arrival time, a comoving frequency off the grid, a
 *           non-finite or negative cell flux: where the reference exits) */
typedef struct {
    double ctheta0;  /* cell polar angle (EjectaID2 ictheta) */
    double ncells;   /* total cell count n^2 (BlastWave Pars ncells) */
    int cil;         /* phi cells of this layer, 2*ilayer+1 (EjectaID2::CellsInLayer) */
    int bw;
    int run;
    int pop_fs;
    int rs;          /* 1: add the reverse-shock term (fsrs blastwave with RS radiation) */
    int i_end_r;
} pba_eats_pw_layer;

int pbagpu_eats_lightcurve_pw(const pba_run_params *params, int n_runs,
                              const pba_eats_pw_layer *layers, int n_layers,
                              const double *d_table, int nt, const double *d_jall, const double *d_aall,
                              int nfreq,
                              int n_pairs, const double *d_tobs, const double *d_nuobs,
                              double *d_out, int *d_status, void *stream);

/* Raw sky-map image arrays, in the order of IMG::m_names (src/image.h:14-26). */
enum pba_img_var {
    PBA_IMG_r = 0, PBA_IMG_ctheta, PBA_IMG_cphi, PBA_IMG_mu, PBA_IMG_xrs, PBA_IMG_yrs, PBA_IMG_intensity,
    PBA_IMG_NV
};

/* Piece-wise sky maps of one run: for every (t_obs, nu_obs) image, every cell's
 * raw entry as EATS::evalSkyMapPW writes it into its `out` arrays
 * (src/eats.h:1163-1184), before Ejecta::computeEjectaSkyMapPW's normalisation
 * (src/ejecta/ejecta_base.h:471-545).  Cells are stacked layer after layer in
 * the order of `layers` (all of run 0, params[0]); ncells = sum of cil.
 *   d_image  [n_images][PBA_IMG_NV][2 * ncells] device, written: primary-jet cells,
 *            then counter-jet cells; cells that do not contribute are 0
 *   d_flux   [n_layers][n_images] device, written: each layer's flux density, mJy
 *            (the value evalSkyMapPW returns)
 *   remove_mu  skymap_remove_mu: 0 divides the intensity by |mu|, as the default */
int pbagpu_eats_skymap_pw(const pba_run_params *params, const pba_eats_pw_layer *layers, int n_layers,
                          const double *d_table, int nt, const double *d_jall, const double *d_aall, int nfreq,
                          int n_images, const double *d_tobs, const double *d_nuobs, int remove_mu,
                          double *d_image, double *d_flux, int *d_status, void *stream);

/* Host evaluation (no GPU) of the routine the electron stage uses on the device for the
 * Bessel synchrotron kernel: K_{1/3}(x[i]) and K_{4/3}(x[i]) by Temme's method.  No
 * counterpart in the reference (which calls std::cyl_bessel_k, kernels.h:402-403): it
 * exists so that the device routine's arithmetic can be checked against libm on the host. */
int pbagpu_bessel_k_thirds(int n, const double *x, double *k13, double *k43);
/* Host evaluation (no GPU) of the device routine behind method_synchrotron = GSL: the
 * synchrotron function F(x[i]) = x int_x^inf K_5/3(t) dt.  The reference takes F from GSL's
 * Chebyshev fits (SynKernel::synchGSL, kernels.h:211-342); this library integrates the
 * integral representation of K_5/3 instead.  No counterpart in the reference: it exists so
 * that the routine can be checked against the defining integral on the host. */
int pbagpu_synchrotron_f(int n, const double *x, double *f);

/* ------------------------------------------- learned SSC emissivity */
/* No counterpart in the reference: a multilayer perceptron standing in for the SSC
 * contraction computeSSCSpectrum (src/microphysics/numeric_model.h:247-285), trained on
 * pairs that the plain contraction itself produces, with the plain contraction as the
 * fallback and the check on a sampled subset of cells every step.
 *
 * pba_mlp: dims[0] -> dims[1] -> ... -> dims[n_layers]; SiLU after every layer but the
 * last; weights w[l] [dims[l]][dims[l+1]] and biases b[l] [dims[l+1]] in FP32 on the
 * device (the master copy); products in BF16 on MFMA with FP32 accumulation.  The
 * gradient (gw, gb) and Adam moment (mw, vw, mb, vb) buffers, of the same shapes, are
 * needed by pbagpu_mlp_train_step only. */
#define PBA_MLP_MAXL 8
typedef struct {
    int n_layers;
    int dims[PBA_MLP_MAXL + 1];
    float *w[PBA_MLP_MAXL], *b[PBA_MLP_MAXL];
    float *gw[PBA_MLP_MAXL], *gb[PBA_MLP_MAXL];
    float *mw[PBA_MLP_MAXL], *vw[PBA_MLP_MAXL], *mb[PBA_MLP_MAXL], *vb[PBA_MLP_MAXL];
} pba_mlp;

/* Inference: d_y [batch][dims[n_layers]] = net(d_x [batch][dims[0]]), FP32 device arrays. */
int pbagpu_mlp_forward(const pba_mlp *net, int batch, const float *d_x, float *d_y, void *stream);
/* One training step on a minibatch: forward, mean square error against d_t
 * [batch][dims[n_layers]] (its value to *d_loss, device, if not NULL), backward (gradients
 * to gw, gb) and an Adam update (beta1 0.9, beta2 0.999, eps 1e-8; `step` >= 1 counts the
 * updates for the bias correction; lr = 0 leaves the weights and moments' effect out). */
int pbagpu_mlp_train_step(const pba_mlp *net, int batch, const float *d_x, const float *d_t, float lr,
                          int step, float *d_loss, void *stream);
/* The product the network is built from, for tests: d_c [M][N] = A B with BF16-rounded
 * operands and FP32 accumulation; d_a holds A [M][K] (trans_a = 0) or A^T [K][M]; d_b holds
 * B [K][N] (trans_b = 0) or B^T [N][K]; FP32 device arrays. */
int pbagpu_gemm_bf16(int M, int N, int K, const float *d_a, int trans_a, const float *d_b, int trans_b,
                     float *d_c, void *stream);

/* Surrogate of the SSC emissivity of one shock's grids (ngam electron bins, nfreq
 * frequencies).  Its input, for a cell with electron spectrum f [ngam] and Compton photon
 * weights w [nfreq] (w_k = n_k / nu_k (nu_{k+1} - nu_k), the weights the plain contraction
 * takes), is x = [ (log10 max(f_j / max f, 1e-30) + 30) / 30 ] (ngam values),
 * [ (log10 max(w_k / max w, 1e-30) + 30) / 30 ] (nfreq), log10(B) / 10 (1 value): dims[0] =
 * ngam + nfreq + 1.  Its output y [nfreq] gives j_ssc[i] = max f max w 10^(y_offset[i] +
 * y_scale y[i]) (the contraction is bilinear in f and w, so the two maxima scale it exactly).
 * Every step, population q is also computed by the plain contraction when
 * (131 q + 37 it) % check_every == 0; where the surrogate is off by more than check_tol
 * (relative, over frequencies above 1e-6 of the spectrum's peak), that population takes the
 * plain contraction for the rest of the call. */
typedef struct {
    pba_mlp net;
    const float *y_offset;   /* [nfreq] device */
    float y_scale;
    float check_tol;
    int check_every;         /* 100: one population in a hundred per step */
    int pad_;
} pba_ssc_surrogate;
typedef struct {
    long long inferred;      /* population-steps whose SSC emissivity the network gave */
    long long checked;       /* population-steps also computed by the plain contraction */
    long long failed;        /* checks above check_tol */
    long long fallback;      /* population-steps computed by the plain contraction after a failed check */
    double max_err;          /* largest relative error seen by a check */
} pba_surrogate_stats;

/* Training pairs: the surrogate's input features of n cells (f [n][ngam], w [n][nfreq],
 * B [n], device) into d_x [n][ngam + nfreq + 1] and the scale max f max w into d_scale [n]. */
int pbagpu_ssc_features(const pba_shock_params *shock, int n, const double *d_f, const double *d_w,
                        const double *d_B, float *d_x, double *d_scale, void *stream);
/* The plain SSC emissivity of n cells (the contraction of computeSSCSpectrum on the shock's
 * grids and Compton table, FP64): d_j [n][nfreq]. */
int pbagpu_ssc_emissivity(const pba_shock_params *shock, int n, const double *d_f, const double *d_w,
                          double *d_j, void *stream);
/* pbagpu_ele_evolve with the SSC emissivity from the surrogate (and the sampled check);
 * d_aall must be NULL (no absorption); stats (host, may be NULL) receives the counts. */
int pbagpu_ele_evolve_surrogate(const pba_run_params *params, int n_runs,
                                const pba_population *pops, int n_pop,
                                const double *d_table, int bw_total, int nt,
                                double *d_jall, double *d_fall, int *d_status,
                                const pba_ssc_surrogate *sur, pba_surrogate_stats *stats, void *stream);

/* ------------------------------------------------- sky-map binning */
/* Histogram image, flux centroid and data extent of raw sky maps: the binning part
 * of ProcessRawSkymap._process_file (package/src/PyBlastAfterglowMag/skymap_process.py:465-485),
 * which calls
 *   _compute_position_of_the_flux_centroid (:98-105): np.average of x and y weighted by I,
 *   _get_data_extend (:107-189): min / max of x and y over the non-empty entries,
 *   np.linspace(min * extend_grid, max * extend_grid, num=nx) for the edges (:480-481),
 *   _total_hist_skymaps (:371-383): np.histogramdd of (x, y) weighted by I.
 * The scattered-data interpolation (scipy griddata), the smoothing filters and the
 * FWHM search of that routine are not part of this entry.
 *
 * An image is a list of entries (ProcessRawSkymap._get_skymap's per-shell principal-
 * and counter-jet arrays, :191-328); entry k holds points [seg_off[k], seg_off[k+1])
 * of d_x / d_y / d_w, and image i holds entries [img_seg[i], img_seg[i+1]).
 *   img_seg   [n_images+1] host, img_seg[0] = 0, non-decreasing
 *   seg_off   [img_seg[n_images]+1] host, seg_off[0] = 0, non-decreasing
 *   d_x, d_y, d_w [seg_off[last]] device: sky positions (mas) and intensities
 *   edges_given  0: the edges are computed from the data extent and written to
 *                d_edges_x / d_edges_y; 1: they are read from there (must increase)
 *   d_edges_x [n_images][nx], d_edges_y [n_images][ny] device
 *   d_hist    [n_images][nx-1][ny-1] device, written (x bin major, as np.histogramdd)
 *   d_count   [n_images][nx-1][ny-1] device or NULL: points per bin
 *   d_moments [n_images][PBA_SKY_NM] device, written
 *   d_status  [n_images] device, written: PBA_SKY_* (an image with a non-zero status
 *             keeps an all-zero histogram; the reference raises an exception there) */
enum pba_sky_moment {
    PBA_SKY_XC = 0, PBA_SKY_YC, PBA_SKY_XMIN, PBA_SKY_XMAX, PBA_SKY_YMIN, PBA_SKY_YMAX,
    PBA_SKY_SUMW, PBA_SKY_NPTS,
    PBA_SKY_NM
};
enum pba_sky_status {
    PBA_SKY_OK = 0,
    PBA_SKY_NO_DATA = 1,      /* edges from the extent, but no entry has two points and a non-zero
                                 intensity sum */
    PBA_SKY_ZERO_WEIGHT = 2,  /* intensities sum to zero: the centroid is undefined */
    PBA_SKY_FLAT = 3,         /* edges from the extent, and an entry has xmin == xmax or
                                 ymin == ymax (:157-160) */
    PBA_SKY_NONFINITE = 4,    /* a non-finite position or intensity */
    PBA_SKY_BAD_EDGES = 5     /* given edges that decrease somewhere */
};
int pbagpu_skymap_hist(int n_images, const int *img_seg, const long *seg_off,
                       const double *d_x, const double *d_y, const double *d_w,
                       int nx, int ny, double extend_grid, int edges_given,
                       double *d_edges_x, double *d_edges_y, double *d_hist, unsigned *d_count,
                       double *d_moments, int *d_status, void *stream);

/* Allocation-free variant of pbagpu_skymap_hist: same arguments, same kernels, same results.
 * It has no counterpart in the reference; it is for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc / hipFree: its work tables (the device copies of img_seg and seg_off,
 * the work-item table, built on the device, and the per-item partial sums) live in a 4 MiB
 * array in device memory that every GPU the library runs on holds until the process ends.
 * A call whose tables need more than 4 MiB is rejected with PBAGPU_E_ARG before any device
 * work (4096 work items of <= 4096 points fit in well under 1 MiB; the item table takes 80
 * bytes per item).  Every call on one GPU shares that array: calls that overlap in time on
 * one GPU (two host threads, or two streams) race and must not be issued; consecutive calls
 * are fine, each one rewrites what it reads. */
int pbagpu_skymap_hist_noalloc(int n_images, const int *img_seg, const long *seg_off,
                               const double *d_x, const double *d_y, const double *d_w,
                               int nx, int ny, double extend_grid, int edges_given,
                               double *d_edges_x, double *d_edges_y, double *d_hist, unsigned *d_count,
                               double *d_moments, int *d_status, void *stream);

#ifdef __cplusplus
}
#endif
#endif
