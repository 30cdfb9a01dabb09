"""Named afterglow workloads (parfile sections + jet structure) used by tests and the benchmark.

* ``tophat_fs_syn``      : single top-hat blastwave, forward shock, synchrotron only,
                           200 gamma bins, 96 comoving freqs, 1000 steps, 96 nu x 128 t_obs.
* ``gauss21_fsrs_ssc``   : 21-layer Gaussian jet, forward + reverse shock, synchrotron + SSC,
                           theta_obs = 0.363 rad, same grids.
* ``tophat64_pw``        : the top-hat jet on a 64-layer piece-wise grid (4096 cells).
* ``kn_c5``              : kilonova-like ejecta, 32 velocity shells x 30 polar layers (960
                           blastwaves), synchrotron + SSC, 128 gamma bins, 128 comoving freqs.
Small variants (fewer steps / bins / pairs) are derived with ``scaled`` (``kn_scaled`` for ejecta).
"""
import copy

import numpy as np


def tophat_fs_syn():
    main = dict(tb0=3e2, tb1=1e14, ntb=1000, n_ism=1e-2, d_l=1.27e26, z=0.0099, theta_obs=0.0,
                lc_freqs="array logspace 1e8 1e29 96", lc_times="array logspace 3e3 1e10 128")
    grb = dict(eps_e_fs=0.1, eps_b_fs=1e-3, p_fs=2.2, ngam_fs=200, gam1_fs=1., gam2_fs=1e8,
               freq1_fs=1e5, freq2_fs=1e32, nfreq_fs=96, method_ssc_fs="none", method_ele_fs="numeric",
               bw_type="fs", do_rs="no", method_eats="adaptive")
    struct = dict(struct="tophat", eiso_c=1e52, gamma0c=300., theta_c=0.1)
    return main, grb, struct


def gauss21_fsrs_ssc():
    main = dict(tb0=3e2, tb1=1e14, ntb=1000, n_ism=10 ** (-1.6), d_l=1.27e26, z=0.0099, theta_obs=0.363,
                lc_freqs="array logspace 1e8 1e29 96", lc_times="array logspace 3e3 1e10 128")
    grb = dict(eps_e_fs=10 ** (-3.42), eps_b_fs=10 ** (-4.02), p_fs=2.1,
               eps_e_rs=10 ** (-3.42), eps_b_rs=10 ** (-4.02), p_rs=2.1,
               ngam_fs=200, gam1_fs=1., gam2_fs=1e8, freq1_fs=1e5, freq2_fs=1e32, nfreq_fs=96,
               ngam_rs=200, gam1_rs=1., gam2_rs=1e8, freq1_rs=1e5, freq2_rs=1e32, nfreq_rs=96,
               method_ssc_fs="numeric", method_ssc_rs="numeric", method_ele_fs="numeric", method_ele_rs="numeric",
               bw_type="fsrs", do_rs="yes", do_rs_radiation="yes", method_eats="adaptive")
    struct = dict(struct="gaussian", eiso_c=10 ** 54.1, gamma0c=300., theta_c=0.061, theta_w=0.436,
                  n_layers_a=21)
    return main, grb, struct


def tophat64_pw():
    """The top-hat jet of tophat_fs_syn on a 64-layer piece-wise grid (2i+1 phi cells per layer,
    4096 cells), same physics and grids: the piece-wise workload with the cell count of C3.

    The 21-layer Gaussian physics on a 64-layer piece-wise grid is not a workload the reference
    completes: its cell fluxes include negative intensities (bilinear extrapolation of the
    emissivity between steps) and the reference exits (microphysics.h:835-841)."""
    main, grb, struct = tophat_fs_syn()
    grb = dict(grb, method_eats="piece-wise")
    struct = dict(struct, n_layers_pw=64)
    return main, grb, struct


# Parameter ranges of the batched-member workload: a Latin hypercube over
# log10 E_iso,c, Gamma0,c, theta_c, log10 n_ism, log10 eps_e, log10 eps_b, p, theta_obs.
LHS_RANGES = (("log10_eiso", 50.0, 54.0), ("gamma0c", 10.0, 1000.0), ("theta_c", 0.02, 0.2),
              ("log10_n_ism", -4.0, 1.0), ("log10_eps_e", -3.0, -0.5), ("log10_eps_b", -5.0, -1.0),
              ("p", 2.05, 2.8), ("theta_obs", 0.0, 0.8))
LHS_SEED = 20261203


def latin_hypercube(n, d, seed=LHS_SEED):
    """n points in [0, 1)^d with exactly one point in each of the n equal strata of every
    dimension (a random permutation of the strata per dimension, a uniform offset inside each)."""
    rng = np.random.default_rng(seed)
    strata = np.stack([rng.permutation(n) for _ in range(d)], axis=1)
    return (strata + rng.random((n, d))) / n


LHS_NTB = (500, 1500)


def lhs_ntb(i, seed=LHS_SEED):
    """Time-grid length of member i: uniform on LHS_NTB (inclusive), from a generator seeded by
    (seed, i), so a member's grid does not depend on the batch it is in."""
    return int(np.random.default_rng([seed, i]).integers(LHS_NTB[0], LHS_NTB[1] + 1))


def lhs_member(i, n=1024, seed=LHS_SEED, theta_w_over_theta_c=4.0, ragged_nt=True):
    """Member i of the n-member Latin-hypercube batch: the 21-layer Gaussian jet of
    gauss21_fsrs_ssc (forward + reverse shock, synchrotron + SSC, same grids) with the eight
    parameters of LHS_RANGES drawn from the hypercube.  The jet's wing is theta_w =
    theta_w_over_theta_c x theta_c (capped at the 0.436 rad of gauss21_fsrs_ssc), so that every
    member keeps its 21 layers (the initial data drop layers with beta0 <= 1e-5).  With ragged_nt
    the member's burster-frame grid has lhs_ntb(i) steps over the same span (else 1000)."""
    u = latin_hypercube(n, len(LHS_RANGES), seed)[i]
    v = {name: float(lo + (hi - lo) * x) for (name, lo, hi), x in zip(LHS_RANGES, u)}
    main, grb, struct = gauss21_fsrs_ssc()
    main = dict(main, n_ism=10 ** v["log10_n_ism"], theta_obs=v["theta_obs"],
                ntb=lhs_ntb(i, seed) if ragged_nt else main["ntb"])
    micro = {}
    for sh in ("fs", "rs"):
        micro.update({f"eps_e_{sh}": 10 ** v["log10_eps_e"], f"eps_b_{sh}": 10 ** v["log10_eps_b"], f"p_{sh}": v["p"]})
    grb = dict(grb, **micro)
    struct = dict(struct, eiso_c=10 ** v["log10_eiso"], gamma0c=v["gamma0c"], theta_c=v["theta_c"],
                  theta_w=min(0.436, theta_w_over_theta_c * v["theta_c"]))
    return main, grb, struct


def lhs_small(i, ragged_nt=True):
    """Member i of the Latin-hypercube workload with the microphysics grids reduced (64 Lorentz-factor
    bins, 48 comoving frequencies) and 6 x 8 observer pairs: the members the batch tests run."""
    main, grb, struct = lhs_member(i, ragged_nt=ragged_nt)
    main = dict(main, lc_freqs="array logspace 1e8 1e29 6", lc_times="array logspace 3e3 1e10 8")
    grb = dict(grb)
    for sh in ("fs", "rs"):
        grb[f"ngam_{sh}"] = 64
        grb[f"nfreq_{sh}"] = 48
    return main, grb, struct


# Per-layer electron grids of the piece-wise Gaussian workload: a seeded uniform draw per layer
C3_NGAM = (128, 256, 384, 512)


def layer_ngam(n_layers, choices=C3_NGAM, seed=LHS_SEED):
    """Chang-Cooper grid size of every layer, uniform over `choices` (a generator seeded by
    (seed, n_layers))."""
    rng = np.random.default_rng([seed, n_layers])
    return tuple(int(x) for x in rng.choice(np.asarray(choices), n_layers))


def gauss64_pw_ng():
    """The 21-layer Gaussian jet's physics (forward + reverse shock, synchrotron + SSC, 96
    comoving frequencies, 1000 steps, 96 nu x 128 t_obs) on a 64-layer piece-wise grid (2i+1 phi
    cells per layer, 4096 cells) with a Chang-Cooper grid of its own for every layer, drawn from
    128, 256, 384, 512 bins (layer_ngam).  The wing is cut at 0.2 rad, as in gauss_pw_small: the
    reference stops on the slower wing layers of the full 0.436 rad jet (negative cell intensities,
    microphysics.h:835-841).  The reference shares one grid per run, so it computes this workload
    one grid size at a time."""
    main, grb, struct = gauss21_fsrs_ssc()
    grb = dict(grb, method_eats="piece-wise")
    struct = dict(struct, n_layers_pw=64, theta_w=0.2, layer_ngam=layer_ngam(64))
    return main, grb, struct


WORKLOADS = {"tophat_fs_syn": tophat_fs_syn, "gauss21_fsrs_ssc": gauss21_fsrs_ssc, "tophat64_pw": tophat64_pw,
             "gauss64_pw_ng": gauss64_pw_ng}


def scaled(name, ntb=None, ngam=None, nfreq=None, n_nu_obs=None, n_t_obs=None, n_layers=None, **over):
    """A smaller variant of a named workload (grids shrunk, physics unchanged)."""
    main, grb, struct = WORKLOADS[name]()
    main, grb, struct = copy.deepcopy(main), copy.deepcopy(grb), copy.deepcopy(struct)
    if ntb:
        main["ntb"] = ntb
    for sh in ("fs", "rs"):
        if ngam and f"ngam_{sh}" in grb:
            grb[f"ngam_{sh}"] = ngam
        if nfreq and f"nfreq_{sh}" in grb:
            grb[f"nfreq_{sh}"] = nfreq
    if n_nu_obs:
        main["lc_freqs"] = f"array logspace 1e8 1e29 {n_nu_obs}"
    if n_t_obs:
        main["lc_times"] = f"array logspace 3e3 1e10 {n_t_obs}"
    if n_layers and struct["struct"] == "gaussian":
        struct["n_layers_a"] = n_layers
    for k, v in over.items():
        if k in struct or k in ("n_layers_pw", "layer_ngam"):
            struct[k] = v
        elif k in ("tb0", "tb1", "ntb", "n_ism", "A0", "s", "r_ej", "r_ism", "d_l", "z", "theta_obs", "lc_freqs",
                   "lc_times", "lc_use_freq_to_time", "rtol", "integrator"):
            main[k] = v
        else:
            grb[k] = v
    return main, grb, struct


SMALL = {
    # reduced grids of the two workloads above, small enough for the CPU oracle in tests
    "tophat_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8),
    "gauss5_small": dict(base="gauss21_fsrs_ssc", ntb=1000, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8, n_layers=5),
    # piece-wise grids (layer i holds 2i+1 phi cells); the Gaussian wing is cut at 0.2 rad,
    # as the piece-wise initial data keep every layer and a slower wing stops short of t_obs.
    # The reverse shock ends on the density criterion (rho4 < 1e-30): with the default 1e-50 it
    # ends when the swept-up fraction M3 -> 1 first rounds to >= 1, a step set by the last ulp
    # of M3 (blastwave.h:487-510), and the piece-wise reverse-shock term (forward-shock
    # emissivity times the reverse-shock mass) is large enough for that step to show.
    "gauss_pw_small": dict(base="gauss21_fsrs_ssc", ntb=1000, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                           method_eats="piece-wise", n_layers_pw=6, theta_w=0.2, rs_shutOff_criterion_rho=1e-30),
    # the jet of gauss_pw_small with a grid size of its own per layer (the GPU path's per-layer
    # grids), and the reference's uniform runs at the other two sizes (one run per size pins it)
    "gauss_pw_ng_small": dict(base="gauss21_fsrs_ssc", ntb=1000, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                              method_eats="piece-wise", n_layers_pw=6, theta_w=0.2, rs_shutOff_criterion_rho=1e-30,
                              layer_ngam=(32, 64, 48, 48, 32, 64)),
    "gauss_pw_ng32": dict(base="gauss21_fsrs_ssc", ntb=1000, ngam=32, nfreq=48, n_nu_obs=6, n_t_obs=8,
                          method_eats="piece-wise", n_layers_pw=6, theta_w=0.2, rs_shutOff_criterion_rho=1e-30),
    "gauss_pw_ng48": dict(base="gauss21_fsrs_ssc", ntb=1000, ngam=48, nfreq=48, n_nu_obs=6, n_t_obs=8,
                          method_eats="piece-wise", n_layers_pw=6, theta_w=0.2, rs_shutOff_criterion_rho=1e-30),
    # the same jet at the default rs_shutOff_criterion_rho = 1e-50: the reverse shock of a layer ends
    # at the step where M3 first rounds to >= 1
    "gauss_pw_default": dict(base="gauss21_fsrs_ssc", ntb=1000, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                             method_eats="piece-wise", n_layers_pw=6, theta_w=0.2),
    "tophat_pw_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                            method_eats="piece-wise", n_layers_pw=4),
    # the 64-layer piece-wise Gaussian jet (4096 cells, both shocks, SSC) of gauss64_pw_ng at the
    # reduced microphysics grids, one grid for every layer, with the reverse shock ending on the
    # density criterion as in gauss_pw_small
    "gauss64_pw_small": dict(base="gauss21_fsrs_ssc", ntb=1000, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                             method_eats="piece-wise", n_layers_pw=64, theta_w=0.2, rs_shutOff_criterion_rho=1e-30),
    # synchrotron self-absorption on (the radio points of the light curve are self-absorbed early on)
    "tophat_ssa_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                             use_ssa_fs="yes"),
    "gauss5_ssa_small": dict(base="gauss21_fsrs_ssc", ntb=1000, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                             n_layers=5, use_ssa_fs="yes", use_ssa_rs="yes"),
    # the Dermer (2009) synchrotron kernel in place of CSYN
    "tophat_dermer_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                                method_synchrotron_fs="Dermer09"),
    # the Bessel-function synchrotron kernel (K_{4/3}, K_{1/3}) in place of CSYN; a shorter grid,
    # since the reference's std::cyl_bessel_k takes up to 0.1 ms per call at large arguments
    "tophat_bessel_small": dict(base="tophat_fs_syn", ntb=300, ngam=48, nfreq=32, n_nu_obs=6, n_t_obs=8,
                                method_synchrotron_fs="Bessel"),
    # GSL's Chebyshev fit of the synchrotron function F(x) in place of CSYN
    "tophat_gsl_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                             method_synchrotron_fs="GSL"),
    # a wind-like ambient medium (n_ism < 0: density A0 R^-2 between r_ej and r_ism, constant inside and
    # outside, RhoISM::evaluateRhoDrhoDrDefault) in place of the uniform ISM
    "tophat_wind_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                              n_ism=-1., A0=3e35, s=2., r_ej=1e16, r_ism=3e18),
    # the classical RK4 integrator, one step per output interval, in place of adaptive DOP853
    "tophat_rk4_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                             integrator="RK4"),
    # the source at z = 0.7 with SSC on (TeV photons), 12 observer frequencies from 1e9 to 1e28 Hz, 3
    # of them inside the frequency range of an EBL optical-depth table (the reference's
    # ebl_tbl_fpath, set by the caller)
    "tophat_ebl_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, n_t_obs=8, z=0.7,
                             lc_freqs="array logspace 1e9 1e28 12", method_ssc_fs="numeric"),
    # one frequency per observer time (lc_use_freq_to_time = yes): 16 pairs, the i-th time with the
    # i-th frequency, a frequency sweep that rises with time
    "tophat_f2t_small": dict(base="tophat_fs_syn", ntb=300, ngam=64, nfreq=48, lc_use_freq_to_time="yes",
                             lc_times="array logspace 3e3 1e10 16", lc_freqs="array logspace 1e8 1e20 16"),
}


def small(name):
    cfg = dict(SMALL[name])
    base = cfg.pop("base")
    return scaled(base, **cfg)


# Kilonova-like ejecta: velocity shells with beta*Gamma log-spaced over KN_MOM and per-shell kinetic
# energies drawn log-uniform over KN_EK (ejecta_structure, seeded)
KN_MOM = (0.1, 3.0)
KN_EK = (1e48, 1e51)
KN_SEED = 20261203


def kn_c5():
    """Kilonova-like ejecta without a jet: 32 velocity shells (beta*Gamma 0.1..3) x 30 polar layers
    (adaptive, up to pi/2) = 960 blastwaves in one ODE system, every one evolved by the
    reference's ejecta right-hand side (bw_type = fs_dense) in a uniform ISM of 1e-2 cm^-3;
    eps_e = 0.1, eps_b = 1e-2, p = 2.5, synchrotron + SSC on 128 Lorentz-factor bins and 128
    comoving frequencies, 1000 steps, 96 nu x 128 t_obs from 1e4 s to 1e10 s.  Returns (main,
    kn, struct) for inputs.write_kn_run."""
    main = dict(tb0=1e3, tb1=1e14, ntb=1000, n_ism=1e-2, d_l=1.27e26, z=0.0099, theta_obs=0.0,
                lc_freqs="array logspace 1e8 1e29 96", lc_times="array logspace 1e4 1e10 128")
    kn = dict(eps_e_fs=0.1, eps_b_fs=1e-2, p_fs=2.5, epsilon_e_rad_fs=0., ngam_fs=128, gam1_fs=1., gam2_fs=1e8,
              freq1_fs=1e5, freq2_fs=1e32, nfreq_fs=128, method_ssc_fs="numeric", method_ele_fs="numeric",
              bw_type="fs_dense", do_rs="no", method_eats="adaptive", method_spread="None",
              fname_ejecta_id="id_kn.h5", fname_dyn="dyn_kn.h5", fname_light_curve="lc_kn.h5")
    struct = dict(n_shells=32, n_layers=30, mom_min=KN_MOM[0], mom_max=KN_MOM[1], ek_min=KN_EK[0], ek_max=KN_EK[1],
                  seed=KN_SEED)
    return main, kn, struct


KN_WORKLOADS = {"kn_c5": kn_c5}


def kn_scaled(name, n_shells=None, n_layers=None, ntb=None, ngam=None, nfreq=None, n_nu_obs=None, n_t_obs=None,
              main=None, **kn_over):
    """A smaller variant of a kilonova workload (fewer blastwaves, steps, bins or pairs; physics
    unchanged).  `main` overrides main-section keys, other keywords kN-section keys."""
    main_over = main or {}
    main, kn, struct = (copy.deepcopy(x) for x in KN_WORKLOADS[name]())
    main.update(main_over)
    if n_shells:
        struct["n_shells"] = n_shells
    if n_layers:
        struct["n_layers"] = n_layers
    if ntb:
        main["ntb"] = ntb
    if ngam:
        kn["ngam_fs"] = ngam
    if nfreq:
        kn["nfreq_fs"] = nfreq
    if n_nu_obs:
        main["lc_freqs"] = f"array logspace 1e8 1e29 {n_nu_obs}"
    if n_t_obs:
        main["lc_times"] = f"array logspace 1e4 1e10 {n_t_obs}"
    kn.update(kn_over)
    return main, kn, struct


KN_SMALL = {
    # 4 velocity shells x 3 layers (12 blastwaves) at reduced grids, small enough for the CPU oracle
    "kn_small": dict(n_shells=4, n_layers=3, ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8),
    # the same ejecta with the observer integral over piece-wise cells (2i+1 per layer)
    "kn_pw_small": dict(n_shells=4, n_layers=3, ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                        method_eats="piece-wise"),
    # method_spread set: the ejecta RHS keeps dtheta/dR = 0 but takes dt_obs/dR and the initial
    # observer time in their spreading form (EQS::evalElapsedTime, init_elapsed_time)
    "kn_spread_small": dict(n_shells=4, n_layers=3, ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                            method_spread="our"),
    # a wind-like ambient medium (n_ism < 0: A0 R^-2 between r_ej and r_ism): the stored density is
    # the one the last RHS call evaluated (set_standard_ism), not the start radius's
    "kn_wind_small": dict(n_shells=4, n_layers=3, ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                          main=dict(n_ism=-1., A0=3e35, s=2., r_ej=1e16, r_ism=3e18)),
    # the classical RK4 integrator: the stored GammaRel is the Gamma of the last RHS call, the
    # state of RK4's last stage rather than the step's end
    "kn_rk4_small": dict(n_shells=4, n_layers=3, ntb=300, ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8,
                         main=dict(integrator="RK4")),
}


def kn_small(name):
    return kn_scaled("kn_c5", **KN_SMALL[name])
