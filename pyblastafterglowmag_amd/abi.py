"""ctypes mirror of include/pbagpu.h and helpers that turn parfile sections into its structures.

The mapping from parfile keys to structure fields follows the readers of the
reference: BlastWaveBase::setBaseParams / setParamsRS
(src/blastwave/blastwave_base.h:95-456), ElectronAndRaidaionBase::setBasePars
(src/microphysics/microphysics_base.h:59-560), ElectronAndRadiation::setPars
(src/microphysics/microphysics.h:72-130), EATS::setEatsPars (src/eats.h:718-770)
and PyBlastAfterglow::setModelPars (src/main.cpp:93-154).
"""
import ctypes as C
import math

import numpy as np

from . import inputs

PBA_NV_NAMES = [
    "tburst", "R", "Rsh", "tt", "tcomov", "Gamma", "beta", "mom", "theta", "ctheta",
    "M2", "Eint2", "Erad2", "Esh2", "Ead2", "rho", "drhodr", "adi", "rho2", "GammaFsh",
    "thickness", "U_p", "gamma_min", "gamma_c", "gamma_max", "B", "ThetaTherm", "z_cool", "nprime", "accel_frac",
    "Eint3", "Erad3", "iEad3", "Esh3", "M3", "deltaR4", "Gamma43", "adi3", "rho4", "rho3",
    "GammaRsh", "Rsh_rs", "thichness_rs", "U_p3", "gamma_min_rs", "gamma_c_rs", "gamma_max_rs", "B_rs",
    "ThetaTherm_rs", "z_cool_rs", "nprime_rs", "accel_frac_rs", "GammaRel", "mflags",
]
PBA_NV = len(PBA_NV_NAMES)
PV = {n: i for i, n in enumerate(PBA_NV_NAMES)}
# raw sky-map arrays (enum pba_img_var; IMG::m_names, src/image.h:14-26)
IMG_NAMES = ["r", "ctheta", "cphi", "mu", "xrs", "yrs", "intensity"]
PBA_IMG_NV = len(IMG_NAMES)
IMG = {n: i for i, n in enumerate(IMG_NAMES)}


class ShockParams(C.Structure):
    _fields_ = [("eps_e", C.c_double), ("eps_b", C.c_double), ("eps_t", C.c_double), ("p", C.c_double),
                ("gam2", C.c_double), ("freq1", C.c_double), ("freq2", C.c_double), ("beta_min", C.c_double),
                ("epsilon_e_rad", C.c_double), ("ngam", C.c_int), ("nfreq", C.c_int),
                ("max_substeps", C.c_int), ("ssc", C.c_int), ("adi_loss_ele", C.c_int), ("ssa", C.c_int),
                ("syn", C.c_int), ("pad_", C.c_int)]


class RunParams(C.Structure):
    _fields_ = [("n_ism", C.c_double), ("d_l", C.c_double), ("z", C.c_double), ("theta_obs", C.c_double),
                ("rtol", C.c_double), ("tb0", C.c_double),
                ("tprompt", C.c_double), ("rs_shutoff_rho", C.c_double), ("min_gamma0_for_rs", C.c_double),
                ("rs_gamma0_frac_no_exceed", C.c_double), ("mom_when_start_spread", C.c_double),
                ("theta_max", C.c_double), ("ism_A0", C.c_double), ("ism_s", C.c_double), ("ism_r_ej", C.c_double),
                ("ism_r_ism", C.c_double), ("rtol_theta", C.c_double), ("rtol_phi", C.c_double),
                ("nmax_theta", C.c_int), ("nmax_phi", C.c_int), ("fsrs", C.c_int), ("do_rs", C.c_int),
                ("do_rs_radiation", C.c_int), ("adiab_loss", C.c_int), ("adiab_loss_rs", C.c_int),
                ("exponential_rho4", C.c_int), ("init_deltaR4", C.c_int), ("counter_jet", C.c_int),
                ("nmax", C.c_int), ("integrator", C.c_int), ("fs", ShockParams), ("rs", ShockParams)]


class LayerInit(C.Structure):
    _fields_ = [("E0", C.c_double), ("M0", C.c_double), ("mom0", C.c_double), ("theta_c_l", C.c_double),
                ("theta_c_h", C.c_double), ("theta_c", C.c_double), ("ctheta0", C.c_double),
                ("theta_wing", C.c_double), ("theta_core", C.c_double), ("theta_b0", C.c_double),
                ("ncells", C.c_double), ("ilayer", C.c_int), ("nlayers", C.c_int), ("ngam_fs", C.c_int),
                ("ngam_rs", C.c_int)]


class BwResult(C.Structure):
    _fields_ = [("status", C.c_int), ("i_end_r", C.c_int), ("fsrs", C.c_int), ("n_steps", C.c_int),
                ("n_rhs", C.c_int), ("pad_", C.c_int)]


class EatsPwLayer(C.Structure):
    _fields_ = [("ctheta0", C.c_double), ("ncells", C.c_double), ("cil", C.c_int), ("bw", C.c_int), ("run", C.c_int),
                ("pop_fs", C.c_int), ("rs", C.c_int), ("i_end_r", C.c_int)]


class Population(C.Structure):
    _fields_ = [("bw", C.c_int), ("run", C.c_int), ("shock", C.c_int), ("ngam", C.c_int)]


PBA_MLP_MAXL = 8
_FP = C.POINTER(C.c_float)


class Mlp(C.Structure):
    """pba_mlp: layer sizes and device pointers (FP32) of weights, biases, gradients and Adam moments."""
    _fields_ = [("n_layers", C.c_int), ("dims", C.c_int * (PBA_MLP_MAXL + 1))] + \
        [(n, C.c_void_p * PBA_MLP_MAXL) for n in ("w", "b", "gw", "gb", "mw", "vw", "mb", "vb")]


class SscSurrogate(C.Structure):
    _fields_ = [("net", Mlp), ("y_offset", C.c_void_p), ("y_scale", C.c_float), ("check_tol", C.c_float),
                ("check_every", C.c_int), ("pad_", C.c_int)]


class Spectra(C.Structure):
    """pba_spectra: device pointers of the dense-output spectra (0 = not written)."""
    _fields_ = [(n, C.c_void_p) for n in ("syn_j", "syn_i", "ssc_j", "total_f", "total_i", "gam_dot_syn", "gam_dot_adi",
                                          "gam_dot_ssc")]


SPECTRA_NU = ("syn_j", "syn_i", "ssc_j", "total_f", "total_i")
SPECTRA_GAM = ("gam_dot_syn", "gam_dot_adi", "gam_dot_ssc")


class SurrogateStats(C.Structure):
    _fields_ = [("inferred", C.c_longlong), ("checked", C.c_longlong), ("failed", C.c_longlong),
                ("fallback", C.c_longlong), ("max_err", C.c_double)]


class EatsLayer(C.Structure):
    _fields_ = [("theta_c_l", C.c_double), ("theta_c_h", C.c_double), ("bw", C.c_int), ("run", C.c_int),
                ("pop_fs", C.c_int), ("pop_rs", C.c_int), ("i_end_r", C.c_int)]


EATS_METHODS = ("adaptive", "piece-wise")
INTEGRATORS = ("DOP853E", "RK4")   # pba_run_params.integrator 0 / 1
SUPPORTED_OPTS = dict(
    method_eats=EATS_METHODS, method_quad="CADRE", method_comp_mode="comovSpec", method_spread="our",
    method_limit_spread="MomValAndFrac", method_dgdr="our", method_eos="Nava13", method_dmdr="usingdthdr",
    integrator=INTEGRATORS, do_average_solution="no",
)
# ebl_tbl_fpath: "none" or an EBL optical-depth table (pyblastafterglowmag_amd.ebl), any path


def _yes(v):
    if v not in ("yes", "no"):
        raise ValueError(f"boolean option expected yes/no, got {v!r}")
    return 1 if v == "yes" else 0


# method_synchrotron values of the numeric electron path and their code in pba_shock_params.syn
SYNCHROTRON_KERNELS = {"CSYN": 0, "Dermer09": 1, "Bessel": 2, "GSL": 3}


def _shock(pars, opts, sh):
    s = ShockParams()
    s.eps_e = pars[f"eps_e_{sh}"]
    s.eps_b = pars[f"eps_b_{sh}"]
    s.eps_t = pars.get(f"eps_t_{sh}", 0.0)
    s.p = pars[f"p_{sh}"]
    s.gam2 = pars[f"gam2_{sh}"]
    s.freq1 = pars[f"freq1_{sh}"]
    s.freq2 = pars[f"freq2_{sh}"]
    s.beta_min = pars.get(f"beta_min_{sh}", 1e-5)
    s.epsilon_e_rad = pars.get(f"epsilon_e_rad_{sh}", 0.0)
    s.ngam = int(pars[f"ngam_{sh}"])
    s.nfreq = int(pars[f"nfreq_{sh}"])
    s.max_substeps = int(pars.get(f"max_substeps_{sh}", 1000))
    s.ssc = 1 if opts.get(f"method_ssc_{sh}", "none") == "numeric" else 0
    s.adi_loss_ele = _yes(opts.get(f"num_ele_use_adi_loss_{sh}", "yes"))
    s.ssa = _yes(opts.get(f"use_ssa_{sh}", "no"))
    syn = opts.get(f"method_synchrotron_{sh}", "CSYN")
    if syn not in SYNCHROTRON_KERNELS:
        raise NotImplementedError(f"method_synchrotron_{sh}={syn} (only {' / '.join(SYNCHROTRON_KERNELS)} is implemented)")
    s.syn = SYNCHROTRON_KERNELS[syn]
    for key, want in ((f"method_ele_{sh}", "numeric"),
                      (f"method_pp_{sh}", "none"), (f"method_ne_{sh}", "useNe"),
                      (f"method_gamma_min_{sh}", "useNumeric"), (f"method_gamma_c_{sh}", "useTcomov"),
                      (f"method_gamma_max_{sh}", "useB"), (f"method_B_{sh}", "useU_b"),
                      (f"method_nonrel_dist_{sh}", "use_Sironi"), (f"method_tau_{sh}", "smooth"),
                      (f"method_Up_{sh}", "useEint2"), (f"method_thickness_{sh}", "useJoh06"),
                      (f"method_vel_{sh}", "sameAsBW")):
        if key in opts and opts[key] != want:
            raise NotImplementedError(f"{key}={opts[key]} (only {want} is implemented)")
    return s


def run_params_from_parfile(path):
    """(RunParams, main_pars, main_opts, grb_pars, grb_opts) from a parfile with a GRB section."""
    mp_, mo = inputs.read_parfile(path, inputs.SEC_MAIN)
    gp, go = inputs.read_parfile(path, inputs.SEC_GRB)
    kp, ko = inputs.read_parfile(path, inputs.SEC_KN)
    if ko.get("run_bws", "no") == "yes":
        # the drop-in driver reads and writes the jet's files; ejecta go through run_params_ejecta,
        # ejecta_inits and pbagpu_dyn_evolve_ejecta (gpu.Pipeline.dynamics_ejecta)
        raise NotImplementedError(f"{path}: kilonova ejecta (kN section, run_bws = yes) are not read by the jet's "
                                  "parfile path; use run_params_ejecta / ejecta_inits and Pipeline.dynamics_ejecta")
    if go.get("run_bws", "no") != "yes":
        raise ValueError(f"{path}: no jet to evolve (GRB section missing or run_bws = no)")
    return run_params(mp_, mo, gp, go), mp_, mo, gp, go


def run_params(mp_, mo, gp, go, opts=None):
    """RunParams of the main section and one afterglow section; `opts` is the table of options the
    path implements (SUPPORTED_OPTS for a jet, EJECTA_OPTS for kilonova ejecta)."""
    ejecta = opts is EJECTA_OPTS
    for key, want in (opts or SUPPORTED_OPTS).items():
        allowed = want if isinstance(want, tuple) else (want,)
        have = mo.get(key, go.get(key, allowed[0]))
        if have not in allowed:
            raise NotImplementedError(f"{key}={have} (only {' / '.join(allowed)} is implemented)")
    if mp_.get("n_ism", -1) <= 0 and (mp_.get("s", -1) < 0 or mp_.get("A0", -1) < 0):
        # RhoISM::evaluateRhoDrhoDrDefault exits there (blastwave_components.h:857-860)
        raise ValueError("if n_ism < 0; positive A0 and s are expected (wind)")
    if int(mp_.get("iout", 1)) != 1:
        raise NotImplementedError("iout != 1")
    P = RunParams()
    P.n_ism = mp_["n_ism"]
    P.ism_A0, P.ism_s = mp_.get("A0", -1.), mp_.get("s", -1.)
    P.ism_r_ej, P.ism_r_ism = mp_.get("r_ej", -1.), mp_.get("r_ism", -1.)
    P.d_l = mp_["d_l"]
    P.z = mp_["z"]
    P.theta_obs = mp_["theta_obs"]
    P.rtol = mp_["rtol"]
    P.tb0 = mp_["tb0"]
    P.nmax = int(mp_.get("nmax", 100000))
    P.integrator = INTEGRATORS.index(mo.get("integrator", "DOP853E"))
    P.tprompt = gp.get("tprompt", 1000.)
    P.rs_shutoff_rho = gp.get("rs_shutOff_criterion_rho", 1e-50)
    P.min_gamma0_for_rs = gp.get("min_Gamma0_for_rs", 0.)
    P.rs_gamma0_frac_no_exceed = gp.get("rs_Gamma0_frac_no_exceed", .98)
    P.mom_when_start_spread = gp.get("mom_when_start_spread", 2.)
    P.theta_max = gp.get("theta_max", np.pi / 2.)
    P.rtol_theta = gp["rtol_theta"]
    P.rtol_phi = gp["rtol_phi"]
    P.nmax_theta = int(gp.get("nmax_theta", 1000))
    P.nmax_phi = int(gp.get("nmax_phi", 1000))
    bw_type = go.get("bw_type", "fs")
    if bw_type not in (("fs_dense",) if ejecta else ("fs", "fsrs")):
        raise NotImplementedError(f"bw_type={bw_type}")
    P.fsrs = 1 if bw_type == "fsrs" else 0
    P.do_rs = _yes(go.get("do_rs", "no")) if P.fsrs else 0
    P.do_rs_radiation = _yes(go.get("do_rs_radiation", "no")) if P.fsrs else 0
    P.adiab_loss = _yes(go.get("use_adiabLoss", "yes"))
    P.adiab_loss_rs = _yes(go.get("use_adiabLoss_rs", "yes"))
    P.exponential_rho4 = _yes(go.get("exponential_rho4", "yes"))
    P.init_deltaR4 = _yes(go.get("init_deltaR4", "no"))
    P.counter_jet = _yes(go.get("counter_jet", "yes"))
    P.fs = _shock(gp, go, "fs")
    P.rs = _shock(gp, go, "rs") if P.fsrs else _shock(gp, go, "fs")
    return P


def layer_inits(idd, method="adaptive"):
    """LayerInit array of the layers of a jet_structure() dict (shell 0).

    The reference's loader builds the adaptive grid and then the piece-wise grid, and the
    latter overwrites the layer edges and centres (EjectaID2::_init_a_grid / _init_pw_grid,
    src/ejecta/ejecta_id.h:142-182, called in that order at :298-300): the edges used by the
    dynamics and the observer integral are 2 asin(i/n sin(theta_w/2)).

    method="piece-wise" (ejecta_id.h:315-322, blastwave_base.h:304-312): E0 and M0 of layer i
    are divided by its 2i+1 cells, every blastwave starts at theta_b0 = theta_wing, and the
    swept-up mass, shock thickness and radiative loss are shared by ncells = n^2 cells.
    """
    if method not in EATS_METHODS:
        raise NotImplementedError(f"method_eats={method}")
    pw = method == "piece-wise"
    n = len(idd["ek"])
    arr = (LayerInit * n)()
    tw = float(idd["theta_wing"])
    dth = tw / n
    th_pw = [2.0 * math.asin((i / n) * math.sin(tw / 2.0)) for i in range(n + 1)]  # libm, as the C++ loader
    for i in range(n):
        L = arr[i]
        cil = float(cells_in_layer(i)) if pw else 1.0
        L.E0 = float(idd["ek"][i]) / cil
        L.M0 = float(idd["mass"][i]) / cil
        L.mom0 = float(idd["mom"][i])
        L.theta_c_l = th_pw[i]
        L.theta_c_h = th_pw[i + 1]
        L.theta_c = i * dth + dth / 2.
        L.ctheta0 = 0.5 * (th_pw[i + 1] + th_pw[i])
        L.theta_wing = tw
        L.theta_core = float(idd["theta_core"])
        L.theta_b0 = tw if pw else th_pw[i + 1]
        L.ncells = float(total_cells(n)) if pw else 1.0
        L.ilayer = i
        L.nlayers = n
    return arr


# Options of the kilonova ejecta (parfile kN section) that the ejecta path implements: every
# ejecta blastwave goes through setStandardISM + rhs_fs_dense (src/model_evolve.h:1627-1651),
# bw_type = fs_dense stores the upstream's state (src/blastwave/blastwave.h:544-560).  method_spread
# only selects the form of dt_obs/dR there (dtheta/dR = 0 in rhs_fs_dense).
EJECTA_OPTS = dict(
    method_eats=EATS_METHODS, method_quad="CADRE", method_comp_mode="comovSpec", method_spread=("None", "our"),
    method_dgdr="our", method_eos="Nava13", method_dmdr="usingdthdr", method_collision="none",
    bw_type="fs_dense", do_nucinj="no", do_eninj_inside_rhs="no", use_1d_id="no", load_r0="no",
    integrator=INTEGRATORS, do_average_solution="no",
)


def run_params_ejecta(mp_, mo, kp, ko):
    """(RunParams, spread) of the ejecta of a kilonova run: parfile main and kN sections."""
    P = run_params(mp_, mo, kp, ko, opts=EJECTA_OPTS)
    return P, ko.get("method_spread", "None") != "None"


def ejecta_inits(id2d, method="adaptive"):
    """LayerInit array of the blastwaves of a 2-D [shells x layers] initial-data dict (r, mom,
    theta, ctheta, ek, mass), in the reference's state-vector order: layer-major, the velocity
    shells of layer 0 first (src/model_evolve.h:225-234).  Returns (inits, order) with order the
    (shell, layer) of each entry.

    Per shell the loader builds the layers' angular grids from theta_wing, the last layer's theta
    of the last shell, as for a jet (EjectaID2::_load_id_file, src/ejecta/ejecta_id.h:221-300).
    A blastwave with ek or mass not above 0, or beta0 <= 1e-6 (its start radius beta0 c t0 then
    also fails the r > 0 test), is not evolved (setEjectaBwPars, src/ejecta/ejecta_base.h:255-265):
    its E0 = M0 = 0 here."""
    ek = np.atleast_2d(np.asarray(id2d["ek"], dtype=np.float64))
    ns, nl = ek.shape
    mass = np.atleast_2d(np.asarray(id2d["mass"], dtype=np.float64))
    mom = np.atleast_2d(np.asarray(id2d["mom"], dtype=np.float64))
    theta = np.atleast_2d(np.asarray(id2d["theta"], dtype=np.float64))
    tw = float(theta[ns - 1, nl - 1])
    rows = [layer_inits(dict(ek=ek[ish], mass=mass[ish], mom=mom[ish], theta_core=tw, theta_wing=tw), method)
            for ish in range(ns)]
    out = (LayerInit * (ns * nl))()
    order = []
    for il in range(nl):
        for ish in range(ns):
            L = out[len(order)]
            C.memmove(C.addressof(L), C.addressof(rows[ish][il]), C.sizeof(LayerInit))
            beta0 = mom[ish, il] / math.sqrt(mom[ish, il] * mom[ish, il] + 1.0)
            if not (ek[ish, il] > 0 and mass[ish, il] > 0 and beta0 > 1e-6):
                L.E0 = L.M0 = 0.0
            order.append((ish, il))
    return out, order


def cells_in_layer(i):
    """phi cells of layer i (EjectaID2::CellsInLayer, src/ejecta/ejecta_id.h:119-121)."""
    return 2 * i + 1


def total_cells(n_layers):
    """Cells of an n-layer jet (EjectaID2::_evalTotalNcells, src/ejecta/ejecta_id.h:149-155)."""
    return sum(cells_in_layer(i) for i in range(n_layers))


def tgrid(mp_):
    return inputs.logspace_vec(np.log10(mp_["tb0"]), np.log10(mp_["tb1"]), int(mp_["ntb"]))


def obs_pairs(mo):
    """(t_obs, nu_obs) pairs (cast_times_freqs, src/utilitites/utils.h:1419-1448): every time with
    every frequency, time-major; with lc_use_freq_to_time = yes the i-th time with the i-th
    frequency (the two lists must have one length)."""
    ts = inputs.vec_from_string(mo["lc_times"])
    fs = inputs.vec_from_string(mo["lc_freqs"])
    one_to_one = _yes(mo.get("lc_use_freq_to_time", "no"))
    if one_to_one:
        if len(ts) != len(fs):
            raise ValueError("lc_use_freq_to_time = yes: lc_times and lc_freqs need one length "
                             f"({len(ts)} != {len(fs)})")
        return ts.copy(), fs.copy()
    return np.repeat(ts, len(fs)), np.tile(fs, len(ts))
