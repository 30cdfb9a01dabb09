"""Run inputs: parfile text and jet initial data, in the reference's file formats.

* Parfile: sections delimited by exact header lines, each with a
  ``* Parameters`` block (numbers) and a ``* Settings`` block (strings), read by
  ``readParFile2`` (src/utilitites/utils.h:1450-1538).  The default key set and
  values are those the reference's Python package writes
  (package/src/PyBlastAfterglowMag/parfile_parameter_list.py:5-180), so a
  parfile written here is accepted unchanged by the reference CPU program.
* Initial data: the adaptive / piece-wise angular grids and the Gaussian and
  top-hat energy / momentum / mass profiles of ``JetStruct``
  (package/src/PyBlastAfterglowMag/id_analytic.py:86-206), written in the
  ``use_1d_id = yes`` layout (1-D datasets plus theta_core / theta_wing
  attributes) that ``EjectaID2::_load_id_file`` reads (src/ejecta/ejecta_id.h:198-221).
"""
import copy
import os

import numpy as np

C_LIGHT = 2.99792458e10

SEC_MAIN = "# -------------------------- main ---------------------------"
SEC_GRB = "# ---------------------- GRB afterglow ----------------------"
SEC_KN = "# ----------------------- kN afterglow ----------------------"
SEC_END = "# --------------------------- END ---------------------------"

# Key order and values follow parfile_parameter_list.py:5-180.
MAIN_DEFAULTS = dict(
    pars=dict(tb0=1.e3, tb1=1.e11, ntb=1000, iout=1, rtol=1e-8, nmax=1000, A0=-1, s=-1, r_ej=-1, r_ism=-1,
              n_ism=1., d_l=3.09e26, z=0.028, theta_obs=0),
    opts=dict(do_average_solution="no", lc_freqs="array 1e9 1e18", lc_times="array logspace 3.e3 1.e10 100",
              lc_use_freq_to_time="no", skymap_freqs="array 1e9 1e18", skymap_times="array logspace 8.e5 2e9 50",
              integrator="DOP853E"),
)
GRB_DEFAULTS = dict(
    pars=dict(eps_e_fs=0.1, eps_b_fs=0.01, eps_t_fs=0., p_fs=2.2, epsilon_e_rad_fs=-1, gamma_max_fs=1e7,
              max_substeps_fs=2000, gam1_fs=1., gam2_fs=1.e8, ngam_fs=451, freq1_fs=1.e5, freq2_fs=1.e32,
              nfreq_fs=601,
              eps_e_rs=0.1, eps_b_rs=0.01, eps_t_rs=0., p_rs=2.2, epsilon_e_rad_rs=-1, gamma_max_rs=1e7,
              max_substeps_rs=2000, gam1_rs=1., gam2_rs=1.e7, ngam_rs=451, freq1_rs=1.e5, freq2_rs=1.e32,
              nfreq_rs=601,
              n_store_substeps=10, tprompt=1.e3, a=1, rs_shutOff_criterion_rho=1e-50, min_Gamma0_for_rs=5.,
              mom_when_start_spread=2, mom0_frac_when_start_spread=0.9, rs_Gamma0_frac_no_exceed=.92,
              save_dyn_every_it=10, rtol_phi=1e-6, rtol_theta=1e-6, nmax_phi=1000, nmax_theta=1000,
              theta_max=3.1415 / 2., beta_min_fs=1e-5, beta_min_rs=1e-5, nsublayers=5, frac_to_increase=1.5,
              max_restarts=12, min_sublayers=3, min_non_zero_cells=5, im_max_theta=1.5708),
    opts=dict(run_bws="yes", save_dynamics="no", load_dynamics="no", do_mphys_in_situ="yes",
              do_mphys_in_ppr="no", save_spec="no", do_lc="yes", do_skymap="no", save_raw_skymap="yes",
              skymap_remove_mu="no", counter_jet="yes",
              do_rs="no", do_rs_radiation="yes", bw_type="fs", init_deltaR4="no", exponential_rho4="yes",
              method_collision="none", method_eats="adaptive", method_quad="CADRE", method_comp_mode="comovSpec",
              allow_termination="no", do_thermrad_loss="no", do_eninj_inside_rhs="no",
              use_1d_id="yes", fname_ejecta_id="id.h5", load_r0="no", fname_dyn="dyn.h5",
              fname_spectrum="spec.h5", fname_light_curve="lc.h5", fname_sky_map="skymap.h5",
              ebl_tbl_fpath="none", do_nucinj="no",
              method_spread="our", method_limit_spread="MomValAndFrac", method_dgdr="our", method_eos="Nava13",
              method_dmdr="usingdthdr", use_dens_prof_behind_jet_for_ejecta="no",
              use_adiabLoss="yes", method_Up_fs="useEint2", method_thickness_fs="useJoh06",
              method_vel_fs="sameAsBW", method_ele_fs="numeric", num_ele_use_adi_loss_fs="yes",
              method_ne_fs="useNe", method_nonrel_dist_fs="use_Sironi", method_gamma_min_fs="useNumeric",
              method_gamma_c_fs="useTcomov", method_gamma_max_fs="useB", method_B_fs="useU_b",
              method_synchrotron_fs="CSYN", incl_th_in_marg21_fs="yes", use_ssa_fs="no", method_ssc_fs="none",
              method_pp_fs="none", method_tau_fs="smooth",
              use_adiabLoss_rs="yes", method_Up_rs="useEint2", method_thickness_rs="useJoh06",
              method_vel_rs="sameAsBW", method_ele_rs="numeric", num_ele_use_adi_loss_rs="yes",
              method_ne_rs="useNe", method_nonrel_dist_rs="use_Sironi", method_gamma_min_rs="useNumeric",
              method_gamma_c_rs="useTcomov", method_gamma_max_rs="useB", method_B_rs="useU_b",
              method_synchrotron_rs="CSYN", incl_th_in_marg21_rs="yes", use_ssa_rs="no", method_ssc_rs="none",
              method_pp_rs="none", method_tau_rs="smooth"),
)


def _section(lines, sep1, default, new):
    d = copy.deepcopy(default)
    for k, v in (new or {}).items():
        if k in d["pars"]:
            d["pars"][k] = v
        elif k in d["opts"]:
            d["opts"][k] = v
        else:
            raise KeyError(f"unknown parfile key {k!r}")
    lines += ["", sep1, "", "* Parameters", ""]
    lines += [f"{k} = {v}" for k, v in d["pars"].items()]
    lines += ["", "* Settings", ""]
    lines += [f"{k} = {v}" for k, v in d["opts"].items()]
    lines += ["", SEC_END, ""]
    return d


def write_parfile(workdir, main=None, grb=None, fname="parfile.par", kn=None):
    """Write a parfile with the main section and (optionally) the GRB and kN sections.  The kN
    section takes the GRB section's key set (the reference package writes both from one
    default table, package/src/PyBlastAfterglowMag/parfile_tools.py:361-426)."""
    os.makedirs(workdir, exist_ok=True)
    lines = []
    _section(lines, SEC_MAIN, MAIN_DEFAULTS, main)
    if grb is not None:
        _section(lines, SEC_GRB, GRB_DEFAULTS, grb)
    if kn is not None:
        _section(lines, SEC_KN, GRB_DEFAULTS, kn)
    with open(os.path.join(workdir, fname), "w") as f:
        for line in lines:
            f.write(line + "\n")
    return os.path.join(workdir, fname)


def read_parfile(path, section):
    """Return (pars: dict[str,float], opts: dict[str,str]) of one section (readParFile2 semantics:
    the first occurrence of a key wins; values are split at '#')."""
    keep_spaces = {"lc_freqs", "lc_times", "skymap_freqs", "skymap_times", "spec_times", "spec_freqs", "spec_gams"}
    pars, opts = {}, {}
    inside = reading_p = reading_o = False
    with open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            if line == section:
                inside = True
            if line == SEC_END:
                inside = False
            if not inside:
                continue
            if line == "* Parameters":
                reading_p, reading_o = True, False
            if line == "* Settings":
                reading_p, reading_o = False, True
            if line[:1] == "#":
                continue
            if reading_p and len(line) > 1 and line != "* Parameters":
                key, _, val = line.partition("=")
                key = "".join(key.split())
                val = "".join(val.split("#")[0].split())
                pars.setdefault(key, float(val))
            if reading_o and len(line) > 1 and line != "* Settings":
                key, _, val = line.partition("=")
                key = "".join(key.split())
                val = val.split("#")[0]
                val = val if key in keep_spaces else "".join(val.split())
                opts.setdefault(key, val)
    return pars, opts


def logspace_vec(start, stop, num, base=10.0):
    """Log grid by repeated multiplication; the last point stays below ``stop``
    (TOOLS::MakeLogspaceVec, src/utilitites/utils.h:91-105)."""
    out = np.empty(int(num))
    cur = base ** start
    ratio = base ** ((stop - start) / num)
    for i in range(int(num)):
        out[i] = cur
        cur *= ratio
    return out


def vec_from_string(s):
    """``array logspace a b n`` or ``array v1 v2 ...`` (makeVecFromString, utils.h:1540-1583)."""
    words = [w for w in s.split(" ")]
    if words and words[0] in ("", " "):
        words = words[1:]
    if words[0] != "array":
        raise ValueError(f"array line expected, got {s!r}")
    if words[1] == "logspace":
        return logspace_vec(np.log10(float(words[2])), np.log10(float(words[3])), int(float(words[4])))
    return np.array([float(w) for w in words[1:]])


def mom_from_gamma(g):
    g = np.asarray(g, dtype=np.float64)
    return g * np.sqrt((g - 1.0) * (g + 1.0)) / g


def beta_from_mom(m):
    m = np.asarray(m, dtype=np.float64)
    return m / np.sqrt(m * m + 1.0)


def jet_structure(struct, eiso_c, gamma0c, theta_c, theta_w=None, n_layers_a=21, n_layers_pw=81,
                  kind="adaptive"):
    """Per-layer initial data of a Gaussian or top-hat jet (JetStruct, id_analytic.py:86-206).

    Returns dict with r, theta, ctheta, mom, mass, ek (1-D over layers) and
    theta_core, theta_wing.
    """
    if struct == "tophat":
        theta_w = theta_c
        n_layers_a = 1
    if kind in ("adaptive", "a"):
        dth = theta_w / n_layers_a
        i = np.arange(n_layers_a, dtype=np.float64)
        th_c = i * dth + dth / 2.0
        th_h = (i + 1) * dth
        if struct == "gaussian":
            frac = 2 * np.sin(0.5 * th_h) * np.sin(0.5 * th_h)
            e0 = eiso_c * np.exp(-0.5 * (th_c * th_c / theta_c / theta_c))
            gam = 1.0 + (gamma0c - 1) * e0 / eiso_c
            mom = mom_from_gamma(gam)
            m0 = e0 / ((gam - 1.0) * C_LIGHT * C_LIGHT)
            e0 = e0 * (frac / 2.0)
            m0 = m0 * (frac / 2.0)
            mask = beta_from_mom(mom) > 1.e-5
            e0, mom, m0, th_c, th_h = e0[mask], mom[mask], m0[mask], th_c[mask], th_h[mask]
        elif struct == "tophat":
            frac = 2 * np.sin(0.5 * theta_c) * np.sin(0.5 * theta_c)
            e0 = np.array([eiso_c * frac / 2.0])
            mom = np.array([mom_from_gamma(gamma0c)])
            gam = np.sqrt(mom * mom + 1.0)
            m0 = eiso_c / ((gam - 1.0) * C_LIGHT * C_LIGHT) * frac / 2.0
        else:
            raise KeyError(struct)
        return dict(r=np.zeros_like(th_h), theta=th_h, ctheta=th_c, mom=mom, mass=m0, ek=e0,
                    theta_core=theta_c, theta_wing=theta_w)
    # piece-wise
    n = n_layers_pw
    th_pw = 2.0 * np.arcsin(np.arange(n + 1) / n * np.sin(theta_w / 2.0))
    cth = 0.5 * (th_pw[1:] + th_pw[:-1])
    cil = 2 * np.arange(n) + 1
    ncells = cil.sum()
    if struct == "gaussian":
        ang = 2.0 * np.pi * (2.0 * np.sin(0.5 * theta_w) * np.sin(0.5 * theta_w))
        e0 = eiso_c * ang / (4.0 * np.pi) * np.exp(-1. * cth * cth / (theta_c * theta_c))
        gam = 1. + (gamma0c - 1.) * np.exp(-1. * cth * cth / (2. * theta_c * theta_c))
        mom = mom_from_gamma(gam)
        m0 = e0 / (gam * C_LIGHT * C_LIGHT)
    else:
        one_min_cos = 2. * np.sin(0.5 * theta_w) * np.sin(0.5 * theta_w)
        ang = 2.0 * np.pi * one_min_cos / (4.0 * np.pi)
        e0 = np.full(n, eiso_c * ang)
        mom = np.full(n, mom_from_gamma(gamma0c))
        m0 = e0 / (gamma0c * C_LIGHT * C_LIGHT)
    e0 = e0 / ncells * cil
    m0 = m0 / ncells * cil
    return dict(r=np.zeros(n), theta=th_pw[1:], ctheta=cth, mom=mom, mass=m0, ek=e0,
                theta_core=theta_c, theta_wing=theta_w)


def write_run(workdir, main, grb, struct):
    """Write parfile.par and the initial-data file for a GRB run; returns parfile path."""
    from . import h5io
    os.makedirs(workdir, exist_ok=True)
    kind = "piece-wise" if grb.get("method_eats", "adaptive") == "piece-wise" else "adaptive"
    if struct.get("layer_ngam") is not None:
        raise ValueError("per-layer electron grids (layer_ngam) have no parfile form: the reference shares one "
                         "grid per run")
    idd = jet_structure(kind=kind, **{k: v for k, v in struct.items() if k != "layer_ngam"})
    fname = grb.get("fname_ejecta_id", "id_a.h5" if kind == "adaptive" else "id_pw.h5")
    grb = dict(grb, fname_ejecta_id=fname, use_1d_id="yes", method_eats=kind)
    h5io.write_id(os.path.join(workdir, fname), idd["r"], idd["mom"], idd["theta"], idd["ctheta"],
                  idd["ek"], idd["mass"], idd["theta_core"], idd["theta_wing"])
    return write_parfile(workdir, main, grb)


def ejecta_structure(n_shells, n_layers, mom_min, mom_max, ek_min, ek_max, seed, theta_wing=np.pi / 2.0):
    """Initial data of kilonova-like ejecta: [n_shells x n_layers] arrays of r, mom, theta, ctheta,
    ek, mass, in the layout EjectaID2::_load_id_file reads with use_1d_id = no
    (src/ejecta/ejecta_id.h:221-273).

    Shells have momenta beta*Gamma log-spaced over [mom_min, mom_max]; shell i carries a kinetic
    energy E_i drawn log-uniform over [ek_min, ek_max] (numpy default_rng(seed)).  The angular
    layers are those of an adaptive jet of half-opening theta_wing (layer j spans the cone up to
    (j + 1) theta_wing / n_layers): as for the jet's layers, a layer holds the energy of its
    whole cone, E_i (1 - cos theta_h) / 2, and mass E / ((Gamma - 1) c^2).  theta is each layer's
    upper edge (the loader takes theta_wing from the last one), ctheta its centre; r = 0 (the
    loader places every shell at beta c t0, load_r0 = no)."""
    mom = np.logspace(np.log10(mom_min), np.log10(mom_max), n_shells)
    e_shell = 10.0 ** np.random.default_rng(seed).uniform(np.log10(ek_min), np.log10(ek_max), n_shells)
    th_h = (np.arange(n_layers) + 1.0) * theta_wing / n_layers
    th_c = th_h - 0.5 * theta_wing / n_layers
    gam = np.sqrt(mom * mom + 1.0)
    ek = e_shell[:, None] * ((1.0 - np.cos(th_h)) / 2.0)[None, :]
    mass = ek / ((gam[:, None] - 1.0) * C_LIGHT * C_LIGHT)
    return dict(r=np.zeros_like(ek), mom=np.repeat(mom[:, None], n_layers, axis=1),
                theta=np.tile(th_h, (n_shells, 1)), ctheta=np.tile(th_c, (n_shells, 1)), ek=ek, mass=mass)


def write_kn_run(workdir, main, kn, struct):
    """Write parfile.par (main + kN sections, no jet) and the ejecta's 2-D initial-data file for a
    kilonova run; returns the parfile path.  struct: keyword arguments of ejecta_structure."""
    from . import h5io
    os.makedirs(workdir, exist_ok=True)
    idd = ejecta_structure(**struct)
    fname = kn.get("fname_ejecta_id", "id_kn.h5")
    kn = dict(kn, fname_ejecta_id=fname, use_1d_id="no")
    tw = float(idd["theta"][-1, -1])
    h5io.write_id(os.path.join(workdir, fname), idd["r"], idd["mom"], idd["theta"], idd["ctheta"], idd["ek"],
                  idd["mass"], tw, tw, one_d=False)
    return write_parfile(workdir, main, None, kn=kn)
