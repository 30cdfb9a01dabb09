"""Write the golden fixtures of the parity tests from the reference CPU program.

Runs oracle/_ref/pba.out (built from /root/reference/src by oracle/build_ref.sh) on the
parfile + initial-data files written by pyblastafterglowmag_amd.inputs, then stores the
reference's outputs (lc.h5 light curves, dyn.h5 dynamics, spec.h5 comoving spectra) as
compressed .npz next to this script.  Only data is stored (inputs and outputs).

    python tests/golden/make_golden.py [name ...]      (no name: every fixture, ALL)

gauss21_fsrs_ssc (the full 21-layer jet, 200 x 96 grids, 12,288 pairs) takes about 25 minutes of
one core.  The kilonova cases (kn_*: ejecta only, written by inputs.write_kn_run) store every
(shell, layer) blastwave's dynamics and light curve, or a subset for the full-size ejecta:
kn_c5_dyn (the C5 parfile, dynamics only) takes about 1 minute, kn_c5_lc (the C5 ejecta at
reduced microphysics grids with light curves) about 6 minutes.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from pyblastafterglowmag_amd import abi, h5io, inputs, workloads  # noqa: E402

PBA = os.path.join(ROOT, "oracle", "_ref", "pba.out")
# the reference's EBL optical-depth table (Franceschini & Rodighiero 2018), for the cases that
# attenuate their light curves; the fixture keeps the table's arrays beside the outputs
EBL_TABLE = "/root/reference/data/EBL/Franceschini18/table.h5"
EBL_CASES = {"tophat_ebl_small"}
DYN_VARS = ["R", "Gamma", "beta", "theta", "M2", "Eint2", "tt", "tcomov", "rho2", "GammaFsh", "thickness", "U_p",
            "gamma_min", "gamma_max", "gamma_c", "B", "accel_frac", "Rsh"]
DYN_VARS_RS = ["M3", "Eint3", "deltaR4", "Gamma43", "rho3", "GammaRsh", "thichness_rs", "U_p3", "gamma_min_rs",
               "gamma_max_rs", "gamma_c_rs", "B_rs", "accel_frac_rs"]


def build_case(name):
    if name in workloads.SMALL:
        return workloads.small(name)
    if name.startswith("lhs_small_"):   # a member of the Latin-hypercube batch at reduced grids
        return workloads.lhs_small(int(name.rsplit("_", 1)[1]))
    return workloads.WORKLOADS[name]()


# spec.h5 datasets beside n_ele / total_j (src/ejecta/ejecta.h:379-440), stored for SPEC_FULL
SPEC_DATASETS = ["gam_dot_syn", "gam_dot_adi", "gam_dot_ssc", "syn_f", "syn_j", "syn_i", "syn_a", "ssc_f", "ssc_j",
                 "ssc_i", "total_f", "total_i", "total_a"]
SPEC_FULL = {"tophat_small": 3, "gauss5_small": 20, "tophat_ssa_small": 3}   # name: step stride

SKY_ATTRS = ["flux", "nshells", "nlayers", "time", "freq", "xmin", "ymin", "xmax", "ymax"]


# light curves of every `stride`-th layer, dynamics of layer 0, no spectra or sky maps: the full-size
# case and the uniform-grid references of the per-layer-grid jet
LARGE = {"tophat64_pw": 7, "gauss_pw_ng32": 1, "gauss_pw_ng48": 1, "gauss21_fsrs_ssc": 4, "gauss64_pw_small": 8,
         "lhs_small_0": 4, "lhs_small_42": 4}
# large cases that keep the dynamics of every stored light-curve layer (not only layer 0): the
# reverse shock's end per layer decides which layers a test holds to the tight bound
LARGE_DYN_ALL = {"gauss21_fsrs_ssc", "gauss64_pw_small", "lhs_small_0", "lhs_small_42"}
# dynamics kept at every n-th step only (dyn_every in the fixture)
DYN_EVERY = {"lhs_small_0": 10, "lhs_small_42": 10}
# large piece-wise cases that keep their raw sky maps
LARGE_SKY = {"gauss64_pw_small"}


def run_reference(name, wd, spec_every=1):
    main, grb, struct = build_case(name)
    large = name in LARGE
    grb = dict(grb, save_dynamics="yes", save_spec="no" if large else "yes")
    if name in EBL_CASES:
        grb["ebl_tbl_fpath"] = EBL_TABLE
    sky = grb.get("method_eats") == "piece-wise" and (not large or name in LARGE_SKY)
    if sky:   # raw sky maps at 2 frequencies x 2 observer times
        main = dict(main, skymap_times="array 1e5 1e7", skymap_freqs="array 1e9 1e18")
        grb = dict(grb, do_skymap="yes", save_raw_skymap="yes")
    inputs.write_run(wd, main, grb, struct)
    subprocess.check_call([PBA, wd + "/", "parfile.par", "err"], stdout=subprocess.DEVNULL)
    P, mp_, mo, gp, go = abi.run_params_from_parfile(os.path.join(wd, "parfile.par"))
    idf = os.path.join(wd, go["fname_ejecta_id"])
    out = {}
    with open(os.path.join(wd, "parfile.par")) as f:
        out["parfile"] = np.frombuffer(f.read().encode(), dtype=np.uint8)
    if name in EBL_CASES:
        for k in ("freq", "z", "tau"):
            out["ebl_" + k] = h5io.read(EBL_TABLE, k)
    for k in ("mom", "ctheta", "ek", "mass"):
        out["id_" + k] = h5io.read(idf, k)
    out["id_theta_wing"] = np.array(h5io.attr(idf, "theta_wing"))
    out["id_theta_core"] = np.array(h5io.attr(idf, "theta_core"))
    nl = len(out["id_mom"])
    lc = os.path.join(wd, "lc.h5")
    out["times"] = h5io.read(lc, "times")
    out["freqs"] = h5io.read(lc, "freqs")
    lc_layers = list(range(0, nl, LARGE[name])) if large else list(range(nl))
    out["lc_layers"] = np.array(lc_layers)
    out["lc"] = np.stack([h5io.read(lc, f"shell=0 layer={il}/fluxdens") for il in lc_layers])
    # file layout of the reference's outputs (names only) and the per-layer light-curve attributes
    out["lc_root_members"] = np.array(h5io.members(lc, "/"))
    out["lc_root_attrs"] = np.array(h5io.attr_names(lc, "/"))
    names_la = h5io.attr_names(lc, "shell=0 layer=0")
    out["lc_layer_attr_names"] = np.array(names_la)
    out["lc_layer_attrs"] = np.array([[h5io.attr(lc, k, f"shell=0 layer={il}") for k in names_la] for il in lc_layers])
    dyn = os.path.join(wd, "dyn.h5")
    names = DYN_VARS + (DYN_VARS_RS if P.fsrs else [])
    nt = int(mp_["ntb"])
    dyn_layers = (lc_layers if name in LARGE_DYN_ALL else [0]) if large else list(range(nl))
    n_dyn = len(dyn_layers)
    D = np.zeros((n_dyn, len(names), nt))
    for i, il in enumerate(dyn_layers):
        for iv, v in enumerate(names):
            ds = f"shell=0 layer={il}/{v}"
            if h5io.has(dyn, ds):
                D[i, iv] = h5io.read(dyn, ds)
    every = DYN_EVERY.get(name, 1)
    out["dyn"] = np.ascontiguousarray(D[:, :, ::every])
    out["dyn_every"] = np.array(every)
    out["dyn_layers"] = np.array(dyn_layers)
    out["dyn_names"] = np.array(names)
    out["dyn_layer_members"] = np.array(h5io.members(dyn, "shell=0 layer=0"))
    names_da = h5io.attr_names(dyn, "shell=0 layer=0")
    out["dyn_layer_attr_names"] = np.array(names_da)
    out["dyn_layer_attrs"] = np.array([[h5io.attr(dyn, k, f"shell=0 layer={il}") for k in names_da]
                                       for il in dyn_layers])
    out["dyn_root_attrs"] = np.array(h5io.attr_names(dyn, "/"))
    out["dyn_root_counts"] = np.array([h5io.attr(dyn, k) for k in ("nshells", "nlayers", "ntimes", "ncells")])
    if large:
        out["spec_every"] = np.array(spec_every)
        if sky:
            _store_sky(out, wd)
        return out
    spec = os.path.join(wd, "spec.h5")
    for tag in ("fs", "rs"):
        J, F = [], []
        for il in range(nl):
            ds = f"shell=0 layer={il}/total_j_{tag}"
            if h5io.has(spec, ds):
                J.append(h5io.read(spec, ds).reshape(nt, -1)[::spec_every])
                F.append(h5io.read(spec, f"shell=0 layer={il}/n_ele_{tag}").reshape(nt, -1)[::spec_every])
            else:
                J.append(None)
        if any(j is not None for j in J):
            ok = [i for i, j in enumerate(J) if j is not None]
            out[f"spec_layers_{tag}"] = np.array(ok)
            out[f"jall_{tag}"] = np.stack([J[i] for i in ok])
            out[f"fall_{tag}"] = np.stack(F)
    out["spec_every"] = np.array(spec_every)
    # every other dataset of spec.h5 (per shock, every SPEC_FULL[name]-th step), for the fixtures
    # that pin the spectra writer
    if name in SPEC_FULL:
        full_every = SPEC_FULL[name]
        out["spec_full_every"] = np.array(full_every)
        for tag in ("fs", "rs"):
            ok = [il for il in range(nl) if h5io.has(spec, f"shell=0 layer={il}/total_j_{tag}")]
            for ds in SPEC_DATASETS:
                rows = [h5io.read(spec, f"shell=0 layer={il}/{ds}_{tag}").reshape(nt, -1)[::full_every]
                        for il in ok if h5io.has(spec, f"shell=0 layer={il}/{ds}_{tag}")]
                if rows:
                    out[f"spec_{ds}_{tag}"] = np.stack(rows)
    if sky:
        _store_sky(out, wd)
    out["spec_root_members"] = np.array(h5io.members(spec, "/"))
    out["spec_layer_members"] = np.array(h5io.members(spec, "shell=0 layer=0"))
    return out


# kilonova ejecta (parfile kN section, no jet): every (shell, layer) blastwave's dynamics and light
# curve; KN_DYN_ONLY runs the dynamics alone (no electrons or light curves) and keeps the
# blastwaves of KN_DYN_KEEP, every KN_DYN_EVERY-th step
KN_CASES = {"kn_small", "kn_pw_small", "kn_spread_small", "kn_wind_small", "kn_rk4_small", "kn_c5_dyn",
            "kn_c5_lc"}
KN_DYN_ONLY = {"kn_c5_dyn"}
KN_DYN_KEEP = {"kn_c5_dyn": ((0, 10, 21, 31), (0, 14, 29)), "kn_c5_lc": ((0, 10, 21, 31), (0, 14, 29))}
KN_DYN_EVERY = {"kn_c5_dyn": 10, "kn_c5_lc": 10}
KN_DYN_VARS = DYN_VARS + ["GammaRel", "rho", "drhodr", "ctheta", "Erad2", "Esh2", "Ead2", "GammaRho", "dGammaRhodr",
                           "dGammaReldGamma", "PCBM", "dPCBMdrho", "MCBM", "CSCBM"]
KN_DYN_VARS_ONLY = ["R", "Gamma", "beta", "theta", "M2", "Eint2", "tt", "tcomov", "rho2", "GammaFsh", "thickness",
                    "U_p", "GammaRel", "rho", "Esh2", "Ead2"]


def build_kn_case(name):
    if name == "kn_c5_dyn":
        main, kn, struct = workloads.kn_c5()
        return main, dict(kn, do_mphys_in_situ="no", do_mphys_in_ppr="no", do_lc="no"), struct
    if name == "kn_c5_lc":   # the C5 ejecta at reduced microphysics grids, without SSC: 6 minutes
        return workloads.kn_scaled("kn_c5", ngam=64, nfreq=48, n_nu_obs=6, n_t_obs=8, method_ssc_fs="none")
    return workloads.kn_small(name)


def run_reference_kn(name, wd):
    main, kn, struct = build_kn_case(name)
    kn = dict(kn, save_dynamics="yes", save_spec="no")
    inputs.write_kn_run(wd, main, kn, struct)
    subprocess.check_call([PBA, wd + "/", "parfile.par", "err"], stdout=subprocess.DEVNULL)
    out = {}
    with open(os.path.join(wd, "parfile.par")) as f:
        out["parfile"] = np.frombuffer(f.read().encode(), dtype=np.uint8)
    idf = os.path.join(wd, kn["fname_ejecta_id"])
    dyn = os.path.join(wd, kn["fname_dyn"])
    ns, nl = int(h5io.attr(dyn, "nshells")), int(h5io.attr(dyn, "nlayers"))
    for k in ("r", "mom", "theta", "ctheta", "ek", "mass"):
        out["id_" + k] = h5io.read(idf, k).reshape(ns, nl)
    shells, layers = KN_DYN_KEEP.get(name, (tuple(range(ns)), tuple(range(nl))))
    names = KN_DYN_VARS_ONLY if name in KN_DYN_ONLY else KN_DYN_VARS
    every = KN_DYN_EVERY.get(name, 1)
    nt = int(main["ntb"])
    D = np.zeros((len(shells), len(layers), len(names), nt))
    for i, ish in enumerate(shells):
        for j, il in enumerate(layers):
            for iv, v in enumerate(names):
                D[i, j, iv] = h5io.read(dyn, f"shell={ish} layer={il}/{v}")
    out["dyn"] = np.ascontiguousarray(D[..., ::every])
    out["dyn_every"] = np.array(every)
    out["dyn_shells"] = np.array(shells)
    out["dyn_layers"] = np.array(layers)
    out["dyn_names"] = np.array(names)
    out["dyn_layer_members"] = np.array(h5io.members(dyn, "shell=0 layer=0"))
    out["dyn_root_counts"] = np.array([h5io.attr(dyn, k) for k in ("nshells", "nlayers", "ntimes", "ncells")])
    out["dyn_root_attrs"] = np.array(h5io.attr_names(dyn, "/"))
    names_da = h5io.attr_names(dyn, "shell=0 layer=0")
    out["dyn_layer_attr_names"] = np.array(names_da)
    out["dyn_layer_attrs"] = np.array([[[h5io.attr(dyn, k, f"shell={ish} layer={il}") for k in names_da]
                                        for il in layers] for ish in shells])
    if name in KN_DYN_ONLY:
        return out
    lc = os.path.join(wd, kn["fname_light_curve"])
    out["times"] = h5io.read(lc, "times")
    out["freqs"] = h5io.read(lc, "freqs")
    out["lc"] = np.stack([np.stack([h5io.read(lc, f"shell={ish} layer={il}/fluxdens") for il in layers])
                          for ish in shells])
    out["lc_root_members"] = np.array(h5io.members(lc, "/"))
    out["lc_root_attrs"] = np.array(h5io.attr_names(lc, "/"))
    names_la = h5io.attr_names(lc, "shell=0 layer=0")
    out["lc_layer_attr_names"] = np.array(names_la)
    out["lc_layer_attrs"] = np.array([[[h5io.attr(lc, k, f"shell={ish} layer={il}") for k in names_la]
                                       for il in layers] for ish in shells])
    return out


def _store_sky(out, wd):
    """The raw sky maps raw_skymap_<n>.h5 of a run, every dataset and attribute."""
    n = 0
    while os.path.exists(os.path.join(wd, f"raw_skymap_{n}.h5")):
        f = os.path.join(wd, f"raw_skymap_{n}.h5")
        for v in h5io.members(f, "shell=0"):
            out[f"sky{n}_{v}"] = h5io.read(f, f"shell=0/{v}")
        out[f"sky{n}_fluxs"] = h5io.read(f, "fluxs")
        out[f"sky{n}_attrs"] = np.array([h5io.attr(f, k) for k in SKY_ATTRS])
        n += 1
    out["sky_n"] = np.array(n)
    out["sky_root_members"] = np.array(h5io.members(os.path.join(wd, "raw_skymap_0.h5"), "/"))
    out["sky_root_attrs"] = np.array(h5io.attr_names(os.path.join(wd, "raw_skymap_0.h5"), "/"))


def main(names):
    if not os.path.exists(PBA):
        raise SystemExit(f"reference program missing: {PBA} (run oracle/build_ref.sh)")
    for name in names:
        wd = tempfile.mkdtemp(prefix=f"golden_{name}_")
        try:
            every = {"tophat_small": 1, "gauss5_small": 5}.get(name, 10)
            if name in KN_CASES:
                out = run_reference_kn(name, wd)
            else:
                out = run_reference(name, wd, spec_every=every)
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
            print(name, {k: v.shape for k, v in out.items()})
        finally:
            shutil.rmtree(wd, ignore_errors=True)


# every fixture the tests load
ALL = ["tophat_small", "gauss5_small", "tophat_fs_syn", "tophat_pw_small", "gauss_pw_small", "gauss_pw_default",
       "tophat64_pw", "gauss_pw_ng32", "gauss_pw_ng48",
       "tophat_ssa_small", "gauss5_ssa_small", "tophat_dermer_small", "tophat_bessel_small", "tophat_gsl_small",
       "tophat_rk4_small", "tophat_wind_small", "gauss21_fsrs_ssc", "gauss64_pw_small", "tophat_ebl_small",
       "lhs_small_0", "lhs_small_42", "tophat_f2t_small", "kn_small", "kn_pw_small", "kn_spread_small",
       "kn_wind_small", "kn_rk4_small", "kn_c5_dyn", "kn_c5_lc"]

if __name__ == "__main__":
    main(sys.argv[1:] or ALL)
