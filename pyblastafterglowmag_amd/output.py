"""HDF5 outputs in the reference program's layout, written from the GPU stages' buffers.

* light curves  ``Ejecta::saveDenseLightCurves``     (src/ejecta/ejecta.h:511-573)
* dynamics      ``Ejecta::saveEjectaBWsDynamics``    (src/ejecta/ejecta.h:165-222)
* spectra       ``Ejecta::computeSaveEjectaSpectrum`` (src/ejecta/ejecta.h:354-509): the electron
  spectrum (``n_ele_*``) and the total comoving emissivity (``total_j_*``) per step, with the
  ``times_freqs/freqs/times_gams/gams`` grids; the GPU path does not keep the separate
  synchrotron/SSC parts, intensities or cooling rates, so those datasets are not written.

Group names are ``"shell=0 layer=<i>"`` for a jet's layers and ``"shell=<i> layer=<j>"`` for kilonova
ejecta (``write_lightcurves_ejecta``, ``write_dynamics_ejecta``); file attributes are nshells, nlayers (and ntimes,
ncells for dynamics/spectra; ncells is the piece-wise cell count n^2 for either EATS method, as
EjectaID2 sets it, ejecta_id.h:180-181) plus every numeric parameter of the main and GRB parfile
sections.
"""
import math
import os

import numpy as np

from . import abi, h5io

C_LIGHT = 2.99792458e10
M_PROTON = 1.6726e-24

# BW::VARS for the forward-shock and forward+reverse-shock blastwave types, in the reference's
# order (src/blastwave/blastwave_pars.h:255-274), with their BW::VARNAMES (:230-250).
DYN_FS = ["R", "Rsh", "rho", "drhodr", "GammaRel", "Gamma", "beta", "mom", "Eint2", "U_p", "theta", "ctheta",
          "Erad2", "Esh2", "Ead2", "M2", "tcomov", "tburst", "tt", "thickness", "adi", "rho2", "GammaFsh",
          "gamma_min", "gamma_c", "gamma_max", "B", "ThetaTherm", "z_cool", "x", "nprime", "accel_frac"]
DYN_FSRS = (DYN_FS[:23] + ["GammaRsh", "Rsh_rs"] + DYN_FS[23:] +
            ["Eint3", "Erad3", "iEad3", "Esh3", "M3", "deltaR4", "Gamma43", "adi3", "rho4", "rho3",
             "thichness_rs", "U_p3", "gamma_min_rs", "gamma_c_rs", "gamma_max_rs", "B_rs", "ThetaTherm_rs",
             "z_cool_rs", "x_rs", "nprime_rs", "accel_frac_rs"])


def file_attrs(mp_, gp, extra):
    out = dict(extra)
    for d in (mp_, gp):
        for k, v in d.items():
            out.setdefault(k, float(v))
    return out


def _rho_at(P, R):
    """Ambient density at R (RhoISM::evaluateRhoDrhoDrDefault, src/blastwave/blastwave_components.h:846-893)."""
    mppme = M_PROTON + 9.1094e-28
    if P.n_ism > 0:
        return P.n_ism * mppme
    if R < P.ism_r_ej:
        return P.ism_A0 * P.ism_r_ej ** (-P.ism_s) * mppme
    if R < P.ism_r_ism:
        return P.ism_A0 * R ** (-P.ism_s) * mppme
    return P.ism_A0 * P.ism_r_ism ** (-P.ism_s) * mppme


def bw_attrs(P, li, tb0, il, dyn=False):
    """Attributes of one blastwave's group (BlastWave Pars: E0, M0, Gamma0, ..., set in
    blastwave.h:73-241; Rd from the density at the start radius); il is its layer."""
    g0 = math.sqrt(li.mom0 * li.mom0 + 1.0)
    beta0 = li.mom0 / g0
    M0 = li.M0 if li.M0 > 0 else li.E0 / (g0 * C_LIGHT * C_LIGHT)
    r0 = beta0 * C_LIGHT * tb0
    rho = _rho_at(P, r0)
    rd = (3. / (4. * math.pi) * 1. / (C_LIGHT * C_LIGHT * M_PROTON) * li.E0 / (rho / M_PROTON * g0 * g0)) ** (1. / 3.)
    out = {"E0": li.E0, "M0": M0, "Gamma0": g0, "beta0": beta0, "mom0": li.mom0, "R0": r0,
           "Eint0": 0.0, "Rd": rd, "t0": tb0, "theta_b0": li.theta_b0, "theta_a": 0.0}
    if dyn:
        out.update({"theta_c": li.theta_c, "theta_c_l": li.theta_c_l, "theta_c_h": li.theta_c_h,
                    "cil": float(abi.cells_in_layer(il))})
    return out


def layer_attrs(run, il, dyn=False):
    """Per-layer attributes of a jet's layer il."""
    return bw_attrs(run.params, run.inits[il], float(run.tgrid[0]), il, dyn)


def write_lightcurves(path, run, flux, mp_, gp):
    """flux: [n_layers, n_pairs] of one run, mJy."""
    nl = run.n_layers
    with h5io.Writer(path) as w:
        for il in range(nl):
            g = f"shell=0 layer={il}"
            w.dataset(g, "fluxdens", flux[il])
            for k, v in layer_attrs(run, il).items():
                w.attr(g, k, v)
        w.dataset("/", "times", run.tobs)
        w.dataset("/", "freqs", run.nuobs)
        for k, v in file_attrs(mp_, gp, {"nshells": 1.0, "nlayers": float(nl)}).items():
            w.attr("/", k, v)


def write_dynamics(path, run, table, fsrs, mp_, gp):
    """table: [n_layers, PBA_NV, nt]; fsrs: per-layer flags from the dynamics results."""
    nl, _, nt = table.shape
    with h5io.Writer(path) as w:
        for il in range(nl):
            g = f"shell=0 layer={il}"
            for name in (DYN_FSRS if fsrs[il] else DYN_FS):
                w.dataset(g, name, table[il, abi.PV[name]] if name in abi.PV else np.zeros(nt))
            for k, v in layer_attrs(run, il, dyn=True).items():
                w.attr(g, k, v)
        extra = {"nshells": 1.0, "nlayers": float(nl), "ntimes": float(nt), "ncells": float(abi.total_cells(nl))}
        for k, v in file_attrs(mp_, gp, extra).items():
            w.attr("/", k, v)


# BW::VARS of the ejecta's blastwave type (iFS_DENSE, src/blastwave/blastwave_pars.h:274-282): the
# forward-shock set and the upstream's state
DYN_DENSE = DYN_FS + ["GammaRho", "dGammaRhodr", "dGammaReldGamma", "PCBM", "dPCBMdrho", "MCBM", "CSCBM"]
# their values in a standard ISM at rest (set_standard_ism, blastwave.h:1282-1295); the swept-up
# mass of the upstream is unset (-1, RhoISM, blastwave_components.h:905) until the first RHS call
_DENSE_UPSTREAM = {"GammaRho": 1.0, "dGammaRhodr": 0.0, "dGammaReldGamma": 1.0, "PCBM": 0.0, "dPCBMdrho": 0.0,
                   "MCBM": 0.0, "CSCBM": 0.0}


def _ejecta_groups(order):
    """(k, group name) of every ejecta blastwave in the reference's file order (shells outer,
    layers inner, src/ejecta/ejecta.h:179-181, 538-541); k indexes the blastwave in `order`."""
    return [(k, f"shell={ish} layer={il}") for k, (ish, il) in sorted(enumerate(order), key=lambda x: x[1])]


def write_lightcurves_ejecta(path, P, inits, order, tgrid, tobs, nuobs, flux, mp_, kp):
    """Light curves of kilonova ejecta (Ejecta::saveDenseLightCurves, src/ejecta/ejecta.h:511-573):
    one group per (shell, layer) blastwave; flux [n_bw, n_pairs] in the order of `order`."""
    ns, nl = 1 + max(o[0] for o in order), 1 + max(o[1] for o in order)
    tb0 = float(tgrid[0])
    with h5io.Writer(path) as w:
        for k, g in _ejecta_groups(order):
            w.dataset(g, "fluxdens", flux[k])
            for key, v in bw_attrs(P, inits[k], tb0, order[k][1]).items():
                w.attr(g, key, v)
        w.dataset("/", "times", tobs)
        w.dataset("/", "freqs", nuobs)
        for key, v in file_attrs(mp_, kp, {"nshells": float(ns), "nlayers": float(nl)}).items():
            w.attr("/", key, v)


def write_dynamics_ejecta(path, P, inits, order, tgrid, table, mp_, kp):
    """Dynamics of kilonova ejecta (Ejecta::saveEjectaBWsDynamics, src/ejecta/ejecta.h:165-222):
    one group per (shell, layer) blastwave with the ejecta type's datasets; table [n_bw, PBA_NV, nt]
    in the order of `order`.  The upstream columns are written where the blastwave was evolved."""
    ns, nl = 1 + max(o[0] for o in order), 1 + max(o[1] for o in order)
    nt = table.shape[2]
    tb0 = float(tgrid[0])
    with h5io.Writer(path) as w:
        for k, g in _ejecta_groups(order):
            T = table[k]
            live = T[abi.PV["R"]] > 0
            for name in DYN_DENSE:
                if name in abi.PV:
                    data = T[abi.PV[name]]
                elif name in _DENSE_UPSTREAM:
                    data = np.where(live, _DENSE_UPSTREAM[name], 0.0)
                    if name == "MCBM" and live[0]:
                        data[0] = -1.0
                else:
                    data = np.zeros(nt)
                w.dataset(g, name, data)
            for key, v in bw_attrs(P, inits[k], tb0, order[k][1], dyn=True).items():
                w.attr(g, key, v)
        extra = {"nshells": float(ns), "nlayers": float(nl), "ntimes": float(nt), "ncells": float(abi.total_cells(nl))}
        for key, v in file_attrs(mp_, kp, extra).items():
            w.attr("/", key, v)


def write_spectra(path, run, pops, jall, fall, ele_grid, nu_grid, mp_, gp, spectra=None, aall=None):
    """spec.h5 as Ejecta::computeSaveEjectaSpectrum writes it (src/ejecta/ejecta.h:354-509).

    jall: [n_pop, nt, nf] total comoving emissivity per electron; fall: [n_pop, nt, ng] (n_ele);
    spectra: None or the dict of pbagpu_ele_evolve_spectra's arrays (syn_j, syn_i, ssc_j, total_f,
    total_i [n_pop, nt, nf]; gam_dot_syn, gam_dot_adi, gam_dot_ssc [n_pop, nt, ng]);
    aall: [n_pop, nt, nf] absorption (use_ssa = yes) or None.
    Per population (forward or reverse shock of a layer) the reference writes n_ele and gam_dot_*,
    syn_f / syn_j / syn_i (+ syn_a with SSA), ssc_f / ssc_j / ssc_i with SSC and total_f / total_j /
    total_i (+ total_a with SSA); syn_f and ssc_f are never filled there (zeros), ssc_i equals ssc_j
    (no SSC absorption without pair production) and syn_a, total_a are the absorption table."""
    nl = run.n_layers
    nt = len(run.tgrid)
    with h5io.Writer(path) as w:
        for q, P in enumerate(pops):
            g = f"shell=0 layer={P.bw}"
            tag = "fs" if P.shock == 0 else "rs"
            sh = run.params.fs if P.shock == 0 else run.params.rs
            if fall is not None:
                w.dataset(g, f"n_ele_{tag}", fall[q])
            if spectra is None:
                w.dataset(g, f"total_j_{tag}", jall[q])
                continue
            for k in ("gam_dot_syn", "gam_dot_adi", "gam_dot_ssc"):
                w.dataset(g, f"{k}_{tag}", spectra[k][q])
            zeros = np.zeros_like(jall[q])
            w.dataset(g, f"syn_f_{tag}", zeros)
            w.dataset(g, f"syn_j_{tag}", spectra["syn_j"][q])
            w.dataset(g, f"syn_i_{tag}", spectra["syn_i"][q])
            if sh.ssa:
                w.dataset(g, f"syn_a_{tag}", aall[q])
            if sh.ssc:
                w.dataset(g, f"ssc_f_{tag}", zeros)
                w.dataset(g, f"ssc_j_{tag}", spectra["ssc_j"][q])
                w.dataset(g, f"ssc_i_{tag}", spectra["ssc_j"][q])
            w.dataset(g, f"total_f_{tag}", spectra["total_f"][q])
            w.dataset(g, f"total_j_{tag}", jall[q])
            w.dataset(g, f"total_i_{tag}", spectra["total_i"][q])
            if sh.ssa:
                w.dataset(g, f"total_a_{tag}", aall[q])
        w.dataset("/", "times_freqs", np.repeat(run.tgrid, len(nu_grid)))
        w.dataset("/", "freqs", np.tile(nu_grid, nt))
        w.dataset("/", "times_gams", np.repeat(run.tgrid, len(ele_grid)))
        w.dataset("/", "gams", np.tile(ele_grid, nt))
        extra = {"nshells": 1.0, "nlayers": float(nl), "ntimes": float(nt), "ncells": float(abi.total_cells(nl))}
        for k, v in file_attrs(mp_, gp, extra).items():
            w.attr("/", k, v)


def ele_grid(shock_params):
    """Chang-Cooper Lorentz-factor grid e_i = step^i, step = exp(ln(gam2)/n)
    (Electrons::build_grid_chang_cooper, base_state.h:217-253)."""
    n = int(shock_params.ngam)
    step = math.exp(math.log(shock_params.gam2) / n)
    return np.array([step ** i for i in range(n)])


def freq_grid(shock_params):
    """Comoving frequency grid by repeated multiplication (MakeLogspaceVec, utils.h:91-105)."""
    from .inputs import logspace_vec
    return logspace_vec(math.log10(shock_params.freq1), math.log10(shock_params.freq2), int(shock_params.nfreq))


DBL_MAX = np.finfo(np.float64).max
DBL_MIN = np.finfo(np.float64).tiny   # std::numeric_limits<double>::min()


def skymap_images(image, flux, nlayers):
    """Normalised and cleaned raw sky maps, as the reference post-processes evalSkyMapPW's arrays.

    image: [n_images][PBA_IMG_NV][2 ncells] raw arrays of one shell (pbagpu_eats_skymap_pw);
    flux:  [n_layers][n_images] per-layer flux densities, mJy.
    Returns one dict per image with the cleaned arrays (IMG::m_names), the shell flux and the
    attributes saveRawImage writes.

    * Ejecta::computeEjectaSkyMapPW (src/ejecta/ejecta_base.h:471-545): the shell flux is the sum
      of the layers' fluxes; a positive flux scales the intensities by flux / (dx dy) / max(I),
      with dx, dy the extent of the cells' sky positions (cells that did not contribute sit at 0
      and count), and sets the image extent.
    * ImageExtend::removeZeroEntries (src/image.h:120-190): unless one jet has every cell above
      1e-10 of the peak intensity, only cells above it in either jet are kept (both jets' entries
      of those cells).
    """
    out = []
    n_img = image.shape[0]
    for k in range(n_img):
        data = np.array(image[k], dtype=np.float64)
        fl_shell = 0.0
        for il in range(flux.shape[0]):
            fl_shell += float(flux[il, k])
        total = 0.0
        bounds = dict(xmin=DBL_MAX, ymin=DBL_MAX, xmax=DBL_MIN, ymax=DBL_MIN)
        if fl_shell <= 0 or not math.isfinite(fl_shell):
            total += fl_shell
        else:
            X, Y, I = data[abi.IMG["xrs"]], data[abi.IMG["yrs"]], data[abi.IMG["intensity"]]
            xmin, xmax, ymin, ymax, imax = DBL_MAX, DBL_MIN, DBL_MAX, DBL_MIN, DBL_MIN
            for i in range(len(I)):   # the reference's sequential scan (strict comparisons)
                if X[i] < xmin: xmin = X[i]
                if X[i] > xmax: xmax = X[i]
                if Y[i] < ymin: ymin = Y[i]
                if Y[i] > ymax: ymax = Y[i]
                if I[i] > imax: imax = I[i]
            if imax <= 0 or not math.isfinite(imax):
                raise RuntimeError(f"sky map {k}: maximum intensity {imax}")
            dfnu = fl_shell / ((xmax - xmin) * (ymax - ymin))
            data[abi.IMG["intensity"]] = I * (dfnu / imax)
            if xmin < bounds["xmin"]: bounds["xmin"] = xmin
            if ymin < bounds["ymin"]: bounds["ymin"] = ymin
            if xmax > bounds["xmax"]: bounds["xmax"] = xmax
            if ymax > bounds["ymax"]: bounds["ymax"] = ymax
            total += fl_shell
        # removeZeroEntries
        nc = data.shape[1] // 2
        I = data[abi.IMG["intensity"]]
        max_pj = max(DBL_MIN, float(np.max(I[:nc]))) if nc else DBL_MIN
        max_cj = max(DBL_MIN, float(np.max(I[nc:]))) if nc else DBL_MIN
        thr = max(max_pj, max_cj) * 1.e-10
        nz_pj = int(np.count_nonzero(I[:nc] > thr))
        nz_cj = int(np.count_nonzero(I[nc:] > thr))
        if not (nz_pj == nc or nz_cj == nc):
            keep = np.nonzero((I[:nc] > thr) | (I[nc:] > thr))[0]
            data = np.concatenate([data[:, keep], data[:, nc + keep]], axis=1)
        out.append(dict(arrays={n: data[i] for i, n in enumerate(abi.IMG_NAMES)}, fluxs=np.array([fl_shell]),
                        attrs=dict(flux=total, nshells=1.0 if data.shape[1] else 0.0, nlayers=float(nlayers),
                                   **bounds)))
    return out


def write_raw_skymaps(workdir, images, times, freqs, mp_, gp):
    """raw_skymap_<i>.h5 files (saveRawImage, src/image.h:250-300): group "shell=0" with the
    IMG::m_names arrays, dataset "fluxs", attributes flux, nshells, nlayers, time, freq and
    the image extent plus every numeric parfile parameter.  Image i is frequency-major:
    i = ifreq * len(times) + it (Ejecta::computeSaveEjectaSkyImages, src/ejecta/ejecta.h:259-312)."""
    paths = []
    for i, im in enumerate(images):
        path = os.path.join(workdir, f"raw_skymap_{i}.h5")
        with h5io.Writer(path) as w:
            if im["attrs"]["nshells"] > 0:
                for n in abi.IMG_NAMES:
                    w.dataset("shell=0", n, im["arrays"][n])
            w.dataset("/", "fluxs", im["fluxs"])
            attrs = dict(im["attrs"], time=float(times[i % len(times)]), freq=float(freqs[i // len(times)]))
            for d in (mp_, gp):   # parfile parameters override, as in the reference's map
                for k, v in d.items():
                    attrs[k] = float(v)
            for k, v in attrs.items():
                w.attr("/", k, v)
        paths.append(path)
    return paths
