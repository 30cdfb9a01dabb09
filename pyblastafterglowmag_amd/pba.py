"""Drop-in driver: ``python -m pyblastafterglowmag_amd.pba <workdir> <parfile> [loglevel] [--precision N]``.

Mirrors the reference CLI ``pba.out <workdir> <parfile> <loglevel>`` (src/main.cpp:208-384) for the
GRB afterglow part of a parfile, or for its kilonova ejecta (kN section, without a jet: ``run_ejecta``): it reads the same parfile and initial-data file, runs the three
stages on the GPU (dynamics, electrons + emissivity, EATS light curves) and writes the files the
reference writes for ``do_lc`` / ``save_dynamics`` / ``save_spec`` / ``do_skymap`` (``fname_light_curve``,
``fname_dyn``, ``fname_spectrum``, ``raw_skymap_<i>.h5``) in the same layout (see ``output``).  Errors end the process
with status 1 after a message, as the reference does.

``--process-skymaps`` then does what the reference package's ``PyBlastAfterglow.run(process_skymaps=True)`` does
after the executable (package/src/PyBlastAfterglowMag/interface.py:1052-1057): ProcessRawSkymap.process_singles over
the raw maps into ``fname_sky_map`` (default ``skymap.h5``), with the package's default configuration or the JSON
given by ``--skymap-conf``.
"""
import json
import argparse
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser(prog="pba", description=__doc__.split("\n")[0])
    ap.add_argument("workdir")
    ap.add_argument("parfile")
    ap.add_argument("loglevel", nargs="?", default="info")
    ap.add_argument("--precision", type=int, default=0,
                    help="0: FP64 everywhere; 1: SSC contraction in BF16 hi/lo on MFMA (FP32 accumulation)")
    ap.add_argument("--process-skymaps", action="store_true",
                    help="bin the raw sky maps into fname_sky_map (ProcessRawSkymap.process_singles)")
    ap.add_argument("--skymap-conf", default=None,
                    help="JSON ProcessRawSkymap configuration (default: nx=64, ny=32, extend_grid=2, fwhm_fac=0.5, "
                         "lat_dist_method=integ, no filters)")
    a = ap.parse_args(argv)
    verbose = a.loglevel == "info"

    from . import abi, gpu, inputs, output
    workdir = a.workdir if a.workdir.endswith("/") else a.workdir + "/"
    kp, ko = inputs.read_parfile(os.path.join(workdir, a.parfile), inputs.SEC_KN)
    if ko.get("run_bws", "no") == "yes":
        _, go_jet = inputs.read_parfile(os.path.join(workdir, a.parfile), inputs.SEC_GRB)
        if go_jet.get("run_bws", "no") == "yes":
            # the reference evolves ejecta behind a jet through prepareDensProfileFromJet, which
            # exits as not implemented (src/blastwave/blastwave.h:1296-1300)
            print("[pba] a jet and kilonova ejecta in one run are not supported", file=sys.stderr)
            return 1
        return run_ejecta(workdir, a.parfile, kp, ko, a.precision, verbose)
    try:
        P, mp_, mo, gp, go = abi.run_params_from_parfile(os.path.join(workdir, a.parfile))
    except (NotImplementedError, KeyError, ValueError) as e:
        print(f"[pba] unsupported or invalid parfile: {e}", file=sys.stderr)
        return 1
    if go.get("run_bws", "yes") != "yes":
        print("[pba] run_bws=no: nothing to evolve", file=sys.stderr)
        return 0
    t0 = time.time()
    run = gpu.Run.from_workdir(workdir, a.parfile)
    save_spec = go.get("save_spec", "no") == "yes"
    pl = gpu.Pipeline([run], precision=a.precision)
    pl.dynamics()
    dst = np.array([r.status for r in pl.res])
    if dst.any():
        print(f"[pba] blastwave evolution failed: status {dst.tolist()}", file=sys.stderr)
        return 1
    pl.electrons(want_f=save_spec, spectra=save_spec)
    pl.observer()
    _, est, lst = pl.status()
    if est.any() or lst.any():
        print(f"[pba] microphysics/EATS failed: electrons {est.tolist()} light curve {lst.tolist()}", file=sys.stderr)
        return 1
    if verbose:
        print(f"[pba] {run.n_layers} layers evolved and integrated in {time.time() - t0:.2f}s", flush=True)
    if go.get("do_lc", "yes") == "yes":
        path = workdir + go.get("fname_light_curve", "lc.h5")
        output.write_lightcurves(path, run, pl.flux()[0], mp_, gp)
        if verbose:
            print(f"[pba] light curves -> {path}", flush=True)
    if go.get("save_dynamics", "no") == "yes":
        path = workdir + go.get("fname_dyn", "dyn.h5")
        output.write_dynamics(path, run, pl.d_table.cpu().numpy(), [r.fsrs for r in pl.res], mp_, gp)
        if verbose:
            print(f"[pba] dynamics -> {path}", flush=True)
    if go.get("do_skymap", "no") == "yes":
        if go.get("save_raw_skymap", "no") != "yes":
            print("[pba] do_skymap=yes needs save_raw_skymap=yes (only raw images are saved)", file=sys.stderr)
            return 1
        if run.method != "piece-wise":
            print("[pba] sky maps are implemented for method_eats=piece-wise only", file=sys.stderr)
            return 1
        from . import inputs
        times = inputs.vec_from_string(mo["skymap_times"])
        freqs = inputs.vec_from_string(mo["skymap_freqs"])
        image, flux, sst = pl.skymaps(times, freqs, go.get("skymap_remove_mu", "no") == "yes")
        if sst.any():
            print(f"[pba] sky map failed: status {sst.tolist()}", file=sys.stderr)
            return 1
        paths = output.write_raw_skymaps(workdir, output.skymap_images(image, flux, run.n_layers), times, freqs,
                                         mp_, gp)
        if verbose:
            print(f"[pba] {len(paths)} raw sky maps -> {workdir}raw_skymap_*.h5", flush=True)
        if a.process_skymaps:
            from .skymap_process import ProcessRawSkymap, DEFAULT_CONF
            conf = dict(DEFAULT_CONF, **(json.loads(a.skymap_conf) if a.skymap_conf else {}))
            out = workdir + go.get("fname_sky_map", "skymap.h5")
            try:
                ProcessRawSkymap(conf=conf).process_singles(infpaths=workdir + "raw_skymap_*.h5", outfpath=out)
            except ValueError as e:
                print(f"[pba] sky-map processing failed: {e}", file=sys.stderr)
                return 1
            if verbose:
                print(f"[pba] processed sky maps -> {out}", flush=True)
    if save_spec:
        path = workdir + go.get("fname_spectrum", "spec.h5")
        npop = len(pl.pops)
        sh = run.params.fs
        spectra = {k: v.cpu().numpy() for k, v in pl.d_spec.items()}
        aall = pl.d_aall[:npop].cpu().numpy() if pl.d_aall is not None else None
        output.write_spectra(path, run, pl.pops, pl.d_jall[:npop].cpu().numpy(), pl.d_fall.cpu().numpy(),
                             output.ele_grid(sh), output.freq_grid(sh), mp_, gp, spectra=spectra, aall=aall)
        if verbose:
            print(f"[pba] spectra -> {path}", flush=True)
    return 0


# datasets of the ejecta's 2-D initial-data file ([nshells][nlayers]; EjectaID2::_load_id_file fills a
# missing one with zeros, src/ejecta/ejecta_id.h:235-245)
_ID_FIELDS = ("r", "mom", "theta", "ctheta", "ek", "mass")


def ejecta_runs(P, inits, order, tgrid, tobs, nuobs, method, ebl_table=None):
    """The ejecta's blastwaves as device runs: one run when they fit one wavefront (64), else one
    run per layer (its velocity shells).  The reference steps all of them together; one step per
    layer moves the step sequence, by 1.8e-12 in C5's light curves."""
    from . import abi, gpu
    n = len(inits)
    ns = 1 + max(o[0] for o in order)
    per = n if n <= 64 else ns
    if per > 64:
        raise NotImplementedError(f"{ns} velocity shells per layer: a run of more than 64 ejecta blastwaves")
    return [gpu.Run(P, (abi.LayerInit * per)(*inits[i:i + per]), tgrid, tobs, nuobs, method, ebl_table=ebl_table)
            for i in range(0, n, per)]


def run_ejecta(workdir, parfile, kp, ko, precision=0, verbose=False):
    """The kilonova ejecta of a parfile (kN section, no jet): the reference's ejecta branch of
    main.cpp (setPars and processEvolved, src/main.cpp:338-347, 378) on the device; writes fname_light_curve (do_lc) and
    fname_dyn (save_dynamics) as Ejecta::saveDenseLightCurves / saveEjectaBWsDynamics do."""
    from . import abi, ebl, gpu, h5io, inputs, output
    mp_, mo = inputs.read_parfile(os.path.join(workdir, parfile), inputs.SEC_MAIN)
    try:
        P, spread = abi.run_params_ejecta(mp_, mo, kp, ko)
        for key in ("do_skymap", "save_spec"):
            if ko.get(key, "no") == "yes":
                raise NotImplementedError(f"{key} = yes for kilonova ejecta")
        idf = os.path.join(workdir, ko["fname_ejecta_id"])
        shape = h5io.shape(idf, "mass")
        if len(shape) != 2:
            raise ValueError(f"{idf}: use_1d_id = no expects [nshells][nlayers] datasets")
        idd = {k: (h5io.read(idf, k).reshape(shape) if h5io.has(idf, k) else np.zeros(shape)) for k in _ID_FIELDS}
        method = ko.get("method_eats", "adaptive")
        inits, order = abi.ejecta_inits(idd, method)
        tgrid = abi.tgrid(mp_)
        tobs, nuobs = abi.obs_pairs(mo)
        tab = ebl.resolve_table_path(ko.get("ebl_tbl_fpath", "none"), workdir)
        runs = ejecta_runs(P, inits, order, tgrid, tobs, nuobs, method,
                           None if tab is None else ebl.EblTable.load(tab))
    except (NotImplementedError, KeyError, ValueError, FileNotFoundError) as e:
        print(f"[pba] unsupported or invalid kilonova parfile: {e}", file=sys.stderr)
        return 1
    t0 = time.time()
    pl = gpu.Pipeline(runs, precision=precision)
    pl.dynamics_ejecta(spread)
    dst = np.array([r.status for r in pl.res])
    if dst.any():
        print(f"[pba] ejecta evolution failed: status {dst.tolist()}", file=sys.stderr)
        return 1
    pl.electrons()
    pl.observer()
    _, est, lst = pl.status()
    if est.any() or lst.any():
        print(f"[pba] microphysics/EATS failed: electrons {est.tolist()} light curve {lst.tolist()}", file=sys.stderr)
        return 1
    if verbose:
        print(f"[pba] {len(inits)} ejecta blastwaves in {len(runs)} run(s) evolved and integrated in "
              f"{time.time() - t0:.2f}s", flush=True)
    if ko.get("do_lc", "yes") == "yes":
        path = workdir + ko.get("fname_light_curve", "lc.h5")
        flux = pl.flux().reshape(len(inits), len(tobs))
        output.write_lightcurves_ejecta(path, P, inits, order, tgrid, tobs, nuobs, flux, mp_, kp)
        if verbose:
            print(f"[pba] light curves -> {path}", flush=True)
    if ko.get("save_dynamics", "no") == "yes":
        path = workdir + ko.get("fname_dyn", "dyn.h5")
        output.write_dynamics_ejecta(path, P, inits, order, tgrid, pl.d_table.cpu().numpy(), mp_, kp)
        if verbose:
            print(f"[pba] dynamics -> {path}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
