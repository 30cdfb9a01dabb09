"""Host side of the GPU path: binds libpbagpu.so (include/pbagpu.h) through ctypes.

Device memory and streams come from PyTorch (ROCm build); every compute entry point of the
library receives raw device pointers.  There is no CPU fallback: if the library or a GPU is
missing, the functions here raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi, h5io, inputs

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpbagpu.so")
_lib = None

EXPORTS = ("pbagpu_device_count", "pbagpu_set_device", "pbagpu_version", "pbagpu_dyn_evolve",
           "pbagpu_ele_evolve", "pbagpu_eats_lightcurve")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"GPU library missing: {_LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.pbagpu_device_count.argtypes = [C.POINTER(C.c_int)]
        L.pbagpu_set_device.argtypes = [C.c_int]
        L.pbagpu_dyn_evolve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.c_int, C.POINTER(abi.LayerInit), C.c_int,
                                        vp, vp, C.POINTER(abi.BwResult), vp]
        L.pbagpu_ele_evolve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.Population), C.c_int, vp,
                                        C.c_int, C.c_int, vp, vp, vp, C.c_int, vp]
        L.pbagpu_eats_lightcurve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.EatsLayer), C.c_int, vp,
                                             C.c_int, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the GPU path needs a ROCm device (torch.cuda.is_available() is False)")
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream_ptr(stream):
    if stream is None:
        return None
    return C.c_void_p(stream.cuda_stream)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed rc={rc}")


class Run:
    """One prepared model: run parameters, layer initial data, time grid and observer pairs.

    Mirrors what PyBlastAfterglow::setModelPars / Ejecta::setPars / processEvolved read from a
    parfile (src/main.cpp:93-154, src/ejecta/ejecta_base.h:141-345, src/ejecta/ejecta.h:35-80).
    """

    def __init__(self, params, inits, tgrid, tobs, nuobs):
        self.params = params
        self.inits = inits
        self.tgrid = np.ascontiguousarray(tgrid, dtype=np.float64)
        self.tobs = np.ascontiguousarray(tobs, dtype=np.float64)
        self.nuobs = np.ascontiguousarray(nuobs, dtype=np.float64)

    @classmethod
    def from_workdir(cls, workdir, parfile="parfile.par"):
        P, mp_, mo, gp, go = abi.run_params_from_parfile(os.path.join(workdir, parfile))
        idf = os.path.join(workdir, go["fname_ejecta_id"])
        if go.get("use_1d_id", "yes") != "yes":
            raise NotImplementedError("only 1-D initial data files are read here")
        idd = dict(mom=h5io.read(idf, "mom"), ctheta=h5io.read(idf, "ctheta"), ek=h5io.read(idf, "ek"),
                   mass=h5io.read(idf, "mass"), theta_wing=h5io.attr(idf, "theta_wing"),
                   theta_core=h5io.attr(idf, "theta_core"))
        tobs, nuobs = abi.obs_pairs(mo)
        return cls(P, abi.layer_inits(idd), abi.tgrid(mp_), tobs, nuobs)

    @classmethod
    def from_workload(cls, main, grb, struct):
        mp_ = dict(inputs.MAIN_DEFAULTS["pars"], **{k: v for k, v in main.items() if k in inputs.MAIN_DEFAULTS["pars"]})
        mo = dict(inputs.MAIN_DEFAULTS["opts"], **{k: v for k, v in main.items() if k in inputs.MAIN_DEFAULTS["opts"]})
        gp = dict(inputs.GRB_DEFAULTS["pars"], **{k: v for k, v in grb.items() if k in inputs.GRB_DEFAULTS["pars"]})
        go = dict(inputs.GRB_DEFAULTS["opts"], **{k: v for k, v in grb.items() if k in inputs.GRB_DEFAULTS["opts"]})
        go["use_1d_id"] = "yes"
        P = abi.run_params(mp_, mo, gp, go)
        idd = inputs.jet_structure(kind="adaptive", **struct)
        tobs, nuobs = abi.obs_pairs(mo)
        return cls(P, abi.layer_inits(idd), abi.tgrid(mp_), tobs, nuobs)

    @property
    def n_layers(self):
        return len(self.inits)


def populations(params_list, results, n_runs, bw_per_run):
    """Electron populations and EATS layer descriptors of a batch of runs (forward shock of every
    blastwave; reverse shock where the blastwave kept it and RS radiation is on)."""
    pops, lays = [], []
    for r in range(n_runs):
        P = params_list[r]
        for i in range(bw_per_run):
            bw = r * bw_per_run + i
            pfs = len(pops)
            pops.append(abi.Population(bw, r, 0, 0))
            prs = -1
            if results[bw].fsrs and P.do_rs_radiation:
                prs = len(pops)
                pops.append(abi.Population(bw, r, 1, 0))
            lays.append((bw, r, pfs, prs, results[bw].i_end_r))
    return pops, lays


def lightcurves(runs, stream=None, want_intermediates=False, precision=0):
    """Run the three GPU stages for a list of Runs with equal grid sizes.

    Returns (flux [n_runs, n_layers, n_pairs] numpy array, extras dict)."""
    torch = _torch()
    L = lib()
    dev = torch.device("cuda", torch.cuda.current_device())
    n_runs = len(runs)
    r0 = runs[0]
    bpr = r0.n_layers
    nt = len(r0.tgrid)
    params = (abi.RunParams * n_runs)(*[r.params for r in runs])
    inits = (abi.LayerInit * (n_runs * bpr))(*[li for r in runs for li in r.inits])
    d_tg = torch.from_numpy(r0.tgrid).to(dev)
    d_table = torch.empty((n_runs * bpr, abi.PBA_NV, nt), dtype=torch.float64, device=dev)
    res = (abi.BwResult * (n_runs * bpr))()
    sp = _stream_ptr(stream)
    _check(L.pbagpu_dyn_evolve(params, n_runs, bpr, inits, nt, _ptr(d_tg), _ptr(d_table), res, sp), "dyn_evolve")
    pops, lays = populations(params, res, n_runs, bpr)
    nf = r0.params.fs.nfreq
    npop = len(pops)
    parr = (abi.Population * npop)(*pops)
    d_jall = torch.empty((npop, nt, nf), dtype=torch.float64, device=dev)
    d_est = torch.zeros(npop, dtype=torch.int32, device=dev)
    _check(L.pbagpu_ele_evolve(params, n_runs, parr, npop, _ptr(d_table), n_runs * bpr, nt, _ptr(d_jall), None,
                               _ptr(d_est), precision, sp), "ele_evolve")
    larr = (abi.EatsLayer * len(lays))()
    for i, (bw, r, pfs, prs, iend) in enumerate(lays):
        li = runs[r].inits[bw - r * bpr]
        larr[i] = abi.EatsLayer(li.theta_c_l, li.theta_c_h, bw, r, pfs, prs, iend)
    npairs = len(r0.tobs)
    d_tobs = torch.from_numpy(r0.tobs).to(dev)
    d_nuobs = torch.from_numpy(r0.nuobs).to(dev)
    d_out = torch.empty((len(lays), npairs), dtype=torch.float64, device=dev)
    d_lst = torch.zeros(len(lays), dtype=torch.int32, device=dev)
    _check(L.pbagpu_eats_lightcurve(params, n_runs, larr, len(lays), _ptr(d_table), nt, _ptr(d_jall), nf, npairs,
                                    _ptr(d_tobs), _ptr(d_nuobs), _ptr(d_out), _ptr(d_lst), sp), "eats_lightcurve")
    flux = d_out.cpu().numpy().reshape(n_runs, bpr, npairs)
    extras = dict(dyn_status=np.array([r.status for r in res]), ele_status=d_est.cpu().numpy(),
                  eats_status=d_lst.cpu().numpy())
    if want_intermediates:
        extras.update(table=d_table.cpu().numpy(), jall=d_jall.cpu().numpy(), pops=pops, layers=lays,
                      results=[(r.status, r.i_end_r, r.fsrs, r.n_steps, r.n_rhs) for r in res])
    return flux, extras
