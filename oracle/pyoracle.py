"""ctypes front-end of the CPU oracle (test infrastructure only: tests/, smoke(), bench cpu_baseline)."""
import ctypes as C
import os

import numpy as np

from pyblastafterglowmag_amd import abi

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise RuntimeError(f"oracle library missing: {_PATH}")
        _lib = C.CDLL(_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        _lib.orc_dyn_evolve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.c_int, C.POINTER(abi.LayerInit),
                                        C.c_int, dp, dp, C.POINTER(abi.BwResult)]
        _lib.orc_dyn_evolve_ejecta.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.c_int, C.POINTER(abi.LayerInit),
                                               C.c_int, C.c_int, dp, dp, C.POINTER(abi.BwResult)]
        if hasattr(_lib, "orc_ele_evolve"):
            _lib.orc_ele_evolve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.Population), C.c_int,
                                            dp, C.c_int, C.c_int, dp, dp, dp, ip]
        if hasattr(_lib, "orc_eats_lightcurve"):
            _lib.orc_eats_lightcurve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.EatsLayer), C.c_int,
                                                 dp, C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, dp, ip]
        if hasattr(_lib, "orc_eats_skymap_pw"):
            _lib.orc_eats_skymap_pw.argtypes = [C.POINTER(abi.RunParams), C.POINTER(abi.EatsPwLayer), C.c_int, dp,
                                                C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, C.c_int, dp, dp, ip]
        if hasattr(_lib, "orc_eats_lightcurve_pw"):
            _lib.orc_eats_lightcurve_pw.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.EatsPwLayer),
                                                    C.c_int, dp, C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, dp, ip]
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def dyn_evolve(params, inits, tgrid, n_runs=1):
    """params: RunParams or array of n_runs; inits: LayerInit array of n_runs*bw_per_run."""
    if isinstance(params, abi.RunParams):
        params = (abi.RunParams * 1)(params)
    nbw_total = len(inits)
    bpr = nbw_total // n_runs
    nt = len(tgrid)
    tg = np.ascontiguousarray(tgrid, dtype=np.float64)
    table = np.zeros((nbw_total, abi.PBA_NV, nt))
    res = (abi.BwResult * nbw_total)()
    rc = lib().orc_dyn_evolve(params, n_runs, bpr, inits, nt, _dp(tg), _dp(table), res)
    if rc != 0:
        raise RuntimeError(f"orc_dyn_evolve rc={rc}")
    return table, res


def dyn_evolve_ejecta(params, inits, tgrid, spread=False, n_runs=1):
    """Ejecta blastwaves of a kilonova run (rhs_fs_dense in a standard ISM); inits in the
    reference's state-vector order, layer-major."""
    if isinstance(params, abi.RunParams):
        params = (abi.RunParams * 1)(params)
    nbw_total = len(inits)
    bpr = nbw_total // n_runs
    nt = len(tgrid)
    tg = np.ascontiguousarray(tgrid, dtype=np.float64)
    table = np.zeros((nbw_total, abi.PBA_NV, nt))
    res = (abi.BwResult * nbw_total)()
    rc = lib().orc_dyn_evolve_ejecta(params, n_runs, bpr, inits, 1 if spread else 0, nt, _dp(tg), _dp(table), res)
    if rc != 0:
        raise RuntimeError(f"orc_dyn_evolve_ejecta rc={rc}")
    return table, res


def _dpn(a):
    return None if a is None else _dp(np.ascontiguousarray(a))


def ele_evolve(params, pops, table, n_runs=1, want_f=False, want_a=False):
    """(jall, fall, status), plus the absorption table aall when want_a (zero where use_ssa is off)."""
    if isinstance(params, abi.RunParams):
        params = (abi.RunParams * 1)(params)
    nbw, nv, nt = table.shape
    npop = len(pops)
    p0 = params[pops[0].run]
    sh = p0.fs if pops[0].shock == 0 else p0.rs
    nf = sh.nfreq
    # f_all rows: the largest electron grid of the populations (its own ngam, or its shock's)
    ng = max(P.ngam if P.ngam > 0 else (params[P.run].rs.ngam if P.shock else params[P.run].fs.ngam) for P in pops)
    tab = np.ascontiguousarray(table)
    jall = np.zeros((npop, nt, nf))
    fall = np.zeros((npop, nt, ng)) if want_f else None
    st = np.zeros(npop, dtype=np.int32)
    aall = np.zeros((npop, nt, nf))
    rc = lib().orc_ele_evolve(params, n_runs, pops, npop, _dp(tab), nbw, nt, _dp(jall), _dp(aall),
                              _dp(fall) if want_f else None, _ip(st))
    if rc != 0:
        raise RuntimeError(f"orc_ele_evolve rc={rc}")
    return (jall, fall, st, aall) if want_a else (jall, fall, st)


def eats_lightcurve(params, layers, table, jall, tobs, nuobs, n_runs=1, aall=None):
    if isinstance(params, abi.RunParams):
        params = (abi.RunParams * 1)(params)
    nbw, nv, nt = table.shape
    nf = jall.shape[2]
    n = len(tobs)
    out = np.zeros((len(layers), n))
    st = np.zeros(len(layers), dtype=np.int32)
    rc = lib().orc_eats_lightcurve(params, n_runs, layers, len(layers), _dp(np.ascontiguousarray(table)), nt,
                                   _dp(np.ascontiguousarray(jall)), _dpn(aall), nf, n,
                                   _dp(np.ascontiguousarray(tobs, dtype=np.float64)),
                                   _dp(np.ascontiguousarray(nuobs, dtype=np.float64)), _dp(out), _ip(st))
    if rc != 0:
        raise RuntimeError(f"orc_eats_lightcurve rc={rc}")
    return out, st


def eats_lightcurve_pw(params, layers, table, jall, tobs, nuobs, n_runs=1, aall=None):
    """Piece-wise EATS light curves; layers: EatsPwLayer array."""
    if isinstance(params, abi.RunParams):
        params = (abi.RunParams * 1)(params)
    nbw, nv, nt = table.shape
    nf = jall.shape[2]
    n = len(tobs)
    out = np.zeros((len(layers), n))
    st = np.zeros(len(layers), dtype=np.int32)
    rc = lib().orc_eats_lightcurve_pw(params, n_runs, layers, len(layers), _dp(np.ascontiguousarray(table)), nt,
                                      _dp(np.ascontiguousarray(jall)), _dpn(aall), nf, n,
                                      _dp(np.ascontiguousarray(tobs, dtype=np.float64)),
                                      _dp(np.ascontiguousarray(nuobs, dtype=np.float64)), _dp(out), _ip(st))
    if rc != 0:
        raise RuntimeError(f"orc_eats_lightcurve_pw rc={rc}")
    return out, st


def eats_skymap_pw(params, layers, table, jall, tobs, nuobs, remove_mu=False, aall=None):
    """Raw piece-wise sky maps of one run: (image [n_images][PBA_IMG_NV][2 ncells], flux [n_layers][n_images],
    status [n_layers])."""
    if isinstance(params, abi.RunParams):
        params = (abi.RunParams * 1)(params)
    nbw, nv, nt = table.shape
    nf = jall.shape[2]
    n = len(tobs)
    ncells = sum(L.cil for L in layers)
    image = np.zeros((n, abi.PBA_IMG_NV, 2 * ncells))
    flux = np.zeros((len(layers), n))
    st = np.zeros(len(layers), dtype=np.int32)
    rc = lib().orc_eats_skymap_pw(params, layers, len(layers), _dp(np.ascontiguousarray(table)), nt,
                                  _dp(np.ascontiguousarray(jall)), _dpn(aall), nf, n,
                                  _dp(np.ascontiguousarray(tobs, dtype=np.float64)),
                                  _dp(np.ascontiguousarray(nuobs, dtype=np.float64)), 1 if remove_mu else 0,
                                  _dp(image), _dp(flux), _ip(st))
    if rc != 0:
        raise RuntimeError(f"orc_eats_skymap_pw rc={rc}")
    return image, flux, st


def ssc_emissivity(shock, f, w):
    """The SSC contraction (computeSSCSpectrum) of cells f [n][ngam], w [n][nfreq] on the host."""
    L = lib()
    f = np.ascontiguousarray(f, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.zeros((f.shape[0], shock.nfreq))
    L.orc_ssc_emissivity.argtypes = [C.POINTER(abi.ShockParams), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rc = L.orc_ssc_emissivity(C.byref(shock), f.shape[0], f.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                              out.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise RuntimeError(f"orc_ssc_emissivity rc={rc}")
    return out
