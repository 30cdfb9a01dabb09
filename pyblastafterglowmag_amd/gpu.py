"""Host side of the GPU path: binds libpbagpu.so (include/pbagpu.h) through ctypes.

Device memory and streams come from PyTorch (ROCm build); every compute entry point of the
library receives raw device pointers.  There is no CPU fallback: if the library or a GPU is
missing, the functions here raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi, ebl, h5io, inputs

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpbagpu.so")
_lib = None

EXPORTS = ("pbagpu_device_count", "pbagpu_set_device", "pbagpu_version", "pbagpu_last_error", "pbagpu_timing_enable",
           "pbagpu_timing_read", "pbagpu_timing_units", "pbagpu_dyn_evolve", "pbagpu_ele_evolve", "pbagpu_eats_lightcurve",
           "pbagpu_eats_lightcurve_pw", "pbagpu_eats_skymap_pw", "pbagpu_skymap_hist", "pbagpu_bessel_k_thirds",
           "pbagpu_release_stream", "pbagpu_stream_create", "pbagpu_stream_destroy",
           "pbagpu_synchrotron_f", "pbagpu_eats_lightcurve_vec", "pbagpu_mlp_forward", "pbagpu_mlp_train_step",
           "pbagpu_gemm_bf16", "pbagpu_ssc_features", "pbagpu_ssc_emissivity", "pbagpu_ele_evolve_surrogate",
           "pbagpu_skymap_hist_noalloc", "pbagpu_ele_evolve_spectra", "pbagpu_dyn_evolve_ragged",
           "pbagpu_set_option", "pbagpu_get_option", "pbagpu_dyn_evolve_ejecta")
OPT_ELE_GRAPH = 0
OPT_DYN_RUNS_PER_WAVE = 1
STAGE_DYN, STAGE_ELE, STAGE_EATS, STAGE_SSC_TABLE, STAGE_EATS_PW, STAGE_SKY_HIST, STAGE_EATS_VEC = 0, 1, 2, 3, 4, 5, 6
N_STAGES = 7
# sky-map binning (enum pba_sky_moment / pba_sky_status)
SKY_MOMENTS = ("xc", "yc", "xmin", "xmax", "ymin", "ymax", "sumw", "npts")
SKY_STATUS = {0: "ok", 1: "no entry with data", 2: "intensities sum to zero", 3: "flat entry (xmin == xmax or ymin == ymax)",
              4: "non-finite position or intensity", 5: "given edges decrease"}


def lib():
    global _lib
    if _lib is None:
        # torch (ROCm build) bundles a HIP runtime with the same SONAME as the system one;
        # importing torch first makes libpbagpu.so bind to that copy, so that device memory
        # and streams from torch and the library's launches share one runtime.
        import torch  # noqa: F401
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"GPU library missing: {_LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.pbagpu_device_count.argtypes = [C.POINTER(C.c_int)]
        L.pbagpu_set_device.argtypes = [C.c_int]
        L.pbagpu_dyn_evolve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.c_int, C.POINTER(abi.LayerInit), C.c_int,
                                        vp, vp, C.POINTER(abi.BwResult), vp]
        L.pbagpu_dyn_evolve_ragged.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.c_int, C.POINTER(abi.LayerInit),
                                               C.c_int, C.POINTER(C.c_int), vp, vp, C.POINTER(abi.BwResult), vp]
        L.pbagpu_dyn_evolve_ejecta.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.c_int, C.POINTER(abi.LayerInit),
                                               C.c_int, C.c_int, vp, vp, C.POINTER(abi.BwResult), vp]
        L.pbagpu_ele_evolve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.Population), C.c_int, vp,
                                        C.c_int, C.c_int, vp, vp, vp, vp, C.c_int, vp]
        L.pbagpu_ele_evolve_spectra.argtypes = L.pbagpu_ele_evolve.argtypes[:-1] + [C.POINTER(abi.Spectra), vp]
        L.pbagpu_set_option.argtypes = [C.c_int, C.c_longlong]
        L.pbagpu_get_option.argtypes = [C.c_int, C.POINTER(C.c_longlong)]
        L.pbagpu_last_error.restype = C.c_char_p
        L.pbagpu_last_error.argtypes = []
        L.pbagpu_timing_enable.argtypes = [C.c_int]
        L.pbagpu_timing_read.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong), C.c_int]
        L.pbagpu_timing_units.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int]
        L.pbagpu_eats_lightcurve.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.EatsLayer), C.c_int, vp,
                                             C.c_int, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.pbagpu_eats_lightcurve_pw.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.EatsPwLayer), C.c_int,
                                                vp, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.pbagpu_eats_skymap_pw.argtypes = [C.POINTER(abi.RunParams), C.POINTER(abi.EatsPwLayer), C.c_int, vp, C.c_int,
                                            vp, vp, C.c_int, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp]
        L.pbagpu_bessel_k_thirds.argtypes = [C.c_int, vp, vp, vp]
        L.pbagpu_release_stream.argtypes = [vp]
        L.pbagpu_synchrotron_f.argtypes = [C.c_int, vp, vp]
        L.pbagpu_mlp_forward.argtypes = [C.POINTER(abi.Mlp), C.c_int, vp, vp, vp]
        L.pbagpu_mlp_train_step.argtypes = [C.POINTER(abi.Mlp), C.c_int, vp, vp, C.c_float, C.c_int, vp, vp]
        L.pbagpu_gemm_bf16.argtypes = [C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp]
        L.pbagpu_ssc_features.argtypes = [C.POINTER(abi.ShockParams), C.c_int, vp, vp, vp, vp, vp, vp]
        L.pbagpu_ssc_emissivity.argtypes = [C.POINTER(abi.ShockParams), C.c_int, vp, vp, vp, vp]
        L.pbagpu_ele_evolve_surrogate.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.Population),
                                                  C.c_int, vp, C.c_int, C.c_int, vp, vp, vp,
                                                  C.POINTER(abi.SscSurrogate), C.POINTER(abi.SurrogateStats), vp]
        L.pbagpu_eats_lightcurve_vec.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.EatsLayer), C.c_int,
                                                 vp, C.c_int, vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_double, vp, vp,
                                                 vp]
        L.pbagpu_stream_create.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.pbagpu_stream_destroy.argtypes = [vp]
        L.pbagpu_skymap_hist.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_long), vp, vp, vp, C.c_int,
                                         C.c_int, C.c_double, C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.pbagpu_skymap_hist_noalloc.argtypes = L.pbagpu_skymap_hist.argtypes
        _lib = L
    return _lib


def set_option(which, value):
    """pbagpu_set_option (OPT_ELE_GRAPH: capture the electron stage's per-step launches in a graph;
    OPT_DYN_RUNS_PER_WAVE: runs per dynamics wavefront, 0 = as many as fit)."""
    _check(lib().pbagpu_set_option(int(which), int(value)), "set_option")


def get_option(which):
    v = C.c_longlong(0)
    _check(lib().pbagpu_get_option(int(which), C.byref(v)), "get_option")
    return v.value


def release_stream(stream):
    """Free the library's device workspace of a torch stream (pbagpu_release_stream)."""
    _check(lib().pbagpu_release_stream(_stream_ptr(stream)), "release_stream")


class CuStream:
    """A HIP stream restricted to compute units [cu_first, cu_first + cu_count) (pbagpu_stream_create);
    cu_count <= 0 gives a plain stream.  Usable wherever the Pipeline takes a stream."""

    def __init__(self, cu_first=0, cu_count=0):
        _torch()
        p = C.c_void_p()
        _check(lib().pbagpu_stream_create(int(cu_first), int(cu_count), C.byref(p)), "stream_create")
        self.cuda_stream = p.value

    def synchronize(self):
        import torch
        torch.cuda.ExternalStream(self.cuda_stream).synchronize()

    def close(self):
        if self.cuda_stream:
            _check(lib().pbagpu_stream_destroy(C.c_void_p(self.cuda_stream)), "stream_destroy")
            self.cuda_stream = None


def device_cu_count(device=0):
    return _torch().cuda.get_device_properties(device).multi_processor_count


def timing_enable(on=True):
    _check(lib().pbagpu_timing_enable(1 if on else 0), "timing_enable")


def timing_read(stage, reset=False):
    """(total device ms, launches) of a stage's main kernel since the last reset."""
    ms, n = C.c_double(0.0), C.c_longlong(0)
    _check(lib().pbagpu_timing_read(stage, C.byref(ms), C.byref(n), 1 if reset else 0), "timing_read")
    return ms.value, n.value


def timing_units(stage, reset=False):
    u = C.c_double(0.0)
    _check(lib().pbagpu_timing_units(stage, C.byref(u), 1 if reset else 0), "timing_units")
    return u.value


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
        msg = lib().pbagpu_last_error()
        raise RuntimeError(f"{what} failed rc={rc}: {msg.decode() if msg else ''}")


def bessel_k_thirds(x):
    """K_{1/3}(x), K_{4/3}(x) by the electron stage's device routine, evaluated on the host."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    k13, k43 = np.empty_like(x), np.empty_like(x)
    _check(lib().pbagpu_bessel_k_thirds(len(x), x.ctypes.data_as(C.c_void_p), k13.ctypes.data_as(C.c_void_p),
                                        k43.ctypes.data_as(C.c_void_p)), "bessel_k_thirds")
    return k13, k43


def synchrotron_f(x):
    """F(x) = x int_x^inf K_5/3 by the electron stage's device routine (GSL kernel option), on the host."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    f = np.empty_like(x)
    _check(lib().pbagpu_synchrotron_f(len(x), x.ctypes.data_as(C.c_void_p), f.ctypes.data_as(C.c_void_p)),
           "synchrotron_f")
    return f


def skymap_hist(images, nx, ny, extend_grid=1.0, edges=None, want_count=False, device=0, noalloc=False):
    """Histogram images, centroids and extents of raw sky maps in one pbagpu_skymap_hist call.

    images: one list of entries per image, an entry being host arrays (x, y, intensity) of one
            shell's principal or counter jet (ProcessRawSkymap._get_skymap's lists).
    edges:  None (edges from the data extent) or (edges_x [n_images][nx], edges_y [n_images][ny]).
    noalloc: call pbagpu_skymap_hist_noalloc (work tables in the library's fixed device scratch)
            instead of pbagpu_skymap_hist.
    Returns a dict of host arrays: hist [n][nx-1][ny-1], count (or None), edges_x, edges_y,
    moments [n][len(SKY_MOMENTS)], status [n]."""
    torch = _torch()
    L = lib()
    dev = torch.device("cuda", device)
    n = len(images)
    img_seg = [0]
    seg_off = [0]
    xs, ys, ws = [], [], []
    for entries in images:
        for (x, y, w) in entries:
            x, y, w = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (x, y, w))
            if not (len(x) == len(y) == len(w)):
                raise ValueError("skymap_hist: x, y and intensity of an entry differ in length")
            xs.append(x); ys.append(y); ws.append(w)
            seg_off.append(seg_off[-1] + len(x))
        img_seg.append(len(seg_off) - 1)
    cat = (lambda a: np.concatenate(a) if a else np.zeros(0))
    d_x, d_y, d_w = (torch.from_numpy(cat(a)).to(dev) for a in (xs, ys, ws))
    if edges is None:
        d_ex = torch.empty((n, nx), dtype=torch.float64, device=dev)
        d_ey = torch.empty((n, ny), dtype=torch.float64, device=dev)
    else:
        ex, ey = (np.array(np.broadcast_to(np.asarray(e, dtype=np.float64), (n, m)), order="C")
                  for e, m in zip(edges, (nx, ny)))
        d_ex, d_ey = torch.from_numpy(ex).to(dev), torch.from_numpy(ey).to(dev)
    d_h = torch.empty((n, max(nx - 1, 0), max(ny - 1, 0)), dtype=torch.float64, device=dev)
    d_c = torch.empty((n, max(nx - 1, 0), max(ny - 1, 0)), dtype=torch.int32, device=dev) if want_count else None
    d_m = torch.empty((n, len(SKY_MOMENTS)), dtype=torch.float64, device=dev)
    d_s = torch.empty(n, dtype=torch.int32, device=dev)
    seg_arr = (C.c_int * len(img_seg))(*img_seg)
    off_arr = (C.c_long * len(seg_off))(*seg_off)
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream()
        fn = L.pbagpu_skymap_hist_noalloc if noalloc else L.pbagpu_skymap_hist
        _check(fn(n, seg_arr, off_arr, _ptr(d_x), _ptr(d_y), _ptr(d_w), nx, ny, float(extend_grid),
                  0 if edges is None else 1, _ptr(d_ex), _ptr(d_ey), _ptr(d_h), _ptr(d_c), _ptr(d_m),
                  _ptr(d_s), _stream_ptr(stream)), "skymap_hist_noalloc" if noalloc else "skymap_hist")
    return dict(hist=d_h.cpu().numpy(), count=None if d_c is None else d_c.cpu().numpy().view(np.uint32),
                edges_x=d_ex.cpu().numpy(), edges_y=d_ey.cpu().numpy(), moments=d_m.cpu().numpy(),
                status=d_s.cpu().numpy())


class Run:
    """One prepared model: run parameters, layer initial data, time grid and observer pairs.

    Mirrors what PyBlastAfterglow::setModelPars / Ejecta::setPars / processEvolved read from a
    parfile (src/main.cpp:93-154, src/ejecta/ejecta_base.h:141-345, src/ejecta/ejecta.h:35-80).
    """

    def __init__(self, params, inits, tgrid, tobs, nuobs, method="adaptive", ebl_table=None):
        self.method = method
        self.params = params
        self.inits = inits
        self.tgrid = np.ascontiguousarray(tgrid, dtype=np.float64)
        self.tobs = np.ascontiguousarray(tobs, dtype=np.float64)
        self.nuobs = np.ascontiguousarray(nuobs, dtype=np.float64)
        # EBL attenuation exp(-tau(nu_obs, z)) per observer pair (ebl_tbl_fpath), or None
        self.ebl_table = ebl_table
        self.attenuation = None if ebl_table is None else ebl_table.attenuation(self.nuobs, params.z)

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
        method = go.get("method_eats", "adaptive")
        tab = ebl.resolve_table_path(go.get("ebl_tbl_fpath", "none"), workdir)
        return cls(P, abi.layer_inits(idd, method), abi.tgrid(mp_), tobs, nuobs, method,
                   ebl_table=None if tab is None else ebl.EblTable.load(tab))

    @classmethod
    def from_workload(cls, main, grb, struct):
        mp_ = dict(inputs.MAIN_DEFAULTS["pars"], **{k: v for k, v in main.items() if k in inputs.MAIN_DEFAULTS["pars"]})
        mo = dict(inputs.MAIN_DEFAULTS["opts"], **{k: v for k, v in main.items() if k in inputs.MAIN_DEFAULTS["opts"]})
        gp = dict(inputs.GRB_DEFAULTS["pars"], **{k: v for k, v in grb.items() if k in inputs.GRB_DEFAULTS["pars"]})
        go = dict(inputs.GRB_DEFAULTS["opts"], **{k: v for k, v in grb.items() if k in inputs.GRB_DEFAULTS["opts"]})
        go["use_1d_id"] = "yes"
        P = abi.run_params(mp_, mo, gp, go)
        method = go.get("method_eats", "adaptive")
        struct = dict(struct)
        ngs = struct.pop("layer_ngam", None)
        idd = inputs.jet_structure(kind=method, **struct)
        tobs, nuobs = abi.obs_pairs(mo)
        inits = abi.layer_inits(idd, method)
        if ngs is not None:   # a Chang-Cooper grid of its own for every layer (both shocks)
            if len(ngs) != len(inits):
                raise ValueError(f"layer_ngam has {len(ngs)} entries for {len(inits)} layers")
            for li, n in zip(inits, ngs):
                li.ngam_fs = li.ngam_rs = int(n)
        tab = ebl.resolve_table_path(go.get("ebl_tbl_fpath", "none"))
        return cls(P, inits, abi.tgrid(mp_), tobs, nuobs, method,
                   ebl_table=None if tab is None else ebl.EblTable.load(tab))

    @property
    def n_layers(self):
        return len(self.inits)


def apply_attenuation(flux, runs):
    """Layer light curves flux [n_runs][n_layers][n_pairs] times each run's EBL attenuation
    (Run.attenuation, per pair; runs without a table are left as they are)."""
    if all(r.attenuation is None for r in runs):
        return flux
    att = np.stack([r.attenuation if r.attenuation is not None else np.ones(flux.shape[-1]) for r in runs])
    return flux * att[:, None, :]


def populations(params_list, results, n_runs, bw_per_run, inits=None):
    """Electron populations and EATS layer descriptors of a batch of runs (forward shock of every
    blastwave; reverse shock where the blastwave kept it and RS radiation is on).  With the
    blastwaves' initial data (inits [n_runs * bw_per_run]) every population carries its
    blastwave's electron-grid size (pba_layer_init ngam_fs / ngam_rs; 0: the run's)."""
    pops, lays = [], []
    for r in range(n_runs):
        P = params_list[r]
        for i in range(bw_per_run):
            bw = r * bw_per_run + i
            ng_fs = inits[bw].ngam_fs if inits is not None else 0
            ng_rs = inits[bw].ngam_rs if inits is not None else 0
            pfs = len(pops)
            pops.append(abi.Population(bw, r, 0, ng_fs))
            prs = -1
            if results[bw].fsrs and P.do_rs_radiation:
                prs = len(pops)
                pops.append(abi.Population(bw, r, 1, ng_rs))
            lays.append((bw, r, pfs, prs, results[bw].i_end_r))
    return pops, lays


class Pipeline:
    """Device buffers and the three stage calls for a fixed batch of Runs with equal grid sizes.

    The dynamics table [n_bw][PBA_NV][nt], the emissivity table [n_pop][nt][nfreq] and the
    light curves [n_layers][n_pairs] stay in device memory between stages; only the small
    per-blastwave results of stage 1 come back to the host (to build the population list,
    as the reference decides per blastwave whether the reverse shock radiates).
    """

    def __init__(self, runs, stream=None, precision=0, rad_stream=None):
        torch = _torch()
        self.torch = torch
        self.L = lib()
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.runs = runs
        self.n_runs = len(runs)
        r0 = runs[0]
        self.bpr = r0.n_layers
        # runs may have time grids of their own (lengths and values): the tables then have
        # nt = the longest grid's steps, and run r fills its first nt_run[r] columns
        self.nt = max(len(r.tgrid) for r in runs)
        self.nt_run = [len(r.tgrid) for r in runs]
        self.ragged = any(len(r.tgrid) != len(r0.tgrid) or not np.array_equal(r.tgrid, r0.tgrid) for r in runs)
        for r in runs:
            if r.n_layers != self.bpr or len(r.tobs) != len(r0.tobs):
                raise ValueError("runs of one batch need equal layer count and pair count")
            if r.method != r0.method:
                raise ValueError("runs of one batch share the EATS method")
            if not np.array_equal(r.tobs, r0.tobs) or not np.array_equal(r.nuobs, r0.nuobs):
                raise ValueError("runs of one batch share the observer pairs")
        self.method = r0.method
        self.nf = r0.params.fs.nfreq
        self.npairs = len(r0.tobs)
        self.precision = precision
        self.stream = stream
        self.sp = _stream_ptr(stream)
        # electrons and observer stage: on rad_stream if given (a pipeline of batches can keep the
        # dynamics of several batches in flight on their own streams and share one stream, and
        # so one set of observer workspaces, for the rest)
        self.rsp = _stream_ptr(rad_stream) if rad_stream is not None else self.sp
        self.params = (abi.RunParams * self.n_runs)(*[r.params for r in runs])
        self.inits = (abi.LayerInit * (self.n_runs * self.bpr))(*[li for r in runs for li in r.inits])
        nbw = self.n_runs * self.bpr
        if self.ragged:
            tg = np.zeros((self.n_runs, self.nt))
            for k, r in enumerate(runs):
                tg[k, :len(r.tgrid)] = r.tgrid
                tg[k, len(r.tgrid):] = r.tgrid[-1]
            self.d_tg = torch.from_numpy(tg).to(self.dev)
        else:
            self.d_tg = torch.from_numpy(r0.tgrid).to(self.dev)
        self.d_table = torch.empty((nbw, abi.PBA_NV, self.nt), dtype=torch.float64, device=self.dev)
        self.res = (abi.BwResult * nbw)()
        # at most two populations per blastwave
        self.d_jall = torch.empty((2 * nbw, self.nt, self.nf), dtype=torch.float64, device=self.dev)
        # absorption tables only when a shock of the batch has synchrotron self-absorption
        self.ssa = any(r.params.fs.ssa or r.params.rs.ssa for r in runs)
        self.d_aall = torch.empty_like(self.d_jall) if self.ssa else None
        self.d_est = torch.zeros(2 * nbw, dtype=torch.int32, device=self.dev)
        self.d_tobs = torch.from_numpy(r0.tobs).to(self.dev)
        self.d_nuobs = torch.from_numpy(r0.nuobs).to(self.dev)
        self.d_out = torch.empty((nbw, self.npairs), dtype=torch.float64, device=self.dev)
        self.d_lst = torch.zeros(nbw, dtype=torch.int32, device=self.dev)
        self.pops = self.layers = None
        # the tensor grid of the frequency-vectorised observer stage, made here: a host-to-device copy
        # later would go through the null stream and wait for other batches' dynamics kernels
        if self.method != "piece-wise":
            try:
                ts, fs = self.tensor_grid()
                self.d_t_u = torch.from_numpy(np.ascontiguousarray(ts)).to(self.dev)
                self.d_nu_u = torch.from_numpy(np.ascontiguousarray(fs)).to(self.dev)
            except ValueError:
                pass
        # the allocations and copies above ran on torch's current stream: wait for that stream only
        # (a device-wide wait would also wait for other batches' dynamics on their own streams)
        torch.cuda.current_stream().synchronize()

    def dynamics(self):
        if self.ragged:
            ntr = (C.c_int * self.n_runs)(*self.nt_run)
            _check(self.L.pbagpu_dyn_evolve_ragged(self.params, self.n_runs, self.bpr, self.inits, self.nt, ntr,
                                                   _ptr(self.d_tg), _ptr(self.d_table), self.res, self.sp),
                   "dyn_evolve_ragged")
        else:
            _check(self.L.pbagpu_dyn_evolve(self.params, self.n_runs, self.bpr, self.inits, self.nt, _ptr(self.d_tg),
                                            _ptr(self.d_table), self.res, self.sp), "dyn_evolve")
        self._set_populations()

    def dynamics_ejecta(self, spread=False):
        """Dynamics of kilonova ejecta (pbagpu_dyn_evolve_ejecta): the runs' blastwaves are ejecta
        blastwaves in the reference's layer-major order (abi.ejecta_inits); spread is method_spread
        != None (abi.run_params_ejecta).  Runs of one time grid, at most 64 blastwaves each."""
        if self.ragged:
            raise ValueError("ejecta runs of one batch share one time grid")
        _check(self.L.pbagpu_dyn_evolve_ejecta(self.params, self.n_runs, self.bpr, self.inits, 1 if spread else 0,
                                               self.nt, _ptr(self.d_tg), _ptr(self.d_table), self.res, self.sp),
               "dyn_evolve_ejecta")
        self._set_populations()

    def _set_populations(self):
        pops, lays = populations(self.params, self.res, self.n_runs, self.bpr, self.inits)
        self.pops = pops
        self.parr = (abi.Population * len(pops))(*pops)
        if self.method == "piece-wise":
            self.larr = (abi.EatsPwLayer * len(lays))()
            for i, (bw, r, pfs, prs, iend) in enumerate(lays):
                li = self.runs[r].inits[bw - r * self.bpr]
                rs = 1 if (self.res[bw].fsrs and self.params[r].do_rs_radiation) else 0
                self.larr[i] = abi.EatsPwLayer(li.ctheta0, li.ncells, abi.cells_in_layer(li.ilayer), bw, r, pfs, rs,
                                               iend)
        else:
            self.larr = (abi.EatsLayer * len(lays))()
            for i, (bw, r, pfs, prs, iend) in enumerate(lays):
                li = self.runs[r].inits[bw - r * self.bpr]
                self.larr[i] = abi.EatsLayer(li.theta_c_l, li.theta_c_h, bw, r, pfs, prs, iend)
        self.layers = lays

    def use_tables(self, table, results):
        """Take the dynamics tables [n_bw][PBA_NV][nt] (host array) and per-blastwave results (the
        oracle's, say) in place of running the dynamics stage, so that the electron and observer
        stages run on given tables."""
        self.d_table.copy_(self.torch.from_numpy(np.ascontiguousarray(table)))
        for i, r in enumerate(results):
            self.res[i] = r
        self._set_populations()

    def ngam_max(self):
        """Largest electron-grid size of the batch's populations (the row stride of f_all)."""
        return max(P.ngam if P.ngam > 0 else (self.params[P.run].rs.ngam if P.shock else self.params[P.run].fs.ngam)
                   for P in self.pops)

    def electrons(self, want_f=False, spectra=False):
        """Electron stage of every population.  spectra=True also keeps the dense-output spectra of
        the reference's save_spec = yes (pbagpu_ele_evolve_spectra) in self.d_spec: a dict of
        [n_pop][nt][nfreq] (syn_j, syn_i, ssc_j, total_f, total_i) and [n_pop][nt][ngam]
        (gam_dot_syn, gam_dot_adi, gam_dot_ssc) device tensors."""
        npop = len(self.pops)
        ng = self.ngam_max()
        self.d_fall = self.torch.empty((npop, self.nt, ng), dtype=self.torch.float64,
                                       device=self.dev) if want_f else None
        args = [self.params, self.n_runs, self.parr, npop, _ptr(self.d_table), self.n_runs * self.bpr, self.nt,
                _ptr(self.d_jall), _ptr(self.d_aall), _ptr(self.d_fall), _ptr(self.d_est), self.precision]
        if not spectra:
            self.d_spec = None
            _check(self.L.pbagpu_ele_evolve(*args, self.rsp), "ele_evolve")
            return
        mk = lambda n: self.torch.empty((npop, self.nt, n), dtype=self.torch.float64, device=self.dev)
        self.d_spec = {k: mk(self.nf) for k in abi.SPECTRA_NU} | {k: mk(ng) for k in abi.SPECTRA_GAM}
        sp = abi.Spectra(**{k: v.data_ptr() for k, v in self.d_spec.items()})
        _check(self.L.pbagpu_ele_evolve_spectra(*args, C.byref(sp), self.rsp), "ele_evolve_spectra")

    def electrons_surrogate(self, sur, want_f=False):
        """Electron stage with the SSC emissivity from a trained surrogate (surrogate.SscSurrogate),
        checked against the plain contraction on a sampled subset every step.  Returns the counts
        (abi.SurrogateStats)."""
        npop = len(self.pops)
        self.d_fall = self.torch.empty((npop, self.nt, self.ngam_max()), dtype=self.torch.float64,
                                       device=self.dev) if want_f else None
        st = abi.SurrogateStats()
        spec = sur.spec()
        _check(self.L.pbagpu_ele_evolve_surrogate(self.params, self.n_runs, self.parr, npop, _ptr(self.d_table),
                                                  self.n_runs * self.bpr, self.nt, _ptr(self.d_jall), _ptr(self.d_fall),
                                                  _ptr(self.d_est), C.byref(spec), C.byref(st), self.rsp),
               "ele_evolve_surrogate")
        return st

    def observer(self, first_layers=None):
        """Observer stage (the reference's per-pair quadrature) of every layer, or of the first
        `first_layers` layers (the runs of a batch are laid out one after the other)."""
        nl = len(self.layers) if first_layers is None else min(int(first_layers), len(self.layers))
        if self.method == "piece-wise":
            _check(self.L.pbagpu_eats_lightcurve_pw(self.params, self.n_runs, self.larr, nl,
                                                    _ptr(self.d_table), self.nt, _ptr(self.d_jall), _ptr(self.d_aall),
                                                    self.nf,
                                                    self.npairs, _ptr(self.d_tobs), _ptr(self.d_nuobs),
                                                    _ptr(self.d_out), _ptr(self.d_lst), self.rsp), "eats_lightcurve_pw")
            return
        _check(self.L.pbagpu_eats_lightcurve(self.params, self.n_runs, self.larr, nl, _ptr(self.d_table),
                                             self.nt, _ptr(self.d_jall), _ptr(self.d_aall), self.nf, self.npairs,
                                             _ptr(self.d_tobs),
                                             _ptr(self.d_nuobs), _ptr(self.d_out), _ptr(self.d_lst), self.rsp),
               "eats_lightcurve")

    def tensor_grid(self):
        """(unique t_obs, unique nu_obs) of the batch's pairs, which are t-major over their product."""
        r0 = self.runs[0]
        n_nu = int(np.count_nonzero(r0.tobs == r0.tobs[0]))
        ts, fs = r0.tobs[::n_nu], r0.nuobs[:n_nu]
        if len(ts) * n_nu != len(r0.tobs) or not (np.array_equal(np.repeat(ts, n_nu), r0.tobs)
                                                    and np.array_equal(np.tile(fs, len(ts)), r0.nuobs)):
            raise ValueError("observer pairs are not a t-major tensor grid")
        return ts, fs

    def observer_vec(self, rtol=1e-4):
        """Adaptive-method light curves by the frequency-vectorised observer kernel
        (pbagpu_eats_lightcurve_vec): one Gauss-Kronrod tree per (layer, t_obs) shared by all nu_obs,
        refined to rtol on every frequency.  Writes the same d_out layout as observer()."""
        if self.method == "piece-wise":
            raise NotImplementedError("the frequency-vectorised observer stage is for the adaptive method")
        if not hasattr(self, "d_t_u"):
            ts, fs = self.tensor_grid()
            self.d_t_u = self.torch.from_numpy(np.ascontiguousarray(ts)).to(self.dev)
            self.d_nu_u = self.torch.from_numpy(np.ascontiguousarray(fs)).to(self.dev)
        _check(self.L.pbagpu_eats_lightcurve_vec(self.params, self.n_runs, self.larr, len(self.layers),
                                                 _ptr(self.d_table), self.nt, _ptr(self.d_jall), self.nf,
                                                 len(self.d_t_u), _ptr(self.d_t_u), len(self.d_nu_u), _ptr(self.d_nu_u),
                                                 float(rtol), _ptr(self.d_out), _ptr(self.d_lst), self.rsp),
               "eats_lightcurve_vec")

    def skymaps(self, times, freqs, remove_mu=False):
        """Raw piece-wise sky maps of a one-run batch at every (freq, time), frequency-major as
        Ejecta::computeSaveEjectaSkyImages orders them (src/ejecta/ejecta.h:259-312).

        Returns (image [n_images][PBA_IMG_NV][2 ncells], flux [n_layers][n_images], status [n_layers])."""
        if self.method != "piece-wise":
            raise NotImplementedError("sky maps are implemented for the piece-wise EATS method only")
        if self.n_runs != 1:
            raise ValueError("sky maps are computed for one run at a time")
        torch = self.torch
        times = np.asarray(times, dtype=np.float64)
        freqs = np.asarray(freqs, dtype=np.float64)
        t_img = np.tile(times, len(freqs))
        nu_img = np.repeat(freqs, len(times))
        n_img = len(t_img)
        nl = len(self.layers)
        ncells = sum(self.larr[i].cil for i in range(nl))
        d_t = torch.from_numpy(t_img).to(self.dev)
        d_nu = torch.from_numpy(nu_img).to(self.dev)
        d_img = torch.empty((n_img, abi.PBA_IMG_NV, 2 * ncells), dtype=torch.float64, device=self.dev)
        d_flux = torch.empty((nl, n_img), dtype=torch.float64, device=self.dev)
        d_st = torch.zeros(nl, dtype=torch.int32, device=self.dev)
        _check(self.L.pbagpu_eats_skymap_pw(self.params, self.larr, nl, _ptr(self.d_table), self.nt, _ptr(self.d_jall),
                                            _ptr(self.d_aall),
                                            self.nf, n_img, _ptr(d_t), _ptr(d_nu), 1 if remove_mu else 0,
                                            _ptr(d_img), _ptr(d_flux), _ptr(d_st), self.rsp), "eats_skymap_pw")
        return d_img.cpu().numpy(), d_flux.cpu().numpy(), d_st.cpu().numpy()

    def step(self):
        self.dynamics()
        self.electrons()
        self.observer()

    def _on_rad_stream(self):
        """torch context on the stream of the electron and observer stages.  Copies to the host
        must not go through the null stream: it waits for every blocking stream, and so for the
        dynamics kernels of other batches that run on CU-masked streams meanwhile."""
        if self.rsp is None or not self.rsp.value:
            return self.torch.cuda.stream(self.torch.cuda.current_stream())
        return self.torch.cuda.stream(self.torch.cuda.ExternalStream(self.rsp.value))

    def status(self):
        npop = len(self.pops)
        with self._on_rad_stream():
            return (np.array([r.status for r in self.res]), self.d_est[:npop].cpu().numpy(),
                    self.d_lst[:len(self.layers)].cpu().numpy())

    def flux(self):
        """Light curves [n_runs][n_layers][n_pairs] of the last observer stage, with each run's EBL
        attenuation applied (EATS::evalFluxDensA, src/eats.h:849-853)."""
        with self._on_rad_stream():
            out = self.d_out.cpu().numpy().reshape(self.n_runs, self.bpr, self.npairs)
        return apply_attenuation(out, self.runs)


def radiate(pl, want_intermediates=False, precision=None):
    """Electron and observer stages of a Pipeline whose dynamics ran (or whose tables were given);
    returns (flux [n_runs, n_layers, n_pairs], extras) as lightcurves() does."""
    if precision is not None:
        pl.precision = precision
    pl.electrons(want_f=want_intermediates)
    pl.observer()
    dst, est, lst = pl.status()
    flux = pl.flux()
    extras = dict(dyn_status=dst, ele_status=est, eats_status=lst)
    if want_intermediates:
        npop = len(pl.pops)
        extras.update(aall=pl.d_aall[:npop].cpu().numpy() if pl.ssa else None)
        extras.update(table=pl.d_table.cpu().numpy(), jall=pl.d_jall[:npop].cpu().numpy(),
                      fall=pl.d_fall.cpu().numpy(), pops=pl.pops, layers=pl.layers,
                      results=[(r.status, r.i_end_r, r.fsrs, r.n_steps, r.n_rhs) for r in pl.res])
    return flux, extras


def lightcurves(runs, stream=None, want_intermediates=False, precision=0):
    """Run the three GPU stages for a list of Runs with equal grid sizes.

    Returns (flux [n_runs, n_layers, n_pairs] numpy array, extras dict)."""
    pl = Pipeline(runs, stream=stream, precision=precision)
    pl.dynamics()
    return radiate(pl, want_intermediates)
