import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm device) and the built HIP library")


def load_golden(name):
    path = os.path.join(GOLDEN, f"{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} missing")
    return dict(np.load(path, allow_pickle=False))


def golden_run(g):
    """Run (pyblastafterglowmag_amd.gpu.Run) rebuilt from a golden fixture's stored parfile + ID."""
    import tempfile
    from pyblastafterglowmag_amd import abi, gpu, inputs
    text = bytes(g["parfile"]).decode()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "parfile.par")
        with open(p, "w") as f:
            f.write(text)
        P, mp_, mo, gp, go = abi.run_params_from_parfile(p)
    idd = dict(mom=g["id_mom"], ctheta=g["id_ctheta"], ek=g["id_ek"], mass=g["id_mass"],
               theta_wing=float(g["id_theta_wing"]), theta_core=float(g["id_theta_core"]))
    tobs, nuobs = abi.obs_pairs(mo)
    method = go.get("method_eats", "adaptive")
    # an EBL case carries the optical-depth table it was run with
    from pyblastafterglowmag_amd import ebl
    tab = ebl.EblTable(g["ebl_freq"], g["ebl_z"], g["ebl_tau"]) if "ebl_tau" in g else None
    return gpu.Run(P, abi.layer_inits(idd, method), abi.tgrid(mp_), tobs, nuobs, method, ebl_table=tab)


def rel_err(a, b, floor_frac=0.0):
    """max |a-b| / |b| over entries where |b| > floor_frac * max|b| (and a != b)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.abs(b) > floor_frac * np.max(np.abs(b)) if b.size else np.zeros(0, bool)
    m &= (a != b)
    if not np.any(m):
        return 0.0
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m])))


def dyn_excess(mine, ref, gamma_ref, rtol=1e-4, ulps=64):
    """Largest |mine - ref| / ((rtol + ulps * eps / (Gamma - 1)) |ref|) over the elements (<= 1
    passes).  In the late, nearly Newtonian phase every quantity formed from Gamma - 1 (beta, the
    internal energy, the swept-up mass through the energy equation) carries the error of Gamma
    divided by Gamma - 1: at Gamma - 1 = 2e-12 a 12-ulp difference of Gamma is a 1e-3 relative
    difference of beta.  The bound is rtol plus a difference of `ulps` ulp of Gamma carried through
    that division; zeros must match exactly."""
    mine, ref, g = np.asarray(mine), np.asarray(ref), np.asarray(gamma_ref)
    tol = rtol + ulps * np.finfo(float).eps / np.maximum(g - 1.0, 1e-300)
    diff = np.abs(mine - ref)
    den = tol * np.abs(ref)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, diff / np.where(den > 0, den, 1.0), np.where(diff > 0, np.inf, 0.0))
    return float(np.max(r)) if r.size else 0.0


def rs_end_steps(gamma43):
    """Per layer: the last step at which the reverse shock is on (Gamma43 != 0), or -1."""
    out = []
    for row in np.atleast_2d(gamma43):
        nz = np.nonzero(row != 0)[0]
        out.append(int(nz[-1]) if len(nz) else -1)
    return out


# Session caches of the GPU tests: a golden fixture's Pipeline after its dynamics stage (the
# reverse-shock runs take tens of seconds of one-wavefront latency each) and the oracle's
# dynamics of the same run; tests rerun the electron and observer stages on them as they need.
# At the start of a session with GPU tests both are computed ahead, in the order the selected
# tests use them: GPU dynamics of up to 4 fixtures at a time, each on its own stream (the library
# runs calls on different streams from different host threads side by side; 4 is the process's
# hardware queue count), and the oracle's dynamics on 6 host threads.
_GPU_PIPELINES = {}
_ORACLE_DYN = {}
_GPU_FUT = {}
_ORACLE_FUT = {}
_POOLS = []

# tests that take a fixture's Pipeline from gpu_pipeline (by their `name` parameter) or its
# oracle dynamics from oracle_dynamics, and the fixed fixtures of tests without that parameter
_PIPELINE_TESTS = {"test_stages_match_oracle", "test_lightcurve_matches_reference", "test_mixed_precision_ssc_lightcurve",
                   "test_skymaps_match_oracle_and_reference", "test_vec_observer_matches_plain_and_reference",
                   "test_cpp_caller_matches_reference_and_python"}
_PIPELINE_FIXED = {"test_full_size_c2_matches_reference_and_oracle": "gauss21_fsrs_ssc",
                   "test_electron_graph_matches_eager": "gauss5_small",
                   "test_surrogate_on_a_run_and_fallback": "gauss5_small"}
_ORACLE_TESTS = {"test_stages_match_oracle", "test_spectra_match_reference"}


def _make_pipeline(name):
    import torch
    from pyblastafterglowmag_amd import gpu
    run = golden_run(load_golden(name))
    pl = gpu.Pipeline([run], stream=torch.cuda.Stream())
    pl.dynamics()
    return run, pl


def _make_oracle(name):
    import pyoracle
    run = golden_run(load_golden(name))
    return pyoracle.dyn_evolve(run.params, run.inits, run.tgrid)


@pytest.fixture(scope="session", autouse=True)
def _prefetch_dynamics(request):
    items = [it for it in request.session.items if it.get_closest_marker("gpu")]
    pipes, orcs = [], []
    for it in items:
        fn = it.originalname if hasattr(it, "originalname") else it.name
        name = getattr(it, "callspec", None) and it.callspec.params.get("name")
        if fn in _PIPELINE_TESTS and name and name not in pipes:
            pipes.append(name)
        if fn in _PIPELINE_FIXED and _PIPELINE_FIXED[fn] not in pipes:
            pipes.append(_PIPELINE_FIXED[fn])
        if fn in _ORACLE_TESTS and name and name not in orcs:
            orcs.append(name)
    started = False
    if items:
        try:
            import torch
            started = torch.cuda.is_available()
        except ImportError:
            started = False
    if started:
        import concurrent.futures as cf
        from pyblastafterglowmag_amd import gpu
        gpu.lib()   # load and declare the library once, before the worker threads call it
        gpool, opool = cf.ThreadPoolExecutor(max_workers=4), cf.ThreadPoolExecutor(max_workers=6)
        _POOLS.extend([gpool, opool])
        for n in pipes:
            _GPU_FUT[n] = gpool.submit(_make_pipeline, n)
        for n in orcs:
            _ORACLE_FUT[n] = opool.submit(_make_oracle, n)
        print(f"[conftest] dynamics computed ahead: GPU {pipes}, oracle {orcs}", flush=True)
    yield
    for p in _POOLS:
        p.shutdown(wait=True, cancel_futures=True)
    _POOLS.clear()


def gpu_pipeline(name):
    """(run, Pipeline) of golden fixture `name` with its GPU dynamics done (computed once, on its
    own stream)."""
    if name not in _GPU_PIPELINES:
        fut = _GPU_FUT.pop(name, None)
        _GPU_PIPELINES[name] = fut.result() if fut is not None else _make_pipeline(name)
    return _GPU_PIPELINES[name]


def oracle_dynamics(name):
    """The oracle's (table, results) of golden fixture `name` (computed once)."""
    if name not in _ORACLE_DYN:
        fut = _ORACLE_FUT.pop(name, None)
        _ORACLE_DYN[name] = fut.result() if fut is not None else _make_oracle(name)
    return _ORACLE_DYN[name]
