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
    return gpu.Run(P, abi.layer_inits(idd, method), abi.tgrid(mp_), tobs, nuobs, method)


def rel_err(a, b, floor_frac=0.0):
    """max |a-b| / |b| over entries where |b| > floor_frac * max|b| (and a != b)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.abs(b) > floor_frac * np.max(np.abs(b)) if b.size else np.zeros(0, bool)
    m &= (a != b)
    if not np.any(m):
        return 0.0
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m])))
