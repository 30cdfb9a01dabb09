"""The device's electron and observer stages on kilonova ejecta blastwaves, against the reference
program's per-(shell, layer) light curves (tests/golden/kn_*.npz).

The device's dynamics stage has no ejecta right-hand side (rhs_fs_dense) yet, so the dynamics
tables come from the CPU oracle, which reproduces the reference's ejecta dynamics bit for bit
(tests/test_kilonova_cpu.py).  From there the device computes the electron spectra, the
emissivities and the observer integral of every blastwave through the C ABI
(pbagpu_ele_evolve, pbagpu_eats_lightcurve / _pw), as for a jet's layers.
"""
import numpy as np
import pytest

from conftest import load_golden, rel_err
from test_kilonova_cpu import kn_golden_run
import pyoracle
from pyblastafterglowmag_amd import abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["kn_small", "kn_pw_small"])
def test_device_radiation_of_ejecta_matches_reference(name):
    """12 ejecta blastwaves (4 velocity shells x 3 layers up to pi/2, SSC, 300 steps): every
    (shell, layer) light curve within 1e-6 of the reference's (the jets' end-to-end bound), and
    the emissivities within 1e-9 of the oracle's on the same tables."""
    from pyblastafterglowmag_amd import gpu
    g = load_golden(name)
    P, spread, inits, order, tgrid, tobs, nuobs, method = kn_golden_run(g)
    table, res = pyoracle.dyn_evolve_ejecta(P, inits, tgrid, spread=spread)
    assert all(r.status == 0 for r in res)
    print(f"[gpu] {name}: electron and observer stages of {len(order)} ejecta blastwaves ({method})", flush=True)
    run = gpu.Run(P, inits, tgrid, tobs, nuobs, method)
    pl = gpu.Pipeline([run])
    pl.use_tables(table, res)
    flux, ex = gpu.radiate(pl, want_intermediates=True, precision=0)
    assert not np.asarray(ex["ele_status"]).any() and not np.asarray(ex["eats_status"]).any()
    pops = (abi.Population * len(order))(*[abi.Population(k, 0, 0, 0) for k in range(len(order))])
    jall_o, _, st = pyoracle.ele_evolve(P, pops, table)
    assert not st.any()
    for q in range(len(order)):
        e = rel_err(ex["jall"][q], jall_o[q], 1e-9)
        assert e <= 1e-9, (name, q, e)
    worst = 0.0
    for k, (ish, il) in enumerate(order):
        e = rel_err(flux[0, k], g["lc"][ish, il], 1e-30)
        worst = max(worst, e)
        assert e <= 1e-6, (name, ish, il, e)
    print(f"[gpu]   light curves vs the reference: {worst:.2e}", flush=True)
