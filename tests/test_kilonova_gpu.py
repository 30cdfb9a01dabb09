"""Kilonova ejecta on the device, against the CPU oracle and the reference program's
per-(shell, layer) light curves (tests/golden/kn_*.npz).

* The electron and observer stages on the oracle's ejecta dynamics tables (bit-exact with the
  reference's, tests/test_kilonova_cpu.py): pbagpu_ele_evolve, pbagpu_eats_lightcurve / _pw as
  for a jet's layers.
* The whole path on the device: pbagpu_dyn_evolve_ejecta (rhs_fs_dense for every blastwave of a
  run, one wavefront per run), then the same two stages.
"""
import numpy as np
import pytest

from conftest import dyn_excess, load_golden, rel_err
from test_kilonova_cpu import kn_golden_run
import pyoracle
from pyblastafterglowmag_amd import abi, inputs, workloads

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


# dynamics compared with the oracle's (the jet tests' set plus the ejecta's upstream columns)
DYN_CHECK = ["R", "Gamma", "beta", "M2", "Eint2", "tt", "tcomov", "rho2", "GammaFsh", "thickness", "U_p",
             "gamma_min", "gamma_max", "B", "accel_frac", "rho", "drhodr", "GammaRel"]
# light curves: the jets' end-to-end bound (measured on one MI355X: 7e-15 .. 6.8e-11); dynamics: the
# jets' conditioning rule at rtol 1e-4 (measured: at most 0.61 % of it)
LC_TOL = 1e-6
DYN_RTOL = 1e-4


@pytest.mark.parametrize("name", ["kn_small", "kn_pw_small", "kn_spread_small", "kn_wind_small", "kn_rk4_small"])
def test_device_ejecta_path_matches_oracle_and_reference(name):
    """The whole ejecta path on the device: pbagpu_dyn_evolve_ejecta, then the electron and
    observer stages.  Dynamics against the oracle's (conftest.dyn_excess: rtol plus 64 ulp of
    Gamma carried into Gamma - 1; gamma_min to the 2e-3 resolution of the reference's bisection),
    equal status bits and last usable steps; light curves against the reference's."""
    from pyblastafterglowmag_amd import gpu
    g = load_golden(name)
    P, spread, inits, order, tgrid, tobs, nuobs, method = kn_golden_run(g)
    table_o, res_o = pyoracle.dyn_evolve_ejecta(P, inits, tgrid, spread=spread)
    print(f"[gpu] {name}: ejecta dynamics, electrons and observer stage of {len(order)} blastwaves "
          f"({method}, spread={spread}, integrator={P.integrator})", flush=True)
    run = gpu.Run(P, inits, tgrid, tobs, nuobs, method)
    pl = gpu.Pipeline([run])
    pl.dynamics_ejecta(spread)
    flux, ex = gpu.radiate(pl, want_intermediates=True, precision=0)
    assert not np.asarray(ex["ele_status"]).any() and not np.asarray(ex["eats_status"]).any()
    worst_dyn = 0.0
    for k in range(len(order)):
        st, iend = ex["results"][k][0], ex["results"][k][1]
        assert st == res_o[k].status == 0 and iend == res_o[k].i_end_r, (name, k, st, iend)
        gam = table_o[k, abi.PV["Gamma"]]
        for n in DYN_CHECK:
            e = dyn_excess(ex["table"][k, abi.PV[n]], table_o[k, abi.PV[n]], gam,
                           rtol=2e-3 if n == "gamma_min" else DYN_RTOL)
            worst_dyn = max(worst_dyn, e)
            assert e <= 1.0, (name, k, n, e)
    worst = 0.0
    for k, (ish, il) in enumerate(order):
        e = rel_err(flux[0, k], g["lc"][ish, il], 1e-30)
        worst = max(worst, e)
        assert e <= LC_TOL, (name, ish, il, e)
    print(f"[gpu]   dynamics vs oracle: {worst_dyn:.3e} of the bound; light curves vs the reference: {worst:.3e}",
          flush=True)


def test_device_ejecta_refuses_runs_beyond_one_wavefront():
    """A run of more than 64 ejecta blastwaves (C5's 960) is refused before any launch."""
    from pyblastafterglowmag_amd import gpu
    main, kn, struct = workloads.kn_small("kn_small")
    d = inputs.ejecta_structure(**dict(struct, n_shells=13, n_layers=5))
    inits, order = abi.ejecta_inits(d)
    P = kn_golden_run(load_golden("kn_small"))[0]
    run = gpu.Run(P, inits, abi.inputs.logspace_vec(3, 14, 20), np.array([1e5]), np.array([1e9]))
    pl = gpu.Pipeline([run])
    with pytest.raises(RuntimeError, match="64 blastwaves"):
        pl.dynamics_ejecta(False)


def test_device_c5_ejecta_as_one_run_per_layer():
    """C5's 960 ejecta blastwaves (kn_c5_lc: 1000 steps, reduced microphysics grids, 6 x 8 pairs) on
    the device as a batch of 30 runs of 32 blastwaves, each run one layer's velocity shells:
    dynamics, electrons and observer stage.  The light curves of the 12 blastwaves the fixture
    keeps against the reference's single-step run, 1e-6; the split alone moves them by at most
    1.8e-12 (tests/test_kilonova_cpu.py)."""
    from pyblastafterglowmag_amd import gpu
    g = load_golden("kn_c5_lc")
    P, spread, inits, order, tgrid, tobs, nuobs, method = kn_golden_run(g)
    assert len(inits) == 960
    print("[gpu] kn_c5_lc: 960 ejecta blastwaves as 30 runs of 32", flush=True)
    runs = [gpu.Run(P, (abi.LayerInit * 32)(*inits[32 * i:32 * (i + 1)]), tgrid, tobs, nuobs, method)
            for i in range(30)]
    pl = gpu.Pipeline(runs)
    pl.dynamics_ejecta(spread)
    flux, ex = gpu.radiate(pl, precision=0)
    assert not np.asarray(ex["dyn_status"]).any() and not np.asarray(ex["ele_status"]).any()
    assert not np.asarray(ex["eats_status"]).any()
    worst = 0.0
    for i, ish in enumerate(g["dyn_shells"]):
        for j, il in enumerate(g["dyn_layers"]):
            r, b = divmod(order.index((int(ish), int(il))), 32)
            e = rel_err(flux[r, b], g["lc"][i, j], 1e-30)
            worst = max(worst, e)
            assert e <= LC_TOL, (int(ish), int(il), e)
    print(f"[gpu]   light curves of 12 blastwaves vs the reference: {worst:.3e}", flush=True)


def test_driver_on_a_kilonova_parfile(tmp_path):
    """pba <workdir> <parfile> on the kn_small parfile (ejecta only): the files it writes against
    the reference's for the same parfile: light curves 1e-6, dynamics by the conditioning rule."""
    from pyblastafterglowmag_amd import h5io, pba
    g = load_golden("kn_small")
    P, spread, inits, order, tgrid, tobs, nuobs, method = kn_golden_run(g)
    main, kn, struct = workloads.kn_small("kn_small")
    wd = str(tmp_path) + "/"
    inputs.write_kn_run(wd, main, dict(kn, save_dynamics="yes", save_spec="no"), struct)
    assert open(wd + "parfile.par").read() == bytes(g["parfile"]).decode()
    print("[gpu] driver on the kn_small parfile", flush=True)
    assert pba.main([wd, "parfile.par", "err"]) == 0
    names = [str(n) for n in g["dyn_names"]]
    for ish, il in order:
        grp = f"shell={ish} layer={il}"
        e = rel_err(h5io.read(wd + kn["fname_light_curve"], grp + "/fluxdens"), g["lc"][ish, il], 1e-30)
        assert e <= LC_TOL, (grp, e)
        gam = g["dyn"][ish, il, names.index("Gamma")]
        for n in ("R", "Gamma", "beta", "M2", "Eint2", "U_p", "B", "GammaRel", "GammaRho", "MCBM"):
            mine = h5io.read(wd + kn["fname_dyn"], grp + "/" + n)
            assert dyn_excess(mine, g["dyn"][ish, il, names.index(n)], gam, rtol=DYN_RTOL) <= 1.0, (grp, n)
