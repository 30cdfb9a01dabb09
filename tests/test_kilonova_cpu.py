"""Kilonova ejecta (the parfile's kN section, no jet) on the CPU oracle against the reference
program (tests/golden/kn_*.npz written by tests/golden/make_golden.py from oracle/_ref/pba.out).

Every ejecta blastwave is evolved by the reference's ejecta right-hand side (setStandardISM +
rhs_fs_dense, src/model_evolve.h:1627-1651) in one ODE system with a shared adaptive step; the
electrons and the observer integral are those of the jet's layers.  The oracle reproduces the
reference bit for bit: every stored dynamical quantity of every (shell, layer) blastwave, and
the light curve of each.  The host-side pieces (initial-data writer, the state-vector order,
option checks) are tested against what the reference's loader reads.
"""
import os
import types
import tempfile

import numpy as np
import pytest

from conftest import dyn_excess, load_golden, rel_err
import pyoracle
from pyblastafterglowmag_amd import abi, h5io, inputs, output, workloads


def kn_golden_run(g):
    """(params, spread, inits, order, tgrid, tobs, nuobs, method) rebuilt from a kilonova fixture's
    stored parfile and 2-D initial data."""
    text = bytes(g["parfile"]).decode()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "parfile.par")
        with open(p, "w") as f:
            f.write(text)
        mp_, mo = inputs.read_parfile(p, inputs.SEC_MAIN)
        kp, ko = inputs.read_parfile(p, inputs.SEC_KN)
        assert inputs.read_parfile(p, inputs.SEC_GRB) == ({}, {})   # no jet in the run
    P, spread = abi.run_params_ejecta(mp_, mo, kp, ko)
    idd = {k: g["id_" + k] for k in ("r", "mom", "theta", "ctheta", "ek", "mass")}
    method = ko.get("method_eats", "adaptive")
    inits, order = abi.ejecta_inits(idd, method)
    tobs, nuobs = abi.obs_pairs(mo) if "lc_times" in mo else (None, None)
    return P, spread, inits, order, abi.tgrid(mp_), tobs, nuobs, method


def _observer_layers(inits, res, method):
    n = len(inits)
    if method == "piece-wise":
        return (abi.EatsPwLayer * n)(*[abi.EatsPwLayer(inits[k].ctheta0, inits[k].ncells,
                                                      abi.cells_in_layer(inits[k].ilayer), k, 0, k, 0,
                                                      res[k].i_end_r) for k in range(n)])
    return (abi.EatsLayer * n)(*[abi.EatsLayer(inits[k].theta_c_l, inits[k].theta_c_h, k, 0, k, -1, res[k].i_end_r)
                                 for k in range(n)])


@pytest.mark.parametrize("name", ["kn_small", "kn_pw_small", "kn_spread_small", "kn_wind_small", "kn_rk4_small"])
def test_oracle_ejecta_matches_reference(name):
    """4 velocity shells x 3 layers (12 blastwaves, 300 steps, SSC on): dynamics of every
    blastwave and the light curve of every (shell, layer) equal the reference's bit for bit.
    Cases: adaptive and piece-wise observer integrals; method_spread set (the spreading form of
    dt_obs/dR); a wind-like medium (the stored density is the last RHS call's); the RK4 integrator
    (the stored GammaRel is the Gamma of RK4's last stage, not the step's end)."""
    g = load_golden(name)
    P, spread, inits, order, tgrid, tobs, nuobs, method = kn_golden_run(g)
    assert spread == (name == "kn_spread_small") and P.fs.ssc == 1
    assert method == ("piece-wise" if "pw" in name else "adaptive") and (P.n_ism < 0) == (name == "kn_wind_small")
    assert P.integrator == (1 if name == "kn_rk4_small" else 0)
    np.testing.assert_array_equal(tobs, g["times"])
    np.testing.assert_array_equal(nuobs, g["freqs"])
    table, res = pyoracle.dyn_evolve_ejecta(P, inits, tgrid, spread=spread)
    assert all(r.status == 0 for r in res)
    names = [str(n) for n in g["dyn_names"]]
    for k, (ish, il) in enumerate(order):
        for iv, n in enumerate(names):
            if n in abi.PV:   # the upstream's constant columns are the writer's (test below)
                np.testing.assert_array_equal(table[k, abi.PV[n]], g["dyn"][ish, il, iv],
                                              err_msg=f"{name} {ish} {il} {n}")
    pops = (abi.Population * len(order))(*[abi.Population(k, 0, 0, 0) for k in range(len(order))])
    jall, _, st, aall = pyoracle.ele_evolve(P, pops, table, want_a=True)
    assert not st.any()
    eats = pyoracle.eats_lightcurve_pw if method == "piece-wise" else pyoracle.eats_lightcurve
    lc, st2 = eats(P, _observer_layers(inits, res, method), table, jall, tobs, nuobs, aall=aall)
    assert not st2.any()
    for k, (ish, il) in enumerate(order):
        assert rel_err(lc[k], g["lc"][ish, il], 1e-30) == 0.0, (name, ish, il)
    # the summed light curve is dominated by the fast shells and is not zero
    assert np.all(g["lc"].sum(axis=(0, 1)) > 0)
    if name == "kn_rk4_small":
        k = order.index((3, 2))
        assert np.sum(table[k, abi.PV["GammaRel"]] != table[k, abi.PV["Gamma"]]) > 290
    if name == "kn_wind_small":   # the stored density follows the blastwave through the profile
        k = order.index((3, 2))
        rho = table[k, abi.PV["rho"]]
        assert len(np.unique(rho)) > 10 and np.any(table[k, abi.PV["drhodr"]] < 0)


def _same_dynamics_inputs(P, Q):
    """Run parameters equal but for the microphysics grids and SSC, which the ejecta RHS never
    reads (rhs_fs_dense has no electron or radiative-loss integral term)."""
    Q2 = abi.RunParams.from_buffer_copy(Q)
    for sh in (Q2.fs, Q2.rs):
        sh.ngam, sh.nfreq, sh.ssc = P.fs.ngam, P.fs.nfreq, P.fs.ssc
    return bytes(Q2) == bytes(P)


def test_oracle_full_size_ejecta_match_reference():
    """The C5 ejecta at full size: 32 shells x 30 layers = 960 blastwaves in one ODE system, 1000
    output steps (about 11,000 shared adaptive steps), against two reference runs:
      - kn_c5_dyn, the C5 parfile itself (128 gamma x 128 nu, SSC), dynamics only;
      - kn_c5_lc, the same ejecta at reduced microphysics grids (64 gamma x 48 nu, no SSC, 6 x 8
        pairs) with every blastwave's electrons and light curve.
    The two share every input of the dynamics.  The oracle evolves all 960 blastwaves once; the
    dynamics of 12 blastwaves spread over the shells and layers (every 10th step, the fixtures'
    subset) equal both runs' bit for bit, and the electron limits and light curves of those 12
    equal kn_c5_lc's (the ejecta RHS has no electron term, so the other blastwaves' electrons do
    not act on them).  About 16 s of one core."""
    g = load_golden("kn_c5_lc")
    P, spread, inits, order, tgrid, tobs, nuobs, method = kn_golden_run(g)
    assert len(inits) == 960 and len(tgrid) == 1000 and P.fs.ssc == 0 and P.fs.ngam == 64
    gd = load_golden("kn_c5_dyn")
    Pd, spread_d, inits_d, order_d, tgrid_d, *_ = kn_golden_run(gd)
    assert Pd.fs.ngam == 128 and Pd.fs.nfreq == 128 and Pd.fs.ssc == 1 and spread_d == spread
    assert _same_dynamics_inputs(P, Pd) and np.array_equal(tgrid_d, tgrid) and order_d == order
    assert bytes(inits_d) == bytes(inits)
    table, res = pyoracle.dyn_evolve_ejecta(P, inits, tgrid, spread=spread)
    assert all(r.status == 0 for r in res) and res[0].n_steps > 10000
    names_d = [str(n) for n in gd["dyn_names"]]
    every_d = int(gd["dyn_every"])
    for i, ish in enumerate(gd["dyn_shells"]):
        for j, il in enumerate(gd["dyn_layers"]):
            k = order.index((int(ish), int(il)))
            for iv, n in enumerate(names_d):
                np.testing.assert_array_equal(table[k, abi.PV[n], ::every_d], gd["dyn"][i, j, iv],
                                              err_msg=f"C5 dynamics: shell {ish} layer {il} {n}")
    names = [str(n) for n in g["dyn_names"]]
    every = int(g["dyn_every"])
    keep = [(int(ish), int(il)) for ish in g["dyn_shells"] for il in g["dyn_layers"]]
    ks = [order.index(sl) for sl in keep]
    for (ish, il), k in zip(keep, ks):
        i, j = list(g["dyn_shells"]).index(ish), list(g["dyn_layers"]).index(il)
        for iv, n in enumerate(names):
            np.testing.assert_array_equal(table[k, abi.PV[n], ::every], g["dyn"][i, j, iv],
                                          err_msg=f"shell {ish} layer {il} {n}")
    pops = (abi.Population * len(ks))(*[abi.Population(k, 0, 0, 0) for k in ks])
    jall, _, st = pyoracle.ele_evolve(P, pops, table)
    assert not st.any()
    lays = (abi.EatsLayer * len(ks))(*[abi.EatsLayer(inits[k].theta_c_l, inits[k].theta_c_h, k, 0, q, -1,
                                                     res[k].i_end_r) for q, k in enumerate(ks)])
    lc, st2 = pyoracle.eats_lightcurve(P, lays, table, jall, tobs, nuobs)
    assert not st2.any()
    for q, (ish, il) in enumerate(keep):
        i, j = list(g["dyn_shells"]).index(ish), list(g["dyn_layers"]).index(il)
        assert rel_err(lc[q], g["lc"][i, j], 1e-30) == 0.0, (ish, il)


def test_ejecta_state_order_is_layer_major():
    """The reference's state vector holds, for each layer, its velocity shells in order
    (src/model_evolve.h:225-234); the shared step's error norm sums in that order."""
    idd = inputs.ejecta_structure(3, 2, 0.1, 3.0, 1e48, 1e51, seed=1)
    inits, order = abi.ejecta_inits(idd)
    assert order == [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]
    for k, (ish, il) in enumerate(order):
        assert inits[k].E0 == idd["ek"][ish, il] and inits[k].M0 == idd["mass"][ish, il]
        assert inits[k].mom0 == idd["mom"][ish, il] and inits[k].ilayer == il and inits[k].nlayers == 2
        assert inits[k].ncells == 1.0
    # piece-wise: energies and masses per cell, every blastwave starts at theta_wing
    pw, _ = abi.ejecta_inits(idd, "piece-wise")
    assert pw[3].E0 == idd["ek"][0, 1] / 3 and pw[3].theta_b0 == pytest.approx(np.pi / 2) and pw[3].ncells == 4.0
    # a blastwave without mass, energy or velocity is not evolved (its E0 = M0 = 0)
    idd["ek"][2, 1] = 0.0
    idd["mom"][1, 0] = 1e-7
    z, _ = abi.ejecta_inits(idd)
    assert z[5].E0 == z[5].M0 == 0.0 and z[1].E0 == z[1].M0 == 0.0 and z[0].E0 > 0


def test_ejecta_structure_layout():
    """The C5 ejecta: 32 shells with beta*Gamma log-spaced over [0.1, 3], each with one energy drawn
    log-uniform over [1e48, 1e51]; layer j holds the energy of the cone up to theta_h, E (1 - cos
    theta_h) / 2, so the cone to pi/2 holds half of it; mass E / ((Gamma - 1) c^2)."""
    main, kn, struct = workloads.kn_c5()
    d = inputs.ejecta_structure(**struct)
    assert d["ek"].shape == (32, 30)
    mom = d["mom"][:, 0]
    assert mom[0] == pytest.approx(0.1) and mom[-1] == pytest.approx(3.0)
    np.testing.assert_allclose(np.diff(np.log(mom)), np.log(30.0) / 31)
    e_shell = 2.0 * d["ek"][:, -1]
    assert np.all((e_shell >= 1e48) & (e_shell <= 1e51))
    assert len(np.unique(np.round(np.log10(e_shell), 6))) == 32   # a draw per shell
    gam = np.sqrt(mom * mom + 1)
    np.testing.assert_allclose(d["mass"] * (gam[:, None] - 1) * inputs.C_LIGHT ** 2, d["ek"], rtol=1e-14)
    np.testing.assert_allclose(d["theta"][0], np.arange(1, 31) * (np.pi / 2) / 30, rtol=1e-15)
    assert np.all(np.diff(d["ek"], axis=1) > 0)
    # the same seed gives the same ejecta
    np.testing.assert_array_equal(inputs.ejecta_structure(**struct)["ek"], d["ek"])


def test_kn_run_files_read_back(tmp_path):
    """write_kn_run: a parfile with the main and kN sections only, and the 2-D initial data in
    the use_1d_id = no layout the reference's loader reads (datasets [nshells][nlayers])."""
    main, kn, struct = workloads.kn_small("kn_small")
    wd = str(tmp_path) + "/"
    inputs.write_kn_run(wd, main, kn, struct)
    kp, ko = inputs.read_parfile(wd + "parfile.par", inputs.SEC_KN)
    assert ko["bw_type"] == "fs_dense" and ko["use_1d_id"] == "no" and ko["method_spread"] == "None"
    assert inputs.read_parfile(wd + "parfile.par", inputs.SEC_GRB) == ({}, {})
    d = inputs.ejecta_structure(**struct)
    for k in ("mom", "ek", "mass", "theta", "ctheta", "r"):
        np.testing.assert_array_equal(h5io.read(wd + ko["fname_ejecta_id"], k).reshape(4, 3), d[k])
    # the fixture was written by this writer (with the dynamics file on)
    g = load_golden("kn_small")
    np.testing.assert_array_equal(g["id_ek"], d["ek"])
    inputs.write_kn_run(wd, main, dict(kn, save_dynamics="yes", save_spec="no"), struct)
    assert bytes(g["parfile"]).decode() == open(wd + "parfile.par").read()


@pytest.mark.parametrize("key,value", [("bw_type", "fs"), ("method_dgdr", "peer"), ("method_collision", "simple"),
                                       ("do_nucinj", "yes"), ("use_1d_id", "yes"), ("method_eos", "Peer12")])
def test_ejecta_options_outside_the_path_are_refused(tmp_path, key, value):
    main, kn, struct = workloads.kn_small("kn_small")
    wd = str(tmp_path) + "/"
    inputs.write_kn_run(wd, main, kn, struct)
    mp_, mo = inputs.read_parfile(wd + "parfile.par", inputs.SEC_MAIN)
    kp, ko = inputs.read_parfile(wd + "parfile.par", inputs.SEC_KN)
    abi.run_params_ejecta(mp_, mo, kp, ko)
    with pytest.raises(NotImplementedError):
        abi.run_params_ejecta(mp_, mo, kp, dict(ko, **{key: value}))


def test_ejecta_entry_refuses_a_reverse_shock():
    g = load_golden("kn_small")
    P, spread, inits, order, tgrid, *_ = kn_golden_run(g)
    P.fsrs = 1
    with pytest.raises(RuntimeError):
        pyoracle.dyn_evolve_ejecta(P, inits, tgrid)


def test_jet_path_refuses_a_kilonova_parfile(tmp_path):
    """The device pipeline evolves jets only: a parfile whose kN section runs ejecta is refused
    with a clear error, not read as a jet."""
    main, kn, struct = workloads.kn_small("kn_small")
    wd = str(tmp_path) + "/"
    inputs.write_kn_run(wd, main, kn, struct)
    with pytest.raises(NotImplementedError, match="kilonova"):
        abi.run_params_from_parfile(wd + "parfile.par")
    inputs.write_kn_run(wd, main, dict(kn, run_bws="no"), struct)
    with pytest.raises(ValueError, match="no jet"):
        abi.run_params_from_parfile(wd + "parfile.par")


def test_c5_split_by_layer_stays_within_the_light_curve_bound():
    """C5's 960 ejecta blastwaves evolved as 30 runs of 32 (one layer's shells each, one shared
    step per run), as a device launch of at most 64 blastwaves per run does, against the
    reference's single shared step (kn_c5_lc): splitting moves the step sequence, so the
    dynamics part at the ODE tolerance (beta: 6.8e-7) and the light curves of the 12 kept
    blastwaves by at most 1.8e-12, far inside the 1e-6 light-curve bound."""
    g = load_golden("kn_c5_lc")
    P, spread, inits, order, tgrid, tobs, nuobs, method = kn_golden_run(g)
    params = (abi.RunParams * 30)(*[P] * 30)
    table, res = pyoracle.dyn_evolve_ejecta(params, inits, tgrid, spread=spread, n_runs=30)
    assert all(r.status == 0 for r in res)
    assert len({res[32 * i].n_steps for i in range(30)}) > 1   # every run takes its own steps
    names = [str(n) for n in g["dyn_names"]]
    keep = [(int(ish), int(il)) for ish in g["dyn_shells"] for il in g["dyn_layers"]]
    ks = [order.index(sl) for sl in keep]
    for (ish, il), k in zip(keep, ks):
        i, j = list(g["dyn_shells"]).index(ish), list(g["dyn_layers"]).index(il)
        gam = g["dyn"][i, j, names.index("Gamma")]
        for n in ("R", "Gamma", "beta", "M2", "Eint2", "U_p", "B"):
            e = dyn_excess(table[k, abi.PV[n], ::10], g["dyn"][i, j, names.index(n)], gam, rtol=1e-5)
            assert e <= 1.0, (ish, il, n, e)
    pops = (abi.Population * len(ks))(*[abi.Population(k, k // 32, 0, 0) for k in ks])
    jall, _, st = pyoracle.ele_evolve(params, pops, table, n_runs=30)
    assert not st.any()
    lays = (abi.EatsLayer * len(ks))(*[abi.EatsLayer(inits[k].theta_c_l, inits[k].theta_c_h, k, k // 32, q, -1,
                                                     res[k].i_end_r) for q, k in enumerate(ks)])
    lc, st2 = pyoracle.eats_lightcurve(params, lays, table, jall, tobs, nuobs, n_runs=30)
    assert not st2.any()
    for q, (ish, il) in enumerate(keep):
        i, j = list(g["dyn_shells"]).index(ish), list(g["dyn_layers"]).index(il)
        assert rel_err(lc[q], g["lc"][i, j], 1e-30) <= 1e-10, (ish, il)


def _sections(g, tmp_path):
    p = tmp_path / "parfile.par"
    p.write_bytes(bytes(g["parfile"]))
    mp_, mo = inputs.read_parfile(str(p), inputs.SEC_MAIN)
    kp, ko = inputs.read_parfile(str(p), inputs.SEC_KN)
    return mp_, mo, kp, ko


@pytest.mark.parametrize("name", ["kn_small", "kn_pw_small"])
def test_ejecta_output_files_follow_the_reference(name, tmp_path):
    """lc_kn.h5 and dyn_kn.h5 as the reference writes them for ejecta: one group per (shell, layer)
    blastwave, the ejecta type's datasets (the upstream's columns included), every attribute
    name and value, and the files' attributes; checked against the reference's files of the same
    run (the light curves written are the reference's own; the dynamics a synthetic table)."""
    g = load_golden(name)
    P, spread, inits, order, tgrid, tobs, nuobs, method = kn_golden_run(g)
    mp_, mo, kp, ko = _sections(g, tmp_path)
    path = str(tmp_path / "lc_kn.h5")
    flux = np.stack([g["lc"][ish, il] for ish, il in order])
    output.write_lightcurves_ejecta(path, P, inits, order, tgrid, tobs, nuobs, flux, mp_, kp)
    assert sorted(h5io.members(path, "/")) == sorted(str(x) for x in g["lc_root_members"])
    assert sorted(h5io.attr_names(path, "/")) == sorted(str(x) for x in g["lc_root_attrs"])
    names = [str(x) for x in g["lc_layer_attr_names"]]
    for ish, il in order:
        grp = f"shell={ish} layer={il}"
        assert sorted(h5io.attr_names(path, grp)) == sorted(names)
        for key, ref in zip(names, g["lc_layer_attrs"][ish, il]):
            assert h5io.attr(path, key, grp) == pytest.approx(ref, rel=1e-12, abs=0), (grp, key)
        np.testing.assert_array_equal(h5io.read(path, grp + "/fluxdens"), g["lc"][ish, il])
    np.testing.assert_array_equal(h5io.read(path, "times"), g["times"])
    assert h5io.attr(path, "nshells") == 4 and h5io.attr(path, "n_ism") == P.n_ism

    path = str(tmp_path / "dyn_kn.h5")
    nt = len(tgrid)
    table = np.zeros((len(order), abi.PBA_NV, nt))
    table[:, abi.PV["R"]] = 1e15
    table[:, abi.PV["Gamma"]] = 2.0
    output.write_dynamics_ejecta(path, P, inits, order, tgrid, table, mp_, kp)
    assert sorted(h5io.attr_names(path, "/")) == sorted(str(x) for x in g["dyn_root_attrs"])
    assert [h5io.attr(path, k) for k in ("nshells", "nlayers", "ntimes", "ncells")] == list(g["dyn_root_counts"])
    names = [str(x) for x in g["dyn_layer_attr_names"]]
    dyn_names = [str(n) for n in g["dyn_names"]]
    for ish, il in order:
        grp = f"shell={ish} layer={il}"
        assert sorted(h5io.members(path, grp)) == sorted(str(x) for x in g["dyn_layer_members"])
        for key, ref in zip(names, g["dyn_layer_attrs"][ish, il]):
            assert h5io.attr(path, key, grp) == pytest.approx(ref, rel=1e-12, abs=0), (grp, key)
        np.testing.assert_array_equal(h5io.read(path, grp + "/Gamma"), np.full(nt, 2.0))
        for n in output._DENSE_UPSTREAM:
            np.testing.assert_array_equal(h5io.read(path, grp + "/" + n), g["dyn"][ish, il, dyn_names.index(n)],
                                          err_msg=f"{grp} {n}")


class _StandInPipeline:
    """The device pipeline's interface with recognisable results: blastwave k of the batch gets
    Gamma = 2 + k and light curve 1000 k + pair index (the driver's placement is what is tested)."""
    made = []

    def __init__(self, runs, precision=0, **kw):
        self.runs = runs
        n = sum(r.n_layers for r in runs)
        self.n, nt, self.npairs = n, len(runs[0].tgrid), len(runs[0].tobs)
        self.res = [abi.BwResult() for _ in range(n)]
        t = np.zeros((n, abi.PBA_NV, nt))
        t[:, abi.PV["R"]] = 1e15
        for k in range(n):
            t[k, abi.PV["Gamma"]] = 2.0 + k
        self._table = t
        self.d_table = types.SimpleNamespace(cpu=lambda: types.SimpleNamespace(numpy=lambda: self._table))
        self.spread = None
        _StandInPipeline.made.append(self)

    def dynamics_ejecta(self, spread):
        self.spread = spread

    def electrons(self, **kw):
        pass

    def observer(self, **kw):
        pass

    def status(self):
        return np.zeros(self.n, int), np.zeros(self.n, int), np.zeros(self.n, int)

    def flux(self):
        k = np.arange(self.n)[:, None]
        return (1000.0 * k + np.arange(self.npairs)[None, :]).reshape(len(self.runs), -1, self.npairs)


@pytest.mark.parametrize("shells,layers,n_runs", [(4, 3, 1), (13, 5, 5)])
def test_driver_places_every_ejecta_blastwave(tmp_path, monkeypatch, shells, layers, n_runs):
    """pba on a kilonova parfile: one device run when the blastwaves fit one wavefront, else one
    run per layer (its shells); every blastwave's light curve and dynamics land in its own
    (shell, layer) group."""
    from pyblastafterglowmag_amd import gpu, pba
    _StandInPipeline.made.clear()
    monkeypatch.setattr(gpu, "Pipeline", _StandInPipeline)
    main, kn, struct = workloads.kn_small("kn_small")
    struct = dict(struct, n_shells=shells, n_layers=layers)
    wd = str(tmp_path) + "/"
    inputs.write_kn_run(wd, main, dict(kn, save_dynamics="yes", method_spread="our"), struct)
    assert pba.main([wd, "parfile.par", "err"]) == 0
    pl = _StandInPipeline.made[-1]
    assert len(pl.runs) == n_runs and pl.spread is True
    if n_runs > 1:   # one layer per run, its shells in order
        for r, run in enumerate(pl.runs):
            assert {run.inits[b].ilayer for b in range(run.n_layers)} == {r}
    d = inputs.ejecta_structure(**struct)
    _, order = abi.ejecta_inits(d)
    for k, (ish, il) in enumerate(order):
        grp = f"shell={ish} layer={il}"
        np.testing.assert_array_equal(h5io.read(wd + kn["fname_light_curve"], grp + "/fluxdens"),
                                      1000.0 * k + np.arange(pl.npairs))
        assert h5io.read(wd + kn["fname_dyn"], grp + "/Gamma")[0] == 2.0 + k
        assert h5io.attr(wd + kn["fname_dyn"], "E0", grp) == d["ek"][ish, il]
    assert h5io.attr(wd + kn["fname_light_curve"], "nshells") == shells


def test_driver_refuses_what_the_ejecta_path_does_not_run(tmp_path, monkeypatch, capsys):
    from pyblastafterglowmag_amd import gpu, pba
    monkeypatch.setattr(gpu, "Pipeline", _StandInPipeline)
    main, kn, struct = workloads.kn_small("kn_small")
    wd = str(tmp_path) + "/"
    # more than 64 velocity shells per layer: no run fits one wavefront
    inputs.write_kn_run(wd, main, kn, dict(struct, n_shells=65, n_layers=2))
    assert pba.main([wd, "parfile.par", "err"]) == 1
    assert "more than 64" in capsys.readouterr().err
    # sky maps of ejecta
    inputs.write_kn_run(wd, main, dict(kn, do_skymap="yes"), struct)
    assert pba.main([wd, "parfile.par", "err"]) == 1
    # a jet and ejecta in one parfile
    inputs.write_parfile(wd, main, dict(run_bws="yes"), kn=dict(kn, run_bws="yes"))
    assert pba.main([wd, "parfile.par", "err"]) == 1
    assert "jet and kilonova ejecta" in capsys.readouterr().err
