"""EBL attenuation of the light curves (pyblastafterglowmag_amd.ebl): the optical-depth lookup
against the reference's own test points and an independent bilinear interpolator, its edge rules,
and the oracle's light curve of a z = 0.7 top-hat with the attenuation applied against the
reference program run with the same table (tests/golden/tophat_ebl_small.npz)."""
import types

import numpy as np
import pytest

from conftest import golden_run, load_golden, rel_err
import pyoracle
from pyblastafterglowmag_amd import abi, ebl, gpu


@pytest.fixture(scope="module")
def table():
    g = load_golden("tophat_ebl_small")
    return ebl.EblTable(g["ebl_freq"], g["ebl_z"], g["ebl_tau"])


def test_reference_test_points(table):
    """EBL::test_cases (src/microphysics/common.h:84-88) states three values of the interpolation;
    the frequencies there have 9 significant digits, so agreement to 1e-8 is what they fix."""
    for nu, z, want in [(4.17852721e+25, 0.5, 0.62492), (4.46885518e+27, 1.0, 414.81),
                        (4.46885518e+27, 2.0, 1776.4800000000002)]:
        assert table.tau(nu, z)[0] == pytest.approx(want, rel=1e-8), (nu, z)


def test_interior_matches_an_independent_bilinear_interpolator(table):
    from scipy.interpolate import RegularGridInterpolator
    f = RegularGridInterpolator((table.freq, table.z), table.tau_table, method="linear")
    rng = np.random.default_rng(3)
    nus = np.exp(rng.uniform(np.log(table.freq[0]), np.log(table.freq[-1]), 200))
    for z in rng.uniform(table.z[0], table.z[-1], 20):
        got = table.tau(nus, z)
        want = f(np.stack([nus, np.full_like(nus, z)], axis=1))
        scale = np.maximum(np.abs(want), 1e-12)
        assert np.all(np.abs(got - want) <= 1e-12 * scale + 1e-14), z
    # at the grid's nodes the table's own values (the cell is the one below the node)
    for i in (1, 10, 30, len(table.freq) - 1):
        for j in (1, 4, len(table.z) - 1):
            assert table.tau(table.freq[i], table.z[j])[0] == pytest.approx(table.tau_table[i, j], rel=1e-14)


def test_edge_rules(table):
    """Below the table's frequencies or redshifts tau = 0; above them the last value is used."""
    assert table.tau(table.freq[0] * 0.999, 1.0)[0] == 0.0
    assert table.tau(1e26, table.z[0] * 0.5)[0] == 0.0
    assert table.tau(table.freq[-1] * 10, 1.0)[0] == table.tau(table.freq[-1], 1.0)[0]
    assert table.tau(1e26, table.z[-1] * 2)[0] == table.tau(1e26, table.z[-1])[0]
    np.testing.assert_array_equal(table.attenuation([1e9, 1e15], 1.0), [1.0, 1.0])
    with pytest.raises(ValueError):
        ebl.EblTable([1.0, 0.5], [0.1, 0.2], np.zeros(4))


def test_table_path_resolution(tmp_path, monkeypatch):
    assert ebl.resolve_table_path("none") is None
    with pytest.raises(FileNotFoundError):
        ebl.resolve_table_path("no/such/table.h5", str(tmp_path))
    (tmp_path / "t.h5").write_bytes(b"")
    assert ebl.resolve_table_path("t.h5", str(tmp_path)) == str(tmp_path / "t.h5")
    monkeypatch.chdir(tmp_path)
    assert ebl.resolve_table_path("t.h5") == "t.h5"


def test_oracle_light_curve_with_ebl_matches_reference():
    """A z = 0.7 top-hat with SSC (12 frequencies from 1e9 to 1e28 Hz, 3 inside the table's range):
    the oracle's light curve times exp(-tau) against the reference program's lc.h5 with the same
    table.  The oracle is bit-exact on this run without the attenuation; the factor is computed in
    numpy here and in C++ there, so the bound is 1e-13."""
    g = load_golden("tophat_ebl_small")
    run = golden_run(g)
    assert run.attenuation is not None and np.any(run.attenuation < 0.9) and np.all(run.attenuation <= 1)
    table, res = pyoracle.dyn_evolve(run.params, run.inits, run.tgrid)
    pops = (abi.Population * 1)(abi.Population(0, 0, 0, 0))
    jall, _, st = pyoracle.ele_evolve(run.params, pops, table)
    assert not st.any()
    li = run.inits[0]
    lays = (abi.EatsLayer * 1)(abi.EatsLayer(li.theta_c_l, li.theta_c_h, 0, 0, 0, -1, res[0].i_end_r))
    lc, st2 = pyoracle.eats_lightcurve(run.params, lays, table, jall, run.tobs, run.nuobs)
    assert not st2.any()
    got = gpu.apply_attenuation(lc[None], [run])[0]
    assert rel_err(got[0], g["lc"][0], 1e-30) <= 1e-13
    # the unattenuated pairs agree bit for bit; pairs inside the comparison (above 1e-30 of the
    # peak: the TeV flux density is ~1e-16 of the radio peak) are attenuated by up to 1e4
    raw, ref = lc[0], g["lc"][0]
    on = run.attenuation < 1.0
    assert np.array_equal(raw[~on], ref[~on])
    strong = on & (run.attenuation < 0.5) & (ref > 1e-30 * ref.max())
    assert strong.any()
    assert np.all(raw[strong] > 2 * ref[strong])


def test_apply_attenuation_leaves_runs_without_a_table():
    runs = [types.SimpleNamespace(attenuation=None), types.SimpleNamespace(attenuation=np.array([0.5, 1.0, 0.0]))]
    flux = np.ones((2, 3, 3))
    out = gpu.apply_attenuation(flux, runs)
    np.testing.assert_array_equal(out[0], flux[0])
    np.testing.assert_array_equal(out[1], np.tile([0.5, 1.0, 0.0], (3, 1)))
    assert gpu.apply_attenuation(flux, runs[:1]) is flux


REF_TABLE = "/root/reference/data/EBL/Franceschini18/table.h5"


def test_run_from_workdir_reads_the_table(tmp_path):
    """A parfile naming a table (the reference package's default ebl_tbl_fpath names its own) is
    read with that table: the run's attenuation equals the fixture's."""
    import os
    from pyblastafterglowmag_amd import inputs, workloads
    if not os.path.exists(REF_TABLE):
        pytest.skip("the reference's EBL table is not on this host")
    main, grb, struct = workloads.small("tophat_ebl_small")
    grb = dict(grb, ebl_tbl_fpath=REF_TABLE)
    inputs.write_run(str(tmp_path) + "/", main, grb, struct)
    run = gpu.Run.from_workdir(str(tmp_path) + "/")
    want = golden_run(load_golden("tophat_ebl_small"))
    np.testing.assert_array_equal(run.nuobs, want.nuobs)
    np.testing.assert_array_equal(run.attenuation, want.attenuation)
