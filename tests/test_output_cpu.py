"""Output files in the reference program's layout (pyblastafterglowmag_amd.output): group and
dataset names, attribute names and the per-layer attribute values, checked against what the
reference program itself wrote for the same parfiles (recorded in the golden fixtures)."""
import os

import numpy as np
import pytest

from conftest import golden_run, load_golden
from pyblastafterglowmag_amd import abi, h5io, output

CASES = ["tophat_small", "gauss5_small", "tophat_fs_syn", "tophat_pw_small", "gauss_pw_small"]


def _parse(g, tmp_path):
    p = tmp_path / "parfile.par"
    p.write_bytes(bytes(g["parfile"]))
    return abi.run_params_from_parfile(str(p))


@pytest.mark.parametrize("name", CASES)
def test_lightcurve_file_layout(name, tmp_path):
    g = load_golden(name)
    run = golden_run(g)
    P, mp_, mo, gp, go = _parse(g, tmp_path)
    path = str(tmp_path / "lc.h5")
    output.write_lightcurves(path, run, g["lc"], mp_, gp)
    assert sorted(h5io.members(path, "/")) == sorted(str(x) for x in g["lc_root_members"])
    assert sorted(h5io.attr_names(path, "/")) == sorted(str(x) for x in g["lc_root_attrs"])
    names = [str(x) for x in g["lc_layer_attr_names"]]
    assert sorted(h5io.attr_names(path, "shell=0 layer=0")) == sorted(names)
    for il in range(run.n_layers):
        for k, ref in zip(names, g["lc_layer_attrs"][il]):
            mine = h5io.attr(path, k, f"shell=0 layer={il}")
            assert mine == pytest.approx(ref, rel=1e-12, abs=0), (name, il, k, mine, ref)
        np.testing.assert_array_equal(h5io.read(path, f"shell=0 layer={il}/fluxdens"), g["lc"][il])
    np.testing.assert_array_equal(h5io.read(path, "times"), g["times"])
    np.testing.assert_array_equal(h5io.read(path, "freqs"), g["freqs"])
    assert h5io.attr(path, "n_ism") == P.n_ism


@pytest.mark.parametrize("name", CASES)
def test_dynamics_file_layout(name, tmp_path):
    g = load_golden(name)
    run = golden_run(g)
    P, mp_, mo, gp, go = _parse(g, tmp_path)
    nt = len(run.tgrid)
    table = np.zeros((run.n_layers, abi.PBA_NV, nt))
    table[:, abi.PV["Gamma"], :] = 2.0
    fsrs = [P.fsrs] * run.n_layers
    path = str(tmp_path / "dyn.h5")
    output.write_dynamics(path, run, table, fsrs, mp_, gp)
    assert sorted(h5io.members(path, "shell=0 layer=0")) == sorted(str(x) for x in g["dyn_layer_members"])
    assert sorted(h5io.attr_names(path, "shell=0 layer=0")) == sorted(str(x) for x in g["dyn_layer_attr_names"])
    assert sorted(h5io.attr_names(path, "/")) == sorted(str(x) for x in g["dyn_root_attrs"])
    np.testing.assert_array_equal(h5io.read(path, "shell=0 layer=0/Gamma"), np.full(nt, 2.0))
    # per-layer values (E0, M0, theta_b0, cil, ...) and the file's counts (ncells = n^2)
    names = [str(x) for x in g["dyn_layer_attr_names"]]
    for il in range(run.n_layers):
        for k, ref in zip(names, g["dyn_layer_attrs"][il]):
            mine = h5io.attr(path, k, f"shell=0 layer={il}")
            assert mine == pytest.approx(ref, rel=1e-12, abs=0), (name, il, k, mine, ref)
    counts = [h5io.attr(path, k) for k in ("nshells", "nlayers", "ntimes", "ncells")]
    np.testing.assert_array_equal(counts, g["dyn_root_counts"])


@pytest.mark.parametrize("name", CASES + ["tophat_ssa_small", "gauss5_ssa_small"])
def test_spectrum_file_grids(name, tmp_path):
    """The spectrum file's root (grids times_freqs/freqs/times_gams/gams and the layer groups) and the
    datasets of layer 0 are exactly the reference's names, for both shocks and with or without SSC
    and SSA."""
    g = load_golden(name)
    run = golden_run(g)
    P, mp_, mo, gp, go = _parse(g, tmp_path)
    nt, nf, ng = len(run.tgrid), P.fs.nfreq, P.fs.ngam
    pops = []
    for i in range(run.n_layers):
        pops.append(abi.Population(i, 0, 0, 0))
        if i in [int(x) for x in g.get("spec_layers_rs", [])]:
            pops.append(abi.Population(i, 0, 1, 0))
    path = str(tmp_path / "spec.h5")
    zn = np.zeros((len(pops), nt, nf))
    spectra = {k: zn for k in abi.SPECTRA_NU} | {k: np.zeros((len(pops), nt, ng)) for k in abi.SPECTRA_GAM}
    output.write_spectra(path, run, pops, zn, np.zeros((len(pops), nt, ng)), output.ele_grid(P.fs),
                         output.freq_grid(P.fs), mp_, gp, spectra=spectra, aall=zn)
    assert sorted(h5io.members(path, "/")) == sorted(str(x) for x in g["spec_root_members"])
    assert sorted(h5io.members(path, "shell=0 layer=0")) == sorted(str(x) for x in g["spec_layer_members"])


@pytest.mark.parametrize("name", ["tophat_pw_small", "gauss_pw_small"])
def test_raw_skymap_file_layout(name, tmp_path):
    """raw_skymap_<i>.h5: root members, root attribute names and the shell group's datasets as the
    reference program wrote them, from the reference's own arrays."""
    g = load_golden(name)
    run = golden_run(g)
    P, mp_, mo, gp, go = _parse(g, tmp_path)
    n = int(g["sky_n"])
    images = [dict(arrays={v: g[f"sky{i}_{v}"] for v in abi.IMG_NAMES}, fluxs=g[f"sky{i}_fluxs"],
                   attrs=dict(zip(["flux", "nshells", "nlayers", "xmin", "ymin", "xmax", "ymax"],
                                  g[f"sky{i}_attrs"][[0, 1, 2, 5, 6, 7, 8]])))
              for i in range(n)]
    paths = output.write_raw_skymaps(str(tmp_path), images, [1e5, 1e7], [1e9, 1e18], mp_, gp)
    assert [os.path.basename(p) for p in paths] == [f"raw_skymap_{i}.h5" for i in range(n)]
    for i, path in enumerate(paths):
        assert sorted(h5io.members(path, "/")) == sorted(str(x) for x in g["sky_root_members"])
        assert sorted(h5io.attr_names(path, "/")) == sorted(str(x) for x in g["sky_root_attrs"])
        assert sorted(h5io.members(path, "shell=0")) == sorted(abi.IMG_NAMES)
        assert h5io.attr(path, "time") == [1e5, 1e7][i % 2] and h5io.attr(path, "freq") == [1e9, 1e18][i // 2]
        np.testing.assert_array_equal(h5io.read(path, "shell=0/intensity"), g[f"sky{i}_intensity"])


def _image(pj, cj):
    """One raw image [1][PBA_IMG_NV][2 nc]: positions spread on a grid, intensities as given."""
    nc = len(pj)
    img = np.zeros((1, abi.PBA_IMG_NV, 2 * nc))
    for v in range(abi.PBA_IMG_NV):
        img[0, v] = np.arange(2 * nc) + 10.0 * v
    img[0, abi.IMG["xrs"]] = np.linspace(-2.0, 3.0, 2 * nc)
    img[0, abi.IMG["yrs"]] = np.linspace(1.0, -1.0, 2 * nc) ** 3
    img[0, abi.IMG["intensity"]] = np.concatenate([pj, cj])
    return img


def test_skymap_zero_cell_removal_and_normalisation():
    """computeEjectaSkyMapPW's normalisation and removeZeroEntries (src/image.h:120-190) on a
    hand-made image: cells at or below 1e-10 of the peak in both jets are dropped (both jets'
    entries of a kept cell stay), intensities scale by flux / (dx dy) / max I, bounds span all cells."""
    pj = np.array([1.0, 0.0, 1e-12, 2.0, 0.0])
    cj = np.array([0.0, 0.0, 3.0, 0.0, 1e-20])
    img = _image(pj, cj)
    flux = np.array([[0.25], [0.5]])   # two layers
    out = output.skymap_images(img, flux, 2)[0]
    keep = np.array([0, 2, 3])
    X, Y = img[0, abi.IMG["xrs"]], img[0, abi.IMG["yrs"]]
    dfnu = 0.75 / ((X.max() - X.min()) * (Y.max() - Y.min()))
    for n in abi.IMG_NAMES:
        raw = img[0, abi.IMG[n]]
        want = np.concatenate([raw[keep], raw[5 + keep]])
        if n == "intensity":
            want = want * (dfnu / 3.0)
        assert np.array_equal(out["arrays"][n], want), n
    assert out["fluxs"][0] == 0.75
    a = out["attrs"]
    assert a["flux"] == 0.75 and a["nshells"] == 1.0 and a["nlayers"] == 2.0
    assert (a["xmin"], a["xmax"], a["ymin"], a["ymax"]) == (X.min(), X.max(), Y.min(), Y.max())


def test_skymap_one_full_jet_keeps_every_cell():
    """If one jet has every cell above the threshold nothing is removed (the reference moves the arrays)."""
    pj = np.array([1.0, 2.0, 3.0])
    cj = np.array([0.0, 0.0, 0.0])
    out = output.skymap_images(_image(pj, cj), np.array([[1.0]]), 1)[0]
    assert len(out["arrays"]["intensity"]) == 6


def test_skymap_without_positive_flux_is_not_normalised():
    """A non-positive (or non-finite) shell flux leaves the intensities as computed and the bounds at
    their start values (numeric_limits<double>::max() / ::min())."""
    pj = np.array([1.0, 0.0])
    cj = np.array([0.5, 0.0])
    img = _image(pj, cj)
    for fl in (0.0, -1.0, np.nan):
        out = output.skymap_images(img, np.array([[fl]]), 1)[0]
        assert np.array_equal(out["arrays"]["intensity"], np.array([1.0, 0.5]))
        a = out["attrs"]
        assert a["xmin"] == np.finfo(float).max and a["xmax"] == np.finfo(float).tiny
        assert a["flux"] == fl or (np.isnan(fl) and np.isnan(a["flux"]))
