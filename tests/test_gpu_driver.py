"""The drop-in driver (python -m pyblastafterglowmag_amd.pba <workdir> <parfile>) on the GPU:
it reads the parfile and initial data the reference reads and writes lc.h5 / dyn.h5 / spec.h5 in
the reference's layout with the reference's values (within the GPU-path tolerances)."""
import os

import numpy as np
import pytest

from conftest import load_golden, rel_err
from pyblastafterglowmag_amd import abi, h5io, inputs, workloads

pytestmark = pytest.mark.gpu


def _case(name):
    return workloads.small(name) if name in workloads.SMALL else workloads.WORKLOADS[name]()


@pytest.mark.parametrize("name", ["tophat_small", "tophat_fs_syn", "tophat_pw_small"])
def test_driver_outputs_match_reference(name, tmp_path):
    from pyblastafterglowmag_amd import pba
    g = load_golden(name)
    main, grb, struct = _case(name)
    grb = dict(grb, save_dynamics="yes", save_spec="yes")
    wd = str(tmp_path) + "/"
    inputs.write_run(wd, main, grb, struct)
    print(f"[gpu] driver on {name}", flush=True)
    assert pba.main([wd, "parfile.par", "err"]) == 0
    lc = wd + "lc.h5"
    assert sorted(h5io.members(lc, "/")) == sorted(str(x) for x in g["lc_root_members"])
    np.testing.assert_array_equal(h5io.read(lc, "times"), g["times"])
    for il in range(len(g["lc"])):
        e = rel_err(h5io.read(lc, f"shell=0 layer={il}/fluxdens"), g["lc"][il], 1e-30)
        assert e <= 1e-6, (name, il, e)
    dyn = wd + "dyn.h5"
    assert sorted(h5io.members(dyn, "shell=0 layer=0")) == sorted(str(x) for x in g["dyn_layer_members"])
    for iv, v in enumerate(str(x) for x in g["dyn_names"]):
        e = rel_err(h5io.read(dyn, f"shell=0 layer=0/{v}"), g["dyn"][0, iv], 1e-300)
        assert e <= 2e-6, (name, v, e)
    spec = wd + "spec.h5"
    every = int(g["spec_every"])
    P = abi.run_params_from_parfile(wd + "parfile.par")[0]
    nt, nf = int(main.get("ntb", 1000)), P.fs.nfreq
    j = h5io.read(spec, "shell=0 layer=0/total_j_fs").reshape(nt, nf)[::every]
    assert rel_err(j, g["jall_fs"][0], 1e-12) <= 1e-6
    # every dataset the reference writes into a layer group, and the values of those the fixture holds
    assert sorted(h5io.members(spec, "shell=0 layer=0")) == sorted(str(x) for x in g["spec_layer_members"])
    for ds in ("syn_j", "syn_i", "total_f", "total_i", "gam_dot_syn", "gam_dot_adi"):
        key = f"spec_{ds}_fs"
        if key in g:
            v = h5io.read(spec, f"shell=0 layer=0/{ds}_fs").reshape(nt, -1)[::int(g["spec_full_every"])]
            assert rel_err(v, g[key][0], 1e-12) <= 1e-6, (name, ds)


def test_driver_raw_skymaps_match_reference(tmp_path):
    """do_skymap = yes with save_raw_skymap = yes on the piece-wise top-hat jet: the driver writes
    raw_skymap_<i>.h5 with the reference's arrays and attributes."""
    from pyblastafterglowmag_amd import pba
    name = "tophat_pw_small"
    g = load_golden(name)
    main, grb, struct = _case(name)
    main = dict(main, skymap_times="array 1e5 1e7", skymap_freqs="array 1e9 1e18")
    grb = dict(grb, do_skymap="yes", save_raw_skymap="yes")
    wd = str(tmp_path) + "/"
    inputs.write_run(wd, main, grb, struct)
    print(f"[gpu] driver sky maps on {name}", flush=True)
    assert pba.main([wd, "parfile.par", "err"]) == 0
    for i in range(int(g["sky_n"])):
        f = wd + f"raw_skymap_{i}.h5"
        assert sorted(h5io.members(f, "/")) == sorted(str(x) for x in g["sky_root_members"])
        assert sorted(h5io.attr_names(f, "/")) == sorted(str(x) for x in g["sky_root_attrs"])
        for v in abi.IMG_NAMES:
            e = rel_err(h5io.read(f, f"shell=0/{v}"), g[f"sky{i}_{v}"], 1e-300)
            assert e <= 1e-6, (i, v, e)
        for k, ref in zip(["flux", "nshells", "nlayers", "time", "freq", "xmin", "ymin", "xmax", "ymax"],
                          g[f"sky{i}_attrs"]):
            assert h5io.attr(f, k) == pytest.approx(ref, rel=1e-6, abs=0), (i, k)


def test_driver_processes_skymaps(tmp_path):
    """--process-skymaps: the raw maps binned into skymap.h5 (one group per map, the histogram image
    on the package's default 64 x 32 grid, the centroid) as the reference package does after a run."""
    import numpy as np
    from pyblastafterglowmag_amd import pba
    from pyblastafterglowmag_amd.skymap_process import entries_of
    name = "tophat_pw_small"
    main, grb, struct = _case(name)
    main = dict(main, skymap_times="array 1e5 1e7", skymap_freqs="array 1e9 1e18")
    grb = dict(grb, do_skymap="yes", save_raw_skymap="yes")
    wd = str(tmp_path) + "/"
    inputs.write_run(wd, main, grb, struct)
    print(f"[gpu] driver sky-map processing on {name}", flush=True)
    assert pba.main([wd, "parfile.par", "err", "--process-skymaps"]) == 0
    out = wd + "skymap.h5"
    assert np.array_equal(h5io.read(out, "times"), [1e5, 1e7, 1e5, 1e7])
    for i, (t, nu) in enumerate([(1e5, 1e9), (1e7, 1e9), (1e5, 1e18), (1e7, 1e18)]):
        grp = "time={:.4e} freq={:.4e}".format(t, nu)
        f = wd + f"raw_skymap_{i}.h5"
        ents = entries_of([tuple(h5io.read(f, f"shell=0/{v}") for v in ("xrs", "yrs", "intensity"))])
        x, y, w = (np.concatenate([e[k] for e in ents]) for k in range(3))
        hist = h5io.read(out, f"{grp}/image_hist")
        assert hist.size == 63 * 31
        # every point lies inside the extent-based grid, so the histogram holds the total intensity
        assert hist.sum() == pytest.approx(w.sum(), rel=1e-12)
        assert h5io.attr(out, "xc", grp) == pytest.approx(np.sum(w * x) / np.sum(w), rel=1e-12, abs=1e-12 * np.abs(x).max())
