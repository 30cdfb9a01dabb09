"""Boundary checks that need no GPU: the C-ABI library loads and exports every function that
include/pbagpu.h declares, the ctypes mirror has the header's exact struct layout, and the
parfile reader fills the structures the way the reference's readers do."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from pyblastafterglowmag_amd import abi, inputs, workloads

HEADER = os.path.join(ROOT, "include", "pbagpu.h")
LIB = os.path.join(ROOT, "pyblastafterglowmag_amd", "_lib", "libpbagpu.so")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pbagpu_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared_functions()
    assert "pbagpu_dyn_evolve" in names and "pbagpu_ele_evolve" in names and "pbagpu_eats_lightcurve" in names
    assert "pbagpu_eats_lightcurve_pw" in names
    assert len(names) >= 6
    from pyblastafterglowmag_amd import gpu
    assert set(names) == set(gpu.EXPORTS)


def test_library_exports_every_declared_function():
    if not os.path.exists(LIB):
        pytest.skip("libpbagpu.so not built")
    lib = C.CDLL(LIB)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(pbagpu_\w+)", out))
    assert set(_declared_functions()) <= exported


def test_version_needs_no_device():
    if not os.path.exists(LIB):
        pytest.skip("libpbagpu.so not built")
    lib = C.CDLL(LIB)
    assert lib.pbagpu_version() > 0



def test_library_options_need_no_device():
    """pbagpu_set_option / pbagpu_get_option: PBAGPU_OPT_ELE_GRAPH takes 0 or 1 and
    PBAGPU_OPT_DYN_RUNS_PER_WAVE 0 .. 64 (both default 0); an unknown option or value is
    PBAGPU_E_ARG and leaves the setting alone."""
    if not os.path.exists(LIB):
        pytest.skip("libpbagpu.so not built")
    lib = C.CDLL(LIB)
    lib.pbagpu_set_option.argtypes = [C.c_int, C.c_longlong]
    lib.pbagpu_get_option.argtypes = [C.c_int, C.POINTER(C.c_longlong)]
    v = C.c_longlong(-5)
    assert lib.pbagpu_get_option(0, C.byref(v)) == 0 and v.value == 0
    assert lib.pbagpu_set_option(0, 2) == -1 and lib.pbagpu_set_option(0, -1) == -1
    assert lib.pbagpu_set_option(2, 0) == -1 and lib.pbagpu_get_option(2, C.byref(v)) == -1
    assert lib.pbagpu_get_option(1, C.byref(v)) == 0 and v.value == 0
    assert lib.pbagpu_set_option(1, 65) == -1 and lib.pbagpu_set_option(1, -1) == -1
    assert lib.pbagpu_set_option(1, 3) == 0 and lib.pbagpu_get_option(1, C.byref(v)) == 0 and v.value == 3
    assert lib.pbagpu_set_option(1, 0) == 0
    assert lib.pbagpu_set_option(0, 1) == 0 and lib.pbagpu_get_option(0, C.byref(v)) == 0 and v.value == 1
    assert lib.pbagpu_set_option(0, 0) == 0 and lib.pbagpu_get_option(0, C.byref(v)) == 0 and v.value == 0

def test_skymap_hist_rejects_bad_arguments_before_any_device_call():
    """pbagpu_skymap_hist checks its host arrays first: PBAGPU_E_ARG (-1) with a message, no HIP call."""
    if not os.path.exists(LIB):
        pytest.skip("libpbagpu.so not built")
    lib = C.CDLL(LIB)
    lib.pbagpu_last_error.restype = C.c_char_p
    f = lib.pbagpu_skymap_hist
    f.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_long)] + [C.c_void_p] * 3 + \
        [C.c_int, C.c_int, C.c_double, C.c_int] + [C.c_void_p] * 7
    dummy = C.c_void_p(8)   # never dereferenced: every case fails validation first
    seg = (C.c_int * 2)(0, 1)
    off = (C.c_long * 2)(0, 4)

    def call(n=1, s=seg, o=off, nx=5, ny=5, eg=1.0):
        return f(n, s, o, dummy, dummy, dummy, nx, ny, eg, 0, dummy, dummy, dummy, None, dummy, dummy, None)
    assert call(n=0) == -1
    assert call(nx=1) == -1 and call(ny=1) == -1 and call(nx=3000, ny=1200) == -1
    assert call(s=(C.c_int * 2)(1, 1)) == -1 and b"img_seg[0]" in lib.pbagpu_last_error()
    assert call(s=(C.c_int * 2)(0, -1)) == -1
    assert call(o=(C.c_long * 2)(0, -3)) == -1 and b"seg_off decreases" in lib.pbagpu_last_error()
    assert call(eg=float("nan")) == -1


def test_bessel_host_routine_argument_checks():
    if not os.path.exists(LIB):
        pytest.skip("libpbagpu.so not built")
    lib = C.CDLL(LIB)
    lib.pbagpu_bessel_k_thirds.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    assert lib.pbagpu_bessel_k_thirds(-1, None, None, None) == -1
    assert lib.pbagpu_bessel_k_thirds(3, None, None, None) == -1
    assert lib.pbagpu_bessel_k_thirds(0, None, None, None) == 0


_STRUCTS = {
    "pba_shock_params": abi.ShockParams,
    "pba_run_params": abi.RunParams,
    "pba_layer_init": abi.LayerInit,
    "pba_bw_result": abi.BwResult,
    "pba_population": abi.Population,
    "pba_eats_layer": abi.EatsLayer,
    "pba_eats_pw_layer": abi.EatsPwLayer,
}


def test_struct_layout_matches_header(tmp_path):
    """sizeof/offsetof of every boundary struct, compiled from the header by the host C compiler."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in _STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append(f'printf("PBA_NV %d\\n", (int)PBA_NV);')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln.strip():
            parts = ln.split()
            got[tuple(parts[:-1])] = int(parts[-1])
    assert got[("PBA_NV",)] == abi.PBA_NV
    for cname, py in _STRUCTS.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_table_variable_order_matches_header():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"enum pba_var \{(.*?)\};", text, flags=re.S).group(1)
    names = [t.strip().split("=")[0].strip() for t in body.split(",") if t.strip()]
    assert names[-1] == "PBA_NV"
    assert len(names) - 1 == abi.PBA_NV


def test_parfile_round_trip(tmp_path):
    """Workload -> parfile text -> reader -> RunParams; the reference reads the same keys
    (src/main.cpp:93-154, src/blastwave/blastwave_base.h:95-456)."""
    main, grb, struct = workloads.gauss21_fsrs_ssc()
    par = inputs.write_parfile(str(tmp_path), main, grb)
    P, mp_, mo, gp, go = abi.run_params_from_parfile(par)
    assert P.fsrs == 1 and P.do_rs == 1 and P.do_rs_radiation == 1
    assert P.n_ism == pytest.approx(10 ** (-1.6), rel=1e-15)
    assert P.theta_obs == 0.363
    assert P.fs.ssc == 1 and P.rs.ssc == 1
    assert P.fs.ngam == 200 and P.fs.nfreq == 96
    assert P.fs.p == 2.1 and P.rs.p == 2.1
    t = abi.tgrid(mp_)
    # MakeLogspaceVec multiplies by (tb1/tb0)^(1/ntb): the grid stops one ratio short of tb1
    ratio = (1e14 / 3e2) ** (1.0 / 1000)
    assert len(t) == 1000 and t[0] == pytest.approx(3e2) and t[-1] == pytest.approx(1e14 / ratio, rel=1e-11)
    assert np.allclose(t[1:] / t[:-1], ratio, rtol=1e-12)
    tobs, nuobs = abi.obs_pairs(mo)
    assert len(tobs) == 96 * 128
    # time-major pairs: the first 96 pairs share the first observer time
    assert np.all(tobs[:96] == tobs[0]) and len(np.unique(nuobs[:96])) == 96


def test_unsupported_options_raise(tmp_path):
    main, grb, struct = workloads.tophat_fs_syn()
    for key, val in (("method_quad", "trap_adapt"), ("method_spread", "None"), ("method_eats", "adaptive-x")):
        par = inputs.write_parfile(str(tmp_path), main, dict(grb, **{key: val}))
        with pytest.raises(NotImplementedError):
            abi.run_params_from_parfile(par)
    par = inputs.write_parfile(str(tmp_path), main, dict(grb, method_eats="piece-wise"))
    abi.run_params_from_parfile(par)


def test_synchrotron_kernel_options(tmp_path):
    """method_synchrotron_fs strings of the numeric electron path map to pba_shock_params.syn
    (microphysics_base.h:310-335); the analytic-path kernels are rejected."""
    main, grb, struct = workloads.tophat_fs_syn()
    for name, code in (("CSYN", 0), ("Dermer09", 1), ("Bessel", 2), ("GSL", 3)):
        par = inputs.write_parfile(str(tmp_path), main, dict(grb, method_synchrotron_fs=name))
        P = abi.run_params_from_parfile(par)[0]
        assert P.fs.syn == code
    for name in ("Joh06", "WSPN99", "Marg21"):
        par = inputs.write_parfile(str(tmp_path), main, dict(grb, method_synchrotron_fs=name))
        with pytest.raises(NotImplementedError):
            abi.run_params_from_parfile(par)


def test_layer_inits_piecewise():
    """Piece-wise blastwaves: E0, M0 per cell (divided by 2i+1, ejecta_id.h:315-322), theta_b0 =
    theta_wing and ncells = n^2 (blastwave_base.h:304-312); adaptive keeps ncells = 1, theta_b0 = theta_c_h."""
    main, grb, struct = workloads.gauss21_fsrs_ssc()
    idd = inputs.jet_structure(kind="piece-wise", **dict(struct, n_layers_pw=7))
    n = len(idd["ek"])
    assert n == 7
    P = abi.layer_inits(idd, "piece-wise")
    A = abi.layer_inits(idd, "adaptive")
    tw = float(idd["theta_wing"])
    assert abi.total_cells(n) == n * n
    for i in range(n):
        assert P[i].E0 * (2 * i + 1) == pytest.approx(idd["ek"][i], rel=1e-15)
        assert P[i].M0 * (2 * i + 1) == pytest.approx(idd["mass"][i], rel=1e-15)
        assert P[i].theta_b0 == tw and P[i].ncells == n * n
        assert A[i].theta_b0 == A[i].theta_c_h and A[i].ncells == 1.0
        assert P[i].ctheta0 == pytest.approx(0.5 * (P[i].theta_c_l + P[i].theta_c_h), rel=1e-15)
    # the jet's energy is shared by the cells: sum over cells = sum over the initial-data layers
    assert sum(P[i].E0 * abi.cells_in_layer(i) for i in range(n)) == pytest.approx(float(np.sum(idd["ek"])), rel=1e-14)
    with pytest.raises(NotImplementedError):
        abi.layer_inits(idd, "something")


def test_layer_edges_piecewise():
    """Layer edges are 2 asin(i/n sin(theta_w/2)) (src/ejecta/ejecta_id.h:142-182)."""
    main, grb, struct = workloads.gauss21_fsrs_ssc()
    idd = inputs.jet_structure(kind="adaptive", **struct)
    L = abi.layer_inits(idd)
    n = len(L)
    tw = float(idd["theta_wing"])
    for i in range(n):
        assert L[i].theta_c_l == pytest.approx(2 * np.arcsin(i / n * np.sin(tw / 2)), rel=1e-15, abs=0)
        assert L[i].theta_c_h > L[i].theta_c_l
    # the layers tile the cap: sum of (cos l - cos h) = 1 - cos(theta_w)
    s = sum(np.cos(L[i].theta_c_l) - np.cos(L[i].theta_c_h) for i in range(n))
    assert s == pytest.approx(1 - np.cos(tw), rel=1e-12)
    # cos(2 asin a) = 1 - 2 a^2: layer i spans (2i+1) equal solid-angle cells of the piece-wise grid
    d = np.array([np.cos(L[i].theta_c_l) - np.cos(L[i].theta_c_h) for i in range(n)])
    assert np.allclose(d / (2 * np.arange(n) + 1), d[0], rtol=1e-9)


def test_latin_hypercube_members():
    """Batched-member workload: one point per stratum in every dimension, parameters inside their
    ranges, a fixed seed gives the same members, and every member keeps the 21 layers and reads
    back through the reference's parfile format."""
    import tempfile
    n = 64
    u = workloads.latin_hypercube(n, len(workloads.LHS_RANGES))
    assert u.shape == (n, 8) and u.min() >= 0 and u.max() < 1
    for j in range(8):
        assert sorted(np.floor(u[:, j] * n).astype(int)) == list(range(n))
    np.testing.assert_array_equal(u, workloads.latin_hypercube(n, len(workloads.LHS_RANGES)))
    assert not np.array_equal(u, workloads.latin_hypercube(n, len(workloads.LHS_RANGES), seed=1))
    for i in (0, 17, 63):
        main, grb, struct = workloads.lhs_member(i, n=n)
        assert 1e50 <= struct["eiso_c"] < 1e54 and 10 <= struct["gamma0c"] < 1000 and 0.02 <= struct["theta_c"] < 0.2
        assert 1e-4 <= main["n_ism"] < 10 and 0 <= main["theta_obs"] < 0.8 and 2.05 <= grb["p_fs"] < 2.8
        assert 1e-3 <= grb["eps_e_fs"] < 10 ** -0.5 and 1e-5 <= grb["eps_b_rs"] < 0.1
        assert len(inputs.jet_structure(kind="adaptive", **struct)["ek"]) == 21
        with tempfile.TemporaryDirectory() as d:
            P = abi.run_params_from_parfile(inputs.write_parfile(d, main, grb))[0]
        assert P.n_ism == pytest.approx(main["n_ism"], rel=1e-12) and P.fs.p == pytest.approx(grb["p_fs"], rel=1e-12)


def test_batch_populations_and_layer_descriptors():
    """gpu.populations: a forward-shock population for every blastwave of every run, a reverse-shock
    one only where the blastwave was evolved with the reverse shock and that run radiates from it
    (considerReverseShock, src/blastwave/blastwave_radiation.h:961-974); layer descriptors point at
    their populations and carry each blastwave's i_end_r."""
    from pyblastafterglowmag_amd import gpu
    params = (abi.RunParams * 2)()
    params[0].do_rs_radiation = 1
    params[1].do_rs_radiation = 0
    res = (abi.BwResult * 6)()
    for bw, (fsrs, iend) in enumerate([(1, 10), (0, 11), (1, 12), (1, 20), (1, 21), (0, 22)]):
        res[bw].fsrs, res[bw].i_end_r = fsrs, iend
    pops, lays = gpu.populations(params, res, 2, 3)
    assert [(p.bw, p.run, p.shock) for p in pops] == [(0, 0, 0), (0, 0, 1), (1, 0, 0), (2, 0, 0), (2, 0, 1),
                                                       (3, 1, 0), (4, 1, 0), (5, 1, 0)]
    assert lays == [(0, 0, 0, 1, 10), (1, 0, 2, -1, 11), (2, 0, 3, 4, 12),
                    (3, 1, 5, -1, 20), (4, 1, 6, -1, 21), (5, 1, 7, -1, 22)]


def test_dyn_entry_rejects_unknown_integrator_before_any_device_call():
    if not os.path.exists(LIB):
        pytest.skip("libpbagpu.so not built")
    lib = C.CDLL(LIB)
    lib.pbagpu_last_error.restype = C.c_char_p
    f = lib.pbagpu_dyn_evolve
    f.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.c_int, C.POINTER(abi.LayerInit), C.c_int, C.c_void_p,
                  C.c_void_p, C.POINTER(abi.BwResult), C.c_void_p]
    P = (abi.RunParams * 1)()
    P[0].integrator = 2
    init = (abi.LayerInit * 1)()
    init[0].ncells, init[0].theta_b0 = 1.0, 0.1
    res = (abi.BwResult * 1)()
    dummy = C.c_void_p(8)   # never dereferenced: the call fails validation first
    assert f(P, 1, 1, init, 10, dummy, dummy, res, None) == -3
    assert b"integrator" in lib.pbagpu_last_error()
    P[0].integrator = 0
    init[0].ncells = 0.5
    assert f(P, 1, 1, init, 10, dummy, dummy, res, None) == -1


def test_eats_entry_rejects_bad_layer_descriptors_before_any_device_call():
    if not os.path.exists(LIB):
        pytest.skip("libpbagpu.so not built")
    lib = C.CDLL(LIB)
    lib.pbagpu_last_error.restype = C.c_char_p
    f = lib.pbagpu_eats_lightcurve
    f.argtypes = [C.POINTER(abi.RunParams), C.c_int, C.POINTER(abi.EatsLayer), C.c_int, C.c_void_p, C.c_int,
                  C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    P = (abi.RunParams * 1)()
    for sh in (P[0].fs, P[0].rs):
        sh.nfreq, sh.freq1, sh.freq2 = 8, 1e5, 1e32
    dummy = C.c_void_p(8)   # never dereferenced: every call fails validation first

    def call(bw=0, pop_fs=0, pop_rs=-1, iend=10, n_runs=1):
        L = (abi.EatsLayer * 1)(abi.EatsLayer(0.0, 0.1, bw, 0, pop_fs, pop_rs, iend))
        return f(P, n_runs, L, 1, dummy, 10, dummy, None, 8, 4, dummy, dummy, dummy, dummy, None)
    assert call(bw=-1) == -1 and b"layer descriptor" in lib.pbagpu_last_error()
    assert call(pop_rs=-2) == -1 and call(pop_fs=-1) == -1 and call(iend=11) == -1
    assert call(n_runs=0) == -1 and call(n_runs=-5) == -1


def test_product_path_fails_loudly_without_gpu_or_library(monkeypatch):
    """No CPU fallback: without a ROCm device the GPU path raises, and without the built library
    loading it raises with the path it looked for."""
    import torch
    from pyblastafterglowmag_amd import gpu
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="ROCm device"):
            gpu.skymap_hist([[(np.zeros(2), np.zeros(2), np.ones(2))]], 5, 5)
        import __graft_entry__
        with pytest.raises(RuntimeError):
            __graft_entry__.smoke()
    monkeypatch.setattr(gpu, "_lib", None)
    monkeypatch.setattr(gpu, "_LIB_PATH", "/nonexistent/libpbagpu.so")
    with pytest.raises(RuntimeError, match="GPU library missing"):
        gpu.lib()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one MI355X: --gpus N > 1 exits with a message instead of running."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr


def test_observer_pairs_grid_and_one_to_one():
    """cast_times_freqs (src/utilitites/utils.h:1419-1448): every time with every frequency, time-major;
    with lc_use_freq_to_time = yes the i-th time with the i-th frequency, lists of one length."""
    mo = {"lc_times": "array 1 2 3", "lc_freqs": "array 10 20"}
    t, f = abi.obs_pairs(mo)
    np.testing.assert_array_equal(t, [1, 1, 2, 2, 3, 3])
    np.testing.assert_array_equal(f, [10, 20, 10, 20, 10, 20])
    t, f = abi.obs_pairs(dict(mo, lc_freqs="array 10 20 30", lc_use_freq_to_time="yes"))
    np.testing.assert_array_equal(t, [1, 2, 3])
    np.testing.assert_array_equal(f, [10, 20, 30])
    with pytest.raises(ValueError):
        abi.obs_pairs(dict(mo, lc_use_freq_to_time="yes"))
