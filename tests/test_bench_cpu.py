"""bench.py's JSON helpers on the CPU, with stand-ins for the GPU pipeline: the roofline entry of
every stage it can name, and the mixed-leg error measure.  A failure here would otherwise surface
only at the end of a full benchmark run, after the timed steps."""
import importlib.util
import os
import sys
import types

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    argv, sys.argv = sys.argv, ["bench.py"]
    try:
        spec.loader.exec_module(mod)
    finally:
        sys.argv = argv
    return mod


def _gpu(units):
    from pyblastafterglowmag_amd import gpu
    return types.SimpleNamespace(STAGE_DYN=gpu.STAGE_DYN, STAGE_ELE=gpu.STAGE_ELE, STAGE_EATS=gpu.STAGE_EATS,
                                 STAGE_SSC_TABLE=gpu.STAGE_SSC_TABLE, STAGE_EATS_PW=gpu.STAGE_EATS_PW,
                                 STAGE_EATS_VEC=gpu.STAGE_EATS_VEC, timing_units=lambda s: units)


def _pipeline(n_runs=176, n_layers=21, n_t=128, n_nu=96):
    run = types.SimpleNamespace(tobs=np.repeat(np.arange(float(n_t)), n_nu), n_layers=n_layers)
    res = [types.SimpleNamespace(n_rhs=3_472_195, fsrs=1) for _ in range(n_runs * n_layers)]
    return types.SimpleNamespace(runs=[run], n_runs=n_runs, bpr=n_layers, res=res)


def test_vectorised_observer_roofline_is_stated_against_fp64(bench):
    """The radiation stream's dominant kernel: FP64 operations per node and per node-frequency over
    the launch time, the FP64 peak, and the measured HBM traffic of the benchmark's batch."""
    g = _gpu(6 * 5.98e8)
    stage_ms = {0: (283_800.0, 6), 1: (21_855.2, 6), 2: (0.0, 0), 3: (2.0, 6), 4: (0.0, 0), 6: (51_767.3, 6)}
    r = bench.roofline(stage_ms, [_pipeline()], g)
    assert r["kernel"] == "eats_vec_kernel" and r["bound"] == "mfma" and r["unit"] == "TFLOP/s"
    fl = 5.98e8 * (bench.FLOPS_PER_NODE + 96 * bench.FLOPS_PER_NODE_NU)
    assert r["achieved"] == pytest.approx(fl / (51_767.3 / 6 / 1e3) / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.PEAK_FP64_TFLOPS)
    assert r["traffic"] == 4_527_263_695   # profiles/traffic.json, eats_vec_kernel|21x176
    assert r["hbm_gbs_from_traffic"] == pytest.approx(4_527_263_695 / (51_767.3 / 6 / 1e3) / 1e9)
    assert r["hbm_gbs_from_traffic"] < 0.01 * bench.PEAK_HBM_GBS


def test_dynamics_roofline_counts_every_rhs(bench):
    g = _gpu(0)
    stage_ms = {0: (283_800.0, 6), 1: (0.0, 0), 2: (0.0, 0), 3: (0.0, 0), 4: (0.0, 0), 6: (0.0, 0)}
    pl = _pipeline(n_runs=4)
    r = bench.roofline(stage_ms, [pl], g, dom=g.STAGE_DYN)
    flops = 4 * 21 * 3_472_195 * bench.FLOPS_PER_RHS[1]
    assert r["kernel"] == "dyn_kernel" and r["achieved"] == pytest.approx(flops / 47.3 / 1e12)
    assert r["traffic"] is None   # no traffic entry for 21 x 4


def test_other_stages_name_their_kernel(bench):
    g = _gpu(2.4e10)
    stage_ms = {0: (0.0, 0), 1: (0.0, 0), 2: (16_300.0, 1), 3: (0.0, 0), 4: (0.0, 0), 6: (0.0, 0)}
    r = bench.roofline(stage_ms, [_pipeline(n_runs=12)], g)
    assert r["kernel"] == "eats_kernel" and r["bound"] == "hbm"
    assert r["achieved"] == pytest.approx(2.4e10 * bench.BYTES_PER_SAMPLE / 16.3 / 1e9)


def test_bench_defaults_fit_one_wavefront_per_simd(bench):
    """Four dynamics launches of one batch each keep one wavefront per SIMD on the dynamics units
    (the kernel needs a SIMD's whole register file), and the radiation stream keeps the rest."""
    sys.argv, argv = ["bench.py"], sys.argv
    try:
        a = bench.parse()
    finally:
        sys.argv = argv
    assert a.depth * a.batch <= 4 * a.dyn_cus
    assert 0 < a.dyn_cus < 256 and a.steps > a.depth and a.warmup >= a.depth
