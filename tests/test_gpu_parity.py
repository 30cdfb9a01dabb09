"""GPU path (libpbagpu.so through its C ABI) against the CPU oracle and the reference's outputs.

Tolerances: the GPU evaluates the same FP64 algorithm, but with fused multiply-adds, the
device math library and FP64 in place of x87 long double; ulp-level differences move the
shared adaptive ODE step.  Forward-shock-only runs agree to 2e-6; with the reverse shock
in any layer of a run the step sequences part after the first ulp difference and the tables agree to 1e-4,
the same bound the oracle keeps against the reference program.  Electron and observer
stages are compared on identical inputs (the GPU's own tables): 1e-9 and 1e-10.  Light
curves end to end against the reference program: 1e-6 relative (the FP64-path target).
"""
import concurrent.futures as cf

import numpy as np
import pytest

from conftest import dyn_excess, golden_run, gpu_pipeline, load_golden, oracle_dynamics, rel_err, rs_end_steps
import pyoracle
from pyblastafterglowmag_amd import abi

pytestmark = pytest.mark.gpu

DYN_CHECK = ["R", "Gamma", "beta", "theta", "M2", "Eint2", "tt", "tcomov", "rho2", "GammaFsh", "thickness", "U_p",
             "gamma_min", "gamma_max", "B", "accel_frac"]


def _gpu(run, name=None, precision=0):
    """Every stage on the GPU; a golden fixture's dynamics come from the session cache."""
    from pyblastafterglowmag_amd import gpu
    if name is not None:
        run, pl = gpu_pipeline(name)
        flux, ex = gpu.radiate(pl, want_intermediates=True, precision=precision)
        flux = flux[:1]
    else:
        flux, ex = gpu.lightcurves([run], want_intermediates=True, precision=precision)
    assert not ex["dyn_status"].any(), ex["dyn_status"]
    assert not ex["ele_status"].any(), ex["ele_status"]
    assert not ex["eats_status"].any(), ex["eats_status"]
    return flux[0], ex


@pytest.mark.parametrize("name", ["tophat_small", "gauss5_small", "tophat_pw_small", "gauss_pw_small",
                                  "gauss_pw_default", "tophat_ssa_small", "gauss5_ssa_small", "tophat_dermer_small", "tophat_bessel_small",
                                  "tophat_gsl_small", "tophat_rk4_small", "tophat_wind_small"])
def test_stages_match_oracle(name):
    g = load_golden(name)
    run = golden_run(g)
    print(f"[gpu] {name}: layers={run.n_layers} nt={len(run.tgrid)} pairs={len(run.tobs)}", flush=True)
    flux, ex = _gpu(run, name)
    table_o, res_o = oracle_dynamics(name)
    # one shared step for all layers of a run: a reverse shock in any layer moves every layer
    tol = 1e-4 if any(r.fsrs for r in res_o) else 2e-6
    for il in range(run.n_layers):
        assert ex["results"][il][1] == res_o[il].i_end_r
        assert ex["results"][il][2] == res_o[il].fsrs
        for n in DYN_CHECK:
            e = rel_err(ex["table"][il, abi.PV[n]], table_o[il, abi.PV[n]], 1e-300)
            assert e <= tol, (name, il, n, e)
    # electron stage on the GPU's own dynamics table vs oracle on the same table
    pops = (abi.Population * len(ex["pops"]))(*ex["pops"])
    jall_o, _, st, aall_o = pyoracle.ele_evolve(run.params, pops, ex["table"], want_a=True)
    assert not st.any()
    for q in range(len(ex["pops"])):
        e = rel_err(ex["jall"][q], jall_o[q], 1e-9)
        assert e <= 1e-9, (name, q, e)
        if ex["aall"] is not None:   # synchrotron self-absorption
            e = rel_err(ex["aall"][q], aall_o[q], 1e-9)
            assert e <= 1e-9, (name, q, "absorption", e)
    # observer stage on the GPU's own tables vs oracle on the same tables
    if run.method == "piece-wise":
        lays = (abi.EatsPwLayer * len(ex["layers"]))()
        for i, (bw, r, pfs, prs, iend) in enumerate(ex["layers"]):
            li = run.inits[bw]
            lays[i] = abi.EatsPwLayer(li.ctheta0, li.ncells, abi.cells_in_layer(li.ilayer), bw, r, pfs,
                                      1 if prs >= 0 else 0, iend)
        lc_o, st2 = pyoracle.eats_lightcurve_pw(run.params, lays, ex["table"], ex["jall"], run.tobs, run.nuobs,
                                                aall=ex["aall"])
    else:
        lays = (abi.EatsLayer * len(ex["layers"]))()
        for i, (bw, r, pfs, prs, iend) in enumerate(ex["layers"]):
            li = run.inits[bw]
            lays[i] = abi.EatsLayer(li.theta_c_l, li.theta_c_h, bw, r, pfs, prs, iend)
        lc_o, st2 = pyoracle.eats_lightcurve(run.params, lays, ex["table"], ex["jall"], run.tobs, run.nuobs,
                                             aall=ex["aall"])
    assert not st2.any()
    for il in range(run.n_layers):
        e = rel_err(flux[il], lc_o[il], 1e-30)
        assert e <= 1e-10, (name, il, e)


# Light curves end to end against the reference program: 1e-6 relative.  With the reverse shock at
# the default rs_shutOff_criterion_rho (1e-50) a layer's reverse shock ends at the first step where
# M3/M0 -> 1 rounds to >= 1 (isToStopReverseShock, src/blastwave/blastwave.h:453-471); a one-ulp
# change of the inputs moves that step and changes the layer's light curve by up to 50 % in the
# reference's own algorithm (tests/test_oracle_golden.py::test_reverse_shock_end_is_decided_by_the_last_ulp).
# A layer whose reverse shock ends at another step than the reference's is therefore held only to
# MOVED_TOL; every other layer to its fixture's bound.
LC_TOL = {"tophat_small": 1e-6, "tophat_fs_syn": 1e-6, "gauss5_small": 1e-6, "tophat_pw_small": 1e-6,
          "gauss_pw_small": 1e-6, "gauss_pw_default": 1e-6, "tophat64_pw": 1e-6, "tophat_ssa_small": 1e-6,
          "gauss5_ssa_small": 1e-6, "tophat_dermer_small": 1e-6, "tophat_bessel_small": 1e-6,
          "tophat_gsl_small": 1e-6, "tophat_rk4_small": 1e-6, "tophat_wind_small": 1e-6,
          # full-size jet: once the shared adaptive step first differs by an ulp (the CPU oracle's
          # too), the reverse-shock layers' dynamics part from the reference's at ~1e-6 (the
          # oracle's layer 4: 6.6e-7 in U_p), and the light curves' high-frequency tails, 13 decades
          # below a layer's peak, amplify that: the oracle's own light curve of layer 4 is 6.3e-6 off
          # (2.2e-6 above 1e-10 of the peak); measured on the GPU: 1.5e-5
          "gauss21_fsrs_ssc": 5e-5}
MOVED_TOL = 1.0
# raw sky-map cells (every cell of every layer at 2 t_obs x 2 nu_obs) against the reference's files;
# in the reverse-shock run a cell's intensity carries the shared ODE step's ulp-level differences
SKY_TOL = {"tophat_pw_small": 1e-6, "gauss_pw_small": 5e-6, "gauss64_pw_small": 5e-6}


@pytest.mark.parametrize("name", ["tophat_small", "gauss5_small", "tophat_fs_syn", "tophat_pw_small", "gauss_pw_small",
                                  "gauss_pw_default", "tophat64_pw", "tophat_ssa_small", "gauss5_ssa_small",
                                  "tophat_dermer_small", "tophat_bessel_small", "tophat_gsl_small",
                                  "tophat_rk4_small", "tophat_wind_small"])
def test_lightcurve_matches_reference(name):
    """Light curves end to end against the reference program; tophat64_pw is the full-size piece-wise
    case (64 layers, 4096 cells, 96 x 128 pairs; the fixture holds every 7th layer)."""
    g = load_golden(name)
    run = golden_run(g)
    print(f"[gpu] {name}: end-to-end light curve", flush=True)
    flux, ex = _gpu(run, name)
    names = [str(n) for n in g["dyn_names"]]
    rs = "Gamma43" in names
    if rs:
        ref_end = rs_end_steps(g["dyn"][:, names.index("Gamma43")])
        gpu_end = rs_end_steps(ex["table"][:, abi.PV["Gamma43"]])
    rows = []
    for k, il in enumerate(g["lc_layers"]):
        e = rel_err(flux[il], g["lc"][k], 1e-30)
        moved = rs and gpu_end[il] != ref_end[il]
        rows.append((int(il), e, moved))
        print(f"  layer {il}: rel err {e:.3e}" + (f", reverse shock ends at step {gpu_end[il]} "
                                                  f"(reference {ref_end[il]})" if rs else ""), flush=True)
    for il, e, moved in rows:
        assert e <= (MOVED_TOL if moved else LC_TOL[name]), (name, il, e, moved)


C2 = "gauss21_fsrs_ssc"


def test_full_size_c2_matches_reference_and_oracle():
    """The 21-layer Gaussian jet at full size (both shocks, SSC on both, 200 gamma x 96 nu,
    1000 steps, 96 x 128 = 12,288 observer pairs), FP64 path.  The fixture holds the reference
    program's light curves and dynamics of every 4th layer.
      - dynamics of those layers against the reference, 1e-4 (the reverse-shock bound) plus 64 ulp of
        Gamma carried into Gamma - 1 (conftest.dyn_excess), for every layer whose reverse shock ends
        at the reference's step;
      - light curves against the reference, LC_TOL for those layers and MOVED_TOL for the others;
      - electron stage against the oracle on the GPU's own table (populations of those layers), 1e-9;
      - observer stage against the oracle on the GPU's own tables at every 16th pair, 1e-10."""
    g = load_golden(C2)
    run = golden_run(g)
    assert run.n_layers == 21 and len(run.tobs) == 12288
    assert run.params.fs.ngam == 200 and run.params.fs.nfreq == 96 and len(run.tgrid) == 1000
    print(f"[gpu] {C2}: full size, 21 layers x 12,288 pairs", flush=True)
    flux, ex = _gpu(run, C2)
    print(f"[gpu] {C2}: GPU stages done", flush=True)
    names = [str(n) for n in g["dyn_names"]]
    layers = [int(x) for x in g["lc_layers"]]
    assert [int(x) for x in g["dyn_layers"]] == layers
    ref_end = rs_end_steps(g["dyn"][:, names.index("Gamma43")])
    gpu_end = rs_end_steps(ex["table"][layers, abi.PV["Gamma43"]])
    rows = []
    for k, il in enumerate(layers):
        moved = gpu_end[k] != ref_end[k]
        e_lc = rel_err(flux[il], g["lc"][k], 1e-30)
        gam = g["dyn"][k, names.index("Gamma")]
        e_dyn = max(dyn_excess(ex["table"][il, abi.PV[n]], g["dyn"][k, iv], gam) for iv, n in enumerate(names))
        rows.append((il, e_lc, e_dyn, moved))
        print(f"  layer {il}: light curve {e_lc:.3e}, dynamics {e_dyn:.3f} of its bound, reverse shock ends at "
              f"step {gpu_end[k]} (reference {ref_end[k]})", flush=True)
    for il, e_lc, e_dyn, moved in rows:
        assert e_lc <= (MOVED_TOL if moved else LC_TOL[C2]), (il, e_lc, moved)
        if not moved:
            assert e_dyn <= 1.0, (il, e_dyn)
    # electron stage: the populations of the fixture's layers, oracle on the GPU's table
    sel = [q for q, P in enumerate(ex["pops"]) if P.bw in layers]
    pops = (abi.Population * len(sel))(*[ex["pops"][q] for q in sel])
    jall_o, _, st = pyoracle.ele_evolve(run.params, pops, ex["table"])
    assert not st.any()
    worst = max(rel_err(ex["jall"][q], jall_o[i], 1e-9) for i, q in enumerate(sel))
    print(f"  electrons, {len(sel)} populations: max rel err vs oracle {worst:.3e}", flush=True)
    assert worst <= 1e-9
    # observer stage: those layers at every 16th pair
    ent = [e for e in ex["layers"] if e[0] in layers]
    lays = (abi.EatsLayer * len(ent))()
    for i, (bw, r, pfs, prs, iend) in enumerate(ent):
        li = run.inits[bw]
        lays[i] = abi.EatsLayer(li.theta_c_l, li.theta_c_h, bw, r, pfs, prs, iend)
    idx = np.arange(0, len(run.tobs), 16)
    lc_o, st2 = pyoracle.eats_lightcurve(run.params, lays, ex["table"], ex["jall"], run.tobs[idx], run.nuobs[idx])
    assert not st2.any()
    worst = max(rel_err(flux[e[0]][idx], lc_o[i], 1e-30) for i, e in enumerate(ent))
    print(f"  observer, {len(ent)} layers x {len(idx)} pairs: max rel err vs oracle {worst:.3e}", flush=True)
    assert worst <= 1e-10


@pytest.mark.parametrize("name", ["gauss5_small", "tophat_small"])
def test_mixed_precision_ssc_lightcurve(name):
    """precision=1 (SSC contraction in BF16 hi/lo on MFMA, FP32 accumulation) against the FP64
    GPU path on the same inputs: light curves within 1e-3 relative wherever the flux is above
    1e-8 of the layer's peak (the mixed path's target is 1 %), spectra within 1e-3."""
    from pyblastafterglowmag_amd import gpu
    g = load_golden(name)
    run = golden_run(g)
    print(f"[gpu] {name}: mixed-precision SSC", flush=True)
    f64, ex64 = _gpu(run, name, precision=0)
    f64 = f64[None]
    mix, exm = _gpu(run, name, precision=1)
    mix = mix[None]
    assert not exm["ele_status"].any() and not exm["eats_status"].any()
    worst_j = 0.0
    for q in range(len(ex64["pops"])):
        e = rel_err(exm["jall"][q], ex64["jall"][q], 1e-8)
        worst_j = max(worst_j, e)
        assert e <= 1e-3, (name, q, e)
    print(f"[gpu] {name}: max relative emissivity difference mixed vs FP64 = {worst_j:.3e}", flush=True)
    worst = 0.0
    for il in range(run.n_layers):
        e = rel_err(mix[0][il], f64[0][il], 1e-8)
        worst = max(worst, e)
        assert e <= 1e-3, (name, il, e)
    print(f"[gpu] {name}: max relative light-curve difference mixed vs FP64 = {worst:.3e}", flush=True)


@pytest.mark.parametrize("name", ["tophat_pw_small", "gauss_pw_small", "gauss64_pw_small"])
def test_skymaps_match_oracle_and_reference(name):
    """Raw piece-wise sky maps (pbagpu_eats_skymap_pw): every cell's entry against the oracle on the
    GPU's own tables (1e-10), and the normalised, cleaned images against the reference program's
    raw_skymap_<i>.h5 (SKY_TOL).  gauss64_pw_small is the 64-layer, 4096-cell jet with both shocks
    and SSC: 4 images of 4096 cells each."""
    from pyblastafterglowmag_amd import gpu, output
    g = load_golden(name)
    run = golden_run(g)
    print(f"[gpu] {name}: raw sky maps", flush=True)
    run, pl = gpu_pipeline(name)
    pl.precision = 0
    pl.electrons()
    pl.observer()
    times, freqs = np.array([1e5, 1e7]), np.array([1e9, 1e18])
    image, flux, st = pl.skymaps(times, freqs)
    assert not st.any(), st
    table = pl.d_table.cpu().numpy()
    jall = pl.d_jall[:len(pl.pops)].cpu().numpy()
    lays = (abi.EatsPwLayer * len(pl.layers))(*[pl.larr[i] for i in range(len(pl.layers))])
    image_o, flux_o, st_o = pyoracle.eats_skymap_pw(run.params, lays, table, jall, np.tile(times, 2),
                                                    np.repeat(freqs, 2))
    assert not st_o.any()
    for k in range(image.shape[0]):
        for v in range(abi.PBA_IMG_NV):
            e = rel_err(image[k, v], image_o[k, v], 1e-300)
            assert e <= 1e-10, (name, k, abi.IMG_NAMES[v], e)
    assert rel_err(flux, flux_o, 1e-300) <= 1e-10
    ims = output.skymap_images(image, flux, run.n_layers)
    for i, im in enumerate(ims):
        for v in abi.IMG_NAMES:
            ref = g[f"sky{i}_{v}"]
            assert im["arrays"][v].shape == ref.shape, (name, i, v)
            e = rel_err(im["arrays"][v], ref, 1e-300)
            assert e <= SKY_TOL[name], (name, i, v, e)


def _lhs_small(i, ragged_nt=False):
    from pyblastafterglowmag_amd import gpu, workloads
    return gpu.Run.from_workload(*workloads.lhs_small(i, ragged_nt=ragged_nt))


def test_heterogeneous_batch_matches_oracle():
    """Two members of the Latin-hypercube workload (different E_iso, Gamma0, theta_c, n_ism, eps_e,
    eps_b, p and theta_obs) in one batch, with the microphysics grids reduced: every stage
    against the oracle run by run (dynamics 1e-4, emissivities 1e-9 and light curves 1e-10 on the
    GPU's own tables), so that per-run parameters reach every kernel correctly."""
    from pyblastafterglowmag_amd import gpu
    runs = [_lhs_small(0), _lhs_small(1)]
    nl = runs[0].n_layers
    print(f"[gpu] heterogeneous batch: 2 runs x {nl} layers", flush=True)
    params = (abi.RunParams * 2)(*[r.params for r in runs])
    inits = (abi.LayerInit * (2 * nl))(*[li for r in runs for li in r.inits])
    with cf.ThreadPoolExecutor(max_workers=1) as pool:   # the oracle's dynamics beside the GPU's
        fut = pool.submit(pyoracle.dyn_evolve, params, inits, runs[0].tgrid, n_runs=2)
        flux, ex = gpu.lightcurves(runs, want_intermediates=True)
        table_o, res_o = fut.result()
    assert not ex["dyn_status"].any() and not ex["ele_status"].any() and not ex["eats_status"].any()
    for bw in range(2 * nl):
        assert ex["results"][bw][1] == res_o[bw].i_end_r and ex["results"][bw][2] == res_o[bw].fsrs, bw
        for n in DYN_CHECK:
            e = rel_err(ex["table"][bw, abi.PV[n]], table_o[bw, abi.PV[n]], 1e-300)
            assert e <= 1e-4, (bw, n, e)
    pops = (abi.Population * len(ex["pops"]))(*ex["pops"])
    jall_o, _, st = pyoracle.ele_evolve(params, pops, ex["table"], n_runs=2)
    assert not st.any()
    for q in range(len(ex["pops"])):
        e = rel_err(ex["jall"][q], jall_o[q], 1e-9)
        assert e <= 1e-9, (q, e)
    lays = (abi.EatsLayer * len(ex["layers"]))()
    for i, (bw, r, pfs, prs, iend) in enumerate(ex["layers"]):
        li = runs[r].inits[bw - r * nl]
        lays[i] = abi.EatsLayer(li.theta_c_l, li.theta_c_h, bw, r, pfs, prs, iend)
    lc_o, st2 = pyoracle.eats_lightcurve(params, lays, ex["table"], ex["jall"], runs[0].tobs, runs[0].nuobs, n_runs=2)
    assert not st2.any()
    got = flux.reshape(2 * nl, -1)
    for i in range(2 * nl):
        e = rel_err(got[i], lc_o[i], 1e-30)
        assert e <= 1e-10, (i, e)


def _run_against_oracle(r, run, ex, got, oracle_dyn, nl, nt_r):
    """Run r of a batch against the oracle on its own grid: the dynamics on the run's steps
    (conftest.dyn_excess: 1e-4 plus 64 ulp of Gamma carried into Gamma - 1; gamma_min to the 2e-3
    resolution of the reference's bisection, which stops at |dm| < 1e-3 |xm|, rootfinders.h:17-58,
    so an ulp-level change of its inputs can end it in a neighbouring interval) and zeros after
    them; the electron and observer stages on the GPU's own table columns of the run (emissivities
    1e-9, light curves 1e-10) with the same status bits as the oracle's."""
    table_o, res_o = oracle_dyn
    for il in range(nl):
        bw = r * nl + il
        assert ex["results"][bw][1] == res_o[il].i_end_r and ex["results"][bw][2] == res_o[il].fsrs, bw
        assert ex["results"][bw][0] == res_o[il].status, (r, il)
        assert not ex["table"][bw, :, nt_r:].any()
        gam = table_o[il, abi.PV["Gamma"]]
        for n in DYN_CHECK:
            e = dyn_excess(ex["table"][bw, abi.PV[n], :nt_r], table_o[il, abi.PV[n]], gam,
                           rtol=2e-3 if n == "gamma_min" else 1e-4)
            assert e <= 1.0, (r, il, n, e)
    sel = [q for q, P in enumerate(ex["pops"]) if P.run == r]
    pops = (abi.Population * len(sel))(*[abi.Population(ex["pops"][q].bw - r * nl, 0, ex["pops"][q].shock, 0)
                                         for q in sel])
    tab = np.ascontiguousarray(ex["table"][r * nl:(r + 1) * nl, :, :nt_r])
    jall_o, _, st = pyoracle.ele_evolve(run.params, pops, tab)
    assert np.array_equal(np.asarray(st), np.asarray(ex["ele_status"])[sel]), r
    for k, q in enumerate(sel):
        assert not ex["jall"][q][nt_r:].any()
        e = rel_err(ex["jall"][q][:nt_r], jall_o[k], 1e-9)
        assert e <= 1e-9, (r, q, e)
    lays = (abi.EatsLayer * nl)()
    for i, (bw, rr, pfs, prs, iend) in enumerate(ex["layers"][r * nl:(r + 1) * nl]):
        li = run.inits[bw - r * nl]
        kfs = sel.index(pfs)
        krs = sel.index(prs) if prs >= 0 else -1
        lays[i] = abi.EatsLayer(li.theta_c_l, li.theta_c_h, bw - r * nl, 0, kfs, krs, iend)
    jsel = np.ascontiguousarray(ex["jall"][sel][:, :nt_r])
    lc_o, st2 = pyoracle.eats_lightcurve(run.params, lays, tab, jsel, run.tobs, run.nuobs)
    assert np.array_equal(np.asarray(st2), np.asarray(ex["eats_status"])[r * nl:(r + 1) * nl]), r
    for il in range(nl):
        if st2[il] == 0:
            e = rel_err(got[r * nl + il], lc_o[il], 1e-30)
            assert e <= 1e-10, (r, il, e)


def test_ragged_time_grids_match_oracle():
    """Two Latin-hypercube members whose burster-frame grids differ in length (their own
    lhs_ntb steps) in one dynamics launch (pbagpu_dyn_evolve_ragged) and one electron and observer
    launch over the longest grid: every run against the oracle on its own grid (_run_against_oracle)."""
    from pyblastafterglowmag_amd import gpu
    members = [1, 2]
    runs = [_lhs_small(i, ragged_nt=True) for i in members]
    nts = [len(r.tgrid) for r in runs]
    assert len(set(nts)) == len(nts), nts
    nl = runs[0].n_layers
    print(f"[gpu] ragged batch: {len(runs)} runs x {nl} layers, steps {nts}", flush=True)
    with cf.ThreadPoolExecutor(max_workers=len(runs)) as pool:   # the oracle's dynamics beside the GPU's
        futs = [pool.submit(pyoracle.dyn_evolve, run.params, run.inits, run.tgrid) for run in runs]
        flux, ex = gpu.lightcurves(runs, want_intermediates=True)
        oracle = [f.result() for f in futs]
    assert not ex["dyn_status"].any() and not ex["ele_status"].any() and not ex["eats_status"].any()
    assert ex["table"].shape[-1] == max(nts)
    got = flux.reshape(len(runs) * nl, -1)
    for r, run in enumerate(runs):
        _run_against_oracle(r, run, ex, got, oracle[r], nl, nts[r])


def test_latin_hypercube_batch_of_64():
    """The Latin-hypercube workload at batch scale: members 0-63 (each with its own E_iso, Gamma0,
    theta_c, n_ism, eps_e, eps_b, p, theta_obs and time-grid length; microphysics grids reduced)
    in one ragged launch per stage.  Members 0, 21, 42 and 63 against the oracle stage by stage
    (_run_against_oracle, status bits included); every member's dynamics end without a status bit,
    and every light curve whose layer has no status bit is finite and non-negative."""
    from pyblastafterglowmag_amd import gpu
    n = 64
    runs = [_lhs_small(i, ragged_nt=True) for i in range(n)]
    nts = [len(r.tgrid) for r in runs]
    nl = runs[0].n_layers
    sample = [0, 21, 42, 63]
    print(f"[gpu] Latin-hypercube batch: {n} runs x {nl} layers, steps {min(nts)}..{max(nts)}", flush=True)
    with cf.ThreadPoolExecutor(max_workers=len(sample)) as pool:   # the oracle's dynamics beside the GPU's
        futs = {r: pool.submit(pyoracle.dyn_evolve, runs[r].params, runs[r].inits, runs[r].tgrid) for r in sample}
        flux, ex = gpu.lightcurves(runs, want_intermediates=True)
        print("[gpu]   GPU stages done", flush=True)
        oracle = {r: f.result() for r, f in futs.items()}
    print(f"[gpu]   status bits: dynamics {int(np.count_nonzero(ex['dyn_status']))}, electrons "
          f"{int(np.count_nonzero(ex['ele_status']))}, observer {int(np.count_nonzero(ex['eats_status']))} "
          f"of {n * nl} layers", flush=True)
    assert not ex["dyn_status"].any()
    got = flux.reshape(n * nl, -1)
    ok = np.asarray(ex["eats_status"]) == 0
    assert np.all(np.isfinite(got[ok])) and np.all(got[ok] >= 0)
    for r in sample:
        _run_against_oracle(r, runs[r], ex, got, oracle[r], nl, nts[r])
        print(f"[gpu]   member {r} matches the oracle", flush=True)


def test_packed_dynamics_bit_identical():
    """Dynamics with several runs per wavefront (library option PBAGPU_OPT_DYN_RUNS_PER_WAVE), one
    pbagpu_dyn_evolve_ragged launch per group, against each fixture's own one-run launch: (a) seven
    one-layer top-hat fixtures with their own parameters, microphysics options, ambient media and
    integrators (DOP853 and RK4 in one wavefront), 3 per wavefront, so the last wavefront holds one
    run; (b) the two 5-layer reverse-shock fixtures in one wavefront.  Every table entry and
    per-blastwave result is bit-identical: a run's arithmetic and its run-wide sums do not depend on
    the lanes beside it."""
    import ctypes as C
    import torch
    from pyblastafterglowmag_amd import gpu
    groups = [(["tophat_small", "tophat_ssa_small", "tophat_dermer_small", "tophat_bessel_small", "tophat_gsl_small",
                "tophat_rk4_small", "tophat_wind_small"], 3),
              (["gauss5_small", "gauss5_ssa_small"], 2)]
    L = gpu.lib()
    for names, rpw in groups:
        single = [gpu_pipeline(n) for n in names]
        runs = [r for r, _ in single]
        n, bpr = len(runs), runs[0].n_layers
        nts = [len(r.tgrid) for r in runs]
        ntm = max(nts)
        tg = np.zeros((n, ntm))
        for k, r in enumerate(runs):
            tg[k, :nts[k]] = r.tgrid
            tg[k, nts[k]:] = r.tgrid[-1]
        d_tg = torch.from_numpy(tg).cuda()
        params = (abi.RunParams * n)(*[r.params for r in runs])
        inits = (abi.LayerInit * (n * bpr))(*[li for r in runs for li in r.inits])
        table = torch.empty((n * bpr, abi.PBA_NV, ntm), dtype=torch.float64, device="cuda")
        res = (abi.BwResult * (n * bpr))()
        stream = torch.cuda.Stream()
        print(f"[gpu] packed dynamics: {n} runs x {bpr} layers, {rpw} per wavefront", flush=True)
        try:
            gpu.set_option(gpu.OPT_DYN_RUNS_PER_WAVE, rpw)
            rc = L.pbagpu_dyn_evolve_ragged(params, n, bpr, inits, ntm, (C.c_int * n)(*nts), C.c_void_p(d_tg.data_ptr()),
                                            C.c_void_p(table.data_ptr()), res, C.c_void_p(stream.cuda_stream))
        finally:
            gpu.set_option(gpu.OPT_DYN_RUNS_PER_WAVE, 0)
        assert rc == 0, L.pbagpu_last_error()
        got = table.cpu().numpy()
        for k, (run, pl) in enumerate(single):
            rows = slice(k * bpr, (k + 1) * bpr)
            assert np.array_equal(got[rows, :, :nts[k]], pl.d_table.cpu().numpy(), equal_nan=True), names[k]
            assert not got[rows, :, nts[k]:].any(), names[k]
            for i in range(bpr):
                a, b = res[k * bpr + i], pl.res[i]
                assert (a.status, a.i_end_r, a.fsrs, a.n_steps, a.n_rhs) == \
                    (b.status, b.i_end_r, b.fsrs, b.n_steps, b.n_rhs), (names[k], i)


def test_per_layer_electron_grids_match_oracle_and_reference():
    """The 6-layer Gaussian piece-wise jet with a Chang-Cooper grid of its own per layer (32, 64, 48,
    48, 32, 64 bins; forward + reverse shock, SSC) in one call of every stage: dynamics against the
    oracle 1e-4 (reverse-shock run), emissivities 1e-9 and light curves 1e-10 against the oracle on
    the GPU's own tables, and every layer's light curve against the reference's uniform run of its
    grid size (LC_TOL).  The mixed-precision SSC contraction (precision 1) over the same mixed
    grids against the FP64 path: 1e-3."""
    from pyblastafterglowmag_amd import gpu, workloads
    refs = {32: load_golden("gauss_pw_ng32"), 48: load_golden("gauss_pw_ng48"), 64: load_golden("gauss_pw_small")}
    main, grb, struct = workloads.small("gauss_pw_ng_small")
    run = gpu.Run.from_workload(main, grb, struct)
    ngs = struct["layer_ngam"]
    print(f"[gpu] per-layer grids {ngs}", flush=True)
    pl = gpu.Pipeline([run])
    pl.dynamics()
    flux, ex = gpu.radiate(pl, want_intermediates=True, precision=0)
    flux = flux[0]
    assert not ex["dyn_status"].any() and not ex["ele_status"].any() and not ex["eats_status"].any()
    table_o, res_o = pyoracle.dyn_evolve(run.params, run.inits, run.tgrid)
    for il in range(run.n_layers):
        for n in DYN_CHECK:
            e = rel_err(ex["table"][il, abi.PV[n]], table_o[il, abi.PV[n]], 1e-300)
            assert e <= 1e-4, (il, n, e)
    assert [P.ngam for P in ex["pops"]] == [ngs[P.bw] for P in ex["pops"]]
    pops = (abi.Population * len(ex["pops"]))(*ex["pops"])
    jall_o, fall_o, st = pyoracle.ele_evolve(run.params, pops, ex["table"], want_f=True)
    assert not st.any()
    for q, P in enumerate(ex["pops"]):
        e = rel_err(ex["jall"][q], jall_o[q], 1e-9)
        assert e <= 1e-9, (q, e)
        # f_all rows have the largest grid as their stride; this population's spectrum comes first
        n = ngs[P.bw]
        assert rel_err(ex["fall"][q][:, :n], fall_o[q][:, :n], 1e-9) <= 1e-9, q
        assert not ex["fall"][q][:, n:].any()
    lays = (abi.EatsPwLayer * len(ex["layers"]))()
    for i, (bw, r, pfs, prs, iend) in enumerate(ex["layers"]):
        li = run.inits[bw]
        lays[i] = abi.EatsPwLayer(li.ctheta0, li.ncells, abi.cells_in_layer(li.ilayer), bw, r, pfs,
                                  1 if prs >= 0 else 0, iend)
    lc_o, st2 = pyoracle.eats_lightcurve_pw(run.params, lays, ex["table"], ex["jall"], run.tobs, run.nuobs)
    assert not st2.any()
    for il in range(run.n_layers):
        assert rel_err(flux[il], lc_o[il], 1e-30) <= 1e-10, il
        e = rel_err(flux[il], refs[ngs[il]]["lc"][il], 1e-30)
        print(f"  layer {il} ({ngs[il]} bins): vs reference {e:.3e}", flush=True)
        assert e <= LC_TOL["gauss_pw_small"], (il, e)
    mix, exm = gpu.radiate(pl, want_intermediates=True, precision=1)
    assert not exm["ele_status"].any()
    for il in range(run.n_layers):
        e = rel_err(mix[0, il], flux[il], 1e-8)
        assert e <= 1e-3, (il, e)
