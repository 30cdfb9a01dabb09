#!/usr/bin/env python3
"""Light curves per second of the radiation hot path on one MI355X.

Workload (default): 21-layer Gaussian jet, forward + reverse shock, synchrotron + SSC on both
shocks, 200 electron Lorentz-factor bins x 96 comoving frequencies, 1000 burster-frame steps,
light curve at 96 nu_obs x 128 t_obs (theta_obs = 0.363 rad, n_ism = 10^-1.6 cm^-3,
E_iso,c = 10^54.1 erg, Gamma0 = 300, theta_c = 0.061, theta_w = 0.436).

Two legs run one after the other, each a pipeline of batches of independent runs of that workload
with every input and intermediate resident in device memory:
  mixed (the value): SSC contraction in BF16 hi/lo on MFMA with FP32 accumulation and the
        frequency-vectorised observer stage (one adaptive Gauss-Kronrod tree per (layer, t_obs)
        shared by all 96 frequencies, rtol 1e-2); the dynamics and the Chang-Cooper update FP64;
  fp64: every stage FP64, the observer stage as the reference's per-pair CADRE quadrature.
The dynamics of one run is a chain of ~3e5 dependent ODE steps on one wavefront (tens of seconds
of latency on a small share of the chip), while the electron and observer stages of a batch fill
the chip, so step s launches the dynamics of batch s on its own stream (compute units
[0, dyn_cus)) and runs the electron and observer stages of batch s - depth, whose dynamics are
done, on the radiation stream (the other units).  Only light curves whose dynamics, electrons and
observer stage all ran inside the timed steps are counted: of K timed steps, K - depth batches;
value = batch x (K - depth) / timed seconds.  After the mixed leg, the FP64 electron and observer
stages are run on 4 of its runs, and the largest relative difference of their layer-summed light
curves is reported (the mixed path's budget is 1 %).

The JSON line carries a roofline entry for the kernel with the most device time in the timed
steps (HIP events recorded by the library on the stream each kernel runs on) and a CPU baseline:
the CPU restatement of the same algorithm (oracle/) timed on this host on a bounded sample of one
run of the same workload, scaled to light curves per second.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# one hardware queue per stream of the pipeline (depth + 1 dynamics streams and the radiation stream):
# streams that share a queue run their kernels one after another, which would put the observer
# stage behind a dynamics kernel; set before the HIP runtime starts
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# algorithmic work per unit, counted from the kernels' source
FLOPS_PER_RHS = {0: 250.0, 1: 330.0}   # one right-hand side of one blastwave (forward shock / with reverse shock)
BYTES_PER_SAMPLE = 144.0               # one EATS integrand sample: 7 state columns at 2 steps + 4 emissivities
FLOPS_PER_CELL = 152.0                 # one piece-wise cell flux, forward-shock term (reverse-shock term not counted)
# one node of the vectorised observer: the surface geometry (mu, the arrival-time search over the
# steps, 5 state columns interpolated in log at the emission time, the Doppler factors and weights);
# per frequency at that node: nu', the grid index, the bilinear blend of 4 emissivities, the
# luminosity and both quadrature sums (forward-shock term; a reverse shock adds 32 where present)
FLOPS_PER_NODE = 130.0
FLOPS_PER_NODE_NU = 36.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed steps of the mixed leg")
    ap.add_argument("--warmup", type=int, default=4, help="warm-up steps of the mixed leg (>= depth fills it)")
    ap.add_argument("--batch", type=int, default=176, help="independent runs per step, mixed leg")
    ap.add_argument("--depth", type=int, default=4,
                    help="pipeline depth: batches whose dynamics run while the radiation stages of an older batch run")
    ap.add_argument("--dyn-cus", type=int, default=176,
                    help="compute units reserved for the dynamics streams (the radiation stream gets the others)")
    ap.add_argument("--vec-rtol", type=float, default=1e-2, help="quadrature tolerance of the vectorised observer")
    ap.add_argument("--mode", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--fp64-batch", type=int, default=12)
    ap.add_argument("--fp64-depth", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--fp64-warmup", type=int, default=2)
    ap.add_argument("--fp64-dyn-cus", type=int, default=32)
    ap.add_argument("--precision", type=int, default=1, help="SSC contraction of the mixed leg: 1 BF16 hi/lo, 0 FP64")
    ap.add_argument("--surrogate", action="store_true",
                    help="mixed leg with the learned SSC emissivity (trained at set-up, untimed, on the FP64 cells of "
                         "Latin-hypercube members other than the workload) in place of the BF16 contraction")
    ap.add_argument("--train-members", default="3,5,6,7", help="Latin-hypercube members whose cells train the surrogate")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--check-tol", type=float, default=0.2,
                    help="surrogate: relative error of a sampled check above which a population falls back")
    ap.add_argument("--workload", default="gauss21_fsrs_ssc",
                    choices=["gauss21_fsrs_ssc", "tophat_fs_syn", "tophat64_pw", "lhs_members"],
                    help="lhs_members: every run a different member of the 1024-member Latin-hypercube batch")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the electron and observer parts of the CPU-baseline sample")
    ap.add_argument("--ele-graph", type=int, default=0, choices=[0, 1],
                    help="capture the electron stage's per-step launches in one graph (PBAGPU_OPT_ELE_GRAPH)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stages", action="store_true", help="also print per-stage device times to stderr")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device-to-device (reference kernel for the FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def _workload(workloads, name, member):
    return workloads.lhs_member(member) if name == "lhs_members" else getattr(workloads, name)()


def rejected_runs(pl, dst, est, lst):
    """Runs of a batch with any non-zero status bit (dynamics, electrons or observer stage)."""
    bad = {i // pl.bpr for i in range(len(dst)) if dst[i]}
    bad |= {pl.pops[q].run for q in range(len(est)) if est[q]}
    bad |= {pl.layers[k][1] for k in range(len(lst)) if lst[k]}
    return bad


def cpu_baseline(workload, budget_s):
    """Time the CPU restatement on a bounded sample of one run of the workload and scale it.

    Sample: the full dynamics of one run (all layers; it is one ODE system), the electron stage
    of a subset of its populations, and the observer quadrature of every layer at a subset of
    (t_obs, nu_obs) pairs spread over the grid;
    light curves/s = 1 / (t_dyn + t_ele / k_pop * n_pop + t_eats / k_pair * n_pair).
    """
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from pyblastafterglowmag_amd import abi, gpu, workloads
    run = gpu.Run.from_workload(*_workload(workloads, workload, 0))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    t0 = time.perf_counter()
    table, res = pyoracle.dyn_evolve(run.params, run.inits, run.tgrid)
    t_dyn = time.perf_counter() - t0
    print(f"[cpu] dynamics {t_dyn:.1f}s", file=sys.stderr, flush=True)
    pops, lays = gpu.populations((abi.RunParams * 1)(run.params), res, 1, run.n_layers)
    n_pop = len(pops)
    k_pop, t_ele = 0, 0.0
    jall = np.zeros((n_pop, len(run.tgrid), run.params.fs.nfreq))
    while k_pop < n_pop and t_ele < budget_s / 2:
        chunk = list(range(k_pop, min(n_pop, k_pop + threads)))
        parr = (abi.Population * len(chunk))(*[pops[q] for q in chunk])
        t0 = time.perf_counter()
        j, _, _ = pyoracle.ele_evolve(run.params, parr, table)
        t_ele += time.perf_counter() - t0
        jall[chunk] = j
        k_pop += len(chunk)
        print(f"[cpu] electrons {k_pop}/{n_pop} populations {t_ele:.1f}s", file=sys.stderr, flush=True)
    for q in range(k_pop, n_pop):       # unmeasured populations reuse a measured emissivity table
        jall[q] = jall[q % max(k_pop, 1)]
    pw = run.method == "piece-wise"
    larr = (abi.EatsPwLayer if pw else abi.EatsLayer) * len(lays)
    larr = larr()
    for i, (bw, r, pfs, prs, iend) in enumerate(lays):
        li = run.inits[bw]
        larr[i] = (abi.EatsPwLayer(li.ctheta0, li.ncells, abi.cells_in_layer(li.ilayer), bw, r, pfs, 1 if prs >= 0 else 0,
                                   iend) if pw else abi.EatsLayer(li.theta_c_l, li.theta_c_h, bw, r, pfs, prs, iend))
    eats = pyoracle.eats_lightcurve_pw if pw else pyoracle.eats_lightcurve
    n_pair = len(run.tobs)
    k_pair, t_eats = 0, 0.0
    stride = 256
    while t_eats < budget_s / 4 and k_pair < n_pair and stride >= 1:
        sel = np.arange(stride // 2, n_pair, stride)
        t0 = time.perf_counter()
        eats(run.params, larr, table, jall, run.tobs[sel], run.nuobs[sel])
        t_eats = time.perf_counter() - t0
        k_pair = len(sel)
        stride //= 2
        print(f"[cpu] observer {k_pair} pairs {t_eats:.1f}s", file=sys.stderr, flush=True)
    per_lc = t_dyn + t_ele / max(k_pop, 1) * n_pop + t_eats / max(k_pair, 1) * n_pair
    # One light curve of this workload's parfile, one thread each, on an 8-core Xeon at 2.7 GHz: the
    # reference program (oracle/_ref/pba.out, built from its sources) 1494.1 s, the oracle 1137.8 s
    # (dynamics 31.3 s, electrons 152.8 s, observer 953.6 s).  The reference is this much slower than
    # its restatement on the same host.
    ref_over_port = 1494.1 / 1137.8
    return {"value": 1.0 / per_lc, "unit": "light curves/s", "cores": threads, "kind": "port",
            "calibration": {"reference_s": 1494.1, "port_s": 1137.8, "reference_over_port": round(ref_over_port, 3),
                            "host": "8-core Xeon 2.7 GHz, one thread each, the gauss21_fsrs_ssc parfile",
                            "implied_reference_value": 1.0 / per_lc / ref_over_port},
            "sample": (f"oracle (CPU restatement, OpenMP over populations and layers, {threads} threads) on one "
                       f"{workload} run: full dynamics {t_dyn:.1f}s, electrons of {k_pop}/{n_pop} populations "
                       f"{t_ele:.1f}s, observer quadrature of all layers at {k_pair}/{n_pair} pairs {t_eats:.1f}s; "
                       f"scaled to {per_lc:.0f}s per light curve")}


def roofline(stage_ms, pls, gpu, dom=None):
    """Roofline entry of one stage (default: the radiation-stream stage with the most device time in
    the timed steps, which sets the step time; the dynamics overlap it on their own units)."""
    import numpy as np
    from pyblastafterglowmag_amd import abi
    names = {gpu.STAGE_DYN: "dyn_kernel", gpu.STAGE_ELE: "ele (per-step kernels)", gpu.STAGE_EATS: "eats_kernel",
             gpu.STAGE_SSC_TABLE: "ssc_table", gpu.STAGE_EATS_PW: "eats_pw_kernel",
             gpu.STAGE_EATS_VEC: "eats_vec_kernel"}
    if dom is None:
        crit = [gpu.STAGE_ELE, gpu.STAGE_EATS, gpu.STAGE_EATS_PW, gpu.STAGE_EATS_VEC]
        dom = max(crit, key=lambda s: stage_ms[s][0])
    ms, launches = stage_ms[dom]
    avg_s = ms / max(launches, 1) / 1e3
    roof = {"kernel": names[dom], "avg_launch_ms": ms / max(launches, 1), "bound": None, "achieved": None,
            "peak": None, "unit": None, "frac": None, "traffic": None}
    if dom == gpu.STAGE_EATS:
        samples = gpu.timing_units(gpu.STAGE_EATS) / max(launches, 1)
        ach = samples * BYTES_PER_SAMPLE / avg_s / 1e9
        roof.update(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                    samples_per_launch=samples, bytes_per_sample=BYTES_PER_SAMPLE)
    elif dom == gpu.STAGE_EATS_VEC:
        # Not HBM-bound: its tables stay in L2 / the infinity cache (counted HBM traffic, FETCH_SIZE +
        # WRITE_SIZE, is ~0.5 GB/s over a launch).  The limit is FP64 issue and the latency of the
        # dependent table gathers, so the line is stated against the FP64 peak.
        nodes = gpu.timing_units(gpu.STAGE_EATS_VEC) / max(launches, 1)
        r0 = pls[0].runs[0]
        nnu = int(np.count_nonzero(r0.tobs == r0.tobs[0]))   # the pairs are t-major
        fl = nodes * (FLOPS_PER_NODE + nnu * FLOPS_PER_NODE_NU)
        ach = fl / avg_s / 1e12
        roof.update(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s", frac=ach / PEAK_FP64_TFLOPS,
                    dtype_peak="fp64", nodes_per_launch=nodes, flops_per_node=FLOPS_PER_NODE,
                    flops_per_node_frequency=FLOPS_PER_NODE_NU, frequencies=nnu,
                    limiter="FP64 issue and dependent L2 gathers (HBM traffic well below peak)")
    elif dom == gpu.STAGE_EATS_PW:
        pl = pls[0]
        cells = gpu.timing_units(gpu.STAGE_EATS_PW) / max(launches, 1)
        fl = cells * FLOPS_PER_CELL
        ach = fl / avg_s / 1e12
        roof.update(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s", frac=ach / PEAK_FP64_TFLOPS,
                    dtype_peak="fp64", cells_per_launch=cells, flops_per_cell=FLOPS_PER_CELL)
    elif dom == gpu.STAGE_DYN:
        pl = pls[0]
        flops = 0.0
        for r in range(pl.n_runs):
            for i in range(pl.bpr):
                res = pl.res[r * pl.bpr + i]
                flops += res.n_rhs * FLOPS_PER_RHS[1 if res.fsrs else 0]
        ach = flops / avg_s / 1e12
        roof.update(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s", frac=ach / PEAK_FP64_TFLOPS,
                    dtype_peak="fp64", flops_per_rhs=FLOPS_PER_RHS)
    elif dom == gpu.STAGE_ELE:
        pl = pls[0]
        mfl = pl.d_table[:, abi.PV["mflags"], :].cpu().numpy().astype(np.int64)
        ng, nf = pl.runs[0].params.fs.ngam, pl.nf
        fl = 0.0
        for P in pl.pops:
            bit = 1 if P.shock == 0 else 2
            n = int(np.count_nonzero(mfl[P.bw] & bit))
            sh = pl.params[P.run].fs if P.shock == 0 else pl.params[P.run].rs
            fl += n * (2.0 * ng * nf + (2.0 * nf * ng * nf if sh.ssc else 0.0))
        ach = fl / avg_s / 1e12
        roof.update(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s", frac=ach / PEAK_FP64_TFLOPS,
                    dtype_peak="fp64")
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        key = f"{roof['kernel']}|{pls[0].runs[0].n_layers}x{pls[0].n_runs}"
        roof["traffic"] = tr.get(key)
        if roof["traffic"]:
            roof["hbm_gbs_from_traffic"] = roof["traffic"] / avg_s / 1e9
    return roof


def train_surrogate(args, gpu, torch, workloads):
    """Untimed set-up of the surrogate leg: the FP64 electron stage of Latin-hypercube members
    (parameters other than the workload's, on its grids) gives the training cells; the network is
    trained by the library's BF16 forward / backward kernels and Adam.  Returns (SscSurrogate, info)."""
    from pyblastafterglowmag_amd import surrogate
    t0 = time.perf_counter()
    members = [int(x) for x in args.train_members.split(",") if x]
    pl = gpu.Pipeline([gpu.Run.from_workload(*workloads.lhs_member(i, ragged_nt=False)) for i in members])
    pl.dynamics()
    pl.electrons(want_f=True)
    f, w, B, sh = surrogate.cells_from_pipeline(pl)
    x, sc, ylog, _ = surrogate.targets(sh, f, w, B)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sur, hist = surrogate.train(x, ylog, steps=args.train_steps, batch=4096)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    sur.check_tol = args.check_tol
    info = {"train_members": members, "train_cells": int(x.shape[0]), "train_steps": args.train_steps,
            "pairs_seconds": t1 - t0, "train_seconds": t2 - t1, "loss_first": hist[0][1], "loss_last": hist[-1][1],
            "check_tol": args.check_tol, "check_every": sur.check_every}
    print(f"[bench:surrogate] {info}", file=sys.stderr, flush=True)
    del pl
    return sur, info


def run_leg(name, args, gpu, torch, make, batch, depth, warmup, steps, dyn_cus, precision, vec, sur=None):
    """One pipelined leg; returns its measurements and the ring of pipelines (for the roofline)."""
    D = max(depth, 1)
    n_cu = gpu.device_cu_count()
    dyn_cus = min(max(dyn_cus, 1), n_cu - 1)
    rad_stream = gpu.CuStream(dyn_cus, n_cu - dyn_cus)
    dyn_stream = [gpu.CuStream(0, dyn_cus) for _ in range(D + 1)]
    ring = [gpu.Pipeline(make(k, batch), stream=dyn_stream[k], precision=precision, rad_stream=rad_stream)
            for k in range(D + 1)]
    threads, rejected = {}, {}
    last_b = warmup + steps - 1 - D   # the last batch whose radiation stages run

    def launch(b):
        pl = ring[b % (D + 1)]
        if args.workload == "lhs_members" and b >= D + 1:
            pl.__init__(make(b, batch), stream=pl.stream, precision=precision, rad_stream=rad_stream)
        th = threading.Thread(target=pl.dynamics)
        th.start()
        threads[b] = th

    sur_stats = {"inferred": 0, "checked": 0, "failed": 0, "fallback": 0, "max_err": 0.0}

    def radiation(b):
        pl = ring[b % (D + 1)]
        threads.pop(b).join()
        if sur is not None:
            st = pl.electrons_surrogate(sur)
            for k in ("inferred", "checked", "failed", "fallback"):
                sur_stats[k] += getattr(st, k)
            sur_stats["max_err"] = max(sur_stats["max_err"], st.max_err)
        else:
            pl.electrons()
        if vec and pl.method != "piece-wise":
            pl.observer_vec(args.vec_rtol)
        else:
            pl.observer()
        dst, est, lst = pl.status()
        if dst.any() or est.any() or lst.any():
            bad = rejected_runs(pl, dst, est, lst)
            if args.workload != "lhs_members":
                print(f"bench.py: non-zero status dyn={sorted(set(dst.tolist()))} ele={sorted(set(est.tolist()))} "
                      f"eats={sorted(set(lst.tolist()))}", file=sys.stderr)
                sys.exit(3)
            # a member the reference would stop on (its status bits mark where) yields no light curve
            rejected[b] = len(bad)

    t_fill = time.perf_counter()

    def step(s):
        if s <= last_b:
            launch(s)
        if s >= D:
            radiation(s - D)
        print(f"[bench:{name}] step {s} done ({'dynamics of batch %d launched, ' % s if s <= last_b else ''}"
              f"{'light curves of batch %d' % (s - D) if s >= D else 'filling the pipeline'}) "
              f"t={time.perf_counter() - t_fill:.1f}s", file=sys.stderr, flush=True)

    for s in range(warmup):
        step(s)
    torch.cuda.synchronize()
    for st in range(gpu.N_STAGES):
        gpu.timing_read(st, reset=True)
        gpu.timing_units(st, reset=True)
    t0 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        step(s)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    for th in threads.values():
        th.join()
    stage_ms = {s: gpu.timing_read(s) for s in range(gpu.N_STAGES)}
    counted = list(range(warmup, last_b + 1))
    n_rej = sum(rejected.get(b, 0) for b in counted)
    lcs = batch * len(counted) - n_rej
    pls = [ring[counted[0] % (D + 1)]] if counted else ring[:1]
    out = {"value": lcs / dt, "ms_per_step": dt / steps * 1e3, "steps": steps, "warmup": warmup, "runs_per_step": batch,
           "pipeline_depth": D, "dynamics_cus": dyn_cus, "radiation_cus": n_cu - dyn_cus,
           "batches_counted": len(counted), "light_curves_counted": lcs,
           "roofline": roofline(stage_ms, pls, gpu),
           "roofline_dynamics": roofline(stage_ms, pls, gpu, dom=gpu.STAGE_DYN) if stage_ms[gpu.STAGE_DYN][1] else None}
    if args.workload == "lhs_members":
        out["members_flagged"] = n_rej
    if sur is not None:
        out["surrogate_population_steps"] = sur_stats
    if args.stages:
        for st_, (ms, n) in stage_ms.items():
            print(f"[stage:{name}] {st_}: {ms:.1f} ms over {n} launches", file=sys.stderr)
    return out, ring, [rad_stream] + dyn_stream


def mixed_error(pl, gpu, n_check=4):
    """Largest relative difference of the layer-summed light curves of the first n_check runs of a
    pipeline between its (mixed) result and the FP64 electron and observer stages on the same
    dynamics tables (those runs only, on a stream over every compute unit)."""
    import numpy as np
    mix = pl.flux()[:n_check].sum(axis=1)
    nb = n_check * pl.bpr
    stream = gpu.CuStream(0, 0)
    ref_pl = gpu.Pipeline(pl.runs[:n_check], stream=stream, precision=0)
    ref_pl.use_tables(pl.d_table[:nb].cpu().numpy(), [pl.res[i] for i in range(nb)])
    ref_pl.electrons()
    ref_pl.observer()
    ref = ref_pl.flux().sum(axis=1)
    del ref_pl
    stream.close()
    worst = 0.0
    for a, b in zip(mix, ref):
        m = np.abs(b) > 1e-12 * np.abs(b).max()
        if m.any():
            worst = max(worst, float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m]))))
    return worst


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: --gpus {args.gpus} requested; the library runs on one GPU only "
              f"(the radiation hot path is measured on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyblastafterglowmag_amd import gpu, workloads
    if not torch.cuda.is_available():
        print("bench.py: no ROCm device visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "lhs_members":     # batch b holds members b*batch ... (b+1)*batch - 1
        make = lambda b, n: [gpu.Run.from_workload(*_workload(workloads, args.workload, b * n + i)) for i in range(n)]
    else:
        one = gpu.Run.from_workload(*_workload(workloads, args.workload, 0))
        make = lambda b, n: [one] * n
    run = make(0, 1)[0]
    gpu.timing_enable(True)
    gpu.set_option(gpu.OPT_ELE_GRAPH, args.ele_graph)
    legs = {}
    sur, sur_info = (None, None)
    if args.surrogate and args.mode in ("both", "mixed"):
        sur, sur_info = train_surrogate(args, gpu, torch, workloads)
    if args.mode in ("both", "mixed"):
        legs["mixed"], ring, streams = run_leg("mixed", args, gpu, torch, make, args.batch, args.depth, args.warmup,
                                               args.steps, args.dyn_cus, args.precision, vec=True, sur=sur)
        if sur_info:
            legs["mixed"]["surrogate"] = sur_info
        t0 = time.perf_counter()
        legs["mixed"]["max_rel_err_vs_fp64"] = mixed_error(ring[0], gpu, n_check=min(4, args.batch))
        print(f"[bench:mixed] light-curve check against the FP64 stages on 4 runs: "
              f"{legs['mixed']['max_rel_err_vs_fp64']:.3e} ({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
        del ring
        for s_ in streams:
            s_.close()
    if args.mode in ("both", "fp64"):
        steps = args.steps if args.mode == "fp64" else args.fp64_steps
        warm = args.warmup if args.mode == "fp64" else args.fp64_warmup
        legs["fp64"], ring, streams = run_leg("fp64", args, gpu, torch, make, args.fp64_batch, args.fp64_depth, warm,
                                              steps, args.fp64_dyn_cus, 0, vec=False)
        del ring
        for s_ in streams:
            s_.close()
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.workload, args.cpu_seconds)
    head = legs["mixed"] if "mixed" in legs else legs["fp64"]
    line = {
        "metric": "light curves/sec (96nu x128t), plain FP64 vs mixed+surrogate GPU @<=1% err",
        "value": head["value"], "unit": "light curves/s", "n_gpus": 1, "steps": head["steps"],
        "warmup": head["warmup"], "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+bf16" if "mixed" in legs else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "layers": run.n_layers, "ngam": run.params.fs.ngam,
                   "nfreq": run.params.fs.nfreq, "nt": len(run.tgrid), "pairs": len(run.tobs),
                   "legs": {k: {kk: vv for kk, vv in v.items() if not kk.startswith("roofline")}
                            for k, v in legs.items()},
                   "value_leg": "mixed" if "mixed" in legs else "fp64",
                   "electron_stage_graph": bool(args.ele_graph),
                   "mixed": ("SSC emissivity from the learned surrogate (BF16 MFMA network, FP32 accumulation; sampled "
                             "checks against the plain contraction with per-population fallback)" if sur is not None
                             else "SSC contraction BF16 hi/lo on MFMA (FP32 accumulation)")
                            + f", frequency-vectorised observer stage rtol {args.vec_rtol:g}; dynamics and Chang-Cooper FP64",
                   **({"members": "Latin hypercube, seed 20261203, runs b*batch .. (b+1)*batch-1 per step"}
                      if args.workload == "lhs_members" else {})},
        "roofline": head["roofline"], "roofline_dynamics": head["roofline_dynamics"],
        "roofline_fp64_leg": legs["fp64"]["roofline"] if ("fp64" in legs and "mixed" in legs) else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
