#!/usr/bin/env python3
"""Light curves per second of the radiation hot path on one MI355X.

Workload (default): 21-layer Gaussian jet, forward + reverse shock, synchrotron + SSC on both
shocks, 200 electron Lorentz-factor bins x 96 comoving frequencies, 1000 burster-frame steps,
light curve at 96 nu_obs x 128 t_obs (theta_obs = 0.363 rad, n_ism = 10^-1.6 cm^-3,
E_iso,c = 10^54.1 erg, Gamma0 = 300, theta_c = 0.061, theta_w = 0.436).

One step = one batch of independent runs of that workload through the three GPU stages, with
every input and intermediate resident in device memory.  The dynamics of one run is a chain of
~3e5 dependent ODE steps on one wavefront (tens of seconds of latency at a small share of the
chip), while the electron and observer stages of a batch fill the whole chip, so batches are
pipelined: step s launches the dynamics of batch s on its own stream (host thread) and runs the
electron and observer stages of batch s - depth, whose dynamics are done, on the shared
radiation stream.  Only light curves whose dynamics, electrons and observer stage all ran inside
the timed steps are counted: of K timed steps, K - depth batches;
value = batch x (K - depth) / timed seconds.

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


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--batch", type=int, default=12, help="independent runs per step")
    ap.add_argument("--depth", type=int, default=4,
                    help="pipeline depth: batches whose dynamics run while the radiation stages of an older batch run")
    ap.add_argument("--dyn-cus", type=int, default=32,
                    help="compute units reserved for the dynamics streams (the radiation stream gets the others)")
    ap.add_argument("--precision", type=int, default=0, help="0: FP64 everywhere")
    ap.add_argument("--workload", default="gauss21_fsrs_ssc",
                    choices=["gauss21_fsrs_ssc", "tophat_fs_syn", "tophat64_pw", "lhs_members"],
                    help="lhs_members: every run a different member of the 1024-member Latin-hypercube batch")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the electron and observer parts of the CPU-baseline sample")
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
    return {"value": 1.0 / per_lc, "unit": "light curves/s", "cores": threads, "kind": "port",
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
             gpu.STAGE_SSC_TABLE: "ssc_table", gpu.STAGE_EATS_PW: "eats_pw_kernel"}
    if dom is None:
        crit = [gpu.STAGE_ELE, gpu.STAGE_EATS, gpu.STAGE_EATS_PW]
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
    return roof


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
    D = max(args.depth, 1)
    if args.workload == "lhs_members":     # batch b holds members b*batch ... (b+1)*batch - 1
        make = lambda b: [gpu.Run.from_workload(*_workload(workloads, args.workload, b * args.batch + i))
                          for i in range(args.batch)]
    else:
        one = gpu.Run.from_workload(*_workload(workloads, args.workload, 0))
        make = lambda b: [one] * args.batch
    run = make(0)[0]
    # D + 1 pipelines in a ring: batch b uses ring[b % (D + 1)]; its dynamics run on that pipeline's own
    # stream, its electron and observer stages on the shared radiation stream
    # the dynamics run on compute units [0, dyn_cus) and the electron and observer stages on the rest, so
    # that the persistent observer kernel, which holds its units until it ends, never keeps a batch's
    # dynamics from starting
    n_cu = gpu.device_cu_count()
    dyn_cus = min(max(args.dyn_cus, 1), n_cu - 1)
    rad_stream = gpu.CuStream(dyn_cus, n_cu - dyn_cus)
    ring = [gpu.Pipeline(make(k), stream=gpu.CuStream(0, dyn_cus), precision=args.precision, rad_stream=rad_stream)
            for k in range(D + 1)]
    gpu.timing_enable(True)
    threads = {}
    rejected = {}
    t_start = {}

    def launch(b):
        """Dynamics of batch b on its pipeline's stream, in a host thread (the call blocks until done)."""
        pl = ring[b % (D + 1)]
        if args.workload == "lhs_members" and b >= D + 1:
            pl.__init__(make(b), stream=pl.stream, precision=args.precision, rad_stream=rad_stream)
        t_start[b] = time.perf_counter()
        th = threading.Thread(target=pl.dynamics)
        th.start()
        threads[b] = th

    def radiation(b):
        pl = ring[b % (D + 1)]
        threads.pop(b).join()
        pl.electrons()
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
            print(f"[bench] batch {b}: {len(bad)} of {pl.n_runs} members flagged (no light curve): {sorted(bad)}",
                  file=sys.stderr, flush=True)

    last_b = args.warmup + args.steps - 1 - D   # the last batch whose radiation stages run

    def step(s):
        if s <= last_b:
            launch(s)
        if s >= D:
            radiation(s - D)
        print(f"[bench] step {s} done ({'dynamics of batch %d launched, ' % s if s <= last_b else ''}"
              f"{'light curves of batch %d' % (s - D) if s >= D else 'filling the pipeline'}) "
              f"t={time.perf_counter() - t_fill:.1f}s", file=sys.stderr, flush=True)

    t_fill = time.perf_counter()
    for s in range(args.warmup):
        step(s)
    torch.cuda.synchronize()
    for st in range(gpu.N_STAGES):
        gpu.timing_read(st, reset=True)
        gpu.timing_units(st, reset=True)
    t0 = time.perf_counter()
    for s in range(args.warmup, args.warmup + args.steps):
        step(s)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    for th in threads.values():
        th.join()
    # batches whose dynamics started inside the timed steps and whose light curves were done by their end
    counted = list(range(args.warmup, last_b + 1))
    pls = [ring[counted[0] % (D + 1)]] if counted else ring[:1]
    stage_ms = {s: gpu.timing_read(s) for s in range(gpu.N_STAGES)}
    if args.stages:
        for s, (ms, n) in stage_ms.items():
            print(f"[stage] {s}: {ms:.1f} ms over {n} launches", file=sys.stderr)
        print(f"[stage] wall per step {dt / args.steps * 1e3:.1f} ms; eats samples "
              f"{gpu.timing_units(gpu.STAGE_EATS):.3e}; piece-wise cells {gpu.timing_units(gpu.STAGE_EATS_PW):.3e}",
              file=sys.stderr)
    roof = roofline(stage_ms, pls, gpu)
    roof_dyn = roofline(stage_ms, pls, gpu, dom=gpu.STAGE_DYN) if stage_ms[gpu.STAGE_DYN][1] else None
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.workload, args.cpu_seconds)
    n_rej = sum(rejected.get(b, 0) for b in counted)
    lcs = (args.batch * len(counted) - n_rej) / dt
    line = {
        "metric": "light curves/sec (96nu x128t), plain FP64 vs mixed+surrogate GPU @<=1% err",
        "value": lcs, "unit": "light curves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f64+bf16", "data": "synthetic",
        "config": {"workload": args.workload, "runs_per_step": args.batch, "layers": run.n_layers,
                   "ngam": run.params.fs.ngam, "nfreq": run.params.fs.nfreq, "nt": len(run.tgrid),
                   "pairs": len(run.tobs), "precision": args.precision, "pipeline_depth": D,
                   "dynamics_cus": dyn_cus, "radiation_cus": n_cu - dyn_cus,
                   "batches_counted": len(counted), "light_curves_counted": args.batch * len(counted) - n_rej,
                   **({"members_flagged": n_rej, "members": "Latin hypercube, seed 20261203, runs b*batch .. "
                       "(b+1)*batch-1 per step"} if args.workload == "lhs_members" else {})},
        "roofline": roof, "roofline_dynamics": roof_dyn, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
