#!/usr/bin/env python3
"""Light curves per second of the radiation hot path on one MI355X.

Workload (default): 21-layer Gaussian jet, forward + reverse shock, synchrotron + SSC on both
shocks, 200 electron Lorentz-factor bins x 96 comoving frequencies, 1000 burster-frame steps,
light curve at 96 nu_obs x 128 t_obs (theta_obs = 0.363 rad, n_ism = 10^-1.6 cm^-3,
E_iso,c = 10^54.1 erg, Gamma0 = 300, theta_c = 0.061, theta_w = 0.436).

One step = the three GPU stages (dynamics, electron kinetics + emissivity, observer
quadrature) over one batch of independent runs of that workload, with every input and
intermediate resident in device memory.  value = batch x steps / timed seconds.

The JSON line carries a roofline entry for the kernel that takes the most device time
(events recorded on the stream it runs on, by the library itself) and a CPU baseline: the
CPU restatement of the same algorithm (oracle/) timed on this host on a bounded sample of the
same workload and scaled to light curves per second.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64, help="independent runs per step")
    ap.add_argument("--precision", type=int, default=0, help="0: FP64 everywhere")
    ap.add_argument("--workload", default="gauss21_fsrs_ssc", choices=["gauss21_fsrs_ssc", "tophat_fs_syn"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stages", action="store_true", help="also print per-stage device times to stderr")
    return ap.parse_args()


def algorithmic_work(pl):
    """Algorithmic FP64 flops per stage of one pipeline step (counted per unit below).

    electron stage, per population step (a step with electrons, table mflags):
        SSC contraction   2 nf ng nf   (photon-density x kernel table x spectrum; only with SSC)
        synchrotron       2 ng nf      (spectrum x synchrotron kernel)
    observer stage: not counted (adaptive; reported as null)
    dynamics stage: not counted (adaptive; reported as null)
    """
    import numpy as np
    from pyblastafterglowmag_amd import abi
    mfl = pl.d_table[:, abi.PV["mflags"], :].cpu().numpy().astype(np.int64)
    nf = pl.nf
    ng = pl.runs[0].params.fs.ngam
    fl = 0.0
    for q, P in enumerate(pl.pops):
        bit = 1 if P.shock == 0 else 2
        nsteps = int(np.count_nonzero(mfl[P.bw] & bit))
        pr = pl.params[P.run]
        sh = pr.fs if P.shock == 0 else pr.rs
        per = 2.0 * ng * nf + (2.0 * nf * ng * nf if sh.ssc else 0.0)
        fl += nsteps * per
    return {"ele": fl}


def cpu_baseline(workload, budget_s):
    """Time the CPU restatement on a bounded sample of one run of the workload and scale it.

    Sample: the full dynamics of one run (all layers), the electron stage of a subset of its
    populations, and the observer quadrature of every layer at a subset of (t_obs, nu_obs)
    pairs; light curves/s = 1 / (t_dyn + t_ele/k_pop * n_pop + t_eats/k_pair * n_pair).
    """
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from pyblastafterglowmag_amd import abi, gpu, workloads
    run = gpu.Run.from_workload(*getattr(workloads, workload)())
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    t0 = time.perf_counter()
    table, res = pyoracle.dyn_evolve(run.params, run.inits, run.tgrid)
    t_dyn = time.perf_counter() - t0
    pops, lays = gpu.populations((abi.RunParams * 1)(run.params), res, 1, run.n_layers)
    n_pop = len(pops)
    # electron stage: populations in chunks of `threads` until a third of the budget is used
    k_pop, t_ele = 0, 0.0
    jall = np.zeros((n_pop, len(run.tgrid), run.params.fs.nfreq))
    while k_pop < n_pop and t_ele < budget_s / 3:
        chunk = list(range(k_pop, min(n_pop, k_pop + threads)))
        parr = (abi.Population * len(chunk))(*[pops[q] for q in chunk])
        t0 = time.perf_counter()
        j, _, _ = pyoracle.ele_evolve(run.params, parr, table)
        t_ele += time.perf_counter() - t0
        jall[chunk] = j
        k_pop += len(chunk)
    for q in range(k_pop, n_pop):       # unmeasured populations reuse a measured spectrum table
        jall[q] = jall[q % max(k_pop, 1)]
    larr = (abi.EatsLayer * len(lays))()
    for i, (bw, r, pfs, prs, iend) in enumerate(lays):
        li = run.inits[bw]
        larr[i] = abi.EatsLayer(li.theta_c_l, li.theta_c_h, bw, r, pfs, prs, iend)
    n_pair = len(run.tobs)
    # observer stage: pairs spread over the grid, in growing sets until the budget is used
    k_pair, t_eats = 0, 0.0
    stride = 64
    while t_eats < budget_s / 2 and k_pair < n_pair and stride >= 1:
        sel = np.arange(0, n_pair, stride)
        t0 = time.perf_counter()
        pyoracle.eats_lightcurve(run.params, larr, table, jall, run.tobs[sel], run.nuobs[sel])
        t_eats = time.perf_counter() - t0
        k_pair = len(sel)
        stride //= 2
    per_lc = t_dyn + t_ele / max(k_pop, 1) * n_pop + t_eats / max(k_pair, 1) * n_pair
    return {"value": 1.0 / per_lc, "unit": "light curves/s", "cores": threads, "kind": "port",
            "sample": (f"oracle (CPU restatement, OpenMP over populations and layers) on one {workload} run: "
                       f"full dynamics {t_dyn:.2f}s, electrons of {k_pop}/{n_pop} populations {t_ele:.2f}s, "
                       f"observer quadrature of all layers at {k_pair}/{n_pair} pairs {t_eats:.2f}s; "
                       f"scaled to {per_lc:.1f}s per light curve")}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: --gpus {args.gpus} requested; this build runs on one GPU only "
              f"(the radiation hot path is measured on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyblastafterglowmag_amd import gpu, workloads
    if not torch.cuda.is_available():
        print("bench.py: no ROCm device visible", file=sys.stderr)
        sys.exit(1)
    run = gpu.Run.from_workload(*getattr(workloads, args.workload)())
    runs = [run] * args.batch
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        pl = gpu.Pipeline(runs, stream=stream, precision=args.precision)
        gpu.timing_enable(True)
        for _ in range(args.warmup):
            pl.step()
        for st in range(4):
            gpu.timing_read(st, reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            pl.step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    dst, est, lst = pl.status()
    if dst.any() or est.any() or lst.any():
        print(f"bench.py: non-zero status dyn={sorted(set(dst.tolist()))} ele={sorted(set(est.tolist()))} "
              f"eats={sorted(set(lst.tolist()))}", file=sys.stderr)
        sys.exit(3)
    names = {gpu.STAGE_DYN: "dyn_kernel", gpu.STAGE_ELE: "ele_kernel", gpu.STAGE_EATS: "eats_kernel",
             gpu.STAGE_SSC_TABLE: "ssc_table"}
    stage_ms = {names[s]: gpu.timing_read(s) for s in range(4)}
    if args.stages:
        for k, (ms, n) in stage_ms.items():
            print(f"[stage] {k}: {ms:.3f} ms over {n} launches", file=sys.stderr)
        print(f"[stage] wall per step {dt / args.steps * 1e3:.1f} ms", file=sys.stderr)
    dom = max(stage_ms, key=lambda k: stage_ms[k][0])
    ms_dom, n_dom = stage_ms[dom]
    avg_s = ms_dom / max(n_dom, 1) / 1e3
    work = algorithmic_work(pl)
    roof = {"kernel": dom, "avg_launch_ms": ms_dom / max(n_dom, 1), "bound": "mfma", "achieved": None,
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": None, "traffic": None,
            "dtype_peak": "fp64"}
    if dom == "ele_kernel":
        ach = work["ele"] / avg_s / 1e12
        roof.update(achieved=ach, frac=ach / PEAK_FP64_TFLOPS)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.workload, args.cpu_seconds)
    lcs = args.batch * args.steps / dt
    line = {
        "metric": "light curves/sec (96nu x128t), plain FP64 vs mixed+surrogate GPU @<=1% err",
        "value": lcs, "unit": "light curves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f64+bf16", "data": "synthetic",
        "config": {"workload": args.workload, "runs_per_step": args.batch, "layers": run.n_layers,
                   "ngam": run.params.fs.ngam, "nfreq": run.params.fs.nfreq, "nt": len(run.tgrid),
                   "pairs": len(run.tobs), "precision": args.precision},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
