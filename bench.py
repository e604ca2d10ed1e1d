#!/usr/bin/env python3
"""Benchmark: FP64 RADM2 cell-integrations/s on one MI355X.

Workload: 1,048,576 synthetic cells (128x128x64), dt = 600 s, rtol 1e-3,
atol 1 molec/cm3, seed 0, zenith U(0,100) deg (10% night cells), 5% plume
cells (NOx 20-500 ppb, VOC x10).  A "step" is one integration of every
cell over dt (one cell-integration per cell), inputs resident in HBM.

Paths timed: the one-shot grid kernel (one lane per cell for the whole
integration) and the persistent work-stealing kernel; the JSON line reports
the persistent path as `value`, with the grid path and the speed-up beside it.
`roofline` is the FP64 compute roofline of the solver kernel (the only
kernel of a step apart from a 1-element counter reset): algorithmic flops
per launch from the per-cell accepted/rejected counts, over the kernel's
HIP-event duration on its stream.  `cpu_baseline` is the C oracle (OpenMP
over all host cores) on a bounded sample of the same cells.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "radm2-solver-multi-gpu_amd")
sys.path.insert(0, PKG)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1048576)
    ap.add_argument("--dt", type=float, default=600.0)
    ap.add_argument("--rtol", type=float, default=1e-3)
    ap.add_argument("--atol", type=float, default=1.0)
    ap.add_argument("--paths", default="grid,persistent")
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (calibrates HBM byte counters)")
    return ap.parse_args()


def run_path(solver, batch, conc0, sched, args, torch):
    import radm2_amd
    solver.set_schedule(sched, args.block, args.refill)
    for _ in range(args.warmup):
        batch.conc.copy_(conc0)
        solver.integrate(batch, args.dt, args.rtol, args.atol)
        solver.sync()
    kernel_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.conc.copy_(conc0)
        torch.cuda.synchronize()
        solver.integrate(batch, args.dt, args.rtol, args.atol)
        solver.sync()
        kernel_ms.append(solver.elapsed_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nacc = batch.nacc.to(torch.int64)
    nrej = batch.nrej.to(torch.int64)
    st = batch.status
    return dict(wall_s=wall, kernel_ms=kernel_ms, nacc=int(nacc.sum()), nrej=int(nrej.sum()),
                nstep_max=int(batch.nstep.max()), nstep_mean=float(batch.nstep.double().mean()),
                ok=int((st == radm2_amd.CELL_OK).sum()))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(cells are independent, so the path shards by contiguous slices -- "
              "see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_probe:
        # a plain copy of known size: 1 GiB read + 1 GiB written
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8}))
        return

    import radm2_amd
    from radm2_amd import synth, work

    n = args.cells
    t = time.perf_counter()
    cells = synth.make_cells(n, 0, 100.0, 0.05)
    gen_s = time.perf_counter() - t
    dev = synth.to_device(cells)
    conc0 = dev["conc"].clone()
    batch = radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"],
                                dev["jrate"], rc_n2o5=dev["rc_n2o5"])
    solver = radm2_amd.Solver(0)
    results = {}
    for name in args.paths.split(","):
        sched = {"grid": radm2_amd.SCHED_GRID, "persistent": radm2_amd.SCHED_PERSISTENT}[name]
        results[name] = run_path(solver, batch, conc0, sched, args, torch)
        r = results[name]
        print("[bench] %s: wall %.3f s for %d steps, kernel ms %s, ok %d/%d, attempts mean %.1f max %d"
              % (name, r["wall_s"], args.steps, ["%.1f" % x for x in r["kernel_ms"]], r["ok"], n,
                 r["nstep_mean"], r["nstep_max"]), file=sys.stderr, flush=True)

    head = "persistent" if "persistent" in results else list(results)[0]
    r = results[head]
    ms_per_step = 1000.0 * r["wall_s"] / args.steps
    value = n / (ms_per_step / 1000.0)
    kms = float(np.median(r["kernel_ms"]))
    flop = work.PER_STEP * r["nacc"] + work.PER_ATTEMPT * (r["nacc"] + r["nrej"]) + work.RCONST * n
    achieved = flop / (kms * 1e-3) / 1e12
    out = {
        "metric": "cell-integrations/s @1M cells FP64; learned-h0 vs plain speedup at rtol 1e-3",
        "value": value,
        "unit": "cell-integrations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, radm2_amd.synth; no reference data shipped)",
        "config": {"workload": "1,048,576 cells (128x128x64), dt 600 s, rtol 1e-3, atol 1, "
                               "seed 0, SZA U(0,100), 5% plume", "cells": n, "dt": args.dt,
                   "rtol": args.rtol, "atol": args.atol, "path": head},
        "paths": {k: {"cells_per_s": n / (v["wall_s"] / args.steps),
                      "kernel_ms_median": float(np.median(v["kernel_ms"])),
                      "attempts_mean": v["nstep_mean"], "attempts_max": v["nstep_max"],
                      "cells_ok": v["ok"]} for k, v in results.items()},
        "roofline": {"bound": "mfma", "compute": "FP64 VALU (no MFMA on this path); "
                     "MI355X FP64 vector peak = 78.6 TFLOP/s",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_bytes(r),
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE per attempt, calibrated "
                                       "with a 1 GiB copy (profiles/radm2_persistent_1m.md), "
                                       "scaled to this launch's attempts",
                     "flop_per_launch": flop, "kernel_ms": kms,
                     "hbm_gbs": (traffic_bytes(r) or 0) / (kms * 1e-3) / 1e9},
    }
    if "grid" in results and "persistent" in results:
        out["persistent_vs_grid"] = (results["grid"]["wall_s"] / results["persistent"]["wall_s"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cells, args)
    print(json.dumps(out), flush=True)


def traffic_bytes(r):
    """HBM bytes of one launch: measured bytes per attempt times attempts."""
    p = os.path.join(ROOT, "profiles", "radm2_persistent_1m_traffic.json")
    if not os.path.exists(p):
        return None
    t = json.load(open(p))
    per_attempt = (t["read_bytes"] + t["write_bytes"]) / t["attempts"]
    return per_attempt * (r["nacc"] + r["nrej"])


def cpu_baseline(cells, args):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ct
    ns = min(args.cpu_sample, cells["conc"].shape[1])
    sub = {k: np.ascontiguousarray(v[..., :ns]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    t = time.perf_counter()
    _, cnt, code = oracle_ct.integrate(sub, args.dt, args.rtol, args.atol, nthreads=cores)
    el = time.perf_counter() - t
    return {"value": ns / el, "unit": "cell-integrations/s", "cores": cores, "kind": "port",
            "sample": "first %d cells of the same 1M-cell workload, C oracle (gcc -O2, OpenMP "
                      "dynamic,64), %.1f s" % (ns, el)}


if __name__ == "__main__":
    main()
