#!/usr/bin/env python3
"""Benchmark: FP64 RADM2 cell-integrations/s on one MI355X.

Workload: 1,048,576 synthetic cells (128x128x64), dt = 600 s, rtol 1e-3,
atol 1 molec/cm3, seed 0, zenith U(0,100) deg (10% night cells), 5% plume
cells (NOx 20-500 ppb, VOC x10).  A "step" is one integration of every
cell over dt (one cell-integration per cell), inputs resident in HBM.

Paths timed, each over the same K steps after W warm-up steps:
  grid        one-shot kernel, one lane per cell for the whole integration;
  persistent  work-stealing kernel (the library default) -- `value`;
  learned     the shipped learned-h0 network (features + BF16-MFMA MLP +
              selection kernels, then the persistent solver), inside the
              timed region; reported as a speed-up over `persistent`, with
              proposals floored at 1e-12 s and, separately, at 1e-6 s;
  end_to_end  one step from pinned host buffers: H2D, solve, D2H.
Also reported: `family` (KPP's other methods and the forced Ros3 on the same
cells) and `batched` (the 64 uneven domains as one HIP-graph launch).
Each path reports histograms of the attempts and rejected steps per cell.
`accuracy` compares the plain and learned end states of the first 16,384
cells with the GPU FP64 path at rtol 1e-10, atol 1e-3 molec/cm3.
`roofline.attainable` holds measured FP64-FMA, device-copy and read peaks.
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
# histogram bin edges: attempts (accepted + rejected steps) and rejected
# steps per cell; bin i counts edges[i] <= x < edges[i+1], the last bin is
# open-ended
ATTEMPT_EDGES = (0, 25, 50, 75, 100, 125, 150, 175, 200, 225, 250, 300, 400, 600, 1000)
REJECT_EDGES = (0, 1, 2, 4, 8, 12, 16, 24, 32, 64)


def histogram(x, edges, torch):
    """Per-bin counts of the integer tensor `x` over `edges`."""
    bounds = torch.tensor(edges[1:], dtype=torch.int64, device=x.device)
    idx = torch.bucketize(x.to(torch.int64), bounds, right=True)
    return {"edges": list(edges), "counts": torch.bincount(idx, minlength=len(edges)).cpu().tolist()}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1048576)
    ap.add_argument("--dt", type=float, default=600.0)
    ap.add_argument("--rtol", type=float, default=1e-3)
    ap.add_argument("--atol", type=float, default=1.0)
    ap.add_argument("--paths", default="grid,persistent,learned")
    ap.add_argument("--acc-cells", type=int, default=16384,
                    help="cells checked against the rtol-1e-10 GPU solve (0 = skip)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the end-to-end, family, batched and accuracy legs (profiling runs)")
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (calibrates HBM byte counters)")
    return ap.parse_args()


def run_path(solver, batch, conc0, sched, args, torch, step=None):
    """W warm-up steps, then K timed steps; `step` replaces the plain
    integrate call (the learned path)."""
    import radm2_amd
    solver.set_schedule(sched, args.block, args.refill)
    call = step or (lambda: solver.integrate(batch, args.dt, args.rtol, args.atol))
    for _ in range(args.warmup):
        batch.conc.copy_(conc0)
        call()
        solver.sync()
    kernel_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.conc.copy_(conc0)
        call()
        solver.sync()
        kernel_ms.append(solver.elapsed_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nacc = batch.nacc.to(torch.int64)
    nrej = batch.nrej.to(torch.int64)
    st = batch.status
    return dict(wall_s=wall, kernel_ms=kernel_ms, nacc=int(nacc.sum()), nrej=int(nrej.sum()),
                nstep_max=int(batch.nstep.max()), nstep_mean=float(batch.nstep.double().mean()),
                nrej_mean=float(batch.nrej.double().mean()),
                ok=int((st == radm2_amd.CELL_OK).sum()),
                attempts_hist=histogram(batch.nstep, ATTEMPT_EDGES, torch),
                rejected_hist=histogram(batch.nrej, REJECT_EDGES, torch))


def learned_step(solver, batch, args, hlo):
    """One learned-h0 step: features -> MLP -> h0 selection -> solver."""
    from radm2_amd import learned_h0
    net, mu, sd = _net()

    def step():
        h0, used = learned_h0.propose(net, batch, mu, sd, args.dt, hlo=hlo)
        batch.h0 = h0
        solver.integrate(batch, args.dt, args.rtol, args.atol)
        step.used = used
    return step


_NET = []


def _net():
    if not _NET:
        from radm2_amd import learned_h0
        _NET.append(learned_h0.load(os.path.join(PKG, "mech", "h0_mlp.npz")))
    return _NET[0]


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU; --gpus %d is not supported "
              "(cells are independent: the path would shard by contiguous slices of "
              "cells, one per GPU)" % args.gpus, file=sys.stderr)
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
    ends = {}
    na = min(args.acc_cells, n)
    names = args.paths.split(",")
    if "learned" in names:
        names = [x for x in names if x != "learned"] + ["learned", "learned_hlo1e-6"]
    for name in names:
        step = None
        if name.startswith("learned"):
            step = learned_step(solver, batch, args, 1e-6 if name.endswith("1e-6") else 1e-12)
        sched = radm2_amd.SCHED_GRID if name == "grid" else radm2_amd.SCHED_PERSISTENT
        results[name] = run_path(solver, batch, conc0, sched, args, torch, step)
        batch.h0 = None
        r = results[name]
        if step is not None:
            r["used"] = float(step.used.double().mean())
        ends[name] = batch.conc[:, :na].clone()
        print("[bench] %s: wall %.3f s for %d steps, kernel ms %s, ok %d/%d, attempts mean %.1f max %d"
              % (name, r["wall_s"], args.steps, ["%.1f" % x for x in r["kernel_ms"]], r["ok"], n,
                 r["nstep_mean"], r["nstep_max"]), file=sys.stderr, flush=True)

    head = "persistent" if "persistent" in results else [k for k in results
                                                          if not k.startswith("learned")][0]
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
                      "rejected_mean": v["nrej_mean"], "cells_ok": v["ok"],
                      "attempts_hist": v["attempts_hist"], "rejected_hist": v["rejected_hist"]}
                  for k, v in results.items()},
        "roofline": {"bound": "mfma", "compute": "FP64 VALU (no MFMA on this path); "
                     "MI355X FP64 vector peak = 78.6 TFLOP/s",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_bytes(r),
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE per attempt, calibrated "
                                       "with a 1 GiB copy (profiles/radm2_persistent_lds160_reg128_1m.md), "
                                       "scaled to this launch's attempts",
                     "flop_per_launch": flop, "kernel_ms": kms,
                     "hbm_gbs": (traffic_bytes(r) or 0) / (kms * 1e-3) / 1e9,
                     "hbm_frac": (traffic_bytes(r) or 0) / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                     "hbm_note": "the measured L2-miss traffic to the memory fabric (the "
                                 "spilled working set) against MI355X's 8 TB/s HBM peak; the "
                                 "256 MB Infinity Cache serves part of it, so the rate exceeds "
                                 "what streaming attains from HBM (roofline.attainable)"},
    }
    if "grid" in results and "persistent" in results:
        out["persistent_vs_grid"] = (results["grid"]["wall_s"] / results["persistent"]["wall_s"])
    for name in ("learned", "learned_hlo1e-6"):
        if name in results and head in results:
            L, P = results[name], results[head]
            out[name] = {"speedup_vs_plain": P["wall_s"] / L["wall_s"],
                         "cells_per_s": n / (L["wall_s"] / args.steps),
                         "attempts_mean": L["nstep_mean"], "plain_attempts_mean": P["nstep_mean"],
                         "rejected_mean": L["nrej_mean"], "plain_rejected_mean": P["nrej_mean"],
                         "proposals_used": L.get("used"), "cells_ok": L["ok"],
                         "timed": "features + MLP + h0 selection + solver, per step",
                         "h0_floor_s": 1e-6 if name.endswith("1e-6") else 1e-12}
    if not args.no_extra:
        # measured attainable peaks beside the published ones
        tf = solver.measure_fp64_peak()
        copy_gbs, read_gbs = attainable_memory(torch)
        rl = out["roofline"]
        rl["attainable"] = {
            "fp64_tflops": tf, "copy_gbs": copy_gbs, "read_gbs": read_gbs,
            "source": "radm2_measure_fp64_peak (8 independent FP64 FMA chains per lane in "
                      "registers); torch's copy and sum over 4 GiB (best of 3)"}
        rl["frac_of_attainable"] = achieved / tf
        print("[bench] attainable: FP64 FMA %.1f TFLOP/s, copy %.0f GB/s, read %.0f GB/s"
              % (tf, copy_gbs, read_gbs), file=sys.stderr, flush=True)
        out["end_to_end"] = end_to_end(solver, cells, args, torch)
        out["family"] = family(solver, batch, conc0, args, torch)
        out["batched"] = batched(solver, args, torch)
    if na > 0 and not args.no_extra:
        out["accuracy"] = accuracy(solver, dev, conc0, ends, na, args, torch)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cells, args)
    print(json.dumps(out), flush=True)


def family(solver, batch, conc0, args, torch):
    """KPP's other Rosenbrock methods and the non-autonomous forcing on the
    family kernel: one warm-up and one timed step each on the same cells
    (the forcing: every cell's J falls linearly to half over the interval)."""
    import radm2_amd
    n = batch.ncells
    out = {}
    cases = [("ros2", 1, False), ("ros4", 3, False), ("rodas3", 4, False), ("rodas4", 5, False),
             ("ros3_forced", 2, True)]
    solver.set_schedule(radm2_amd.SCHED_PERSISTENT, args.block, args.refill)
    for name, method, forced in cases:
        batch.jrate_end = (0.5 * batch.jrate).contiguous() if forced else None
        opts = radm2_amd.Opts(method=method)
        for rep in range(2):
            batch.conc.copy_(conc0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            solver.integrate(batch, args.dt, args.rtol, args.atol, opts=opts)
            solver.sync()
            wall = time.perf_counter() - t0
        out[name] = {"cells_per_s": n / wall, "kernel_ms": solver.elapsed_ms(),
                     "attempts_mean": float(batch.nstep.double().mean()),
                     "cells_ok": int((batch.status == radm2_amd.CELL_OK).sum())}
        print("[bench] family %s: %.1f ms, attempts mean %.1f, ok %d" % (
            name, out[name]["kernel_ms"], out[name]["attempts_mean"], out[name]["cells_ok"]),
            file=sys.stderr, flush=True)
    batch.jrate_end = None
    return out


def attainable_memory(torch, gib=4):
    """Streaming rates from HBM: a device copy (bytes read + written) and a
    read-only sum, best of three after a warm-up, GB/s."""
    n = (gib << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def best(fn):
        t = []
        for _ in range(4):
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            t.append(e0.elapsed_time(e1))
        return min(t[1:])
    copy_gbs = 2 * n * 8 / (best(lambda: b.copy_(a)) * 1e-3) / 1e9
    read_gbs = n * 8 / (best(lambda: a.sum()) * 1e-3) / 1e9
    del a, b
    return copy_gbs, read_gbs


def batched(solver, args, torch):
    """The 64-domain workload (synth.make_domains(): log-uniform 1k-200k
    cells per domain, per-domain dt in {60, 300, 600, 3600} s and rtol in
    {1e-2, 1e-3, 1e-4}) captured as one HIP-graph plan: one warm-up launch,
    then one timed (the graph's own events: counter reset + solver)."""
    import radm2_amd
    from radm2_amd import synth
    doms = synth.make_domains()
    batches, conc0 = [], []
    for c, _, _, _ in doms:
        d = synth.to_device(c)
        conc0.append(d["conc"].clone())
        batches.append(radm2_amd.CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"],
                                           rc_n2o5=d["rc_n2o5"]))
    total = sum(b.ncells for b in batches)
    solver.set_schedule(radm2_amd.SCHED_PERSISTENT, args.block, args.refill)
    plan = solver.plan([b.domain(dt, rtol, atol) for b, (_, dt, rtol, atol) in zip(batches, doms)])
    try:
        for _ in range(2):
            for b, c0 in zip(batches, conc0):
                b.conc.copy_(c0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.launch()
            solver.sync()
            wall = time.perf_counter() - t0
    finally:
        plan.close()
    ms = solver.elapsed_ms()
    attempts = sum(int(b.nstep.to(torch.int64).sum()) for b in batches)
    out = {"domains": len(doms), "cells": total, "graph_ms": ms, "wall_ms": 1000.0 * wall,
           "cells_per_s": total / (ms * 1e-3), "attempts_mean": attempts / total,
           "cells_ok": sum(int((b.status == radm2_amd.CELL_OK).sum()) for b in batches)}
    print("[bench] batched: %d domains, %d cells, graph %.1f ms, attempts mean %.1f, ok %d" % (
        len(doms), total, ms, out["attempts_mean"], out["cells_ok"]), file=sys.stderr, flush=True)
    return out


def end_to_end(solver, cells, args, torch):
    """One step from pinned host buffers: H2D of the inputs, the solve, D2H
    of concentrations and per-cell counters (a warm-up step first)."""
    import radm2_amd
    keys = ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")
    host = {k: torch.from_numpy(np.ascontiguousarray(cells[k])).pin_memory() for k in keys}
    dev = {k: torch.empty(host[k].shape, dtype=torch.float64, device="cuda") for k in keys}
    back = torch.empty(host["conc"].shape, dtype=torch.float64).pin_memory()
    cnt = torch.empty((4, host["temp"].shape[0]), dtype=torch.int32).pin_memory()
    b = radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"], dev["jrate"],
                            rc_n2o5=dev["rc_n2o5"])
    solver.set_schedule(radm2_amd.SCHED_PERSISTENT, args.block, args.refill)
    times = []
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in keys:
            dev[k].copy_(host[k], non_blocking=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        solver.integrate(b, args.dt, args.rtol, args.atol)
        solver.sync()
        t2 = time.perf_counter()
        back.copy_(b.conc, non_blocking=True)
        for i, x in enumerate((b.nstep, b.nacc, b.nrej, b.status)):
            cnt[i].copy_(x, non_blocking=True)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        times.append((t1 - t0, t2 - t1, t3 - t2, t3 - t0))
    h2d, solve, d2h, tot = times[-1]
    n = host["temp"].shape[0]
    return {"cells_per_s": n / tot, "h2d_ms": 1e3 * h2d, "solve_ms": 1e3 * solve,
            "d2h_ms": 1e3 * d2h, "note": "pinned host buffers; second of two runs"}


def accuracy(solver, dev, conc0, ends, na, args, torch):
    """Max and median relative error (species above 1e3 molec/cm3) of each
    path's end state against the GPU FP64 path at rtol 1e-10, atol 1e-3 on
    the first `na` cells."""
    import radm2_amd
    keys = ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")
    sub = {k: dev[k][..., :na].clone() for k in keys}
    for k in ("temp", "cair", "h2o", "rc_n2o5"):
        sub[k] = sub[k].contiguous()
    sub["conc"] = conc0[:, :na].contiguous().clone()
    b = radm2_amd.CellBatch(sub["conc"], sub["temp"], sub["cair"], sub["h2o"],
                            sub["jrate"].contiguous(), rc_n2o5=sub["rc_n2o5"])
    solver.set_schedule(radm2_amd.SCHED_PERSISTENT, 0, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.integrate(b, args.dt, 1e-10, 1e-3)
    solver.sync()
    el = time.perf_counter() - t0
    ref = b.conc
    big = ref.abs() > 1e3
    out = {"reference": "GPU FP64 path, rtol 1e-10, atol 1e-3, first %d cells (%.1f s, "
                        "%d/%d cells ok)" % (na, el, int((b.status == 1).sum()), na)}
    for name, c in ends.items():
        e = torch.where(big, (c - ref).abs() / ref.abs().clamp_min(1e-300), torch.zeros_like(ref))
        per_cell = e.max(0).values
        out[name] = {"max_rel_err": float(e.max()), "median_cell_max_rel_err":
                     float(per_cell.median())}
    return out


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
