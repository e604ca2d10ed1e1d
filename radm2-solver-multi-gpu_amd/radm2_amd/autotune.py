"""Autotuner for the solver's launch configuration.

Axes (all compiled in; see include/radm2.h):
  schedule   one-shot grid (lockstep blocks of 64..1024 cells) or
             persistent work stealing (blocks of 64 or 256)
  placement  working arrays compiler-managed (VGPR/AGPR, spills to
             scratch), an explicit wave-interleaved HBM workspace, the
             workspace with the most accessed entries in LDS (160 or 312
             doubles per lane) and optionally registers (the next 128), or
             a cell spread over 2 or 4 lanes with a register tier split
             across them (persistent only); the LDS and split placements
             run 64-lane blocks only
  waves      register budget: at least 1, 2, 4 or 8 resident waves/SIMD
  refill     persistent: idle lanes per wave that trigger a refill

A cell's whole working set (1,028 doubles, 8.2 KB; 526 KB per wave of 64
cells) does not fit on chip, so only the most accessed part is placed there.
Unroll depth is not a run-time axis: the generated Jacobian, decomposition
and solves are fully unrolled for the mechanism's fixed sparsity, and their
one free depth, the load-batch size, is a compile-time choice.

The search is coordinate descent on a sample of the caller's own cells.
Every candidate integrates the same cells, and its end states and step
counts are compared with the first candidate's: the variants compute the
same operations in the same order, so any difference is reported (as
`mismatches`).  Results are cached in JSON keyed by device and workload
signature.
"""
import json
import os
import time

from . import (PLACE_LDS160, PLACE_LDS160_REG112, PLACE_LDS312, PLACE_PRIVATE, PLACE_SPLIT2,
               PLACE_SPLIT4, PLACE_WORKSPACE, SCHED_GRID, SCHED_PERSISTENT, Radm2Error)

KERNELS = [(PLACE_PRIVATE, 1), (PLACE_PRIVATE, 2), (PLACE_PRIVATE, 4), (PLACE_PRIVATE, 8),
           (PLACE_WORKSPACE, 1), (PLACE_WORKSPACE, 2), (PLACE_WORKSPACE, 4),
           (PLACE_LDS160, 1), (PLACE_LDS312, 1), (PLACE_LDS160_REG112, 1),
           (PLACE_SPLIT2, 1), (PLACE_SPLIT4, 1)]
SCHEDULES = ((SCHED_PERSISTENT, 64), (SCHED_PERSISTENT, 256), (SCHED_GRID, 64),
             (SCHED_GRID, 256), (SCHED_GRID, 512), (SCHED_GRID, 1024))
REFILLS = (1, 2, 3, 4, 8, 16)
DEFAULT = dict(schedule=SCHED_PERSISTENT, placement=PLACE_LDS160_REG112, waves=1, block=64, refill=3)
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "radm2_amd", "autotune.json")


def apply(solver, cfg):
    solver.set_schedule(cfg["schedule"], cfg["block"], cfg["refill"])
    solver.set_kernel(cfg["placement"], cfg["waves"])


def _time(solver, batch, conc0, cfg, dt, rtol, atol, ncells):
    import torch
    apply(solver, cfg)
    batch.conc[:, :ncells].copy_(conc0[:, :ncells])
    solver.integrate(batch, dt, rtol, atol, ncells=ncells)
    solver.sync()                       # warm-up (code objects, workspace)
    batch.conc[:, :ncells].copy_(conc0[:, :ncells])
    torch.cuda.synchronize()
    solver.integrate(batch, dt, rtol, atol, ncells=ncells)
    solver.sync()
    return solver.elapsed_ms()


def signature(device_name, cus, ncells, dt, rtol):
    return "%s/%dcu/n%d/dt%g/rtol%g" % (device_name, cus, ncells, dt, rtol)


def autotune(solver, batch, dt, rtol=1e-3, atol=1.0, sample=262144, cache=CACHE,
             log=None, use_cache=True):
    """Pick the fastest configuration for `batch`'s cells.  Returns a dict:
    best (the configuration, applied to the solver), timings ({config key:
    ms}, inf for a combination the library refuses), mismatches (keys of
    candidates whose end states or step counts differ from the first
    candidate's), sample (cells timed) and cached (taken from the cache).
    The batch's concentrations are restored before returning."""
    import torch
    props = torch.cuda.get_device_properties(0)
    n = min(sample, batch.ncells)
    sig = signature(props.name, props.multi_processor_count, n, dt, rtol)
    store = {}
    if use_cache and cache and os.path.exists(cache):
        try:
            store = json.load(open(cache))
        except (OSError, ValueError):
            store = {}
        if sig in store:
            cfg = store[sig]["best"]
            apply(solver, cfg)
            return dict(best=cfg, timings=store[sig]["timings"], mismatches=[], sample=n,
                        cached=True)
    conc0 = batch.conc.clone()
    timings = {}
    ref = []
    mismatches = []

    def run(cfg):
        key = json.dumps(cfg, sort_keys=True)
        if key not in timings:
            try:
                timings[key] = _time(solver, batch, conc0, cfg, dt, rtol, atol, n)
                out = (batch.conc[:, :n].clone(), batch.nstep[:n].clone())
                if not ref:
                    ref.append(out)
                elif not (torch.equal(out[0], ref[0][0]) and torch.equal(out[1], ref[0][1])):
                    mismatches.append(key)
            except Radm2Error:
                timings[key] = float("inf")
            if log:
                log("autotune %s: %.1f ms" % (key, timings[key]))
        return timings[key]

    best = dict(DEFAULT)
    # 1) kernel variant (placement x register budget)
    for pl, w in KERNELS:
        cand = dict(best, placement=pl, waves=w)
        if run(cand) < run(best):
            best = cand
    # 2) schedule and block
    for sched, block in SCHEDULES:
        cand = dict(best, schedule=sched, block=block)
        if run(cand) < run(best):
            best = cand
    # 3) refill threshold
    if best["schedule"] == SCHED_PERSISTENT:
        for refill in REFILLS:
            cand = dict(best, refill=refill)
            if run(cand) < run(best):
                best = cand
    batch.conc.copy_(conc0)
    apply(solver, best)
    if cache and not mismatches:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        store[sig] = {"best": best, "timings": timings, "when": time.time()}
        with open(cache, "w") as f:
            json.dump(store, f, indent=1)
    return dict(best=best, timings=timings, mismatches=mismatches, sample=n, cached=False)
