"""Autotuner for the solver's launch configuration.

Axes (all compiled in; see include/radm2.h):
  schedule   one-shot grid (lockstep blocks of 64..1024 cells) or
             persistent work stealing (blocks of 64 or 256)
  placement  working arrays compiler-managed (VGPR/AGPR, spills to
             scratch), an explicit wave-interleaved HBM workspace, or the
             workspace with the most accessed entries in LDS (160 or 312
             doubles per lane) and optionally registers (the next 128);
             the LDS placements run 64-lane blocks only
  waves      register budget: at least 1, 2, 4 or 8 resident waves/SIMD
  refill     persistent: idle lanes per wave that trigger a refill

A cell's whole working set (1,110 doubles, 8.9 KB; 568 KB per wave) does not
fit on chip, so only the most accessed part is placed there.

The search is coordinate descent on a sample of the caller's own cells
(every candidate integrates the same cells, so the results also double as a
bit-identity check across variants).  Results are cached in JSON keyed by
device and workload signature.
"""
import json
import os
import time

from . import (PLACE_LDS160, PLACE_LDS160_REG128, PLACE_LDS312, PLACE_PRIVATE, PLACE_WORKSPACE,
               SCHED_GRID, SCHED_PERSISTENT, CellBatch, Radm2Error)

KERNELS = [(PLACE_PRIVATE, 1), (PLACE_PRIVATE, 2), (PLACE_PRIVATE, 4), (PLACE_PRIVATE, 8),
           (PLACE_WORKSPACE, 1), (PLACE_WORKSPACE, 2), (PLACE_WORKSPACE, 4),
           (PLACE_LDS160, 1), (PLACE_LDS312, 1), (PLACE_LDS160_REG128, 1)]
DEFAULT = dict(schedule=SCHED_PERSISTENT, placement=PLACE_LDS160_REG128, waves=1, block=64, refill=1)
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
    """Pick the fastest configuration for `batch`'s cells; returns
    (config dict, {config key: ms}).  The batch's concentrations are restored
    before returning."""
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
            return cfg, store[sig]["timings"]
    conc0 = batch.conc.clone()
    timings = {}

    def run(cfg):
        key = json.dumps(cfg, sort_keys=True)
        if key not in timings:
            try:
                timings[key] = _time(solver, batch, conc0, cfg, dt, rtol, atol, n)
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
    for sched, block in ((SCHED_PERSISTENT, 64), (SCHED_PERSISTENT, 256), (SCHED_GRID, 64),
                         (SCHED_GRID, 256), (SCHED_GRID, 512)):
        cand = dict(best, schedule=sched, block=block)
        if run(cand) < run(best):
            best = cand
    # 3) refill threshold
    if best["schedule"] == SCHED_PERSISTENT:
        for refill in (1, 2, 4, 8, 16):
            cand = dict(best, refill=refill)
            if run(cand) < run(best):
                best = cand
    batch.conc.copy_(conc0)
    apply(solver, best)
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        store[sig] = {"best": best, "timings": timings, "when": time.time()}
        with open(cache, "w") as f:
            json.dump(store, f, indent=1)
    return best, timings
