"""Learned initial step size for the FP64 Rosenbrock path.

A 96 -> 128 -> 128 -> 16 SiLU MLP (83 live features, 2 live outputs) runs on
the project's BF16-MFMA kernels.  Output 0 is log10 of a proposed initial
step h0, output 1 a confidence logit.  The FP64 solver always does the
integration: a proposal only replaces the options' Hstart for the cells
whose confidence exceeds the threshold (the solver treats it as that cell's
user Hstart: min(h0, dt, Hmax)).

Training data come from the solver itself.  Every training cell is
integrated once with the plain call (INTEGRATE's Hstart = 0.01 dt) and once
per candidate h0 on a half-decade grid (1e-11 s .. 1 s); its attempt counts
over the grid, smoothed over +-1/2 decade (a window of three candidates, so
that a proposal off by the network's typical error still lands on a good
value), give
  target 0  log10 of the candidate with the fewest smoothed attempts;
  target 1  1 if that smoothed minimum is below the plain call's count.
"""
import numpy as np

from . import Opts, mlp

DIMS = [96, 128, 128, 16]
GRID = tuple(np.arange(-11.0, 0.01, 0.5))


def _cells_to_batch(cells, device="cuda", hfirst=False):
    from . import CellBatch, synth
    d = synth.to_device(cells, device)
    return CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"], rc_n2o5=d["rc_n2o5"],
                     hfirst=hfirst)


def fit_normalizer(batch):
    """Per-feature mean and spread of the batch's raw features."""
    import torch
    mu = torch.zeros(mlp.NFEAT, dtype=torch.float32, device=batch.conc.device)
    sd = torch.ones_like(mu)
    _, Xf = mlp.features(batch, mu, sd, want_f32=True)
    f = Xf[:, :mlp.NFEAT]
    mu = f.mean(0)
    sd = f.std(0).clamp_min(1e-3)
    return mu.contiguous(), sd.contiguous()


def sweep(solver, batch, dt, rtol=1e-3, atol=1.0, grid=GRID):
    """Attempt counts of every cell: the plain call and each constant h0 of
    `grid` (log10 s).  Returns (plain [n] int32, counts [len(grid) x n]
    float32 with failed cells at +inf); the batch's state is restored."""
    import torch
    conc0 = batch.conc.clone()
    h0_saved = batch.h0
    n = batch.ncells
    batch.h0 = None
    solver.integrate(batch, dt, rtol, atol)
    plain = torch.where(batch.status == 1, batch.nstep, torch.iinfo(torch.int32).max)
    counts = torch.empty((len(grid), n), dtype=torch.float32, device=conc0.device)
    h0 = torch.empty(n, dtype=torch.float64, device=conc0.device)
    batch.h0 = h0
    for i, lh in enumerate(grid):
        batch.conc.copy_(conc0)
        h0.fill_(10.0 ** lh)
        solver.integrate(batch, dt, rtol, atol)
        counts[i] = torch.where(batch.status == 1, batch.nstep.float(),
                                torch.full_like(counts[i], float("inf")))
    batch.h0 = h0_saved
    batch.conc.copy_(conc0)
    solver.sync()
    return plain, counts


def targets(plain, counts, grid=GRID):
    """Regression and confidence targets from a sweep (see module doc)."""
    import torch
    c = counts
    pad = torch.cat([c[:1], c, c[-1:]])
    smooth = (pad[:-2] + pad[1:-1] + pad[2:]) / 3.0
    best = smooth.argmin(0)
    g = torch.tensor(grid, dtype=torch.float32, device=c.device)
    sbest = smooth.gather(0, best[None])[0]
    valid = torch.isfinite(sbest) & (plain < torch.iinfo(torch.int32).max)
    label = (sbest < plain.float()).float()
    tgt = torch.stack([g[best], label], 1)
    mask = torch.stack([valid, valid], 1).float()
    return tgt.contiguous(), mask.contiguous()


def generate(solver, batch, dt, rtol=1e-3, atol=1.0, log=None):
    """Solver-generated targets for every cell of `batch` (state restored)."""
    plain, counts = sweep(solver, batch, dt, rtol, atol)
    tgt, mask = targets(plain, counts)
    info = dict(plain=plain.double().mean().item(),
                per_h0={"%.1f" % g: counts[i].double().mean().item() for i, g in enumerate(GRID)},
                cell_best=counts.min(0).values.double().mean().item())
    if log:
        log("targets: %d cells, valid %.4f, helped %.4f, attempts plain %.2f, best constant h0 "
            "%.2f, per-cell best %.2f" % (batch.ncells, mask[:, 0].double().mean().item(),
                                          tgt[:, 1].double().mean().item(), info["plain"],
                                          min(info["per_h0"].values()), info["cell_best"]))
    return tgt, mask, info


def train(net, X, tgt, mask, epochs=3, batch=8192, lr=2e-3, log=None, seed=0):
    """Mini-batch Adam on the project's kernels; returns per-epoch mean loss."""
    import torch
    n = (X.shape[0] // 16) * 16
    g = torch.Generator(device=X.device)
    g.manual_seed(seed)
    losses = []
    loss = torch.zeros(1, dtype=torch.float32, device=X.device)
    for ep in range(epochs):
        perm = torch.randperm(n, device=X.device, generator=g)
        loss.zero_()
        seen = 0
        for s in range(0, n - batch + 1, batch):
            idx = perm[s:s + batch]
            xb, tb, mb = X[idx].contiguous(), tgt[idx].contiguous(), mask[idx].contiguous()
            out = net.forward(xb, keep=True)
            dout = torch.empty_like(out)
            mlp._ck(mlp.lib().radm2_mlp_h0_loss(mlp._st(), batch, out.shape[1], mlp._p(out), mlp._p(tb),
                                                mlp._p(mb), mlp._p(dout), mlp._p(loss)))
            net.zero_grad()
            net.backward(dout)
            net.adam(lr=lr)
            seen += batch
        losses.append(loss.item() / max(seen, 1))
        if log:
            log("epoch %d: loss %.4f" % (ep, losses[-1]))
    return losses


HLO = 1e-12   # lowest proposal passed on (the solver's own floor is 10 roundoff)


def propose(net, batch, mu, sd, dt, threshold=0.5, ncells=None, fallback=0.0, hlo=HLO):
    """Features -> MLP -> h0 per cell (device tensor) and the taken mask;
    proposals are clamped to [hlo, dt]."""
    import torch
    n = batch.ncells if ncells is None else ncells
    X, _ = mlp.features(batch, mu, sd, ncells=n)
    pad = (-n) % 16
    if pad:
        X = torch.cat([X, torch.zeros((pad, X.shape[1]), dtype=X.dtype, device=X.device)])
    out = net.forward(X)
    h0 = torch.empty(n, dtype=torch.float64, device=X.device)
    used = torch.empty(n, dtype=torch.int32, device=X.device)
    mlp._ck(mlp.lib().radm2_mlp_h0_select(mlp._st(), n, out.shape[1], mlp._p(out), float(hlo),
                                          float(dt), float(threshold), float(fallback),
                                          mlp._p(h0), mlp._p(used)))
    return h0, used


def build(solver, cells, dt=600.0, rtol=1e-3, atol=1.0, epochs=3, seed=1, log=None):
    """Train a fresh network on solver-generated data for `cells`."""
    import torch
    b = _cells_to_batch(cells)
    mu, sd = fit_normalizer(b)
    tgt, mask, info = generate(solver, b, dt, rtol, atol, log=log)
    X, _ = mlp.features(b, mu, sd)
    net = mlp.MLP(DIMS, seed=seed)
    losses = train(net, X, tgt, mask, epochs=epochs, log=log, seed=seed)
    del X, b
    torch.cuda.empty_cache()
    return net, mu, sd, losses, info


def save(path, net, mu, sd):
    d = net.state()
    d["mu"] = mu.cpu().numpy()
    d["sd"] = sd.cpu().numpy()
    np.savez(path, **d)


def load(path, device="cuda"):
    import torch
    d = np.load(path)
    net = mlp.MLP([int(x) for x in d["dims"]])
    net.load_state(d)
    mu = torch.from_numpy(d["mu"]).to(device)
    sd = torch.from_numpy(d["sd"]).to(device)
    return net, mu, sd


def evaluate(solver, net, mu, sd, cells, dt=600.0, rtol=1e-3, atol=1.0, threshold=0.5,
             reps=1, hlo=HLO):
    """Plain vs learned-h0 on the same cells: attempts, device time (solver
    launch, plus feature/MLP/selection time for the learned path) and the
    final states of both runs (host arrays)."""
    import time

    import torch
    b = _cells_to_batch(cells, hfirst=False)
    conc0 = b.conc.clone()
    res = {}
    for name in ("plain", "learned"):
        best = None
        for r in range(reps + 1):                     # first rep warms up
            b.conc.copy_(conc0)
            b.h0 = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            used = None
            if name == "learned":
                h0, used = propose(net, b, mu, sd, dt, threshold, hlo=hlo)
                b.h0 = h0
            solver.integrate(b, dt, rtol, atol)
            solver.sync()
            t = time.perf_counter() - t0
            if r > 0 or reps == 0:
                best = t if best is None else min(best, t)
        res[name] = dict(seconds=best, attempts=b.nstep.double().mean().item(),
                         rejected=b.nrej.double().mean().item(),
                         ok=int((b.status == 1).sum()), conc=b.conc.cpu().numpy(),
                         nstep=b.nstep.cpu().numpy(),
                         used=None if used is None else used.double().mean().item())
    b.conc.copy_(conc0)
    return res


def main():
    import argparse
    import json
    import os
    import sys

    import torch

    from . import Solver, synth
    ap = argparse.ArgumentParser(description="train the learned-h0 network on solver data")
    ap.add_argument("--train-cells", type=int, default=1048576)
    ap.add_argument("--eval-cells", type=int, default=1048576)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), "mech", "h0_mlp.npz"))
    a = ap.parse_args()
    log = lambda m: print("[learned_h0] " + m, file=sys.stderr, flush=True)
    s = Solver(0)
    train_cells = synth.make_cells(a.train_cells, 1, 100.0, 0.05)
    net, mu, sd, losses, info = build(s, train_cells, epochs=a.epochs, log=log)
    del train_cells
    save(a.out, net, mu, sd)
    ev = synth.make_cells(a.eval_cells, 2, 100.0, 0.05)
    r = evaluate(s, net, mu, sd, ev)
    out = {k: {kk: vv for kk, vv in v.items() if kk not in ("conc", "nstep")} for k, v in r.items()}
    out["losses"] = losses
    out["train_sweep"] = info
    print(json.dumps(out))


if __name__ == "__main__":
    main()
