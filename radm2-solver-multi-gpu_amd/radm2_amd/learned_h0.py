"""Learned initial step size for the FP64 Rosenbrock path.

A 96 -> 128 -> 128 -> 16 SiLU MLP (83 live features, 2 live outputs) runs on
the project's BF16-MFMA kernels.  Output 0 is log10 of a proposed initial
step h0, output 1 a confidence logit.  The FP64 solver always does the
integration: a proposal only replaces the options' Hstart for cells whose
confidence exceeds the threshold, clamped to [1e-6 s, dt].

Training data come from the solver itself, two runs over the same cells:
  run A  plain call (INTEGRATE's Hstart = 0.01 dt): per cell the attempt
         count and `hfirst`, the controller's step proposal after the first
         accepted step, FacSafe/err^(1/3) times that step, uncapped -- an
         estimate of the largest step the local error allows at t = 0;
  run B  the same cells with h0 = clamp(hfirst, 1e-6, dt).
Targets: log10 clamp(hfirst, 1e-6, dt) (cells that succeeded), and the label
"run B took fewer attempts than run A" for the confidence.
"""
import numpy as np

from . import Opts, mlp

DIMS = [96, 128, 128, 16]


def _cells_to_batch(cells, device="cuda", hfirst=True):
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


def generate(solver, batch, dt, rtol=1e-3, atol=1.0, log=None):
    """Solver-generated targets for every cell of `batch` (state restored)."""
    import torch
    conc0 = batch.conc.clone()
    n = batch.ncells
    solver.integrate(batch, dt, rtol, atol)
    solver.sync()
    ns_a = batch.nstep.clone()
    st_a = batch.status.clone()
    hf = batch.hfirst.clone()
    h_target = hf.clamp(1e-6, dt)
    valid = (st_a == 1) & (hf > 0) & torch.isfinite(hf)
    batch.conc.copy_(conc0)
    h0_saved = batch.h0
    batch.h0 = h_target.contiguous()
    solver.integrate(batch, dt, rtol, atol)
    solver.sync()
    ns_b = batch.nstep.clone()
    batch.h0 = h0_saved
    batch.conc.copy_(conc0)
    tgt = torch.stack([torch.log10(h_target), (ns_b < ns_a).to(torch.float64)], 1).float()
    mask = torch.stack([valid, valid], 1).float()
    if log:
        log("targets: %d cells, valid %.4f, helped %.4f, attempts plain %.2f -> oracle h0 %.2f"
            % (n, valid.double().mean().item(), tgt[:, 1].double().mean().item(),
               ns_a.double().mean().item(), ns_b.double().mean().item()))
    return tgt.contiguous(), mask.contiguous(), dict(nstep_plain=ns_a, nstep_oracle_h0=ns_b)


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
            mlp._ck(mlp.lib().radm2_mlp_h0_loss(None, batch, out.shape[1], mlp._p(out), mlp._p(tb),
                                                mlp._p(mb), mlp._p(dout), mlp._p(loss)))
            net.zero_grad()
            net.backward(dout)
            net.adam(lr=lr)
            seen += batch
        losses.append(loss.item() / max(seen, 1))
        if log:
            log("epoch %d: loss %.4f" % (ep, losses[-1]))
    return losses


def propose(net, batch, mu, sd, dt, threshold=0.5, ncells=None, fallback=0.0):
    """Features -> MLP -> h0 per cell (device tensor) and the taken mask."""
    import torch
    n = batch.ncells if ncells is None else ncells
    X, _ = mlp.features(batch, mu, sd, ncells=n)
    pad = (-n) % 16
    if pad:
        X = torch.cat([X, torch.zeros((pad, X.shape[1]), dtype=X.dtype, device=X.device)])
    out = net.forward(X)
    h0 = torch.empty(n, dtype=torch.float64, device=X.device)
    used = torch.empty(n, dtype=torch.int32, device=X.device)
    mlp._ck(mlp.lib().radm2_mlp_h0_select(None, n, out.shape[1], mlp._p(out), float(dt),
                                          float(threshold), float(fallback), mlp._p(h0),
                                          mlp._p(used)))
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
             reps=1):
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
                h0, used = propose(net, b, mu, sd, dt, threshold)
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
    ap.add_argument("--train-cells", type=int, default=2097152)
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
    print(json.dumps(out))


if __name__ == "__main__":
    main()
