"""Learned end-state surrogate with the FP64 solver as judge.

A two-head MLP, 96 -> 256 -> 256 -> 256 -> 2x64 (SiLU hidden layers, 83 live
features, 59 live outputs per head), on the project's BF16-MFMA kernels,
predicts for each variable species the change of log10 max(c, 1 molec/cm3)
over the interval dt.  The heads share the trunk and are trained on
different bootstrap weightings of the same cells, so they agree where the
data pin the answer and drift apart where they do not.

Inference (`solve`): features -> network -> judge kernel.  A cell whose
heads disagree by more than `thr` in log10 space (over the species
predicted above 1e3 molec/cm3), plus a random `audit` fraction of all
cells, is re-solved by the FP64 Rosenbrock path; every other cell takes the
heads' mean.  The re-solve runs as one dense batch: flagged cells are
compacted, gathered, integrated and scattered back by the project's kernels.

Training data are the solver's own end states for the training cells.
"""
import numpy as np

from . import CellBatch, mlp

DIMS = [96, 256, 256, 256, 128]
WIDTH = 64
HEADS = 2
FLOOR = 1.0
CMIN = 1e3


def _zeros_like_batch(n, device):
    import torch
    return dict(conc=torch.empty((59, n), dtype=torch.float64, device=device),
                temp=torch.empty(n, dtype=torch.float64, device=device),
                cair=torch.empty(n, dtype=torch.float64, device=device),
                h2o=torch.empty(n, dtype=torch.float64, device=device),
                jrate=torch.empty((21, n), dtype=torch.float64, device=device),
                rc_n2o5=torch.empty(n, dtype=torch.float64, device=device))


def dlog_targets(c0, c1, tmu=None, tsd=None):
    """[n x 59] FP32 targets from start/end states [59 x n] (device)."""
    import torch
    n = c0.shape[1]
    t = torch.empty((n, 59), dtype=torch.float32, device=c0.device)
    mlp._ck(mlp.lib().radm2_mlp_dlog_targets(mlp._st(), n, n, mlp._p(c0), mlp._p(c1), FLOOR,
                                             mlp._p(tmu), mlp._p(tsd), mlp._p(t)))
    return t


def end_states(solver, batch, dt, rtol, atol):
    """The FP64 solver's end state for every cell (the batch is restored)."""
    c0 = batch.conc.clone()
    solver.integrate(batch, dt, rtol, atol)
    c1 = batch.conc.clone()
    batch.conc.copy_(c0)
    solver.sync()
    return c0, c1


def train(net, X, tgt, epochs=6, batch=8192, lr=1e-3, seed=0, keep=0.75, log=None):
    """Mini-batch Adam on the project's kernels; each head sees its own
    Bernoulli(keep) subset of the rows (fixed per cell).  Returns per-epoch
    mean loss."""
    import torch
    n = (X.shape[0] // 16) * 16
    g = torch.Generator(device=X.device)
    g.manual_seed(seed)
    hmask = (torch.rand((X.shape[0], HEADS), generator=g, device=X.device) < keep).float()
    loss = torch.zeros(1, dtype=torch.float32, device=X.device)
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(n, device=X.device, generator=g)
        loss.zero_()
        seen = 0
        for s in range(0, n - batch + 1, batch):
            idx = perm[s:s + batch]
            xb, tb, mb = X[idx].contiguous(), tgt[idx].contiguous(), hmask[idx].contiguous()
            out = net.forward(xb, keep=True)
            dout = torch.empty_like(out)
            mlp._ck(mlp.lib().radm2_mlp_mse_loss(mlp._st(), batch, out.shape[1], HEADS, WIDTH, 59,
                                                 mlp._p(out), mlp._p(tb), mlp._p(mb), mlp._p(dout),
                                                 mlp._p(loss)))
            net.zero_grad()
            net.backward(dout)
            net.adam(lr=lr)
            seen += batch
        losses.append(loss.item() / max(seen * 59 * keep * HEADS, 1))
        if log:
            log("epoch %d: loss %.5f" % (ep, losses[-1]))
    return losses


def build(solver, cells, dt=600.0, rtol=1e-3, atol=1.0, epochs=6, seed=1, batch=8192, lr=1e-3,
          log=None):
    """Train a fresh surrogate on the solver's end states for `cells`."""
    import torch
    from . import learned_h0
    b = learned_h0._cells_to_batch(cells)
    mu, sd = learned_h0.fit_normalizer(b)
    c0, c1 = end_states(solver, b, dt, rtol, atol)
    raw = dlog_targets(c0, c1)
    tmu = raw.mean(0).contiguous()
    tsd = raw.std(0).clamp_min(1e-3).contiguous()
    tgt = dlog_targets(c0, c1, tmu, tsd)
    del raw, c1
    X, _ = mlp.features(b, mu, sd)
    net = mlp.MLP(DIMS, seed=seed)
    losses = train(net, X, tgt, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    del X, b, tgt
    torch.cuda.empty_cache()
    return net, dict(mu=mu, sd=sd, tmu=tmu, tsd=tsd), losses


def save(path, net, norm):
    d = net.state()
    for k, v in norm.items():
        d[k] = v.cpu().numpy()
    np.savez(path, **d)


def load(path, device="cuda"):
    import torch
    d = np.load(path)
    net = mlp.MLP([int(x) for x in d["dims"]])
    net.load_state(d)
    norm = {k: torch.from_numpy(d[k]).to(device) for k in ("mu", "sd", "tmu", "tsd")}
    return net, norm


# per-cell status of a cell whose end state is the surrogate's prediction
# (not a KPP code: those are 1 and -1..-8); its nstep/nacc/nrej and hexit
# are 0
STATUS_PREDICTED = 2


def solve(solver, net, norm, batch, dt, rtol=1e-3, atol=1.0, thr=1e-3, audit=0.05, seed=0):
    """Surrogate pass over `batch` (in place): predicted cells take the
    heads' mean, flagged and audited cells the FP64 solver's result.  The
    per-cell outputs: the solver's for re-solved cells; STATUS_PREDICTED,
    zero counts and hexit 0 for predicted cells.  Returns the flags (device
    int32, 1 = re-solved) and the re-solved count."""
    import torch
    n = batch.ncells
    dev = batch.conc.device
    for name, v in (("nstep", 0), ("nacc", 0), ("nrej", 0), ("status", STATUS_PREDICTED),
                    ("hexit", 0.0)):
        dst = getattr(batch, name)
        if dst is not None:
            dst[:n].fill_(v)
    X, _ = mlp.features(batch, norm["mu"], norm["sd"])
    pad = (-n) % 16
    if pad:
        X = torch.cat([X, torch.zeros((pad, X.shape[1]), dtype=X.dtype, device=dev)])
    out = net.forward(X)
    flag = torch.empty(n, dtype=torch.int32, device=dev)
    L = mlp.lib()
    mlp._ck(L.radm2_mlp_surrogate_apply(mlp._st(), n, batch.ld, out.shape[1], WIDTH, mlp._p(out),
                                        mlp._p(norm["tmu"]), mlp._p(norm["tsd"]), FLOOR, CMIN,
                                        float(thr), float(audit), int(seed), mlp._p(batch.conc),
                                        mlp._p(flag), None))
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    mlp._ck(L.radm2_mlp_compact(mlp._st(), n, mlp._p(flag), mlp._p(idx), mlp._p(cnt)))
    m = int(cnt.item())
    if m == 0:
        return flag, 0
    sub = _zeros_like_batch(m, dev)
    st = mlp._st()
    for name, rows in (("conc", 59), ("temp", 1), ("cair", 1), ("h2o", 1), ("jrate", 21),
                       ("rc_n2o5", 1)):
        src = getattr(batch, name)
        if src is None:
            sub[name] = None
            continue
        mlp._ck(L.radm2_mlp_gather(st, 8, rows, batch.ld, mlp._p(src), m, mlp._p(idx), m,
                                   mlp._p(sub[name])))
    sb = CellBatch(sub["conc"], sub["temp"], sub["cair"], sub["h2o"], sub["jrate"],
                   rc_n2o5=sub["rc_n2o5"])
    solver.integrate(sb, dt, rtol, atol)
    st = mlp._st()
    mlp._ck(L.radm2_mlp_scatter(st, 8, 59, m, mlp._p(sb.conc), m, mlp._p(idx), batch.ld,
                                mlp._p(batch.conc)))
    for name, size in (("nstep", 4), ("nacc", 4), ("nrej", 4), ("status", 4), ("hexit", 8)):
        dst = getattr(batch, name)
        if dst is not None:
            mlp._ck(L.radm2_mlp_scatter(st, size, 1, m, mlp._p(getattr(sb, name)), m,
                                        mlp._p(idx), batch.ld, mlp._p(dst)))
    return flag, m


def evaluate(solver, net, norm, cells, dt=600.0, rtol=1e-3, atol=1.0, thr=1e-3, audit=0.05,
             reps=1):
    """Plain FP64 vs surrogate pass on the same cells: wall time (device
    work synchronised), fallback rate, and both end states (host arrays)."""
    import time

    import torch
    from . import learned_h0
    b = learned_h0._cells_to_batch(cells)
    conc0 = b.conc.clone()
    res = {}
    for name in ("plain", "surrogate"):
        best = None
        for r in range(reps + 1):
            b.conc.copy_(conc0)
            b.status.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            flag, m = None, None
            if name == "plain":
                solver.integrate(b, dt, rtol, atol)
            else:
                flag, m = solve(solver, net, norm, b, dt, rtol, atol, thr, audit)
            solver.sync()
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if r > 0 or reps == 0:
                best = t if best is None else min(best, t)
        res[name] = dict(seconds=best, conc=b.conc.cpu().numpy(),
                         flag=None if flag is None else flag.cpu().numpy(),
                         resolved=m, fallback_rate=None if m is None else m / b.ncells,
                         ok_resolved=int((b.status == 1).sum()))
    b.conc.copy_(conc0)
    return res


def threshold_report(solver, net, norm, cells, thrs=(1e-3, 3e-3, 1e-2, 3e-2, 0.1, 0.3),
                     dt=600.0, rtol=1e-3, atol=1.0, nref=16384, audit=0.05):
    """The gate's trade-off on `cells`: for each disagreement threshold, the
    fallback rate (gate plus audit) and the error of the cells the surrogate
    would keep, against the GPU FP64 path at rtol 1e-10, atol 1e-3 on the
    first `nref` cells; the plain rtol-`rtol` solve's error beside it."""
    import torch
    from . import learned_h0
    b = learned_h0._cells_to_batch(cells)
    n = b.ncells
    c0 = b.conc.clone()
    X, _ = mlp.features(b, norm["mu"], norm["sd"])
    out = net.forward(X)
    flag = torch.empty(n, dtype=torch.int32, device=c0.device)
    dis = torch.empty(n, dtype=torch.float32, device=c0.device)
    mlp._ck(mlp.lib().radm2_mlp_surrogate_apply(mlp._st(), n, b.ld, out.shape[1], WIDTH, mlp._p(out),
                                                mlp._p(norm["tmu"]), mlp._p(norm["tsd"]), FLOOR,
                                                CMIN, float("inf"), 0.0, 0, mlp._p(b.conc),
                                                mlp._p(flag), mlp._p(dis)))
    pred = b.conc[:, :nref].clone()
    b.conc.copy_(c0)
    solver.integrate(b, dt, rtol, atol)
    plain = b.conc[:, :nref].clone()
    r = learned_h0._cells_to_batch({k: np.ascontiguousarray(v[..., :nref]) for k, v in cells.items()
                                    if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")})
    solver.integrate(r, dt, 1e-10, 1e-3)
    solver.sync()
    ref = r.conc
    big = ref.abs() > CMIN

    def cell_err(c):
        e = torch.where(big, (c - ref).abs() / ref.abs().clamp_min(1e-300), torch.zeros_like(ref))
        return e.max(0).values
    ep, eq = cell_err(plain), cell_err(pred)
    d = dis[:nref]
    rows = []
    for t in thrs:
        keep_all = dis <= t
        keep = keep_all[:nref]
        kept = eq[keep]
        rows.append({"thr": t, "fallback_rate": 1.0 - (1.0 - audit) * keep_all.double().mean().item(),
                     "kept_max_rel_err": float(kept.max()) if kept.numel() else None,
                     "kept_median_rel_err": float(kept.median()) if kept.numel() else None})
    return {"plain_max_rel_err": float(ep.max()), "plain_median_rel_err": float(ep.median()),
            "disagreement_quantiles": {str(q): float(torch.quantile(d.float(), q))
                                       for q in (0.01, 0.1, 0.5, 0.9)},
            "gates": rows, "reference": "GPU FP64, rtol 1e-10, atol 1e-3, first %d cells" % nref}


def main():
    import argparse
    import json
    import os
    import sys

    from . import Solver, synth
    ap = argparse.ArgumentParser(description="train the end-state surrogate on solver data")
    ap.add_argument("--train-cells", type=int, default=1048576)
    ap.add_argument("--eval-cells", type=int, default=1048576)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), "mech", "surrogate_mlp.npz"))
    a = ap.parse_args()
    log = lambda m: print("[surrogate] " + m, file=sys.stderr, flush=True)
    s = Solver(0)
    train_cells = synth.make_cells(a.train_cells, 1, 100.0, 0.05)
    net, norm, losses = build(s, train_cells, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
    del train_cells
    save(a.out, net, norm)
    ev = synth.make_cells(a.eval_cells, 3, 100.0, 0.05)
    r = evaluate(s, net, norm, ev)
    out = {k: {kk: vv for kk, vv in v.items() if kk not in ("conc", "flag")} for k, v in r.items()}
    out["losses"] = losses
    out["gates"] = threshold_report(s, net, norm, ev)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
