"""Learned initial step: trained on solver-generated data by the project's
kernels, it must not cost attempts or accuracy, and its proposals go through
the FP64 solver unchanged (the solver stays the judge)."""
import os

import numpy as np
import pytest

import oracle_ct as O

pytestmark = pytest.mark.gpu
NSUB = 256


def _max_rel_err(held, runs):
    """Max relative error of each run's end state against the FP64 oracle at
    rtol 1e-8 over the first NSUB cells (species above 1e3 molec/cm3)."""
    sub = {k: np.ascontiguousarray(v[..., :NSUB]) for k, v in held.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    ref, _, code = O.integrate(sub, 600.0, 1e-8, 1e-3)
    assert code == 0
    big = np.abs(ref) > 1e3
    return [np.max(np.abs(r["conc"][:, :NSUB][big] - ref[big]) / np.abs(ref[big])) for r in runs]


def _check(r, n, label):
    plain, learned = r["plain"], r["learned"]
    assert plain["ok"] == learned["ok"] == n
    return plain, learned


def test_learned_h0_gain_and_accuracy(solver):
    from radm2_amd import learned_h0, synth
    train = synth.make_cells(65536, 1, 100.0, 0.05)
    net, mu, sd, losses, info = learned_h0.build(solver, train, epochs=3)
    assert losses[-1] < losses[0]
    # the sweep's own bounds: per-cell best <= best constant <= plain
    assert info["cell_best"] <= min(info["per_h0"].values()) <= info["plain"]
    held = synth.make_cells(16384, 2, 100.0, 0.05)
    plain, learned = _check(learned_h0.evaluate(solver, net, mu, sd, held, reps=0), 16384, "fresh")
    # the gain the network buys over INTEGRATE's Hstart = 0.01*dt
    assert learned["attempts"] < plain["attempts"]
    assert learned["used"] > 0.5
    e_plain, e_learned = _max_rel_err(held, [plain, learned])
    print("attempts %.2f -> %.2f (sweep: best constant %.2f, per-cell best %.2f), used %.3f, "
          "max rel err vs rtol-1e-8 oracle %.2e -> %.2e"
          % (plain["attempts"], learned["attempts"], min(info["per_h0"].values()),
             info["cell_best"], learned["used"], e_plain, e_learned))
    assert e_learned <= 1.25 * e_plain


def _oracle_ref(seed):
    """The CPU oracle's rtol-1e-10 end states of the first cells of a
    1,048,576-cell workload (tests/golden/accuracy_ref.npz) and the digest of
    their initial states."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "accuracy_ref.npz"))
    return d["s%d_conc" % seed], str(d["s%d_conc0_sha256" % seed])


def _cell_err(c, ref):
    big = np.abs(ref) > 1e3
    return np.where(big, np.abs(c - ref) / np.maximum(np.abs(ref), 1e-300), 0.0).max(0)


def test_shipped_network_on_the_held_out_million(solver):
    # the workload as specified: the shipped network (trained on 2,097,152
    # seed-1 cells) on all 1,048,576 held-out seed-2 cells, proposals
    # clamped to [1e-6 s, dt]; every cell succeeds, attempts fall, and on
    # the first cells the error against the CPU oracle at rtol 1e-10 is no
    # worse than the plain call's
    import hashlib
    from radm2_amd import learned_h0, synth
    path = os.path.join(os.path.dirname(learned_h0.__file__), "..", "mech", "h0_mlp.npz")
    net, mu, sd = learned_h0.load(path)
    held = synth.make_cells(1048576, 2, 100.0, 0.05)
    ref, digest = _oracle_ref(2)
    m = ref.shape[1]
    assert hashlib.sha256(np.ascontiguousarray(held["conc"][:, :m]).tobytes()).hexdigest() == digest
    r = learned_h0.evaluate(solver, net, mu, sd, held, reps=0, hlo=1e-6)
    plain, learned = _check(r, 1048576, "held-out million")
    e_plain = _cell_err(plain["conc"][:, :m], ref)
    e_learned = _cell_err(learned["conc"][:, :m], ref)
    print("held-out 1M cells: attempts %.2f -> %.2f, rejected %.2f -> %.2f, used %.3f; first %d "
          "cells vs rtol-1e-10 oracle: max %.3e -> %.3e, median %.3e -> %.3e"
          % (plain["attempts"], learned["attempts"], plain["rejected"], learned["rejected"],
             learned["used"], m, e_plain.max(), e_learned.max(), np.median(e_plain),
             np.median(e_learned)))
    assert learned["attempts"] < plain["attempts"]
    assert learned["rejected"] < plain["rejected"]
    assert e_learned.max() <= 1.25 * e_plain.max()
    assert np.median(e_learned) <= 1.25 * np.median(e_plain)


def test_shipped_network_on_unseen_cells(solver):
    from radm2_amd import learned_h0, synth
    path = os.path.join(os.path.dirname(learned_h0.__file__), "..", "mech", "h0_mlp.npz")
    net, mu, sd = learned_h0.load(path)
    held = synth.make_cells(65536, 3, 100.0, 0.05)
    plain, learned = _check(learned_h0.evaluate(solver, net, mu, sd, held, reps=0), 65536, "shipped")
    e_plain, e_learned = _max_rel_err(held, [plain, learned])
    print("shipped network: attempts %.2f -> %.2f, used %.3f, max rel err %.2e -> %.2e"
          % (plain["attempts"], learned["attempts"], learned["used"], e_plain, e_learned))
    assert learned["attempts"] < plain["attempts"]
    assert e_learned <= 1.25 * e_plain
