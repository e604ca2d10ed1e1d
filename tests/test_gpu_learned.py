"""Learned initial step: trained on solver-generated data by the project's
kernels, it must not cost attempts or accuracy, and its proposals go through
the FP64 solver unchanged (the solver stays the judge)."""
import numpy as np
import pytest

import oracle_ct as O

pytestmark = pytest.mark.gpu


def test_learned_h0_gain_and_accuracy(solver):
    from radm2_amd import learned_h0, synth
    train = synth.make_cells(65536, 1, 100.0, 0.05)
    net, mu, sd, losses, info = learned_h0.build(solver, train, epochs=3)
    assert losses[-1] < losses[0]
    held = synth.make_cells(16384, 2, 100.0, 0.05)
    r = learned_h0.evaluate(solver, net, mu, sd, held, reps=0)
    plain, learned = r["plain"], r["learned"]
    assert plain["ok"] == learned["ok"] == 16384
    # the gain the network buys over INTEGRATE's Hstart = 0.01*dt
    assert learned["attempts"] < plain["attempts"]
    assert learned["used"] > 0.5
    # accuracy judged against the FP64 oracle at rtol 1e-8 on a subsample
    sub = {k: np.ascontiguousarray(v[..., :64]) for k, v in held.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    ref, _, code = O.integrate(sub, 600.0, 1e-8, 1e-3)
    assert code == 0
    big = np.abs(ref) > 1e3

    def err(c):
        return np.max(np.abs(c[:, :64][big] - ref[big]) / np.abs(ref[big]))
    e_plain, e_learned = err(plain["conc"]), err(learned["conc"])
    print("attempts %.2f -> %.2f, used %.3f, max rel err vs rtol-1e-8 oracle %.2e -> %.2e"
          % (plain["attempts"], learned["attempts"], learned["used"], e_plain, e_learned))
    assert e_learned <= max(2.0 * e_plain, 1e-3)
