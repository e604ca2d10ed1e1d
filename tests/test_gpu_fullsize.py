"""The 1,048,576-cell workload (128x128x64 cells, zenith U(0,100) degrees,
5% plume cells, seed 0; dt 600 s, rtol 1e-3, atol 1) at full size.  The
oracle cannot integrate all of it within a test, so the test checks
properties that hold at any size:
- every cell ends with status 1 and consistent step statistics;
- the persistent and one-shot schedules agree bit for bit;
- a cell's result does not depend on the batch it is integrated in;
- a seeded sample of 512 cells matches the oracle."""
import numpy as np
import pytest

import oracle_ct as O

pytestmark = pytest.mark.gpu

KEYS = ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")
OUT = ("conc", "nstep", "nacc", "nrej", "status", "hexit")


@pytest.fixture(scope="module")
def cells():
    from radm2_amd import synth
    return synth.make_cells(**synth.CONFIGS["c1_1m"])


def _run(solver, cells, schedule, block):
    import radm2_amd
    from radm2_amd import synth
    dev = synth.to_device(cells)
    b = radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"], dev["jrate"],
                            rc_n2o5=dev["rc_n2o5"])
    solver.set_schedule(schedule, block, 1)
    solver.set_kernel(-1, 0)
    try:
        solver.integrate(b, 600.0, 1e-3, 1.0)
        solver.sync()
    finally:
        solver.set_schedule(1, 0, 0)
    return {k: getattr(b, k).cpu().numpy() for k in OUT}


def test_full_size_schedules_agree_and_sample_matches_oracle(solver, cells):
    import radm2_amd
    n = cells["conc"].shape[1]
    assert n == 1 << 20
    print("full size: persistent schedule, %d cells" % n, flush=True)
    p = _run(solver, cells, radm2_amd.SCHED_PERSISTENT, 64)
    assert (p["status"] == 1).all()
    assert (p["nacc"] >= 1).all() and (p["nstep"] >= p["nacc"] + p["nrej"]).all()
    assert (p["hexit"] > 0).all() and np.isfinite(p["conc"]).all()
    print("mean attempts per cell %.3f, max %d" % (p["nstep"].mean(), p["nstep"].max()), flush=True)

    print("full size: one-shot grid schedule", flush=True)
    g = _run(solver, cells, radm2_amd.SCHED_GRID, 64)
    for k in OUT:
        np.testing.assert_array_equal(g[k], p[k], err_msg=k)
    del g

    # a seeded sample, integrated on its own: the same bits as inside the
    # full batch, and the oracle's results within 1e-6
    idx = np.sort(np.random.default_rng(7).choice(n, 512, replace=False))
    sub = {k: np.ascontiguousarray(cells[k][..., idx]) for k in KEYS}
    print("sample of %d cells: alone on the GPU, then the oracle" % idx.size, flush=True)
    s = _run(solver, sub, radm2_amd.SCHED_PERSISTENT, 64)
    for k in OUT:
        np.testing.assert_array_equal(s[k], p[k][..., idx], err_msg=k)
    co, oo, code = O.integrate(sub, 600.0, 1e-3, 1.0, nthreads=0)
    assert code == 0 and (oo["status"] == 1).all()
    big = np.abs(co) > 1e3
    rel = np.where(big, np.abs(s["conc"] - co) / np.maximum(np.abs(co), 1e-300), 0.0).max(axis=0)
    same = (s["nstep"] == oo["nstep"]) & (s["nacc"] == oo["nacc"]) & (s["nrej"] == oo["nrej"])
    print("within 1e-6: %d of %d, same step counts: %d, max rel %.3e"
          % ((rel <= 1e-6).sum(), idx.size, same.sum(), rel.max()), flush=True)
    assert (rel <= 1e-6).all()
    assert same.all()


@pytest.mark.parametrize("method", [1, 5])
def test_full_size_family_methods_sample_matches_oracle(solver, cells, method):
    # KPP's Ros2 (the fused stages) and Rodas4 (the general stage loop) on
    # all 1,048,576 cells through the family kernel: every cell ends with
    # status 1, a seeded sample integrated on its own gives the same bits as
    # inside the full batch, and it matches the oracle within 1e-6 with the
    # oracle's step counts
    import radm2_amd
    from radm2_amd import synth
    n = cells["conc"].shape[1]
    opts = radm2_amd.Opts(method=method)

    def run(c):
        dev = synth.to_device(c)
        b = radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"], dev["jrate"],
                                rc_n2o5=dev["rc_n2o5"])
        solver.integrate(b, 600.0, 1e-3, 1.0, opts=opts)
        solver.sync()
        return {k: getattr(b, k).cpu().numpy() for k in OUT}

    print("full size: method %d, %d cells" % (method, n), flush=True)
    p = run(cells)
    assert (p["status"] == 1).all()
    assert (p["nacc"] >= 1).all() and (p["nstep"] >= p["nacc"] + p["nrej"]).all()
    assert np.isfinite(p["conc"]).all()
    print("mean attempts per cell %.3f, max %d" % (p["nstep"].mean(), p["nstep"].max()), flush=True)
    idx = np.sort(np.random.default_rng(10 + method).choice(n, 256, replace=False))
    sub = {k: np.ascontiguousarray(cells[k][..., idx]) for k in KEYS}
    s = run(sub)
    for k in OUT:
        np.testing.assert_array_equal(s[k], p[k][..., idx], err_msg=k)
    co, oo, code = O.integrate(sub, 600.0, 1e-3, 1.0, opts=O.OptsIn(method=method), nthreads=0)
    assert code == 0 and (oo["status"] == 1).all()
    big = np.abs(co) > 1e3
    rel = np.where(big, np.abs(s["conc"] - co) / np.maximum(np.abs(co), 1e-300), 0.0).max(axis=0)
    same = (s["nstep"] == oo["nstep"]) & (s["nacc"] == oo["nacc"]) & (s["nrej"] == oo["nrej"])
    print("method %d sample: within 1e-6 %d of %d, same step counts %d, max rel %.3e"
          % (method, (rel <= 1e-6).sum(), idx.size, same.sum(), rel.max()), flush=True)
    assert (rel <= 1e-6).all()
    assert same.all()
