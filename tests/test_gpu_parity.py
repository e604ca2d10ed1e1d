"""HIP path vs the CPU oracle on identical seeded inputs (through the C ABI)."""
import numpy as np
import pytest

import oracle_ct as O

pytestmark = pytest.mark.gpu


def _run_gpu(solver, cells, dt, rtol, atol, schedule, h0=None, opts=None):
    import torch
    import radm2_amd
    from radm2_amd import synth
    dev = synth.to_device(cells)
    h0t = None if h0 is None else torch.from_numpy(np.ascontiguousarray(h0)).cuda()
    b = radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"], dev["jrate"],
                            rc_n2o5=dev["rc_n2o5"], h0=h0t)
    solver.set_schedule(schedule)
    solver.integrate(b, dt, rtol, atol, opts=opts)
    solver.sync()
    out = dict(nstep=b.nstep.cpu().numpy(), nacc=b.nacc.cpu().numpy(),
               nrej=b.nrej.cpu().numpy(), status=b.status.cpu().numpy(),
               hexit=b.hexit.cpu().numpy())
    return b.conc.cpu().numpy(), out


def _compare(cg, og, co, oo, label):
    same_counts = (og["nstep"] == oo["nstep"]) & (og["nacc"] == oo["nacc"]) & (og["nrej"] == oo["nrej"])
    big = np.abs(co) > 1e3
    rel = np.where(big, np.abs(cg - co) / np.maximum(np.abs(co), 1e-300), 0.0)
    cell_rel = rel.max(axis=0)
    ok = cell_rel <= 1e-6
    print("%s: cells %d, same step/accept/reject counts %d, within 1e-6 %d, max rel %.3e"
          % (label, cg.shape[1], same_counts.sum(), ok.sum(), cell_rel.max()))
    return same_counts, ok, cell_rel


@pytest.mark.parametrize("schedule", [0, 1])
def test_parity_config0_subset(solver, schedule):
    from radm2_amd import synth
    cells = synth.make_cells(1024, 0, 80.0, 0.0)
    co, oo, code = O.integrate(cells, 600.0, 1e-3, 1.0, nthreads=0)
    assert code == 0
    cg, og = _run_gpu(solver, cells, 600.0, 1e-3, 1.0, schedule)
    assert (og["status"] == 1).all()
    same, ok, rel = _compare(cg, og, co, oo, "sched %d" % schedule)
    # identical rate constants would give identical bits; libm vs device
    # exp/pow differ by an ulp, which can flip a rare accept/reject decision
    assert ok.mean() >= 0.99
    assert same.mean() >= 0.99


def test_parity_tight_tolerance_reference_run(solver):
    # the GPU FP64 path at rtol 1e-10, atol 1e-3 serves as the accuracy
    # reference over many cells; it must agree with the oracle's own run
    from radm2_amd import synth
    cells = synth.make_cells(64, 11, 100.0, 0.05)
    co, oo, code = O.integrate(cells, 600.0, 1e-10, 1e-3, nthreads=0)
    assert code == 0
    cg, og = _run_gpu(solver, cells, 600.0, 1e-10, 1e-3, 1)
    assert (og["status"] == 1).all()
    same, ok, rel = _compare(cg, og, co, oo, "rtol 1e-10")
    # both are converged solutions: they agree far below the rtol-1e-3 error
    assert rel.max() <= 1e-6


@pytest.mark.parametrize("method", [1, 3, 4, 5])
def test_other_kpp_methods_match_oracle(solver, method):
    # KPP's Ros2, Ros4, Rodas3 and Rodas4 (IPAR[3] 1, 3, 4, 5) on the family
    # kernel (the default placement), persistent and grid schedules; the
    # other placements are built for Ros3 alone and refuse them
    import radm2_amd
    from radm2_amd import synth
    cells = synth.make_cells(256, 30 + method, 100.0, 0.05)
    oo_in = O.OptsIn(method=method)
    co, oo, code = O.integrate(cells, 600.0, 1e-3, 1.0, opts=oo_in, nthreads=0)
    assert code == 0 and (oo["status"] == 1).all()
    try:
        for sched in (radm2_amd.SCHED_PERSISTENT, radm2_amd.SCHED_GRID):
            cg, og = _run_gpu(solver, cells, 600.0, 1e-3, 1.0, sched, opts=radm2_amd.Opts(method=method))
            assert (og["status"] == 1).all()
            same, ok, rel = _compare(cg, og, co, oo, "method %d schedule %d" % (method, sched))
            assert ok.mean() >= 0.99 and same.mean() >= 0.99
        for pl, w in ((radm2_amd.PLACE_PRIVATE, 4), (radm2_amd.PLACE_WORKSPACE, 1),
                      (radm2_amd.PLACE_LDS160, 1)):
            solver.set_kernel(pl, w)
            with pytest.raises(radm2_amd.Radm2Error) as e:
                _run_gpu(solver, cells, 600.0, 1e-3, 1.0, 1, opts=radm2_amd.Opts(method=method))
            assert e.value.code == radm2_amd.E_UNSUPPORTED
    finally:
        solver.set_kernel(-1, 0)
        solver.set_schedule(radm2_amd.SCHED_PERSISTENT, 0, 0)
