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
def test_parity_config0_full(solver, schedule):
    # the whole 4,096-cell workload (16x16x16, seed 0, zenith U(0,80)), both
    # schedules: every cell within 1e-6 of the oracle with the same step,
    # accept and reject counts (device and host exp/pow may differ by an
    # ulp in a rate constant; on these seeded inputs no decision flips)
    from radm2_amd import synth
    cells = synth.make_cells(**synth.CONFIGS["c0_4k"])
    assert cells["temp"].shape == (4096,)
    co, oo, code = O.integrate(cells, 600.0, 1e-3, 1.0, nthreads=0)
    assert code == 0
    cg, og = _run_gpu(solver, cells, 600.0, 1e-3, 1.0, schedule)
    assert (og["status"] == 1).all()
    same, ok, rel = _compare(cg, og, co, oo, "sched %d" % schedule)
    assert ok.all(), np.flatnonzero(~ok)
    assert same.all(), np.flatnonzero(~same)


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
            assert ok.all() and same.all()
        for pl, w in ((radm2_amd.PLACE_PRIVATE, 4), (radm2_amd.PLACE_WORKSPACE, 1),
                      (radm2_amd.PLACE_LDS160, 1)):
            solver.set_kernel(pl, w)
            with pytest.raises(radm2_amd.Radm2Error) as e:
                _run_gpu(solver, cells, 600.0, 1e-3, 1.0, 1, opts=radm2_amd.Opts(method=method))
            assert e.value.code == radm2_amd.E_UNSUPPORTED
    finally:
        solver.set_kernel(-1, 0)
        solver.set_schedule(radm2_amd.SCHED_PERSISTENT, 0, 0)


def test_mixed_method_domains_on_the_family_kernel(solver):
    # one batched call with a Ros3 domain and a Rodas3 domain runs on the
    # family kernel; its Ros3 domain is bit-identical to a Ros3-only call on
    # the Ros3 kernel, its Rodas3 domain to a Rodas3-only call
    import radm2_amd
    from radm2_amd import synth
    ca = synth.make_cells(300, 41, 100.0, 0.05)
    cb = synth.make_cells(200, 42, 100.0, 0.05)

    def batch(c):
        d = synth.to_device(c)
        return radm2_amd.CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"],
                                   rc_n2o5=d["rc_n2o5"])

    def out(b):
        return [t.cpu().numpy() for t in (b.conc, b.nstep, b.nacc, b.nrej, b.status, b.hexit)]

    a1, b1 = batch(ca), batch(cb)
    o4 = radm2_amd.Opts(method=4)
    solver.integrate_batched([a1.domain(600.0, 1e-3, 1.0), b1.domain(300.0, 1e-3, 1.0, opts=o4)])
    a2, b2 = batch(ca), batch(cb)
    solver.integrate(a2, 600.0, 1e-3, 1.0)
    solver.integrate(b2, 300.0, 1e-3, 1.0, opts=o4)
    solver.sync()
    for x, y in zip(out(a1) + out(b1), out(a2) + out(b2)):
        np.testing.assert_array_equal(x, y)
    assert (b1.status.cpu().numpy() == 1).all()
    co, oo, code = O.integrate(cb, 300.0, 1e-3, 1.0, opts=O.OptsIn(method=4), nthreads=0)
    same, ok, rel = _compare(b1.conc.cpu().numpy(), dict(nstep=out(b1)[1], nacc=out(b1)[2],
                                                         nrej=out(b1)[3]), co, oo, "mixed")
    assert ok.all() and same.all()


def _forced_cells(n, seed):
    from radm2_amd import synth
    cells = synth.make_cells(n, seed, 100.0, 0.05)
    rng = np.random.default_rng(seed)
    # J at the end of the interval: each cell's rates scaled by U(0, 2)
    # (sunset to brightening), the same factor for all 21 rates
    jend = np.ascontiguousarray(cells["jrate"] * rng.uniform(0.0, 2.0, n)[None, :])
    return cells, jend


@pytest.mark.parametrize("method", [2, 4])
def test_nonautonomous_forcing_matches_oracle(solver, method):
    # time-varying photolysis J(t) = J0 + (t/dt)(J1 - J0) with dF/dT in every
    # stage, on the family kernel, persistent and grid schedules
    import torch
    import radm2_amd
    from radm2_amd import synth
    cells, jend = _forced_cells(256, 50 + method)
    co, oo, code = O.integrate(cells, 600.0, 1e-3, 1.0, opts=O.OptsIn(method=method),
                               jrate_end=jend, nthreads=0)
    assert code == 0 and (oo["status"] == 1).all()
    # the forcing matters: the autonomous run ends elsewhere
    ca, _, _ = O.integrate(cells, 600.0, 1e-3, 1.0, opts=O.OptsIn(method=method), nthreads=0)
    assert np.abs(ca - co).max() > 1e-3 * np.abs(co).max()
    for sched in (radm2_amd.SCHED_PERSISTENT, radm2_amd.SCHED_GRID):
        dev = synth.to_device(cells)
        b = radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"], dev["jrate"],
                                rc_n2o5=dev["rc_n2o5"],
                                jrate_end=torch.from_numpy(jend).cuda())
        solver.set_schedule(sched)
        solver.integrate(b, 600.0, 1e-3, 1.0, opts=radm2_amd.Opts(method=method))
        solver.sync()
        og = dict(nstep=b.nstep.cpu().numpy(), nacc=b.nacc.cpu().numpy(),
                  nrej=b.nrej.cpu().numpy())
        assert (b.status.cpu().numpy() == 1).all()
        same, ok, rel = _compare(b.conc.cpu().numpy(), og, co, oo,
                                 "forced method %d schedule %d" % (method, sched))
        assert ok.all() and same.all()
    solver.set_schedule(radm2_amd.SCHED_PERSISTENT)


@pytest.mark.parametrize("method", [2, 4])
def test_singular_matrix_retries_match_oracle(solver, method):
    # exact zero pivots at H = 500 s (all three cells) and 250 s (cell 0):
    # -8 at once without retries, cell 0 alone with one, none with KPP's five
    import torch
    import radm2_amd
    from radm2_amd import synth
    from test_integrator_oracle import _singular_cells
    cells = _singular_cells(1, method)
    h0 = np.full(3, 500.0)
    for retries, want in ((0, [-8, -8, -8]), (1, [-8, 1, 1]), (5, [1, 1, 1])):
        co, oo, code = O.integrate(cells, 600.0, 1e-3, 1.0, h0=h0, nthreads=0,
                                   opts=O.OptsIn(method=method, sing_retries=retries))
        assert code == 0 and list(oo["status"]) == want
        dev = synth.to_device(cells)
        b = radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"], dev["jrate"],
                                rc_n2o5=dev["rc_n2o5"], h0=torch.from_numpy(h0).cuda())
        solver.integrate(b, 600.0, 1e-3, 1.0,
                         opts=radm2_amd.Opts(method=method, sing_retries=retries))
        solver.sync()
        st = b.status.cpu().numpy()
        assert list(st) == want, (retries, st)
        for key in ("nstep", "nacc", "nrej"):
            np.testing.assert_array_equal(getattr(b, key).cpu().numpy(), oo[key], err_msg=key)
        okc = st == 1
        hx = b.hexit.cpu().numpy()
        np.testing.assert_array_equal(hx[~okc], oo["hexit"][~okc])    # 500 or 250 s exactly
        np.testing.assert_allclose(hx[okc], oo["hexit"][okc], rtol=1e-6)
        if okc.any():
            big = np.abs(co[:, okc]) > 1e3
            cg = b.conc.cpu().numpy()[:, okc]
            assert (np.abs(cg - co[:, okc])[big] <= 1e-6 * np.abs(co[:, okc])[big]).all()


def test_kpp_frontend_nonautonomous(solver):
    # IPAR[0] = 0 over the domain's jrate_end: the same integration as the
    # batched entry point with the forcing, one more Fun per step for dF/dT
    import torch
    import radm2_amd
    from radm2_amd import synth
    cells, jend = _forced_cells(200, 60)

    def batch():
        dev = synth.to_device(cells)
        return radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"],
                                   dev["jrate"], rc_n2o5=dev["rc_n2o5"],
                                   jrate_end=torch.from_numpy(jend).cuda())

    a, b = batch(), batch()
    RPAR = [0.0] * 20
    IPAR = [0] * 20
    IPAR[3] = 2
    RPAR[2] = 6.0
    assert solver.rosenbrock(a, 0.0, 600.0, [1.0] * 59, [1e-3] * 59, RPAR, IPAR) == 1
    solver.integrate(b, 600.0, 1e-3, 1.0, opts=radm2_amd.Opts(hstart=6.0))
    solver.sync()
    np.testing.assert_array_equal(a.conc.cpu().numpy(), b.conc.cpu().numpy())
    nacc, nstp = int(b.nacc.sum()), int(b.nstep.sum())
    assert IPAR[12] == nstp and IPAR[13] == nacc
    assert IPAR[10] == 2 * nacc + nstp


def test_mixed_forcing_methods_and_retries_in_one_call(solver):
    # one batched call on the family kernel whose waves mix a forced Ros3
    # domain, an autonomous Rodas4 domain and an autonomous Ros3 domain with
    # retries enabled: each domain bit-identical to its own single-domain call
    import torch
    import radm2_amd
    from radm2_amd import synth
    ca, ja = _forced_cells(150, 71)
    cb = synth.make_cells(130, 72, 100.0, 0.05)
    cc = synth.make_cells(110, 73, 100.0, 0.05)

    def batch(c, jend=None):
        d = synth.to_device(c)
        return radm2_amd.CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"],
                                   rc_n2o5=d["rc_n2o5"],
                                   jrate_end=None if jend is None else torch.from_numpy(jend).cuda())

    oa, ob, oc = radm2_amd.Opts(), radm2_amd.Opts(method=5), radm2_amd.Opts(sing_retries=5)
    for sched in (radm2_amd.SCHED_PERSISTENT, radm2_amd.SCHED_GRID):
        solver.set_schedule(sched)
        a1, b1, c1 = batch(ca, ja), batch(cb), batch(cc)
        solver.integrate_batched([a1.domain(600.0, 1e-3, 1.0, opts=oa),
                                  b1.domain(600.0, 1e-3, 1.0, opts=ob),
                                  c1.domain(600.0, 1e-3, 1.0, opts=oc)])
        a2, b2, c2 = batch(ca, ja), batch(cb), batch(cc)
        solver.integrate(a2, 600.0, 1e-3, 1.0, opts=oa)
        solver.integrate(b2, 600.0, 1e-3, 1.0, opts=ob)
        solver.integrate(c2, 600.0, 1e-3, 1.0, opts=radm2_amd.Opts())   # Ros3 kernel
        solver.sync()
        for x, y in ((a1, a2), (b1, b2), (c1, c2)):
            for key in ("conc", "nstep", "nacc", "nrej", "status", "hexit"):
                np.testing.assert_array_equal(getattr(x, key).cpu().numpy(),
                                              getattr(y, key).cpu().numpy(), err_msg=key)
            assert (x.status.cpu().numpy() == 1).all()
    solver.set_schedule(radm2_amd.SCHED_PERSISTENT)


def test_forced_ros3_fused_stages_match_the_general_loop(solver):
    # forced Ros3 runs the fused Ros3 stages when every active lane of a wave
    # is a forced Ros3 cell, else the general stage loop.  One forced domain
    # (whole waves: the fused stages) against the same cells as one-cell
    # domains alternating with autonomous cells (mixed waves: the general
    # loop for the forced cells), bit for bit
    import torch
    import radm2_amd
    from radm2_amd import synth
    n = 96
    ca, ja = _forced_cells(n, 81)
    cb = synth.make_cells(n, 82, 100.0, 0.05)
    keys = ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")

    def batch(c, jend=None):
        d = synth.to_device(c)
        return radm2_amd.CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"],
                                   rc_n2o5=d["rc_n2o5"],
                                   jrate_end=None if jend is None else torch.from_numpy(jend).cuda())

    def cell(c, i):
        return {k: np.ascontiguousarray(c[k][..., i:i + 1]) for k in keys}

    o = radm2_amd.Opts()
    for sched in (radm2_amd.SCHED_PERSISTENT, radm2_amd.SCHED_GRID):
        solver.set_schedule(sched)
        whole = batch(ca, ja)
        solver.integrate(whole, 600.0, 1e-3, 1.0, opts=o)
        singles, others, doms = [], [], []   # every batch held for the call
        for i in range(n):
            a = batch(cell(ca, i), np.ascontiguousarray(ja[:, i:i + 1]))
            b = batch(cell(cb, i))
            singles.append(a)
            others.append(b)
            doms += [a.domain(600.0, 1e-3, 1.0, opts=o), b.domain(600.0, 1e-3, 1.0, opts=o)]
        solver.integrate_batched(doms)
        solver.sync()
        assert (whole.status.cpu().numpy() == 1).all()
        assert all((x.status.cpu().numpy() == 1).all() for x in others)
        for key in ("conc", "nstep", "nacc", "nrej", "status", "hexit"):
            got = np.concatenate([getattr(a, key).cpu().numpy() for a in singles], axis=-1)
            np.testing.assert_array_equal(got, getattr(whole, key).cpu().numpy(),
                                          err_msg="%s sched %d" % (key, sched))
    solver.set_schedule(radm2_amd.SCHED_PERSISTENT)


def test_ros2_fused_stages_match_the_general_loop(solver):
    # autonomous Ros2 runs the first two fused stages and its own combination
    # when every active lane of a wave is an autonomous Ros2 cell, else the
    # general stage loop.  One Ros2 domain (whole waves: the fused stages)
    # against the same cells as one-cell domains alternating with Ros3 cells
    # (mixed waves: the general loop), bit for bit, on both schedules; the
    # oracle's step counts and end states within 1e-6 for the first
    import radm2_amd
    from radm2_amd import synth
    n = 96
    ca = synth.make_cells(n, 91, 100.0, 0.05)
    cb = synth.make_cells(n, 92, 100.0, 0.05)
    keys = ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")

    def batch(c):
        d = synth.to_device(c)
        return radm2_amd.CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"],
                                   rc_n2o5=d["rc_n2o5"])

    def cell(c, i):
        return {k: np.ascontiguousarray(c[k][..., i:i + 1]) for k in keys}

    o2, o3 = radm2_amd.Opts(method=1), radm2_amd.Opts()
    co, oo, code = O.integrate(ca, 600.0, 1e-3, 1.0, opts=O.OptsIn(method=1), nthreads=0)
    assert code == 0 and (oo["status"] == 1).all()
    for sched in (radm2_amd.SCHED_PERSISTENT, radm2_amd.SCHED_GRID):
        solver.set_schedule(sched)
        whole = batch(ca)
        solver.integrate(whole, 600.0, 1e-3, 1.0, opts=o2)
        singles, others, doms = [], [], []   # every batch held for the call
        for i in range(n):
            a, b = batch(cell(ca, i)), batch(cell(cb, i))
            singles.append(a)
            others.append(b)
            doms += [a.domain(600.0, 1e-3, 1.0, opts=o2), b.domain(600.0, 1e-3, 1.0, opts=o3)]
        solver.integrate_batched(doms)
        solver.sync()
        assert (whole.status.cpu().numpy() == 1).all()
        assert all((x.status.cpu().numpy() == 1).all() for x in others)
        for key in ("conc", "nstep", "nacc", "nrej", "status", "hexit"):
            got = np.concatenate([getattr(a, key).cpu().numpy() for a in singles], axis=-1)
            np.testing.assert_array_equal(got, getattr(whole, key).cpu().numpy(),
                                          err_msg="%s sched %d" % (key, sched))
        og = dict(nstep=whole.nstep.cpu().numpy(), nacc=whole.nacc.cpu().numpy(),
                  nrej=whole.nrej.cpu().numpy())
        same, ok, rel = _compare(whole.conc.cpu().numpy(), og, co, oo, "Ros2 sched %d" % sched)
        assert ok.all() and same.all()
    solver.set_schedule(radm2_amd.SCHED_PERSISTENT)


def test_family_plan_replays_match_eager(solver):
    # a captured plan of a forced Rodas3 domain replays bit-identically to the
    # eager call, with eager work between the replays
    import torch
    import radm2_amd
    from radm2_amd import synth
    cells, jend = _forced_cells(200, 80)
    d = synth.to_device(cells)
    b = radm2_amd.CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"],
                            rc_n2o5=d["rc_n2o5"], jrate_end=torch.from_numpy(jend).cuda())
    conc0 = b.conc.clone()
    o = radm2_amd.Opts(method=4)
    solver.integrate(b, 600.0, 1e-3, 1.0, opts=o)
    solver.sync()
    eager = [t.cpu().numpy() for t in (b.conc, b.nstep, b.nacc, b.nrej, b.status, b.hexit)]
    assert (eager[4] == 1).all()
    plan = solver.plan([b.domain(600.0, 1e-3, 1.0, opts=o)])
    for rep in range(3):
        b.conc.copy_(conc0)
        b.nstep.zero_()
        torch.cuda.synchronize()
        plan.launch()
        solver.sync()
        got = [t.cpu().numpy() for t in (b.conc, b.nstep, b.nacc, b.nrej, b.status, b.hexit)]
        for x, y in zip(got, eager):
            np.testing.assert_array_equal(x, y, err_msg="replay %d" % rep)
        torch.cuda.synchronize()
        _ = (b.conc * 2.0).sum().item()
    plan.close()


def test_kpp_frontend_nonautonomous_without_forcing_is_autonomous(solver):
    # IPAR[0] = 0 with no jrate_end: J is constant, dF/dT = 0, so the
    # integration is the autonomous one; Nfun still counts the derivative's
    # evaluation once per step
    import radm2_amd
    from radm2_amd import synth
    cells = synth.make_cells(128, 90, 100.0, 0.05)

    def batch():
        d = synth.to_device(cells)
        return radm2_amd.CellBatch(d["conc"], d["temp"], d["cair"], d["h2o"], d["jrate"],
                                   rc_n2o5=d["rc_n2o5"])

    out = []
    for ipar0 in (0, 1):
        b = batch()
        RPAR = [0.0] * 20
        IPAR = [0] * 20
        IPAR[0], IPAR[1], IPAR[3] = ipar0, 1, 2
        RPAR[2] = 6.0
        assert solver.rosenbrock(b, 0.0, 600.0, [1.0], [1e-3], RPAR, IPAR) == 1
        out.append((b.conc.cpu().numpy(), IPAR))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    i0, i1 = out[0][1], out[1][1]
    assert i0[12] == i1[12] and i0[13] == i1[13]
    assert i0[10] == i1[10] + i1[13]
