"""GPU tests through the C ABI: every kernel variant and schedule against the
oracle or against each other, batched domains, graph plans, options, edge
cases and per-cell failure isolation."""
import ctypes

import numpy as np
import pytest

import oracle_ct as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cells():
    from radm2_amd import synth
    return synth.make_cells(2048, 5, 100.0, 0.05)


def _batch(cells, h0=None, hfirst=False, pad=0):
    import torch
    import radm2_amd
    from radm2_amd import synth
    dev = synth.to_device(cells)
    if pad:
        n = cells["conc"].shape[1]
        ld = n + pad
        for k in ("conc", "jrate"):
            t = torch.full((dev[k].shape[0], ld), -7.0, dtype=torch.float64, device="cuda")
            t[:, :n] = dev[k]
            dev[k] = t
        for k in ("temp", "cair", "h2o", "rc_n2o5"):
            t = torch.full((ld,), 250.0, dtype=torch.float64, device="cuda")
            t[:n] = dev[k]
            dev[k] = t
    h0t = None if h0 is None else torch.from_numpy(np.ascontiguousarray(h0)).cuda()
    return radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"], dev["jrate"],
                               rc_n2o5=dev["rc_n2o5"], h0=h0t, hfirst=hfirst)


def _out(b):
    return dict(conc=b.conc.cpu().numpy(), nstep=b.nstep.cpu().numpy(), nacc=b.nacc.cpu().numpy(),
                nrej=b.nrej.cpu().numpy(), status=b.status.cpu().numpy(),
                hexit=b.hexit.cpu().numpy())


def _close(cg, co, tol=1e-6):
    big = np.abs(co) > 1e3
    rel = np.where(big, np.abs(cg - co) / np.maximum(np.abs(co), 1e-300), 0.0).max(axis=0)
    return rel <= tol


def test_all_variants_and_schedules_bit_identical(solver, cells):
    import radm2_amd
    sub = {k: np.ascontiguousarray(v[..., :512]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    ref = None
    configs = []
    for pl, w in ((0, 1), (0, 2), (0, 4), (0, 8), (1, 1), (1, 2), (1, 4), (2, 1), (3, 1), (4, 1),
                  (5, 1), (6, 1)):
        for sched, block in ((radm2_amd.SCHED_GRID, 64), (radm2_amd.SCHED_GRID, 256),
                             (radm2_amd.SCHED_GRID, 512), (radm2_amd.SCHED_GRID, 1024),
                             (radm2_amd.SCHED_PERSISTENT, 64), (radm2_amd.SCHED_PERSISTENT, 256)):
            if pl >= radm2_amd.PLACE_LDS160 and block != 64:
                continue          # LDS placements run 64-lane blocks only
            if pl >= radm2_amd.PLACE_SPLIT2 and sched == radm2_amd.SCHED_GRID:
                continue          # cells over several lanes: persistent only
            if block >= 512 and pl != radm2_amd.PLACE_WORKSPACE:
                continue          # the wide lockstep blocks on one placement suffice
            configs.append((pl, w, sched, block))
    for pl, w, sched, block in configs:
        b = _batch(sub)
        solver.set_schedule(sched, block, 3)
        solver.set_kernel(pl, w)
        solver.integrate(b, 600.0, 1e-3, 1.0)
        solver.sync()
        out = _out(b)
        if ref is None:
            ref = out
            co, oo, code = O.integrate(sub, 600.0, 1e-3, 1.0)
            assert code == 0 and (out["status"] == 1).all()
            assert _close(out["conc"], co).all()
        for k in ref:
            np.testing.assert_array_equal(out[k], ref[k], err_msg=str((pl, w, sched, block, k)))
    solver.set_schedule(1, 0, 0)
    solver.set_kernel(-1, 0)


def test_single_domain_entry_point_matches_batched(solver, cells):
    # radm2_integrate (explicit arguments) and radm2_integrate_batched (a
    # domain table) run the same kernels: bit-identical results
    sub = {k: np.ascontiguousarray(v[..., :300]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    a, b = _batch(sub), _batch(sub)
    solver.integrate(a, 600.0, 1e-3, 1.0)
    solver.integrate_single(b, 600.0, 1e-3, 1.0)
    solver.sync()
    oa, ob = _out(a), _out(b)
    for k in oa:
        np.testing.assert_array_equal(oa[k], ob[k], err_msg=k)
    assert (ob["status"] == 1).all()


def test_split_placements_reject_the_grid_schedule(solver, cells):
    import radm2_amd
    sub = {k: np.ascontiguousarray(v[..., :64]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    b = _batch(sub)
    for pl in (radm2_amd.PLACE_SPLIT2, radm2_amd.PLACE_SPLIT4):
        solver.set_kernel(pl, 1)
        solver.set_schedule(radm2_amd.SCHED_GRID, 64, 0)
        with pytest.raises(radm2_amd.Radm2Error) as e:
            solver.integrate(b, 600.0, 1e-3, 1.0)
        assert e.value.code == radm2_amd.E_UNSUPPORTED
    solver.set_schedule(1, 0, 0)
    solver.set_kernel(-1, 0)


def test_lds_placements_reject_wide_blocks(solver, cells):
    import radm2_amd
    sub = {k: np.ascontiguousarray(v[..., :64]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    b = _batch(sub)
    for pl in (radm2_amd.PLACE_LDS160, radm2_amd.PLACE_LDS312, radm2_amd.PLACE_LDS160_REG112):
        solver.set_kernel(pl, 1)
        solver.set_schedule(radm2_amd.SCHED_PERSISTENT, 256, 0)
        with pytest.raises(radm2_amd.Radm2Error) as e:
            solver.integrate(b, 600.0, 1e-3, 1.0)
        assert e.value.code == radm2_amd.E_ARG
        # a plan checks the same launch shape before it captures anything
        with pytest.raises(radm2_amd.Radm2Error) as e:
            solver.plan([b.domain(600.0, 1e-3, 1.0)])
        assert e.value.code == radm2_amd.E_ARG
    solver.set_schedule(1, 0, 0)
    solver.set_kernel(-1, 0)


def test_batched_domains_match_oracle_per_domain(solver):
    import radm2_amd
    from radm2_amd import synth
    specs = [(100, 60.0, 1e-2, 1.0, 21), (513, 600.0, 1e-3, 1.0, 22), (777, 3600.0, 1e-4, 10.0, 23)]
    batches, hosts = [], []
    for n, dt, rtol, atol, seed in specs:
        c = synth.make_cells(n, seed, 100.0, 0.05)
        hosts.append(c)
        batches.append(_batch(c))
    doms = [b.domain(dt, rtol, atol) for b, (n, dt, rtol, atol, seed) in zip(batches, specs)]
    solver.integrate_batched(doms)
    solver.sync()
    for b, c, (n, dt, rtol, atol, seed) in zip(batches, hosts, specs):
        co, oo, code = O.integrate(c, dt, rtol, atol)
        out = _out(b)
        assert (out["status"] == 1).all()
        assert _close(out["conc"], co).all(), (n, dt)
        same = (out["nstep"] == oo["nstep"]) & (out["nrej"] == oo["nrej"])
        assert same.all()


def test_plan_graph_replays_match_eager(solver, cells):
    import torch
    sub = {k: np.ascontiguousarray(v[..., :700]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    b = _batch(sub)
    conc0 = b.conc.clone()
    solver.integrate(b, 600.0, 1e-3, 1.0)
    solver.sync()
    eager = _out(b)
    plan = solver.plan([b.domain(600.0, 1e-3, 1.0)])
    for rep in range(3):
        b.conc.copy_(conc0)
        b.nstep.zero_()
        torch.cuda.synchronize()
        plan.launch()
        solver.sync()
        got = _out(b)
        for k in eager:
            np.testing.assert_array_equal(got[k], eager[k], err_msg="replay %d %s" % (rep, k))
        # unrelated eager work between replays
        torch.cuda.synchronize()
        _ = (b.conc * 2.0).sum().item()
    plan.close()


def test_padding_columns_untouched_and_prefix_stable(solver, cells):
    sub = {k: np.ascontiguousarray(v[..., :257]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    b = _batch(sub, pad=63)
    solver.integrate(b, 300.0, 1e-3, 1.0, ncells=257)
    solver.sync()
    out = _out(b)
    assert (out["conc"][:, 257:] == -7.0).all()
    assert (out["nstep"][257:] == 0).all()
    # the first 256 cells do not depend on the presence of cell 257
    sub2 = {k: np.ascontiguousarray(v[..., :256]) for k, v in sub.items()}
    b2 = _batch(sub2)
    solver.integrate(b2, 300.0, 1e-3, 1.0)
    solver.sync()
    out2 = _out(b2)
    np.testing.assert_array_equal(out2["conc"], out["conc"][:, :256])
    np.testing.assert_array_equal(out2["nstep"], out["nstep"][:256])


def test_zero_and_single_cell(solver, cells):
    sub = {k: np.ascontiguousarray(v[..., :1]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    b = _batch(sub)
    solver.integrate(b, 600.0, 1e-3, 1.0, ncells=0)
    solver.sync()
    assert (b.nstep.cpu().numpy() == 0).all()
    solver.integrate(b, 600.0, 1e-3, 1.0)
    solver.sync()
    co, oo, code = O.integrate(sub, 600.0, 1e-3, 1.0)
    out = _out(b)
    assert out["status"][0] == 1 and _close(out["conc"], co).all()


def test_options_and_h0_match_oracle(solver, cells):
    import radm2_amd
    sub = {k: np.ascontiguousarray(v[..., :300]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    kw = dict(hmin=0.0, hmax=120.0, hstart=0.5, facmin=0.3, facmax=4.0, facrej=0.2,
              facsafe=0.85, max_steps=0, method=2)
    b = _batch(sub)
    solver.integrate(b, 600.0, 1e-3, 1.0, opts=radm2_amd.Opts(**kw))
    solver.sync()
    co, oo, code = O.integrate(sub, 600.0, 1e-3, 1.0, opts=O.OptsIn(**kw))
    out = _out(b)
    assert code == 0 and (out["status"] == 1).all()
    assert _close(out["conc"], co).all()
    assert (out["nstep"] == oo["nstep"]).all()
    # per-cell Hstart: min(h0, dt, Hmax), DeltaMin below 10 roundoff
    h0 = np.geomspace(1e-16, 1e3, 300)
    b = _batch(sub, h0=h0)
    solver.integrate(b, 600.0, 1e-3, 1.0)
    solver.sync()
    co, oo, code = O.integrate(sub, 600.0, 1e-3, 1.0, h0=h0)
    out = _out(b)
    assert _close(out["conc"], co).all()
    assert (out["nstep"] == oo["nstep"]).all()


def test_failure_statuses_isolated(solver, cells):
    import radm2_amd
    sub = {k: np.array(v[..., :128]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    sub["conc"][57, 5] = np.nan                       # one poisoned cell
    opts = radm2_amd.Opts(max_steps=175)
    b = _batch(sub)
    solver.integrate(b, 600.0, 1e-3, 1.0, opts=opts)
    solver.sync()
    out = _out(b)
    co, oo, code = O.integrate(sub, 600.0, 1e-3, 1.0, opts=O.OptsIn(max_steps=175))
    assert out["status"][5] == -7 and oo["status"][5] == -7
    assert (out["status"] == -6).sum() > 0
    np.testing.assert_array_equal(out["status"], oo["status"])
    ok = out["status"] == 1
    assert ok.sum() > 10
    assert _close(out["conc"][:, ok], co[:, ok]).all()


def test_call_level_errors(solver, cells):
    import radm2_amd
    sub = {k: np.ascontiguousarray(v[..., :8]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    b = _batch(sub)
    for opts, code in ((radm2_amd.Opts(max_steps=-1), -1), (radm2_amd.Opts(method=6), -2),
                       (radm2_amd.Opts(hmin=-1.0), -3), (radm2_amd.Opts(facmax=-1.0), -4)):
        with pytest.raises(radm2_amd.Radm2Error) as e:
            solver.integrate(b, 600.0, 1e-3, 1.0, opts=opts)
        assert e.value.code == code
    with pytest.raises(radm2_amd.Radm2Error) as e:
        solver.integrate(b, 600.0, 1.5, 1.0)
    assert e.value.code == -5
    d = b.domain(600.0, 1e-3, 1.0)
    d.ld = 4                                            # ld < ncells
    with pytest.raises(radm2_amd.Radm2Error) as e:
        solver.integrate_batched([d])
    assert e.value.code == radm2_amd.E_ARG


def test_kpp_frontend_statistics(solver, cells):
    sub = {k: np.ascontiguousarray(v[..., :200]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    b = _batch(sub)
    RPAR = [0.0] * 20
    IPAR = [0] * 20
    IPAR[0] = 1
    IPAR[3] = 2
    RPAR[2] = 0.01 * 240.0
    ierr = solver.rosenbrock(b, 0.0, 240.0, [1.0] * 59, [1e-3] * 59, RPAR, IPAR)
    assert ierr == 1
    co, oo, code = O.integrate(sub, 240.0, 1e-3, 1.0)
    assert abs(IPAR[12] - oo["nstep"].sum()) <= 0.01 * oo["nstep"].sum()
    assert IPAR[13] + IPAR[14] == IPAR[12]
    assert IPAR[15] == IPAR[12] and IPAR[16] == 3 * IPAR[12]
    assert RPAR[10] == 240.0
    # the reference rejects IPAR[3] = 0 (its own default) with -2
    IPAR2 = [0] * 20
    IPAR2[0] = 1
    assert solver.rosenbrock(b, 0.0, 240.0, [1.0] * 59, [1e-3] * 59, [0.0] * 20, IPAR2) == -2
    # Rodas4 (IPAR[3] = 5): six solves per attempt, five new F evaluations
    b5 = _batch(sub)
    RPAR5 = [0.0] * 20
    IPAR5 = [0] * 20
    IPAR5[0] = 1
    IPAR5[3] = 5
    RPAR5[2] = 0.01 * 240.0
    assert solver.rosenbrock(b5, 0.0, 240.0, [1.0] * 59, [1e-3] * 59, RPAR5, IPAR5) == 1
    co5, oo5, code5 = O.integrate(sub, 240.0, 1e-3, 1.0, opts=O.OptsIn(hstart=2.4, method=5))
    assert code5 == 0
    assert abs(IPAR5[12] - oo5["nstep"].sum()) <= 0.01 * oo5["nstep"].sum()
    assert IPAR5[16] == 6 * IPAR5[12] and IPAR5[10] == IPAR5[13] + 5 * IPAR5[12]


def test_batched_64_uneven_domains_plan_matches_oracle(solver):
    # the batched workload: 64 domains, log-uniform 1k..200k cells (about
    # 2.5M), per-domain dt in {60, 300, 600, 3600} s and rtol in {1e-2, 1e-3,
    # 1e-4}, run as one captured graph; 32 cells of every domain checked
    # against the oracle with the domain's own dt and rtol
    import time
    import torch
    from radm2_amd import synth
    doms = synth.make_domains()
    total = sum(c["conc"].shape[1] for c, _, _, _ in doms)
    assert 2_000_000 <= total <= 3_500_000
    batches = [_batch(c) for c, _, _, _ in doms]
    conc0 = [b.conc.clone() for b in batches]
    plan = solver.plan([b.domain(dt, rtol, atol) for b, (_, dt, rtol, atol) in zip(batches, doms)])
    for rep in range(2):
        for b, c0 in zip(batches, conc0):
            b.conc.copy_(c0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan.launch()
        solver.sync()
        el = time.perf_counter() - t
    print("64 domains, %d cells: %.3f s, %.0f cell-integrations/s" % (total, el, total / el))
    nsub = 32
    good = 0
    for b, (c, dt, rtol, atol) in zip(batches, doms):
        st = b.status.cpu().numpy()
        assert (st == 1).all(), (dt, rtol)
        sub = {k: np.ascontiguousarray(v[..., :nsub]) for k, v in c.items()
               if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
        co, oo, code = O.integrate(sub, dt, rtol, atol)
        assert code == 0
        ok = _close(b.conc[:, :nsub].cpu().numpy(), co)
        ok &= b.nstep[:nsub].cpu().numpy() == oo["nstep"]
        good += int(ok.sum())
    assert good == nsub * len(doms), good
    plan.close()


def test_autotuner_picks_a_compiled_configuration(solver, cells):
    # coordinate descent over placement x register budget, schedule x block
    # and refill on a sample: every candidate the library accepts is timed,
    # every one gives the first candidate's end states and step counts bit
    # for bit, the batch comes back restored, and the pick runs
    import json
    import torch
    import radm2_amd
    from radm2_amd import autotune
    sub = {k: np.ascontiguousarray(v[..., :2048]) for k, v in cells.items()
           if k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    b = _batch(sub)
    c0 = b.conc.clone()
    try:
        r = autotune.autotune(solver, b, 600.0, sample=1024, cache=None, use_cache=False)
        assert r["mismatches"] == [] and r["sample"] == 1024 and not r["cached"]
        assert torch.equal(b.conc, c0)
        finite = {k: v for k, v in r["timings"].items() if np.isfinite(v)}
        assert json.dumps(r["best"], sort_keys=True) in finite
        # every kernel variant, and at least one other schedule
        assert len(finite) >= len(autotune.KERNELS) + 1
        # every kernel variant and every schedule of the search was timed
        seen = [json.loads(k) for k in finite]
        assert {(c["placement"], c["waves"]) for c in seen} >= set(autotune.KERNELS)
        assert {(c["schedule"], c["block"]) for c in seen} >= {
            (radm2_amd.SCHED_PERSISTENT, 64), (radm2_amd.SCHED_GRID, 64)}
        print("autotune best %s of %d candidates" % (r["best"], len(finite)))
        # the chosen configuration is applied and integrates like the oracle
        solver.integrate(b, 600.0, 1e-3, 1.0)
        solver.sync()
        out = _out(b)
        co, oo, code = O.integrate(sub, 600.0, 1e-3, 1.0)
        assert code == 0 and _close(out["conc"], co).all()
        assert (out["nstep"] == oo["nstep"]).all()
    finally:
        solver.set_schedule(1, 0, 0)
        solver.set_kernel(-1, 0)


def test_measured_fp64_peak_is_plausible(solver):
    # the roofline report's attainable FP64 peak (FMA chains in registers):
    # positive and not above MI355X's published 78.6 TFLOP/s by more than
    # measurement noise
    tflops = solver.measure_fp64_peak()
    print("attainable FP64 FMA %.1f TFLOP/s" % tflops)
    assert 10.0 < tflops <= 1.05 * 78.6
