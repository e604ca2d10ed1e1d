"""Pin the oracle's step-size control and its error-estimator weights.

* The accept/reject block of the reference's acceptStep (mgpu.cu:3238-3260:
  Err, Fac, Hnew, the RejectLastH / RejectMoreH state machine, FacRej after
  repeated rejections, the Hnew <= H cap after a rejection, acceptance when
  H <= Hmin) was evaluated as written on scripted error sequences by
  tests/golden/make_golden.py (radm2_golden_control.npz); the oracle's
  oracle_step_control, which both of its cell integrators call, must
  reproduce every attempt bit for bit.  Mutated controllers (the two flags
  swapped in either arm, the flag updates reordered) disagree with the
  fixtures, so they pin each of those mistakes.
* The error estimate sum_j E_j K_j of one step is the difference between
  the method's solution and its embedded one, so it falls as h^ELO (Ros2 2,
  Ros3 3, Ros4 4, Rodas3 3, Rodas4 4): observed at halving steps on a
  nonlinear two-reaction problem, for the oracle's Ros3 stages (the path of
  oracle_ros3_cell) and its generic stage loop.  Ros3's E2
  (mgpu.cu:1046) perturbed in the 8th digit breaks the observed order.
"""
import os

import numpy as np
import pytest

import oracle_ct as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "radm2_golden_control.npz"))
ROWS = G["CTRL_ROWS"]
SETS = G["CTRL_SETS"]
NO, NO2, O3 = 57, 58, 45


def _opts(s):
    hmin, facmin, facmax, facrej, facsafe, elo, h0 = SETS[s]
    o = O.OptsOut()
    o.hmin, o.facmin, o.facmax, o.facrej, o.facsafe = hmin, facmin, facmax, facrej, facsafe
    o.roundoff = 2.0 ** -52
    return o, elo, h0


def _sequences():
    keys = sorted({(int(r[0]), int(r[1])) for r in ROWS})
    return [(s, q, ROWS[(ROWS[:, 0] == s) & (ROWS[:, 1] == q)]) for s, q in keys]


def _run(control, s, rows):
    """Replay one scripted sequence through `control`; per attempt
    (accepted, H, T, rej_last, rej_more)."""
    o, elo, h0 = _opts(s)
    H, T, rl, rm = float(h0), 0.0, 0, 0
    out = []
    for r in rows:
        acc, H, T, rl, rm = control(float(r[3]), elo, o, H, T, rl, rm)
        out.append((acc, H, T, rl, rm))
    return out


def _expected(rows):
    return [(int(r[8]), r[4], r[5], int(r[6]), int(r[7])) for r in rows]


def test_fixture_covers_every_arm():
    acc = ROWS[:, 8] == 1
    assert acc.any() and (~acc).any()
    assert (ROWS[:, 7] == 1).any()                       # RejectMoreH set
    # FacRej taken: a rejection whose previous attempt left RejectMoreH set
    prev_more = np.r_[0, ROWS[:-1, 7]] * (np.r_[-1, ROWS[:-1, 1]] == ROWS[:, 1])
    assert ((~acc) & (prev_more == 1)).any()
    # accepted with Err > 1 because H <= Hmin (set 2 has Hmin = 1e-3)
    err = np.sqrt(ROWS[:, 3] / 59.0)
    assert (acc & (err > 1.0)).any()
    # Err exactly 1 is accepted
    assert (acc & (ROWS[:, 3] == 59.0)).any()


@pytest.mark.parametrize("s,q,rows", _sequences(), ids=lambda v: str(v) if np.isscalar(v) else "")
def test_oracle_controller_matches_reference_text(s, q, rows):
    got = _run(O.step_control, s, rows)
    assert got == _expected(rows)


def _mutant(kind):
    """The reference's controller with one planted mistake (test-side)."""
    def control(err_sum, elo, o, H, T, rl, rm):
        err = np.sqrt(err_sum / 59.0)
        with np.errstate(divide="ignore"):
            fac = min(o.facmax, max(o.facmin, o.facsafe / np.float64(err) ** (1.0 / elo)))
        hn = H * fac
        if err <= 1.0 or H <= o.hmin:
            T = T + H
            if (rm if kind == "cap_uses_more" else rl):
                hn = min(hn, H)
            return 1, hn, T, 0, 0
        if (rl if kind == "facrej_uses_last" else rm):
            hn = H * o.facrej
        if kind == "flags_reordered":
            rl = 1
            rm = rl
        else:
            rm, rl = rl, 1
        return 0, hn, T, rl, rm
    return control


@pytest.mark.parametrize("kind", ["cap_uses_more", "facrej_uses_last", "flags_reordered"])
def test_fixtures_reject_planted_controller_mistakes(kind):
    bad = 0
    for s, q, rows in _sequences():
        if _run(_mutant(kind), s, rows) != _expected(rows):
            bad += 1
    assert bad > 0


def _problem():
    # NO + O3 -> NO2 (reaction 25) and NO2 photolysis (reaction 0): a
    # nonlinear system with time scales of ~100 s
    k = np.zeros(156)
    k[25] = 1.8e-14
    k[0] = 8e-3
    y = np.zeros(59)
    y[O3] = y[NO] = 5e11
    y[NO2] = 1e11
    return y, k


def _observed_orders(method, table=None):
    y, k = _problem()
    hs = [2.0 ** -i for i in range(8)]
    norms = [np.linalg.norm(O.ros_attempt(method, y, k, 1e17, 2.5e19, h, table=table)[2])
             for h in hs]
    return [np.log2(norms[i] / norms[i + 1]) for i in range(len(norms) - 1)]


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5])
def test_error_estimate_falls_as_h_to_the_elo(method):
    elo = O.method_coeffs(method)["g_elo"][1]
    orders = _observed_orders(method)
    assert all(abs(p - elo) < 0.1 for p in orders), (elo, orders)


def _ros3_table():
    t = O.method_coeffs(2)
    t["newf"] = [1, 1, 0, 0, 0, 0]   # stage 3 reuses F2 (mgpu.cu:1035-1037)
    return t


def test_generic_loop_error_estimate_equals_ros3_stages():
    y, k = _problem()
    for h in (1.0, 0.01):
        e1, yn1, ye1 = O.ros_attempt(2, y, k, 1e17, 2.5e19, h)
        e2, yn2, ye2 = O.ros_attempt(2, y, k, 1e17, 2.5e19, h, table=_ros3_table())
        assert e1 == e2
        np.testing.assert_array_equal(yn1, yn2)
        np.testing.assert_array_equal(ye1, ye2)


def test_perturbed_e2_breaks_the_observed_order():
    t = _ros3_table()
    assert _observed_orders(2, t) == pytest.approx(_observed_orders(2), abs=1e-12)
    t["E"][1] *= 1.0 + 1e-8
    orders = _observed_orders(2, t)
    assert not all(abs(p - 3.0) < 0.1 for p in orders), orders
