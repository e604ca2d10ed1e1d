"""Pin the CPU oracle to the reference's own arithmetic.

The fixtures in tests/golden/radm2_golden.npz were written by
tests/golden/make_golden.py, which evaluates the reference's rate-constant,
tendency, Jacobian and solve statements (mgpu.cu:708-918, 1103-1385,
1685-2651, 2873-3090) literally, without this package's mechanism table.
"""
import os

import numpy as np
import pytest

import oracle_ct as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "radm2_golden.npz"))
CELLS = range(G["T"].shape[0])


def relerr(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


@pytest.mark.parametrize("c", CELLS)
def test_rate_constants_match_reference_text(c):
    k = O.rconst(G["T"][c], G["M"][c], G["J"][c], G["RC_N2O5"][c])
    # same expressions, same libm, no contraction: identical bits
    np.testing.assert_array_equal(k, G["RCONST"][c])


@pytest.mark.parametrize("c", CELLS)
def test_tendency_matches_reference_text(c):
    f = O.fun(G["V"][c], G["H2O"][c], G["M"][c], G["RCONST"][c])
    ref = G["VDOT"][c]
    # the oracle accumulates with fma in table order; the reference rounds
    # each product: agreement to a few ulp of the largest term
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(f - ref)) <= 1e-13 * scale


@pytest.mark.parametrize("c", CELLS)
def test_jacobian_matches_reference_text(c):
    j = O.jac(G["V"][c], G["H2O"][c], G["M"][c], G["RCONST"][c])
    ref = G["JVS"][c]
    assert relerr(j, ref, 1e-300) <= 1e-13
    # the 91 entries the reference writes as literal 0 stay exactly 0
    zero = ref == 0.0
    assert zero.sum() >= 87
    assert np.all(j[zero] == 0.0)


@pytest.mark.parametrize("c", CELLS)
def test_lu_matches_reference_decomposition(c):
    G0 = O.jac_prep(G["V"][c], G["H2O"][c], G["M"][c], G["RCONST"][c], G["hinv"][c])
    rc, lu = O.decomp(G0, ref_division=True)
    assert rc == 0
    ref = G["LU"][c]
    # entry-wise relative error; cancellation in eliminated entries loosens it
    assert np.max(np.abs(lu - ref)) <= 1e-12 * np.max(np.abs(ref))
    rc2, lui = O.decomp(G0, ref_division=False)
    assert rc2 == 0
    diag = G["LU_DIAG"]
    off = np.ones(659, bool)
    off[diag] = False
    assert np.max(np.abs(lui[off] - lu[off])) <= 1e-12 * np.max(np.abs(ref))
    # pivots formed by cancellation keep fewer digits than the entries that
    # formed them: compare on the scale of the largest entry of the matrix
    assert np.max(np.abs(1.0 / lui[diag] - lu[diag])) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("c", CELLS)
def test_solve_matches_reference_text(c):
    x = O.solve(G["LU"][c], G["SOLVE_B"][c], ref_division=True)
    np.testing.assert_array_equal(x, G["SOLVE_X"][c])


def test_lu_tables_match_reference_solve_pattern():
    import json
    m = json.load(open(os.path.join(O.ROOT, "radm2-solver-multi-gpu_amd", "mech", "radm2.json")))
    np.testing.assert_array_equal(m["lu"]["crow"], G["LU_CROW"])
    np.testing.assert_array_equal(m["lu"]["diag"], G["LU_DIAG"])
    np.testing.assert_array_equal(m["lu"]["icol"], G["LU_ICOL"])
