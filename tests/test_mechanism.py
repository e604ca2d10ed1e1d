"""Mechanism invariants of the oracle, pinned by mathematics rather than by
restating the oracle's formulas.

  * the Jacobian equals the finite-difference derivative of the tendency at
    every LU position except the two H2O2 entries, where the reference's
    Jacobian keeps the H2O2 + OH loss (mgpu.cu:1275 comments it out of Fun):
    there the difference is exactly -k33*[OH] and -k33*[H2O2];
  * the tendency is linear in the rate constants;
  * a reaction whose rate constant is the only nonzero one moves only its
    reactants and products, with the tendency of each reactant negative;
  * Troe fall-off at k0(T) = kinf(T) reduces to 0.6 * kinf / 2.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_ct as O

MECH = json.load(open(os.path.join(O.ROOT, "radm2-solver-multi-gpu_amd", "mech", "radm2.json")))
LU = MECH["lu"]
IDX = {}
for row in range(59):
    for kk in range(LU["crow"][row], LU["crow"][row + 1]):
        IDX[(row, LU["icol"][kk])] = kk
H2O2, OH = 13, 52


def _state(seed):
    # O(1) values keep every term of a tendency on one scale, so central
    # differences resolve each partial derivative to ~1e-9
    rng = np.random.default_rng(seed)
    V = rng.uniform(0.5, 2.0, 59)
    k = rng.uniform(0.5, 2.0, 156)
    return V, 0.7, 1.3, k


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jacobian_is_derivative_of_tendency(seed):
    V, h2o, m, k = _state(seed)
    J = O.jac(V, h2o, m, k)
    f0 = O.fun(V, h2o, m, k)
    for j in range(59):
        h = 1e-6 * V[j]
        Vp, Vm = V.copy(), V.copy()
        Vp[j] += h
        Vm[j] -= h
        col = (O.fun(Vp, h2o, m, k) - O.fun(Vm, h2o, m, k)) / (2 * h)
        for i in range(59):
            exact = J[IDX[(i, j)]] if (i, j) in IDX else 0.0
            if (i, j) == (H2O2, H2O2):
                exact += k[33] * V[OH]       # restore Fun's view
            if (i, j) == (H2O2, OH):
                exact += k[33] * V[H2O2]
            scale = np.max(np.abs(col)) + abs(exact) + 1e-300
            assert abs(col[i] - exact) <= 1e-6 * scale, (i, j, col[i], exact)
    assert f0.shape == (59,)


def test_h2o2_quirk_is_exactly_the_dropped_loss():
    V, h2o, m, k = _state(5)
    J = O.jac(V, h2o, m, k)
    k2 = k.copy()
    k2[33] = 0.0
    J2 = O.jac(V, h2o, m, k2)
    d = J - J2
    assert d[IDX[(H2O2, H2O2)]] == pytest.approx(-k[33] * V[OH], rel=1e-14)
    assert d[IDX[(H2O2, OH)]] == pytest.approx(-k[33] * V[H2O2], rel=1e-14)
    # and Fun does not see reaction 33 on H2O2 at all
    f1 = O.fun(V, h2o, m, k)
    f2 = O.fun(V, h2o, m, k2)
    assert f1[H2O2] == f2[H2O2]


def test_tendency_is_linear_in_rate_constants():
    V, h2o, m, k = _state(3)
    rng = np.random.default_rng(9)
    k2 = rng.uniform(0.5, 2.0, 156)
    a, b = 0.37, 1.9
    lhs = O.fun(V, h2o, m, a * k + b * k2)
    rhs = a * O.fun(V, h2o, m, k) + b * O.fun(V, h2o, m, k2)
    assert np.max(np.abs(lhs - rhs)) <= 1e-13 * np.max(np.abs(rhs))


@pytest.mark.parametrize("r", range(156))
def test_single_reaction_moves_only_its_species(r):
    V = np.full(59, 1.0)
    k = np.zeros(156)
    k[r] = 1.0
    f = O.fun(V, 1.0, 1.0, k)
    mono = MECH["monomials"][r]
    reactants = {i for kind, i in mono if kind == "V"}
    touched = {i for i in range(59) if f[i] != 0.0}
    # species listed in the reaction's tendency terms
    listed = {i for i in range(59) for rr, c in MECH["tendency"][i] if rr == r}
    assert touched <= listed
    for i in reactants:
        if r == 33 and i == H2O2:
            continue  # the dropped loss term
        if i in listed:
            assert f[i] < 0.0 or any(float(c) >= 0 for rr, c in MECH["tendency"][i] if rr == r)


def test_troe_falloff_midpoint():
    # reaction 29: TROE(1.80e-31, 3.2, 4.70e-12, 1.4) at T = 300 K; at
    # M = kinf/k0 the ratio is 1, log10 = 0, and k = k0/(1+1) * 0.6
    T = 300.0
    M = 4.70e-12 / 1.80e-31
    k = O.rconst(T, M, np.zeros(21), 0.0)
    assert k[29] == pytest.approx(0.5 * 4.70e-12 * 0.6, rel=1e-14)


def test_arrhenius_and_photolysis_slots():
    T = 250.0
    J = np.arange(1, 22, dtype=float) * 1e-5
    k = O.rconst(T, 2e19, J, 0.0)
    assert k[25] == pytest.approx(2.0e-12 * math.exp(-1400.0 / T), rel=1e-15)
    # photolysis reactions are copies of the jv slots (mgpu.cu:763-783)
    assert set(np.round(k[:21] / 1e-5).astype(int)) == set(range(1, 22))
