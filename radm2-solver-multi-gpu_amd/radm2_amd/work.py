"""Algorithmic work per cell-integration, counted from the mechanism table.

Counting rules: a multiply, add or divide is 1 FP64 flop, an fma is 2, a
transcendental (exp, pow, log10, sqrt) is 1.  The counts are the minimum
work of the reference's algorithm (mgpu.cu CS3 step), not of the kernels'
schedule: one Fun and one Jacobian per time step (accepted attempt), and per
attempt one matrix preparation, one LU decomposition, one more Fun, three
solves and the stage/combination/error vector work.  The kernels recompute
Fun(y) and the Jacobian on rejected attempts instead of storing them; that
extra work is not counted.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
MECH = os.path.join(os.path.dirname(HERE), "mech", "radm2.json")


def _counts():
    m = json.load(open(MECH))
    # tendency: monomial products, then one op per term (fma for |c| != 1)
    fun = sum(len(f) for f in m["monomials"])
    for terms in m["tendency"]:
        for q, (r, c) in enumerate(terms):
            cv = abs(float(c))
            if q == 0:
                fun += 0 if cv == 1 else 1
            else:
                fun += 1 if cv == 1 else 2
    # Jacobian as the reference forms it (mgpu.cu:1685-2651): each distinct
    # partial dA_r/dV_j once (a square contributes one partial with a factor
    # 2), then each LU entry accumulates its partials times coefficients
    lu = m["lu"]
    pos = {}
    for row in range(59):
        for kk in range(lu["crow"][row], lu["crow"][row + 1]):
            pos[(row, lu["icol"][kk])] = kk
    partials = {}
    entries = {}
    for i in range(59):
        full = list(m["tendency"][i]) + [[r, c] for (ii, r, c) in m["fun_dropped_terms"] if ii == i]
        for r, c in full:
            facs = m["monomials"][r]
            vj = [j for kind, j in facs if kind == "V"]
            for j in sorted(set(vj)):
                partials[(r, j)] = len(facs) - 1 + (1 if vj.count(j) > 1 else 0)
                entries.setdefault(pos[(i, j)], []).append(abs(float(c)))
    jac = sum(partials.values())
    for cs in entries.values():
        jac += (0 if cs[0] == 1 else 1) + sum(1 if cv == 1 else 2 for cv in cs[1:])
    prep = 59 + 2  # diagonal shift and 1/(H*gamma)
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    dec_fma = 0
    nl = 0
    for k in range(59):
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            nl += 1
            dec_fma += crow[j + 1] - diag[j] - 1
    decomp = nl + 2 * dec_fma + 59      # multiplier, updates, inverse pivots
    solve = 2 * (659 - 59) + 59         # fma per off-diagonal, scale per row
    # stage vectors: Y+K1 (59), K2 (fma), K3 (2 fma), combination (mul + 2 fma
    # + add), error (mul + 2 fma), scale (fma), |.| max, divide, square-add
    vec = 59 * (1 + 2 + 4 + (1 + 4 + 1) + (1 + 4) + 2 + 1 + 1 + 2) + 6
    return dict(fun=fun, jac=jac, prep=prep, decomp=decomp, solve=solve, vec=vec,
                decomp_fma=dec_fma, lower=nl)


COUNTS = _counts()
PER_STEP = COUNTS["fun"] + COUNTS["jac"]
PER_ATTEMPT = COUNTS["prep"] + COUNTS["decomp"] + COUNTS["fun"] + 3 * COUNTS["solve"] + COUNTS["vec"]
# rate constants, once per cell-integration: 57 shared exp/pow, 5 Troe
# fall-off terms (log10 + pow + ~10 ops), ~2 ops per remaining law
RCONST = 57 + 5 * 12 + 2 * 156
# compulsory HBM bytes per cell: in 59 conc + T, M, H2O, rc + 21 J; out 59
# conc + 4 int32 counters/status + exit step
BYTES_PER_CELL = 8 * (59 + 4 + 21) + 8 * 59 + 4 * 4 + 8


def flops(nacc, nrej):
    """Algorithmic FP64 flops of cells with the given accepted/rejected counts."""
    return PER_STEP * nacc + PER_ATTEMPT * (nacc + nrej) + RCONST
