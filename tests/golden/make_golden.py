#!/usr/bin/env python3
"""Write golden vectors by evaluating the reference's own arithmetic text.

The reference (a single CUDA file) cannot be compiled here, so this script
reads the per-cell arithmetic of ``mgpu.cu`` as text and evaluates each
statement literally in IEEE double (Python floats, the platform libm for
exp/pow/log10), in the order the reference writes it.  It does not use this
package's mechanism table or code generator, so the fixtures pin both.

Blocks evaluated (reference line numbers):
  rate-law helpers ARR2/TROE/TROEE/THERMAL_T2/k46   mgpu.cu:708-750
  rate constants RCONST[0..155]                     mgpu.cu:763-918
  rates A[] and tendencies Vdot[]                   mgpu.cu:1103-1385
  Jacobian partials B[] and entries JVS[]           mgpu.cu:1685-2651
  unrolled forward/backward substitution            mgpu.cu:2873-3090

The decomposition kernel (mgpu.cu:2740-2779) indexes tables that the
reference never fills (LU_CROW/LU_DIAG/LU_ICOL); its loop is restated here
with those tables rebuilt from the solve text, so the LU fixture pins the
unrolled decomposition against an independent loop.

Output: ``radm2_golden.npz`` (plain arrays, no pickles) beside this script.
Run only where the reference is mounted:
    python make_golden.py /root/reference/mgpu.cu
"""
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NCELL = 8


def stmts(src, lo, hi):
    text = "\n".join(src[lo - 1:hi])
    text = re.sub(r"//[^\n]*", "", text)
    out = []
    for st in text.split(";"):
        st = " ".join(st.split())
        if st:
            out.append(st)
    return out


def c2py(expr):
    expr = re.sub(r"MY_I_A\((\d+)\)", r"\1", expr)
    expr = re.sub(r"MY_I\((\w+)\)", r"\1", expr)
    expr = re.sub(r"\bpow\(", "math.pow(", expr)
    expr = re.sub(r"\bexp\(", "math.exp(", expr)
    expr = re.sub(r"\blog10\(", "math.log10(", expr)
    expr = re.sub(r"\(double\)", "", expr)
    return expr


def helper_functions(src):
    """Translate the five rate-law helpers (mgpu.cu:708-750) into Python."""
    text = "\n".join(src[707:750])
    ns = {"math": math}
    for m in re.finditer(r"__device__\s+double\s+(\w+)\(([^)]*)\)\s*\{(.*?)\n\}", text, re.S):
        name, args, body = m.group(1), m.group(2), m.group(3)
        argnames = [a.split()[-1] for a in args.split(",")]
        lines = []
        body = re.sub(r"//[^\n]*", "", body)
        for st in body.split(";"):
            st = " ".join(st.split())
            if not st or st.startswith("double "):
                continue
            if st.startswith("return"):
                lines.append("    return " + c2py(st[len("return"):].strip()))
            else:
                lines.append("    " + c2py(st))
        code = "def %s(%s):\n%s\n" % (name, ", ".join(argnames), "\n".join(lines))
        exec(code, ns)
    for need in ("ARR2", "TROE", "TROEE", "THERMAL_T2", "k46"):
        assert need in ns, need
    return ns


def eval_block(statements, ns, lhs_names):
    """Execute assignments NAME[i] = expr in order inside namespace ns."""
    for st in statements:
        m = re.match(r"(\w+)\[(?:MY_I(?:_A)?\()?(\d+)\)?\]\s*=\s*(.*)$", st)
        assert m and m.group(1) in lhs_names, st
        ns[m.group(1)][int(m.group(2))] = eval(c2py(m.group(3)), ns)


def solve_text(src):
    out = []
    for st in stmts(src, 2873, 3090):
        st = re.sub(r"JVS\((\d+)\)", r"JVS[\1]", st)
        st = re.sub(r"X\((\d+)\)", r"X[\1]", st)
        out.append(st)
    return out


def lu_tables(solve_statements):
    pos = {}
    for st in solve_statements:
        i = int(re.match(r"X\[(\d+)\]", st).group(1))
        rhs = st.split("=", 1)[1]
        for k, j in re.findall(r"JVS\[(\d+)\]\*X\[(\d+)\]", rhs.replace(" ", "")):
            pos[int(k)] = (i, int(j))
        d = re.search(r"/\(?JVS\[(\d+)\]\)?$", rhs.replace(" ", ""))
        if d:
            pos[int(d.group(1))] = (i, i)
    lst = [pos[k] for k in range(659)]
    crow = [min(k for k in range(659) if lst[k][0] == r) for r in range(59)] + [659]
    diag = [lst.index((r, r)) for r in range(59)]
    icol = [j for (_, j) in lst]
    return crow, diag, icol


def kpp_decomp(JVS, crow, diag, icol):
    """mgpu.cu:2740-2779 restated (row-wise no-pivot LU with a dense row W)."""
    W = [0.0] * 59
    for k in range(59):
        if JVS[diag[k]] == 0.0:
            return k + 1
        for kk in range(crow[k], crow[k + 1]):
            W[icol[kk]] = JVS[kk]
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            a = -W[j] / JVS[diag[j]]
            W[j] = -a
            for jj in range(diag[j] + 1, crow[j + 1]):
                W[icol[jj]] += a * JVS[jj]
        for kk in range(crow[k], crow[k + 1]):
            JVS[kk] = W[icol[kk]]
    return 0


def main(path):
    src = open(path).read().split("\n")
    helpers = helper_functions(src)
    rc_st = stmts(src, 763, 918)
    fun_st = stmts(src, 1103, 1385)
    jac_st = stmts(src, 1685, 2651)
    sol_st = solve_text(src)
    crow, diag, icol = lu_tables(sol_st)

    rng = np.random.default_rng(20240611)
    T = rng.uniform(240.0, 310.0, NCELL)
    P = rng.uniform(300.0, 1050.0, NCELL) * 100.0
    M = P / (1.380649e-23 * T) * 1e-6
    H2O = M * 10 ** rng.uniform(-3.5, -1.5, NCELL)
    J = 10 ** rng.uniform(-7, -2, (NCELL, 21))
    J[1] = 0.0                                  # a night cell
    RC = 10 ** rng.uniform(-8, -4, NCELL)
    V = M[:, None] * 10 ** rng.uniform(-12, -6, (NCELL, 59))
    hinv = 10 ** rng.uniform(-3, 1, NCELL)      # 1/(H*gamma) for the prepared matrix

    RCONST = np.zeros((NCELL, 156))
    VDOT = np.zeros((NCELL, 59))
    JAC = np.zeros((NCELL, 659))
    LU = np.zeros((NCELL, 659))
    SOLB = np.zeros((NCELL, 59))
    SOLX = np.zeros((NCELL, 59))
    for c in range(NCELL):
        ns = dict(helpers)
        ns.update(RCONST=[0.0] * 156, jv=list(J[c]) + [0.0] * 7, C_M=float(M[c]),
                  TEMP=float(T[c]), rc_n2o5=float(RC[c]))
        eval_block(rc_st, ns, {"RCONST"})
        RCONST[c] = ns["RCONST"]

        ns = dict(A=[0.0] * 156, Vdot=[0.0] * 59, V=list(V[c]),
                  F=[float(H2O[c]), float(M[c])], RCT=list(RCONST[c]), math=math)
        eval_block(fun_st, ns, {"A", "Vdot"})
        VDOT[c] = ns["Vdot"]

        ns = dict(B=[0.0] * 280, JVS=[0.0] * 659, V=list(V[c]),
                  F=[float(H2O[c]), float(M[c])], RCT=list(RCONST[c]), math=math)
        eval_block(jac_st, ns, {"B", "JVS"})
        JAC[c] = ns["JVS"]

        # prepared matrix -J + hinv*I on the true diagonal, then the KPP LU
        G = [-x for x in JAC[c]]
        for r in range(59):
            G[diag[r]] += hinv[c]
        assert kpp_decomp(G, crow, diag, icol) == 0
        LU[c] = G

        b = rng.standard_normal(59)
        ns = dict(X=list(b), JVS=list(G))
        for st in sol_st:
            exec(st, ns)
        SOLB[c] = b
        SOLX[c] = ns["X"]

    dst = os.path.join(HERE, "radm2_golden.npz")
    np.savez_compressed(dst, T=T, M=M, H2O=H2O, J=J, RC_N2O5=RC, V=V, hinv=hinv,
                        RCONST=RCONST, VDOT=VDOT, JVS=JAC, LU=LU, SOLVE_B=SOLB,
                        SOLVE_X=SOLX, LU_CROW=np.array(crow), LU_DIAG=np.array(diag),
                        LU_ICOL=np.array(icol))
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/mgpu.cu")
