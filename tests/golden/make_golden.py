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
  step-size control of acceptStep                   mgpu.cu:3238-3260
    (with MAX/MIN/ONE as mgpu.cu:11, 191-192 define them), run on
    scripted sequences of error sums, option sets and ELO values

The decomposition kernel (mgpu.cu:2740-2779) indexes tables that the
reference never fills (LU_CROW/LU_DIAG/LU_ICOL); its loop is restated here
with those tables rebuilt from the solve text, so the LU fixture pins the
unrolled decomposition against an independent loop.

Output: ``radm2_golden.npz`` and ``radm2_golden_control.npz`` (plain
arrays, no pickles) beside this script.
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


def macro_fn(src, name):
    """A function-like macro of the reference (``#define MAX(a,b) ...``) as a
    Python function: its single ternary ``c ? x : y`` becomes ``x if c else y``."""
    line = next(ln for ln in src if re.match(r"#define\s+%s\(" % name, ln))
    m = re.match(r"#define\s+\w+\(([^)]*)\)\s*(.*)$", line)
    args = [a.strip() for a in m.group(1).split(",")]
    body = m.group(2).strip()
    assert body[0] == "(" and body[-1] == ")", body
    cond, rest = body[1:-1].split("?")
    a, b = rest.split(":")
    return eval("lambda %s: (%s) if (%s) else (%s)" % (", ".join(args), a, cond, b))


def macro_const(src, name):
    line = next(ln for ln in src if re.match(r"#define\s+%s\s" % name, ln))
    return float(eval(line.split(None, 2)[2].replace("(double)", "")))


def c_block_to_python(lines):
    """Translate a small block of C statements (assignments, if / else with
    or without braces, counted for loops) into Python source.  Lines inside
    #ifdef ... #endif are dropped (the reference builds without those
    flags)."""
    out, ind, skip = [], 0, 0
    for raw in lines:
        ln = raw.strip()
        if ln.startswith("#ifdef") or ln.startswith("#if "):
            skip += 1
            continue
        if ln.startswith("#endif"):
            skip -= 1
            continue
        if skip or not ln:
            continue
        ln = re.sub(r"MY_I\((\w+)\)", r"\1", ln)
        ln = ln.replace("(double)", "").replace("||", " or ").replace("&&", " and ")
        ln = re.sub(r"\bsqrt\(", "math.sqrt(", ln)
        ln = re.sub(r"\bpow\(", "math.pow(", ln)
        if ln == "}":
            ind -= 1
            continue
        if re.match(r"\}\s*else\s*\{$", ln):
            out.append("    " * (ind - 1) + "else:")
            continue
        m = re.match(r"for\s*\(\s*(\w+)\s*=\s*(\w+)\s*;\s*\w+\s*<\s*(\w+)\s*;\s*\w+\+\+\s*\)\s*\{$", ln)
        if m:
            out.append("    " * ind + "for %s in range(%s, %s):" % m.groups())
            ind += 1
            continue
        m = re.match(r"if\s*(\(.*\))\s*\{$", ln)
        if m:
            out.append("    " * ind + "if %s:" % m.group(1))
            ind += 1
            continue
        m = re.match(r"if\s*(\(.*?\))\s*([^()]+?=.*);$", ln)
        if m:
            out.append("    " * ind + "if %s:" % m.group(1))
            out.append("    " * (ind + 1) + m.group(2).strip())
            continue
        assert ln.endswith(";"), ln
        out.append("    " * ind + ln[:-1].strip())
    assert ind == 0, ind
    return "\n".join(out)


# option sets (Hmin, FacMin, FacMax, FacRej, FacSafe) and ELO for the
# controller fixtures: KPP's defaults at Ros3's order, other factors at
# order 2, and a positive Hmin at order 4 (its "accept when H <= Hmin" arm)
CONTROL_SETS = (
    ((0.0, 0.2, 6.0, 0.1, 0.9), 3.0, 6.0),
    ((0.0, 0.5, 3.0, 0.3, 0.8), 2.0, 60.0),
    ((1.0e-3, 0.2, 6.0, 0.1, 0.9), 4.0, 0.05),
)


def control_sequences(rng):
    """Scripted scaled errors Err per attempt: fixed patterns (accepts,
    runs of one to five rejections, Err exactly 1, just above and below
    1, 0, tiny and huge values) and random runs."""
    fixed = [
        [0.5, 0.5, 0.5],
        [2.0, 0.5, 0.7],
        [2.0, 3.0, 0.5, 0.4],
        [2.0, 3.0, 4.0, 0.5, 0.5],
        [2.0, 3.0, 4.0, 5.0, 6.0, 0.5, 2.0, 0.5],
        [1.0, 1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53, 0.999, 1.001],
        [0.0, 1e-300, 1e-12, 1e6, 1e300, 0.5],
        [50.0, 0.3, 9.0, 9.0, 9.0, 9.0, 9.0, 9.0, 9.0, 9.0, 9.0, 9.0, 9.0, 9.0, 0.01],
    ]
    rand = []
    for _ in range(24):
        n = int(rng.integers(10, 40))
        e = 10 ** rng.uniform(-3.0, 1.5, n)
        # runs of rejections: force a few stretches above 1
        for a in rng.integers(0, n, 3):
            e[a:a + int(rng.integers(1, 5))] = 10 ** rng.uniform(0.01, 1.0)
        rand.append(list(e))
    return fixed + rand


def control_fixture(src, rng):
    """acceptStep's control block (mgpu.cu:3238-3260) evaluated as written
    on every (option set, sequence): per attempt the input error sum and
    H, T, RejectLastH, RejectMoreH and the accepted flag after it."""
    MAX, MIN = macro_fn(src, "MAX"), macro_fn(src, "MIN")
    ONE = macro_const(src, "ONE")
    code = c_block_to_python(src[3237:3260])
    rows = []   # set, sequence, step, err_sum, H, T, rej_last, rej_more, accepted
    seqs = control_sequences(rng)
    for si, (fac, elo, h0) in enumerate(CONTROL_SETS):
        hmin, facmin, facmax, facrej, facsafe = (np.float64(v) for v in fac)
        for qi, errs in enumerate(seqs):
            st = dict(H=[np.float64(h0)], T=[np.float64(0.0)], RejectLastH=[0], RejectMoreH=[0],
                      vecMask2=[0], Y=[0.0] * 59, Ynew=[0.0] * 59)
            for t, e in enumerate(errs):
                err_sum = np.float64(59.0) * np.float64(e) * np.float64(e)
                ns = dict(st, math=math, MAX=MAX, MIN=MIN, ONE=ONE, NVAR=59, p=0,
                          Err=err_sum, Hmin=hmin, FacMin=facmin, FacMax=facmax, FacRej=facrej,
                          FacSafe=facsafe, ros_ELO=np.float64(elo), Direction=1)
                with np.errstate(divide="ignore", over="ignore"):
                    exec(code, ns)
                rows.append((si, qi, t, float(err_sum), float(ns["H"][0]), float(ns["T"][0]),
                             ns["RejectLastH"][0], ns["RejectMoreH"][0],
                             1 if ns["vecMask2"][0] == 0 else 0))
    a = np.array(rows, dtype=np.float64)
    sets = np.array([list(f) + [elo, h0] for f, elo, h0 in CONTROL_SETS])
    return dict(CTRL_ROWS=a, CTRL_SETS=sets, CTRL_CODE=np.frombuffer(code.encode(), np.uint8))


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

    dst = os.path.join(HERE, "radm2_golden_control.npz")
    np.savez_compressed(dst, **control_fixture(src, np.random.default_rng(20261204)))
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/mgpu.cu")
