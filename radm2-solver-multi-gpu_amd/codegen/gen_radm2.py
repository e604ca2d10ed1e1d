#!/usr/bin/env python3
"""Build-time generator: mechanism table -> straight-line HIP device code.

Reads ``mech/radm2.json`` and writes ``csrc/generated/radm2_gen.h`` holding
the per-cell numerics of the RADM2 Rosenbrock step as fully unrolled,
branch-free device functions with compile-time array indices:

  rconst    rate constants (reference RCONST, mgpu.cu:752-919) with the
            exponentials shared across reactions of equal activation
            temperature and the Troe powers shared across equal exponents
  fun       156 mass-action rates -> 59 tendencies (mgpu.cu:1084-1386)
  jac_prep  G = -J + ghinv*I written straight into the LU layout
            (mgpu.cu:1672-2652 plus ros_PrepareMatrix, mgpu.cu:3128-3149,
            with the shift on the true diagonal)
  decomp    row-wise (IKJ) no-pivot sparse LU in the 659-position pattern,
            inverse pivots stored on the diagonal (mgpu.cu:2726-2780)
  solve     forward (unit L) and backward (U) substitution (mgpu.cu:2854-3096)

Arithmetic order contract (shared with the CPU oracle, which implements the
same order with table-driven loops, so both produce identical bits for the
same rate constants):
  * a monomial rate is k[r] times its factors left to right;
  * every accumulation is acc = fma(c, term, acc) starting from 0, in the
    order the mechanism table lists the terms;
  * a Jacobian term is c * (k[r] * the other factors left to right), taken
    while walking species i, its tendency terms (dropped terms last), and
    the V factors of each monomial in order; G accumulates -c*term, and the
    diagonal shift is added after all terms;
  * decomposition and solves update in increasing column order.

Array arguments are generic "indexables" (raw pointers or strided views),
so the caller decides whether a working array lives in VGPRs, LDS or a
global SoA workspace.
"""
import ast
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)


def lit(c):
    """C double literal for a coefficient string such as '0.65' or '2'."""
    s = c.strip()
    v = float(s)
    if v == int(v) and "e" not in s.lower() and "." not in s:
        return "%d.0" % int(v)
    return s


def load_mech(path=None):
    path = path or os.path.join(PKG, "mech", "radm2.json")
    with open(path) as f:
        return json.load(f)


# --------------------------------------------------------------------------
# rate constants
# --------------------------------------------------------------------------
class RateEmitter:
    """Translate the rate-law notation into C with shared subexpressions."""

    def __init__(self):
        self.pre = []          # (name, C expression)
        self.cache = {}

    def _shared(self, key, expr):
        if key not in self.cache:
            name = "rc_s%d" % len(self.cache)
            self.cache[key] = name
            self.pre.append((name, expr))
        return self.cache[key]

    def exp_over_T(self, num):
        # exp(num/T); num is a C literal string (its value decides sharing)
        return self._shared(("expT", float(num)), "exp(%s / T)" % num)

    def zt_pow(self, n):
        return self._shared(("ztpow", float(n)), "pow(zt, %s)" % n)

    def emit(self, node):
        if isinstance(node, ast.Expression):
            return self.emit(node.body)
        if isinstance(node, ast.Constant):
            return self.src_of(node)
        if isinstance(node, ast.Name):
            return {"T": "T", "M": "M", "RC_N2O5": "RC"}[node.id]
        if isinstance(node, ast.Subscript):
            assert node.value.id == "J"
            return "J[%d]" % node.slice.value
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
            return "(-%s)" % self.emit(node.operand)
        if isinstance(node, ast.BinOp):
            op = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}[type(node.op)]
            return "(%s %s %s)" % (self.emit(node.left), op, self.emit(node.right))
        if isinstance(node, ast.Call):
            fn = node.func.id
            a = [self.src_of(x) for x in node.args]
            if fn == "ARR2":
                # A*exp(-B/T): -B/T == (-B)/T exactly, so share by -B
                return "(%s * %s)" % (a[0], self.exp_over_T(neg(a[1])))
            if fn == "exp":
                # exp(c/T) as written
                arg = node.args[0]
                assert isinstance(arg, ast.BinOp) and isinstance(arg.op, ast.Div) \
                    and isinstance(arg.right, ast.Name) and arg.right.id == "T"
                return self.exp_over_T(self.src_of(arg.left))
            if fn == "pow":
                return "pow(%s, %s)" % (self.emit(node.args[0]), self.emit(node.args[1]))
            if fn == "TROE":
                k0, n, ki, m = a
                return "radm2_troe(%s, %s, %s, %s, M)" % (k0, self.zt_pow(n), ki, self.zt_pow(m))
            if fn == "TROEE":
                A, B, k0, n, ki, m = a
                return "(%s * %s * radm2_troe(%s, %s, %s, %s, M))" % (
                    A, self.exp_over_T(neg(B)), k0, self.zt_pow(n), ki, self.zt_pow(m))
            if fn == "THERMAL_T2":
                c, d = a
                return "((T * T) * %s * %s)" % (c, self.exp_over_T(neg(d)))
            if fn == "K46":
                return "radm2_k46(T, M, %s, %s, %s)" % (
                    self.exp_over_T("785.0"), self.exp_over_T("1440.0"), self.exp_over_T("725.0"))
            raise ValueError(fn)
        raise ValueError(ast.dump(node))

    def src_of(self, node):
        if isinstance(node, ast.Constant):
            return self._src[node.col_offset:node.end_col_offset]
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub) \
                and isinstance(node.operand, ast.Constant):
            return "-" + self.src_of(node.operand)
        return self.emit(node)


def neg(s):
    s = s.strip()
    return s[1:] if s.startswith("-") else "-" + s


def photo_j(mech):
    """J index of each photolysis reaction: the rate laws that are a bare
    J[s] are exactly the first ones, in reaction order."""
    import re
    out = []
    for e in mech["rates"]:
        mt = re.fullmatch(r"\(J\[(\d+)\]\)", e.strip())
        if not mt:
            break
        out.append(int(mt.group(1)))
    assert all(not re.search(r"\bJ\[", e) for e in mech["rates"][len(out):])
    assert sorted(out) == list(range(21))
    return out


def rate_exprs(mech):
    """C expression of every rate constant (reference RCONST, mgpu.cu:752-919)
    and the shared exp/pow subexpressions they use: (exprs, pre)."""
    em = RateEmitter()
    exprs = []
    for e in mech["rates"]:
        em._src = e
        exprs.append(em.emit(ast.parse(e, mode="eval")))
    return exprs, em.pre


def rate_forms(mech):
    """Per-cell storage of the rate constants.  Each reaction's expression is
    one of: a literal (not stored: written into the code at each use); A * e
    with e one of the shared exponentials exp(c/T) (e stored once for every
    reaction with that activation temperature, the product formed at each
    use: the same multiplication, so the same bits); anything else -- a bare
    J, a Troe fall-off, sums, T^2 forms, k46, the N2O5 hydrolysis input --
    stored whole.  Returns (prims, forms): prims[p] = C expression of stored
    value p, the 21 photolysis rate constants first (value r = reaction r,
    the index the forcing code writes); forms[r] = ("lit", c), ("mul", A, p)
    or ("prim", p)."""
    import re
    exprs, _ = rate_exprs(mech)
    nph = len(photo_j(mech))
    prims, index, forms = [], {}, []

    def prim(key, ex):
        if key not in index:
            index[key] = len(prims)
            prims.append(ex)
        return index[key]
    for r in range(nph):
        assert re.fullmatch(r"J\[\d+\]", exprs[r]), exprs[r]
        assert prim(("r", r), exprs[r]) == r
    num = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
    for r, ex in enumerate(exprs):
        if r < nph:
            forms.append(("prim", r))
        elif re.fullmatch(num, ex):
            forms.append(("lit", ex))
        else:
            mo = re.fullmatch(r"\((%s) \* (rc_s\d+)\)" % num, ex)
            if mo:
                forms.append(("mul", mo.group(1), prim(("e", mo.group(2)), mo.group(2))))
            else:
                forms.append(("prim", prim(("r", r), ex)))
    return prims, forms


def kexpr(forms, r, rd):
    """Code of rate constant r at a use; rd(p) reads stored value p."""
    f = forms[r]
    if f[0] == "lit":
        return f[1]
    if f[0] == "mul":
        return "(%s * %s)" % (f[1], rd(f[2]))
    return rd(f[1])


def gen_rconst(mech):
    exprs, pre = rate_exprs(mech)
    prims, forms = rate_forms(mech)
    out = []
    out.append("// rate constants: %d reactions, %d shared exp/pow subexpressions; %d values"
               " stored per cell (rate_forms)" % (len(mech["rates"]), len(pre), len(prims)))
    out.append("template <class KA>")
    out.append("__device__ __forceinline__ void rconst(double T, double M, const double* __restrict__ J,"
               " int64_t sJ, double RC, KA k) {")
    out.append("  (void)RC;")
    out.append("  const double zt = 300.0 / T;")
    # J slots are addressed with a stride
    for name, ex in pre:
        out.append("  const double %s = %s;" % (name, ex))
    for p, ex in enumerate(prims):
        out.append("  WR(k, %d, %s);" % (p, ex.replace("J[", "JJ(")))
    out.append("}")
    text = "\n".join(out)
    # JJ(i] -> J[i*sJ]
    import re
    text = re.sub(r"JJ\((\d+)\]", r"J[\1 * sJ]", text)
    return text, len(pre)


# --------------------------------------------------------------------------
# tendency
# --------------------------------------------------------------------------
def factor_src(kind, i):
    return "RD(v, %d)" % i if kind == "V" else ("h2o" if i == 0 else "m")


def gen_fun(mech):
    """Tendency in reaction-major (scatter) order: each rate is formed once
    and added into the accumulators of the species it moves, so only the 59
    accumulators stay live.  Every species' terms are listed in increasing
    reaction order, so each accumulator sums in exactly the table order.
    RF > 0 launders the input views every RF reactions."""
    nterms = sum(len(t) for t in mech["tendency"])
    _, forms = rate_forms(mech)
    out = ["// tendency: %d rates, %d terms (reaction-major scatter)" % (mech["nreact"], nterms)]
    out.append("template <int RF, class KA, class VA, class FA>")
    out.append("__device__ __forceinline__ void fun(KA k, VA v, double h2o, double m, FA f) {")
    by_r = [[] for _ in range(mech["nreact"])]
    for i, terms in enumerate(mech["tendency"]):
        for r, c in terms:
            by_r[r].append((i, c))
    started = set()
    for i in range(59):
        if not mech["tendency"][i]:
            out.append("  double f%d = 0.0;" % i)
            started.add(i)
    for r, facs in enumerate(mech["monomials"]):
        if r > 0:
            out.append("  if (RF > 0 && %d %% (RF > 0 ? RF : 1) == 0) { launder(k); launder(v); }" % r)
        if not by_r[r]:
            continue
        expr = kexpr(forms, r, lambda q: "RD(k, %d)" % q)
        for kind, i in facs:
            expr = "(%s * %s)" % (expr, factor_src(kind, i))
        out.append("  { const double a = %s;" % expr)
        for i, c in by_r[r]:
            cv = float(c)
            if i not in started:
                rhs = "a" if cv == 1 else ("(-a)" if cv == -1 else "(%s * a)" % lit(c))
                out.append("    double f%d = %s;" % (i, rhs) if False else "    f%d_init = %s;" % (i, rhs))
                started.add(i)
            elif cv == 1:
                out.append("    f%d = f%d + a;" % (i, i))
            elif cv == -1:
                out.append("    f%d = f%d - a;" % (i, i))
            else:
                out.append("    f%d = fma(%s, a, f%d);" % (i, lit(c), i))
        out.append("  }")
    for i in range(59):
        out.append("  WR(f, %d, f%d);" % (i, i))
    out.append("}")
    # declare accumulators up front; first contributions initialise them
    text = "\n".join(out)
    decl = "  double " + ", ".join("f%d" % i for i in range(59) if mech["tendency"][i]) + ";"
    text = text.replace("__device__ __forceinline__ void fun(KA k, VA v, double h2o, double m, FA f) {",
                        "__device__ __forceinline__ void fun(KA k, VA v, double h2o, double m, FA f) {\n" + decl, 1)
    import re
    text = re.sub(r"f(\d+)_init = ", r"f\1 = ", text)
    # species with no terms were declared separately above
    return text


# --------------------------------------------------------------------------
# Jacobian in LU layout, negated, diagonal shifted
# --------------------------------------------------------------------------
def jac_terms(mech):
    """Per LU position: list of (coef string, r, [other factors]) in the
    contract order (species i, its tendency terms then dropped terms,
    V factors of the monomial in order)."""
    lu = mech["lu"]
    pos = {}
    for row in range(59):
        for kk in range(lu["crow"][row], lu["crow"][row + 1]):
            pos[(row, lu["icol"][kk])] = kk
    terms = [[] for _ in range(659)]
    for i in range(59):
        full = list(mech["tendency"][i]) + [[r, c] for (ii, r, c) in mech["fun_dropped_terms"] if ii == i]
        for r, c in full:
            facs = mech["monomials"][r]
            for q, (kind, j) in enumerate(facs):
                if kind != "V":
                    continue
                others = [f for q2, f in enumerate(facs) if q2 != q]
                terms[pos[(i, j)]].append((c, r, others))
    return terms


def gen_jac_prep(mech):
    terms = jac_terms(mech)
    lu = mech["lu"]
    diag = set(lu["diag"])
    nt = sum(len(t) for t in terms)
    out = ["// G = -J + ghinv*I in LU layout: %d Jacobian terms over %d positions" % (nt, 659)]
    out.append("template <class KA, class VA, class GA>")
    out.append("__device__ __forceinline__ void jac_prep(KA k, VA v, double h2o, double m,"
               " double ghinv, GA g) {")
    _, forms = rate_forms(mech)
    for p in range(659):
        acc = None
        for c, r, others in terms[p]:
            t = kexpr(forms, r, lambda q: "RD(k, %d)" % q)
            for kind, j in others:
                t = "(%s * %s)" % (t, factor_src(kind, j))
            cv = float(c)
            if acc is None:
                acc = ("(-%s)" % t) if cv == 1 else (t if cv == -1 else "(%s * %s)" % (lit(neg(c)), t))
            elif cv == 1:
                acc = "(%s - %s)" % (acc, t)
            elif cv == -1:
                acc = "(%s + %s)" % (acc, t)
            else:
                acc = "fma(%s, %s, %s)" % (lit(neg(c)), t, acc)
        if p in diag:
            acc = "(%s + ghinv)" % acc if acc else "ghinv"
        out.append("  WR(g, %d, %s);" % (p, acc if acc else "0.0"))
    out.append("}")
    return "\n".join(out)


# --------------------------------------------------------------------------
# LU decomposition and solve
# --------------------------------------------------------------------------
def _entry_expr(terms_p, forms):
    """-(sum of Jacobian terms) for one LU entry, in the contract order."""
    acc = None
    for c, r, others in terms_p:
        t = kexpr(forms, r, lambda q: "RD(k, %d)" % q)
        for kind, j in others:
            t = "(%s * %s)" % (t, factor_src(kind, j))
        cv = float(c)
        if acc is None:
            acc = ("(-%s)" % t) if cv == 1 else (t if cv == -1 else "(%s * %s)" % (lit(neg(c)), t))
        elif cv == 1:
            acc = "(%s - %s)" % (acc, t)
        elif cv == -1:
            acc = "(%s + %s)" % (acc, t)
        else:
            acc = "fma(%s, %s, %s)" % (lit(neg(c)), t, acc)
    return acc


def gen_decomp(mech, fused):
    """Row-wise (IKJ) no-pivot LU in the reference's 659-position pattern.

    fused = False: decomp(g) factors a prepared matrix in place.
    fused = True:  jac_decomp(k, v, h2o, m, ghinv, g) forms each row of
                   G = -J + ghinv*I right before eliminating it, so only one
                   row (at most 47 entries) is live at a time.
    Both store 1/U_kk on the diagonal and return 0, or k+1 for the first row
    whose prepared diagonal is 0 (the reference's check, mgpu.cu:2753).
    R > 0 launders the views every R rows (earlier rows are re-read from
    memory instead of held in registers)."""
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    terms = jac_terms(mech) if fused else None
    _, forms = rate_forms(mech)
    out = []
    if fused:
        out.append("template <int R, class KA, class VA, class GA>")
        out.append("__device__ __forceinline__ int jac_decomp(KA k, VA v, double h2o, double m,"
                   " double ghinv, GA g) {")
    else:
        out.append("template <int R, class GA>")
        out.append("__device__ __forceinline__ int decomp(GA g) {")
    out.append("  int sing = 0;")
    nfma = 0
    for k in range(59):
        cols = icol[crow[k]:crow[k + 1]]
        if k > 0:
            extra = " launder(k); launder(v);" if fused else ""
            out.append("  if (R > 0 && %d %% (R > 0 ? R : 1) == 0) { launder(g);%s }" % (k, extra))
        out.append("  { // row %d" % k)
        for kk in range(crow[k], crow[k + 1]):
            c = icol[kk]
            if fused:
                e = _entry_expr(terms[kk], forms)
                if kk == diag[k]:
                    e = "(%s + ghinv)" % e if e else "ghinv"
                out.append("    double w%d = %s;" % (c, e if e else "0.0"))
            else:
                out.append("    double w%d = RD(g, %d);" % (c, kk))
        out.append("    sing = (sing == 0 && w%d == 0.0) ? %d : sing;" % (k, k + 1))
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            out.append("    w%d = w%d * RD(g, %d);" % (j, j, diag[j]))
            for jj in range(diag[j] + 1, crow[j + 1]):
                c = icol[jj]
                assert c in cols
                out.append("    w%d = fma(-w%d, RD(g, %d), w%d);" % (c, j, jj, c))
                nfma += 1
        for kk in range(crow[k], crow[k + 1]):
            c = icol[kk]
            if kk == diag[k]:
                out.append("    WR(g, %d, 1.0 / w%d);" % (kk, c))
            else:
                out.append("    WR(g, %d, w%d);" % (kk, c))
        out.append("  }")
    out.append("  return sing;")
    out.append("}")
    return "\n".join(out), nfma


def gen_solve(mech):
    """x <- G^-1 x: forward (unit L) then backward (U, inverse pivots) with
    the 59 entries of x held in registers and the LU streamed once."""
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    out = ["// x <- G^-1 x with unit-lower L and U (inverse pivots on the diagonal)"]
    out.append("template <class GA, class XA>")
    out.append("__device__ __forceinline__ void solve(GA g, XA x) {")
    for i in range(59):
        out.append("  double x%d = RD(x, %d);" % (i, i))
    for i in range(59):
        if diag[i] > crow[i]:
            acc = "x%d" % i
            for kk in range(crow[i], diag[i]):
                acc = "fma(-RD(g, %d), x%d, %s)" % (kk, icol[kk], acc)
            out.append("  x%d = %s;" % (i, acc))
    for i in range(58, -1, -1):
        acc = "x%d" % i
        for kk in range(diag[i] + 1, crow[i + 1]):
            acc = "fma(-RD(g, %d), x%d, %s)" % (kk, icol[kk], acc)
        out.append("  x%d = %s * RD(g, %d);" % (i, acc, diag[i]))
    for i in range(59):
        out.append("  WR(x, %d, x%d);" % (i, i))
    out.append("}")
    return "\n".join(out)


def gen_solve_pf(mech, G):
    """solve() with the LU streamed in groups of G entries, each group's
    loads issued one group ahead of the multiply-adds that consume them
    (scheduling fences keep the compiler from sinking the loads back to
    their uses).  Same operations in the same order as solve(): a nested
    fma(a, b, fma(c, d, x)) is written as two accumulating statements."""
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    ops = []          # (lu entry, statement using lu_<entry>)
    for i in range(59):
        for kk in range(crow[i], diag[i]):
            ops.append((kk, "x%d = fma(-V(lu%d), x%d, x%d);" % (i, kk, icol[kk], i)))
    for i in range(58, -1, -1):
        for kk in range(diag[i] + 1, crow[i + 1]):
            ops.append((kk, "x%d = fma(-V(lu%d), x%d, x%d);" % (i, kk, icol[kk], i)))
        ops.append((diag[i], "x%d = x%d * V(lu%d);" % (i, i, diag[i])))
    assert sorted(e for e, _ in ops) == list(range(lu["nnz"]))
    groups = [ops[a:a + G] for a in range(0, len(ops), G)]
    fence = "  __builtin_amdgcn_sched_barrier(0);"
    out = ["// solve() with the LU streamed %d entries at a time, one group ahead" % G]
    out.append("template <class GA, class XA>")
    out.append("__device__ __forceinline__ void solve_pf%d(GA g, XA x) {" % G)
    for i in range(59):
        out.append("  double x%d = RD(x, %d);" % (i, i))

    def loads(grp):
        return ["  const auto lu%d = RDP<%d>(g);" % (e, e) for e, _ in grp]
    out += loads(groups[0])
    for m, grp in enumerate(groups):
        if m + 1 < len(groups):
            out += loads(groups[m + 1])
        out.append(fence)
        out += ["  " + st for _, st in grp]
        out.append(fence)
    for i in range(59):
        out.append("  WR(x, %d, x%d);" % (i, i))
    out.append("}")
    return "\n".join(out)


def recomputed_l(mech, S=160, R=112, P=0, L=1, maxmod=0, maxterms=1 << 30):
    """L positions (row k, column j < k) that the solves re-form instead of
    reading back, with the decomposition's own operations: L(k, j) =
    (Jacobian expression - sum over the row's earlier eliminations i that
    reach column j of L(k, i) U(i, j), in order) times row j's inverse
    pivot.  Eligible: at most `maxmod` such eliminations, at most
    `maxterms` Jacobian terms, the entry's rate constants, concentrations,
    U(i, j) and inverse pivot all on chip under hot set (S, R, P, L), and
    no photolysis rate constant among them: a non-autonomous integration
    moves those between the decomposition and the later stages' solves,
    while the other rate constants and the state stay fixed through an
    attempt.  The L(k, i) are at hand, formed or loaded by the same row's
    forward substitution just before."""
    import re
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    terms = jac_terms(mech)
    _, forms = rate_forms(mech)
    code = hot_codes(mech, S, R, P, L)
    pos = lu_positions(mech)
    nph = len(photo_j(mech))

    def onchip(a, i):
        return code[a][i] > HBM_BASE
    out = set()
    for k in range(59):
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            mods = l_updates(mech, k, j, pos)
            e = _entry_expr(terms[kk], forms) or "0.0"
            if len(mods) > maxmod or len(terms[kk]) > maxterms:
                continue
            refs = re.findall(r"RD\((k|v), (\d+)\)", e)
            if any(a == "k" and int(i) < nph for a, i in refs):
                continue
            if (all(onchip(0 if a == "k" else 1, int(i)) for a, i in refs) and onchip(2, diag[j])
                    and all(onchip(2, pos[(i, j)]) for i in mods)):
                out.add(kk)
    return frozenset(out)


def recomputed_u(mech, S=160, R=112, P=0, L=1, maxterms=1 << 30):
    """Off-diagonal U positions (row k, column c > k) left in the workspace
    under hot set (S, R, P, L) that no elimination of their row updates:
    U(k, c) is then the Jacobian expression itself.  Eligible when that
    expression has at most `maxterms` terms, reads only on-chip rate
    constants and concentrations and no photolysis rate constant (which the
    forcing moves within an attempt): the decomposition does not store
    these, and the later rows that eliminate with row k and the backward
    substitutions form them at their use."""
    import re
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    terms = jac_terms(mech)
    _, forms = rate_forms(mech)
    code = hot_codes(mech, S, R, P, L)
    pos = lu_positions(mech)
    nph = len(photo_j(mech))

    def onchip(a, i):
        return code[a][i] > HBM_BASE
    out = set()
    for k in range(59):
        lcols = [icol[q] for q in range(crow[k], diag[k])]
        for kk in range(diag[k] + 1, crow[k + 1]):
            c = icol[kk]
            if onchip(2, kk) or len(terms[kk]) > maxterms:
                continue
            if any(pos.get((i, c), -1) > diag[i] for i in lcols):
                continue
            e = _entry_expr(terms[kk], forms)
            if e is None:
                continue
            refs = re.findall(r"RD\((k|v), (\d+)\)", e)
            if any(a == "k" and int(i) < nph for a, i in refs):
                continue
            if all(onchip(0 if a == "k" else 1, int(i)) for a, i in refs):
                out.add(kk)
    return frozenset(out)


def lu_positions(mech):
    """(row, column) -> LU position."""
    lu = mech["lu"]
    return {(r, lu["icol"][kk]): kk for r in range(59) for kk in range(lu["crow"][r], lu["crow"][r + 1])}


def l_updates(mech, k, j, pos):
    """The earlier eliminations of row k that update its column j < k: the
    columns i < j of row k's L part whose pivot row has U(i, j), in the
    decomposition's order."""
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    return [icol[q] for q in range(crow[k], diag[k])
            if icol[q] < j and pos.get((icol[q], j), -1) > diag[icol[q]]]


def gen_solve_rc_pf_reg(mech, G, rc, name, formed_u=frozenset()):
    """solve_pf<G>_reg() for an LU whose positions `rc` were not stored: each
    such L(k, j) is formed at its use as the decomposition formed it
    (recomputed_l()); every other entry is read as before.  The same
    operations in the same order; the LU loads are issued G at a time ahead
    of their use, the on-chip inputs of the formed entries at their use
    (schedule_loads)."""
    import re
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    terms = jac_terms(mech)
    _, forms = rate_forms(mech)
    pos = lu_positions(mech)
    loads, ops = [], []

    def load(nm, src):
        loads.append((nm, src, None))
        return len(loads) - 1
    for i in range(59):
        lval = {}          # column -> (name of L(i, column), its load index or None)
        for kk in range(crow[i], diag[i]):
            j = icol[kk]
            if kk in rc:
                used = []
                mine = {}

                def rd(nm, src, used=used, mine=mine):
                    if nm not in mine:
                        mine[nm] = load(nm, src)
                        used.append(mine[nm])
                    return nm

                def sub(mo, kk=kk, rd=rd):
                    return rd("%s%s_e%d" % (mo.group(1), mo.group(2), kk),
                              "RD(%s, %s)" % (mo.group(1), mo.group(2)))
                # a fill-in position starts from 0.0, as in the decomposition
                e = re.sub(r"RD\((k|v), (\d+)\)", sub, _entry_expr(terms[kk], forms) or "0.0")
                for c in l_updates(mech, i, j, pos):
                    ln, li = lval[c]
                    if li is not None:
                        used.append(li)
                    u = rd("u%d_e%d" % (pos[(c, j)], kk), "RD(g, %d)" % pos[(c, j)])
                    e = "fma(-%s, %s, %s)" % (ln, u, e)
                p = rd("p%d_e%d" % (diag[j], kk), "RD(g, %d)" % diag[j])
                ops.append(("const double l%d = %s * %s; x[%d] = fma(-l%d, x[%d], x[%d]);"
                            % (kk, e, p, i, kk, j, i), used, None))
                lval[j] = ("l%d" % kk, None)
            else:
                li = load("lu%d" % kk, "RD(g, %d)" % kk)
                ops.append(("x[%d] = fma(-lu%d, x[%d], x[%d]);" % (i, kk, j, i), [li], None))
                lval[j] = ("lu%d" % kk, li)
    for i in range(58, -1, -1):
        for kk in range(diag[i] + 1, crow[i + 1]):
            if kk in formed_u:
                used = []
                mine = {}

                def fsub(mo, used=used, mine=mine, kk=kk):
                    nm = "f%s%s_u%d" % (mo.group(1), mo.group(2), kk)
                    if nm not in mine:
                        mine[nm] = load(nm, "RD(%s, %s)" % (mo.group(1), mo.group(2)))
                        used.append(mine[nm])
                    return nm
                eu = re.sub(r"RD\((k|v), (\d+)\)", fsub, _entry_expr(terms[kk], forms))
                ops.append(("x[%d] = fma(-(%s), x[%d], x[%d]);" % (i, eu, icol[kk], i), used, None))
                continue
            ops.append(("x[%d] = fma(-lu%d, x[%d], x[%d]);" % (i, kk, icol[kk], i),
                        [load("lu%d" % kk, "RD(g, %d)" % kk)], None))
        ops.append(("x[%d] = x[%d] * lu%d;" % (i, i, diag[i]),
                    [load("lu%d" % diag[i], "RD(g, %d)" % diag[i])], None))
    body, _ = schedule_loads(ops, loads, G, lambda nm: nm.startswith("lu"))
    out = ["// solve_pf%d_reg() with %d L entries formed at their use (recomputed_l())"
           % (G, len(rc)),
           "template <class KA, class VA, class GA>",
           "__device__ __forceinline__ void %s(KA k, VA v, double h2o, double m, GA g,"
           " double (&x)[59]) {" % name]
    out += body
    out.append("}")
    return "\n".join(out)


def deferred_src(src):
    """RD(a, i) -> RDP<i>(a): a read issued with its batch, except for an
    entry in a register tier split over a cell's lanes, whose cross-lane
    move RDP defers to the use (V(name)), so that it holds no register
    between the batch and the use."""
    import re
    mo = re.fullmatch(r"RD\((\w+), (\d+)\)", src)
    assert mo, src
    return "RDP<%s>(%s)" % (mo.group(2), mo.group(1))


def wrap_uses(stmt, names):
    """Each use of a scheduled load's name in `stmt` as V(name)."""
    import re
    for nm in sorted(set(names), key=len, reverse=True):
        stmt = re.sub(r"\b%s\b" % re.escape(nm), "V(%s)" % nm, stmt)
    return stmt


def schedule_loads(ops, loads, G, batched_if=None, frontier_rule=True):
    """Body lines for `ops` (statement, [load indices], stored entry or None)
    with the `loads` (name, source, producing entry or None) issued in
    batches of G behind scheduling fences, so that between G and 2G loads
    run ahead of their first use; a load waits for the op storing its
    entry.  batched_if: a predicate on a load's name; the loads it rejects
    are emitted right before the first op that uses them, inside that op's
    scheduling region; with frontier_rule an op without batched loads also
    tops the look-ahead up.  Returns (lines, number of empty batches)."""
    ops = [(wrap_uses(st, [loads[u][0] for u in used]), used, sto) for st, used, sto in ops]
    store_op = {st: i for i, (_, _, st) in enumerate(ops) if st is not None}
    fence = "  __builtin_amdgcn_sched_barrier(0);"
    batched = [u for u in range(len(loads)) if batched_if is None or batched_if(loads[u][0])]
    pos = {u: n for n, u in enumerate(batched)}
    out = []
    issued = 0
    stalls = 0
    defined = set()
    frontier = -1
    for i, (st, used, _) in enumerate(ops):
        mine = [pos[u] for u in used if u in pos]
        if batched_if is None or not frontier_rule:
            nm = max(mine) if mine else -1
        else:
            nm = frontier = max(mine + [frontier])
        while issued <= nm or (issued - nm - 1 < G and issued < len(batched)):
            batch = []
            while len(batch) < G and issued < len(batched):
                name, src, dep = loads[batched[issued]]
                if dep is not None and store_op[dep] >= i:
                    break            # its producer is not emitted yet
                batch.append("  const auto %s = %s;" % (name, deferred_src(src)))
                issued += 1
            if not batch:
                assert issued > nm
                stalls += 1
                break
            out.append(fence)
            out += batch
            out.append(fence)
        for u in used:
            if u not in pos and u not in defined:
                defined.add(u)
                out.append("  const auto %s = %s;" % (loads[u][0], deferred_src(loads[u][1])))
        out.append("  " + st)
    assert issued == len(batched)
    return out, stalls


def gen_fun_pf(mech, G, window=16):
    """fun() with its loads scheduled like jac_decomp_pf(): the same
    statements in the same order; stored rate-constant values are loaded
    once each, concentrations once per window of reactions."""
    _, forms = rate_forms(mech)
    by_r = [[] for _ in range(mech["nreact"])]
    for i, terms in enumerate(mech["tendency"]):
        for r, c in terms:
            by_r[r].append((i, c))
    loads, load_ix, ops = [], {}, []

    def need(name, src):
        if name not in load_ix:
            load_ix[name] = len(loads)
            loads.append((name, src, None))
        return load_ix[name]
    started = set(i for i in range(59) if not mech["tendency"][i])
    for r, facs in enumerate(mech["monomials"]):
        if not by_r[r]:
            continue
        used = []

        def rdk(q, used=used):
            used.append(need("k%d" % q, "RD(k, %d)" % q))
            return "k%d" % q
        expr = kexpr(forms, r, rdk)
        for kind, i in facs:
            if kind == "V":
                nm = "v%d_%d" % (i, r // window)
                used.append(need(nm, "RD(v, %d)" % i))
                expr = "(%s * %s)" % (expr, nm)
            else:
                expr = "(%s * %s)" % (expr, factor_src(kind, i))
        st = ["{ const double a = %s;" % expr]
        for i, c in by_r[r]:
            cv = float(c)
            if i not in started:
                rhs = "a" if cv == 1 else ("(-a)" if cv == -1 else "(%s * a)" % lit(c))
                st.append("f%d = %s;" % (i, rhs))
                started.add(i)
            elif cv == 1:
                st.append("f%d = f%d + a;" % (i, i))
            elif cv == -1:
                st.append("f%d = f%d - a;" % (i, i))
            else:
                st.append("f%d = fma(%s, a, f%d);" % (i, lit(c), i))
        ops.append((" ".join(st) + " }", used, None))
    out = ["// fun() with loads issued %d at a time ahead of their use" % G]
    out.append("template <class KA, class VA, class FA>")
    out.append("__device__ __forceinline__ void fun_pf%d(KA k, VA v, double h2o, double m, FA f) {" % G)
    out.append("  double " + ", ".join("f%d" % i for i in range(59) if mech["tendency"][i]) + ";")
    for i in range(59):
        if not mech["tendency"][i]:
            out.append("  double f%d = 0.0;" % i)
    body, _ = schedule_loads(ops, loads, G)
    out += body
    for i in range(59):
        out.append("  WR(f, %d, f%d);" % (i, i))
    out.append("}")
    return "\n".join(out)


def gen_solve_pf_reg(mech, G):
    """solve_pf<G>() on a register-resident right-hand side: x is a
    59-entry local array of the caller (indexed by constants only), solved
    in place; the LU is streamed as in solve_pf<G>()."""
    text = gen_solve_pf(mech, G)
    lines = text.split("\n")
    out = ["// solve_pf%d() with x held by the caller (registers), solved in place" % G,
           "template <class GA>",
           "__device__ __forceinline__ void solve_pf%d_reg(GA g, double (&x)[59]) {" % G]
    import re
    for ln in lines[3:]:
        if re.match(r"\s*double x\d+ = RD\(x, \d+\);", ln) or re.match(r"\s*WR\(x, \d+, x\d+\);", ln):
            continue
        out.append(re.sub(r"\bx(\d+)\b", r"x[\1]", ln))
    return "\n".join(out)


def gen_fun_pf_reg(mech, G, window=16):
    """fun_pf<G>() leaving the tendency in a 59-entry local array of the
    caller (registers) instead of storing it."""
    import re
    text = gen_fun_pf(mech, G, window)
    lines = text.split("\n")
    out = ["// fun_pf%d() with the tendency left in the caller's registers" % G,
           "template <class KA, class VA>",
           "__device__ __forceinline__ void fun_pf%d_reg(KA k, VA v, double h2o, double m,"
           " double (&f)[59]) {" % G]
    for ln in lines[3:]:
        if re.match(r"\s*double f\d+(, f\d+)*;", ln) or re.match(r"\s*WR\(f, \d+, f\d+\);", ln):
            continue
        ln = re.sub(r"\s*double f(\d+) = 0\.0;", r"  f[\1] = 0.0;", ln)
        out.append(re.sub(r"\bf(\d+)\b", r"f[\1]", ln))
    return "\n".join(out)


def gen_jac_decomp_pf(mech, G, skip=frozenset(), name=None, formed_u=frozenset()):
    """jac_decomp() with every load scheduled explicitly: the same
    operations in the same order, but rate constants, concentrations and
    earlier LU rows are loaded in batches of G, issued so that between G and
    2G loads run ahead of their first use (scheduling fences keep the
    compiler from sinking them).  Within a row each rate constant and
    concentration is loaded once.  A load of an LU entry is never issued
    before the store that produces it.  skip: LU positions not stored (the
    solves form them, recomputed_l(), recomputed_u()); formed_u: U
    positions whose reads by later rows are formed from their Jacobian
    expression (recomputed_u()), their inputs read at the use."""
    import re
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    terms = jac_terms(mech)
    _, forms = rate_forms(mech)
    loads = []       # (name, source text, producing LU entry or None)
    load_ix = {}
    ops = []         # (statement, [load indices], stored LU entry or None)

    def need(name, src, dep):
        if name not in load_ix:
            load_ix[name] = len(loads)
            loads.append((name, src, dep))
        return load_ix[name]

    for k in range(59):
        for kk in range(crow[k], crow[k + 1]):
            c = icol[kk]
            e = _entry_expr(terms[kk], forms)
            used = []
            if e:
                def sub(mo, used=used):
                    arr, i = mo.group(1), int(mo.group(2))
                    nm = "%s%d_%d" % (arr, i, k)
                    used.append(need(nm, "RD(%s, %d)" % (arr, i), None))
                    return nm
                e = re.sub(r"RD\((k|v), (\d+)\)", sub, e)
            if kk == diag[k]:
                e = "(%s + ghinv)" % e if e else "ghinv"
            ops.append(("double w%d_%d = %s;" % (c, k, e if e else "0.0"), used, None))
        ops.append(("sing = (sing == 0 && w%d_%d == 0.0) ? %d : sing;" % (k, k, k + 1), [], None))
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            li = need("u%d_%d" % (diag[j], k), "RD(g, %d)" % diag[j], diag[j])
            ops.append(("w%d_%d = w%d_%d * u%d_%d;" % (j, k, j, k, diag[j], k), [li], None))
            for jj in range(diag[j] + 1, crow[j + 1]):
                c = icol[jj]
                if jj in formed_u:
                    used = []

                    def fsub(mo, used=used, jj=jj, k=k):
                        nm = "f%s%s_u%d_%d" % (mo.group(1), mo.group(2), jj, k)
                        used.append(need(nm, "RD(%s, %s)" % (mo.group(1), mo.group(2)), None))
                        return nm
                    eu = re.sub(r"RD\((k|v), (\d+)\)", fsub, _entry_expr(terms[jj], forms))
                    ops.append(("w%d_%d = fma(-w%d_%d, %s, w%d_%d);" % (c, k, j, k, eu, c, k),
                                used, None))
                    continue
                li = need("u%d_%d" % (jj, k), "RD(g, %d)" % jj, jj)
                ops.append(("w%d_%d = fma(-w%d_%d, u%d_%d, w%d_%d);" % (c, k, j, k, jj, k, c, k),
                            [li], None))
        for kk in range(crow[k], crow[k + 1]):
            if kk in skip:
                continue
            c = icol[kk]
            v = ("1.0 / w%d_%d" % (c, k)) if kk == diag[k] else ("w%d_%d" % (c, k))
            ops.append(("WR(g, %d, %s);" % (kk, v), [], kk))
    out = ["// jac_decomp() with loads issued %d at a time ahead of their use" % G]
    out.append("template <class KA, class VA, class GA>")
    out.append("__device__ __forceinline__ int %s(KA k, VA v, double h2o, double m,"
               " double ghinv, GA g) {" % (name or "jac_decomp_pf%d" % G))
    out.append("  int sing = 0;")
    body, stalls = schedule_loads(ops, loads, G,
                                  (lambda nm: not nm.startswith("f")) if formed_u else None,
                                  frontier_rule=False)
    out += body
    out.append("  return sing;")
    out.append("}")
    return "\n".join(out), len(loads), stalls


def hot_ranking(mech):
    """(array, index) entries of the rate constants (0), state (1) and LU (2)
    ranked by memory accesses per Rosenbrock attempt of the load-scheduled
    kernels, most accessed first: an LU entry is stored once and read by
    the decomposition of every later row that eliminates with it and by the
    three solves; rate constants and state entries count the loads of the
    Jacobian/decomposition (once per row) and of the tendency (the state
    once; rate constants twice, weighted 3 as measured best), plus the
    state's three element-wise passes.  Ties keep array order."""
    import collections
    import re
    lu = mech["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    acc = collections.Counter()
    for q in range(lu["nnz"]):
        acc[(2, q)] += 4
    for k in range(59):
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            acc[(2, diag[j])] += 1
            for jj in range(diag[j] + 1, crow[j + 1]):
                acc[(2, jj)] += 1
    text = gen_jac_decomp_pf(mech, 16)[0]
    for mo in re.finditer(r"RDP<(\d+)>\((k|v)\)", text):
        acc[(0 if mo.group(2) == "k" else 1, int(mo.group(1)))] += 1
    for mo in re.finditer(r"RDP<(\d+)>\((k|v)\)", gen_fun_pf(mech, 16)):
        if mo.group(2) == "k":
            # two tendencies per attempt; weighted 3, the measured best
            # (2: 770.7 ms, 3: 758.9 ms, 4: 773.2 ms at 1M cells)
            acc[(0, int(mo.group(1)))] += 3
        else:
            acc[(1, int(mo.group(1)))] += 1
    for i in range(59):
        acc[(1, i)] += 3
    keys = [(0, r) for r in range(len(rate_forms(mech)[0]))] + [(1, i) for i in range(59)] + \
        [(2, q) for q in range(lu["nnz"])]
    return sorted(keys, key=lambda key: -acc[key]), acc


# Hot-set configurations (S, R, P, L): S LDS slots per cell, R registers per
# lane holding the same entries on every lane of a cell, P registers per lane
# of a cell split over its L lanes (register r of lane q of the cell holds
# split entry r*L + q, so a cell holds P*L entries there); L = 1: one cell
# per lane
HOT_BUDGETS = ((160, 0, 0, 1), (312, 0, 0, 1), (160, 112, 0, 1), (160, 0, 128, 2),
               (320, 0, 128, 4))
# code of a split-tier entry: SPLIT_BASE - (its index in the split tier);
# of a workspace entry: HBM_BASE - (its rank among the workspace entries,
# 0 the most accessed)
SPLIT_BASE = -(1 << 14)
HBM_BASE = -(1 << 13)


def hot_name(S, R, P, L):
    return "HOT%d_%d_%dx%d" % (S, R, P, L)


def hot_codes(mech, S, R, P, L):
    """Per array (rate constants, state, LU), the code of each entry under
    configuration (S, R, P, L): s >= 0 for LDS slot s (the first S of the
    ranking), -2 - r for replicated register r (the next R), SPLIT_BASE - j
    for split entry j (the next P*L; lane j % L of the cell, register
    j // L), HBM_BASE - h for the h-th most accessed of the entries left in
    the HBM workspace."""
    ranked, _ = hot_ranking(mech)
    sizes = (len(rate_forms(mech)[0]), 59, mech["lu"]["nnz"])
    code = [[None] * n for n in sizes]
    for s_, (a, i) in enumerate(ranked[:S]):
        code[a][i] = s_
    for r, (a, i) in enumerate(ranked[S:S + R]):
        code[a][i] = -2 - r
    for j, (a, i) in enumerate(ranked[S + R:S + R + P * L]):
        code[a][i] = SPLIT_BASE - j
    for h, (a, i) in enumerate(ranked[S + R + P * L:]):
        code[a][i] = HBM_BASE - h
    assert all(c is not None for arr in code for c in arr)
    return code


def gen_hot_tables(mech):
    """The code of every rate constant, state and LU entry per hot-set
    configuration (hot_codes), and hot_code<S, R, P, L>(array, index)."""
    sizes = (len(rate_forms(mech)[0]), 59, mech["lu"]["nnz"])
    out = ["// hot set per (LDS slots S, replicated registers R, split registers P, lanes per"
           " cell L): code per entry (s >= 0 LDS slot s, -2 - r replicated register r,"
           " %d - j split entry j, %d - h the h-th most accessed HBM entry)"
           % (SPLIT_BASE, HBM_BASE)]
    out.append("constexpr int HOT_SPLIT_BASE = %d;" % SPLIT_BASE)
    out.append("constexpr int HOT_HBM_BASE = %d;" % HBM_BASE)
    for S, R, P, L in HOT_BUDGETS:
        code = hot_codes(mech, S, R, P, L)
        for a, name in enumerate(("K", "Y", "G")):
            out.append("__device__ constexpr int16_t %s_%s[%d] = {%s};"
                       % (hot_name(S, R, P, L), name, sizes[a], ",".join(map(str, code[a]))))
    out.append("template <int S, int R, int P, int L>")
    out.append("__device__ __forceinline__ constexpr int hot_code(int arr, int i) {")
    for S, R, P, L in HOT_BUDGETS:
        n = hot_name(S, R, P, L)
        out.append("  if constexpr (S == %d && R == %d && P == %d && L == %d)" % (S, R, P, L))
        out.append("    return arr == 0 ? %s_K[i] : arr == 1 ? %s_Y[i] : %s_G[i];" % (n, n, n))
    out.append("  return -1;")
    out.append("}")
    return "\n".join(out)


def gen_wrf_tables(mech):
    """dev_convert's species shuffle (mgpu.cu:585-645) and unit constants
    (mgpu.cu:30-44): WRF source array (0 chemistry, 1 radicals) and field
    per variable species."""
    w = mech["wrf"]
    c = w["constants"]
    src = [a for a, _ in w["var_source"]]
    fld = [f for _, f in w["var_source"]]
    return "\n".join([
        "// WRF-Chem input shuffle: source array (0 chemistry, 1 radicals) and field per species",
        "constexpr int WRF_SRC[59] = {%s};" % ",".join(map(str, src)),
        "constexpr int WRF_FIELD[59] = {%s};" % ",".join(map(str, fld)),
        "constexpr double WRF_NAVGDRO = %r;" % c["navgdro"],
        "constexpr double WRF_MWAIR = %r;" % c["mwair"],
        "constexpr double WRF_MWH2O = 2 * %r + %r;" % (c["mwh"], c["mwo"]),
    ])


# LU entries the decomposition and the solves form instead of storing and
# reading back (recomputed_l(), recomputed_u()): none that earlier
# eliminations of its row update, at most RC_MAXTERMS Jacobian terms (the
# _rc pair, which the family kernel runs) or RC_MAXTERMS_WIDE (the _rcw
# pair, which the Ros3 kernels run).  Measured at 1M cells, default kernel: L only, updates
# allowed 0 / 1 / 2 / 8: 627.7 / 647 / 655 / 660 ms (all terms); terms at
# most 1 / 2 / any: 623.5 / 621.8 / 627.7 (no updates); none formed: 647.
# With the U entries and the current build: terms at most 2 / 3 / 5:
# 540.7 / 537.1 / 544.4 ms; one update allowed: 559.  The family kernel is
# faster with 2 (2-5%); so was the one-shot grid (775 against 798 ms) until
# the load batches of 20, with which it gains from 3 (751.9 -> 738.9 ms).
RC_MAXMOD = 0
RC_MAXTERMS = 2
RC_MAXTERMS_WIDE = 3
# Load batch of the _rcw pair and of the tendency every kernel runs in
# registers (fun_pf20_reg), with the default placement's register tier at
# 112 doubles: measured at 1M cells, default kernel, batches 16 / 18 / 20 /
# 22 / 24 with 112 registers 532 (16 at 128) / 525 / 519 / 531 / 526 ms; 20
# with 104 / 108 / 116 / 120 registers 526 / 526 / 534 / 548; 20 with one
# of the decomposition, the solves or the tendency at 16 or 24: 523-535;
# 24 with 96 registers 522, 32 with 64 registers 525 (all bit-identical).
# The family kernel: the tendency at 20 against 16, Ros4, Rodas3 and Rodas4
# 2.3% faster, Ros2 1.2% and forced Ros3 0.4% slower; 20 in the general
# stage loop only was slower on every method.  The _rc pair (the one-shot
# grid and the family kernel) at 20 against 16: the grid 758.5 -> 751.5 ms,
# Ros2, Ros4, Rodas3, Rodas4 and forced Ros3 2.0-4.7% faster.
RC_BATCH_WIDE = 20
RC_BATCH = 20
FUN_REG_BATCH = 20


def generate(dst=None, mech_path=None):
    mech = load_mech(mech_path)
    dst = dst or os.path.join(PKG, "csrc", "generated", "radm2_gen.h")
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    rc, nshared = gen_rconst(mech)
    rc_l = recomputed_l(mech, maxmod=RC_MAXMOD, maxterms=RC_MAXTERMS)
    rc_u = recomputed_u(mech, maxterms=RC_MAXTERMS)
    rcw_l = recomputed_l(mech, maxmod=RC_MAXMOD, maxterms=RC_MAXTERMS_WIDE)
    rcw_u = recomputed_u(mech, maxterms=RC_MAXTERMS_WIDE)
    dec, nfma = gen_decomp(mech, False)
    jdec, _ = gen_decomp(mech, True)
    parts = [
        "// GENERATED by codegen/gen_radm2.py from mech/radm2.json -- do not edit.",
        "#pragma once",
        "#include <hip/hip_runtime.h>",
        "#include <stdint.h>",
        "namespace radm2 {",
        "namespace gen {",
        "constexpr int NVAR = %d;" % mech["nvar"],
        "constexpr int NREACT = %d;" % mech["nreact"],
        "// rate-constant values stored per cell (rate_forms): the working array k",
        "constexpr int NKSTORE = %d;" % len(rate_forms(mech)[0]),
        "constexpr int LU_NNZ = %d;" % mech["lu"]["nnz"],
        "constexpr int NJ = 21;",
        "// photolysis: reaction r < NPHOTO has rate constant J[PHOTO_J[r]]",
        "constexpr int NPHOTO = %d;" % len(photo_j(mech)),
        "constexpr int PHOTO_J[%d] = {%s};" % (len(photo_j(mech)), ",".join(map(str, photo_j(mech)))),
        "constexpr int DECOMP_FMA = %d;" % nfma,
        "constexpr int LU_DIAG[%d] = {%s};" % (mech["nvar"], ",".join(map(str, mech["lu"]["diag"]))),
        "",
        "// Troe fall-off with Fc = 0.6 (Stockwell et al. 1997), powers of 300/T precomputed",
        "__device__ __forceinline__ double radm2_troe(double k0_300, double ztn, double kinf_300,"
        " double ztm, double M) {",
        "  const double k0 = k0_300 * ztn * M;",
        "  const double kinf = kinf_300 * ztm;",
        "  const double kr = k0 / kinf;",
        "  const double lg = log10(kr);",
        "  return k0 / (1.0 + kr) * pow(0.6, 1.0 / (1.0 + lg * lg));",
        "}",
        "// HNO3 + OH three-channel rate: k0 + k3/(1 + k3/k2)",
        "__device__ __forceinline__ double radm2_k46(double T, double M, double e785,"
        " double e1440, double e725) {",
        "  (void)T;",
        "  const double k0 = 7.2e-15 * e785;",
        "  const double k2 = 4.1e-16 * e1440;",
        "  const double k3 = 1.9e-33 * e725 * M;",
        "  return k0 + k3 / (1.0 + k3 / k2);",
        "}",
        "",
        rc,
        "",
        gen_fun(mech),
        "",
        gen_jac_prep(mech),
        "",
        dec,
        "",
        jdec,
        "",
        gen_solve(mech),
        "",
        gen_solve_pf(mech, 8),
        "",
        gen_solve_pf(mech, 16),
        "",
        gen_jac_decomp_pf(mech, 16)[0],
        "",
        gen_fun_pf(mech, 16),
        "",
        gen_solve_pf_reg(mech, 16),
        "",
        gen_fun_pf_reg(mech, FUN_REG_BATCH),
        "",
        "// L and U entries formed at their use instead of stored and read back",
        "constexpr int RECOMPUTED_L = %d;" % len(rc_l),
        "constexpr int RECOMPUTED_U = %d;" % len(rc_u),
        gen_jac_decomp_pf(mech, RC_BATCH, rc_l | rc_u, "jac_decomp_rc_pf%d" % RC_BATCH, rc_u)[0],
        "",
        gen_solve_rc_pf_reg(mech, RC_BATCH, rc_l, "solve_rc_pf%d_reg" % RC_BATCH, rc_u),
        "",
        "constexpr int RECOMPUTED_L_WIDE = %d;" % len(rcw_l),
        "constexpr int RECOMPUTED_U_WIDE = %d;" % len(rcw_u),
        gen_jac_decomp_pf(mech, RC_BATCH_WIDE, rcw_l | rcw_u, "jac_decomp_rcw_pf%d" % RC_BATCH_WIDE, rcw_u)[0],
        "",
        gen_solve_rc_pf_reg(mech, RC_BATCH_WIDE, rcw_l, "solve_rcw_pf%d_reg" % RC_BATCH_WIDE, rcw_u),
        "",
        gen_hot_tables(mech),
        "",
        gen_wrf_tables(mech),
        "",
        "}  // namespace gen",
        "}  // namespace radm2",
        "",
    ]
    with open(dst, "w") as f:
        f.write("\n".join(parts))
    return dst


if __name__ == "__main__":
    print(generate(*(sys.argv[1:2] or [None])))
