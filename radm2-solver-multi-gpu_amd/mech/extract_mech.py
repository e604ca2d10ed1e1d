#!/usr/bin/env python3
"""One-off extraction of the RADM2 mechanism table from the reference's text.

This script reads ``mgpu.cu`` of the reference *as text* (nothing is compiled
or run) and writes ``radm2.json`` next to it: the species order, the rate law
of each of the 156 reactions, the rate monomials, the net stoichiometry that
the tendency function uses, the Jacobian terms, and the sparse-LU pattern.
Every solver path of this package (the GPU code generator and the CPU oracle)
is generated from that JSON; the JSON itself is data, not code.

Where each table comes from (reference line numbers):
  species order        mgpu.cu:125-189   (#define ind_* / indf_*)
  rate laws            mgpu.cu:763-918   (RCONST[...] = ...)
  rate monomials       mgpu.cu:1103-1258 (A[r] = RCT[r] * V[..] * F[..])
  tendency terms       mgpu.cu:1260-1385 (Vdot[i] = sum c*A[r])
  Jacobian partials    mgpu.cu:1685-1964 (B[b] = RCT[r] * ...)
  Jacobian entries     mgpu.cu:1966-2651 (JVS[k] = sum c*B[b])
  LU pattern           mgpu.cu:2873-3090 (unrolled forward/backward solve)

The extraction cross-checks itself: the JVS entries as written must equal the
exact derivative of the tendency with the H2O2 + OH loss term restored
(mgpu.cu:1275 comments that term out of Vdot[13] while the Jacobian keeps it),
and the LU pattern must equal the symbolic no-pivot LU of the Jacobian
pattern.

Usage (only where the reference is mounted):
    python extract_mech.py /root/reference/mgpu.cu
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _lines(src, lo, hi):
    """Return reference lines lo..hi (1-based, inclusive) joined."""
    return "\n".join(src[lo - 1:hi])


def _statements(text):
    """Split C text into ';'-terminated statements with // comments removed."""
    text = re.sub(r"//[^\n]*", "", text)
    out = []
    for st in text.split(";"):
        st = " ".join(st.split())
        if st:
            out.append(st)
    return out


def _unmacro(s):
    s = re.sub(r"MY_I_A\((\d+)\)", r"\1", s)
    s = re.sub(r"MY_I\((\w+)\)", r"\1", s)
    return s


def parse_species(src):
    var, fix = {}, {}
    for ln in src[124:189]:
        m = re.match(r"#define\s+ind_(\w+)\s+(\d+)", ln)
        if m and int(m.group(2)) < 59:
            var[int(m.group(2))] = m.group(1)
        m = re.match(r"#define\s+indf_(\w+)\s+(\d+)", ln)
        if m:
            fix[int(m.group(2))] = m.group(1)
    assert sorted(var) == list(range(59)), var
    assert sorted(fix) == [0, 1], fix
    return [var[i] for i in range(59)], [fix[i] for i in range(2)]


def parse_rates(src):
    """RCONST[r] = (expr) -> expression in the package's rate-law notation.

    Notation: T (K), M (molec/cm3), J[i] (photolysis slot i of the 21-slot
    input, reference jv order), RC_N2O5 (s-1); helper laws ARR2(A,B),
    TROE(k0,n,kinf,m), TROEE(A,B,k0,n,kinf,m), THERMAL_T2(c,d), K46().
    """
    rates = {}
    for st in _statements(_unmacro(_lines(src, 763, 918))):
        m = re.match(r"RCONST\[(\d+)\]\s*=\s*(.*)$", st)
        assert m, st
        e = m.group(2)
        e = re.sub(r"jv\[(\d+)\]", r"J[\1]", e)
        e = re.sub(r",\s*TEMP\s*,\s*C_M\s*\)", ")", e)      # TROE/TROEE
        e = re.sub(r"ARR2\(([^,]+),\s*([^,]+),\s*TEMP\s*\)", r"ARR2(\1, \2)", e)
        e = re.sub(r"THERMAL_T2\(([^,]+),\s*([^,]+),\s*TEMP\s*\)", r"THERMAL_T2(\1, \2)", e)
        e = e.replace("k46(C_M, TEMP)", "K46()")
        e = re.sub(r"\bTEMP\b", "T", e)
        e = re.sub(r"\bC_M\b", "M", e)
        e = re.sub(r"\brc_n2o5\b", "RC_N2O5", e)
        rates[int(m.group(1))] = e
    assert sorted(rates) == list(range(156))
    return [rates[r] for r in range(156)]


def parse_fun(src):
    monomials, tend = {}, {}
    for st in _statements(_unmacro(_lines(src, 1103, 1385))):
        m = re.match(r"A\[(\d+)\]\s*=\s*RCT\[(\d+)\]\s*((?:\*\s*[VF]\s*\[\d+\]\s*)+)$", st)
        if m:
            r = int(m.group(1))
            assert int(m.group(2)) == r
            facs = re.findall(r"([VF])\s*\[(\d+)\]", m.group(3))
            monomials[r] = [[k, int(i)] for k, i in facs]
            continue
        m = re.match(r"Vdot\[(\d+)\]\s*=\s*(.*)$", st)
        assert m, st
        terms = []
        body = m.group(2).replace(" ", "")
        for sgn, coef, r in re.findall(r"([+-]?)(?:([0-9.eE+-]+?)\*)?A\[(\d+)\]", body):
            c = coef if coef else "1"
            c = ("-" if sgn == "-" else "") + c
            terms.append([int(r), c])
        # the regex must consume the whole body
        rebuilt = "".join(
            ("" if (i == 0 and not c.startswith("-")) else ("" if c.startswith("-") else "+"))
            + (c if c not in ("1", "-1") else c[:-1]) + ("*" if c not in ("1", "-1") else "")
            + "A[%d]" % r for i, (r, c) in enumerate(terms))
        assert rebuilt == body, (rebuilt, body)
        tend[int(m.group(1))] = terms
    assert sorted(monomials) == list(range(156))
    assert sorted(tend) == list(range(59))
    return [monomials[r] for r in range(156)], [tend[i] for i in range(59)]


def parse_jac(src):
    bdefs, jvs = {}, {}
    for st in _statements(_unmacro(_lines(src, 1685, 2651))):
        m = re.match(r"B\[(\d+)\]\s*=\s*RCT\[(\d+)\](.*)$", st)
        if m:
            b, r, rest = int(m.group(1)), int(m.group(2)), m.group(3).replace(" ", "")
            facs = re.findall(r"\*(?:(\d+)\*)?([VF])\[(\d+)\]", rest)
            scale = 1
            flist = []
            for mult, k, i in facs:
                if mult:
                    scale *= int(mult)
                flist.append([k, int(i)])
            bdefs[b] = {"r": r, "scale": scale, "f": flist, "src": rest}
            continue
        m = re.match(r"JVS\[(\d+)\]\s*=\s*(.*)$", st)
        assert m, st
        body = m.group(2).replace(" ", "")
        if body == "0":
            jvs[int(m.group(1))] = []
            continue
        terms = []
        for sgn, coef, b in re.findall(r"([+-]?)(?:([0-9.eE]+?)\*)?B\[(\d+)\]", body):
            c = (coef if coef else "1")
            terms.append([int(b), ("-" if sgn == "-" else "") + c])
        jvs[int(m.group(1))] = terms
    assert sorted(jvs) == list(range(659))
    return bdefs, [jvs[k] for k in range(659)]


def parse_solve_pattern(src):
    """LU positions from the unrolled solve: L in the forward sweep, U in the
    backward sweep (X(i) = (X(i) - sum JVS(k)*X(j)) / JVS(d))."""
    pos = {}
    text = _lines(src, 2873, 3090)
    for st in _statements(text):
        m = re.match(r"X\((\d+)\)\s*=\s*(.*)$", st)
        assert m, st
        i = int(m.group(1))
        body = m.group(2).replace(" ", "")
        for k, j in re.findall(r"JVS\((\d+)\)\*X\((\d+)\)", body):
            pos[int(k)] = (i, int(j))
        d = re.search(r"/\(?JVS\((\d+)\)\)?$", body)
        if d:
            pos[int(d.group(1))] = (i, i)
    assert sorted(pos) == list(range(659)), len(pos)
    return [pos[k] for k in range(659)]


def symbolic_lu(pattern, n=59):
    """No-pivot symbolic LU fill of a sparse pattern (set of (i,j))."""
    rows = [set(j for (i, j) in pattern if i == r) | {r} for r in range(n)]
    for k in range(n):
        j = -1
        while True:
            # next column left of the diagonal, including fill created so far
            left = [c for c in rows[k] if j < c < k]
            if not left:
                break
            j = min(left)
            # eliminating with row j adds row j's upper part to row k
            rows[k] |= set(c for c in rows[j] if c > j)
    return [(i, j) for i in range(n) for j in sorted(rows[i])]


def parse_wrf_map(src):
    """dev_convert's species shuffle (mgpu.cu:585-645): for each variable
    species, the WRF source array (0 = the chemistry array, 1 = the
    radical array) and its field; the chemistry fields' P_* names
    (mgpu.cu:81-122); and the molecular-weight constants (mgpu.cu:30-44)."""
    text = "\n".join(src[512:646])
    pairs = re.findall(r"VAR\[MY_I\((\d+)\)\]\s*=\s*conv \* MAX\(\((wrfchemdata|wrfraddata)\[\(\((\d+)\)",
                       text)
    vmap = {}
    for v, arr, fld in pairs:
        assert int(v) not in vmap
        vmap[int(v)] = [0 if arr == "wrfchemdata" else 1, int(fld)]
    assert sorted(vmap) == list(range(59)), sorted(set(range(59)) - set(vmap))
    pnames = {}
    for ln in src[80:123]:
        mo = re.match(r"#define P_(\w+)\s+\((\d+)\)", ln.strip())
        if mo:
            pnames[mo.group(1)] = int(mo.group(2))
    consts = {}
    for name in ("navgdro", "mwh", "mwo", "mwair"):
        for ln in src[25:50]:
            mo = re.match(r"#define %s\s+\(([-0-9.eE+]+)\)" % name, ln.strip())
            if mo:
                consts[name] = float(mo.group(1))
    assert set(consts) == {"navgdro", "mwh", "mwo", "mwair"}, consts
    return {"var_source": [vmap[v] for v in range(59)], "chem_names": pnames,
            "constants": consts, "nph": 28, "halo": [4, 0, 4]}


def main(path):
    src = open(path).read().split("\n")
    species, fixed = parse_species(src)
    rates = parse_rates(src)
    monomials, tend = parse_fun(src)
    bdefs, jvs = parse_jac(src)
    lu_pos = parse_solve_pattern(src)

    # --- LU layout checks --------------------------------------------------
    assert lu_pos == sorted(lu_pos), "LU positions must be row-major sorted"
    crow, diag = [], []
    for r in range(59):
        idx = [k for k, (i, j) in enumerate(lu_pos) if i == r]
        crow.append(idx[0])
        diag.append(lu_pos.index((r, r)))
    crow.append(659)
    icol = [j for (i, j) in lu_pos]

    # --- full stoichiometry: Vdot as written plus the dropped H2O2 loss ----
    # mgpu.cu:1275 writes Vdot[13] = -A[8]+A[31]+A[32]; //-A[33]
    dropped = [[13, 33, "-1"]]

    # Jacobian pattern (structurally nonzero positions of J = dF/dV, full)
    stoich_full = [list(t) for t in tend]
    for i, r, c in dropped:
        stoich_full[i].append([r, c])
    jpat = set()
    for i in range(59):
        for r, c in stoich_full[i]:
            for k, j in monomials[r]:
                if k == "V":
                    jpat.add((i, j))
    nonzero = len(jpat | {(i, i) for i in range(59) if (i, i) in jpat})
    sym = symbolic_lu(jpat)
    assert sym == lu_pos, "symbolic LU of the Jacobian pattern != reference LU pattern"

    # --- Jacobian as written == exact derivative of the full tendency ------
    rng = np.random.default_rng(0)
    V = rng.uniform(0.5, 2.0, 59)
    F = rng.uniform(0.5, 2.0, 2)
    K = rng.uniform(0.5, 2.0, 156)
    B = {}
    for b, d in bdefs.items():
        v = K[d["r"]] * d["scale"]
        for k, i in d["f"]:
            v *= V[i] if k == "V" else F[i]
        B[b] = v
    jw = np.array([sum(float(c) * B[b] for b, c in jvs[k]) for k in range(659)])
    jd = np.zeros(659)
    lu_index = {p: k for k, p in enumerate(lu_pos)}
    for i in range(59):
        for r, c in stoich_full[i]:
            fac = monomials[r]
            for q, (kk, j) in enumerate(fac):
                if kk != "V":
                    continue
                prod = K[r] * float(c)
                for q2, (kk2, j2) in enumerate(fac):
                    if q2 != q:
                        prod *= V[j2] if kk2 == "V" else F[j2]
                jd[lu_index[(i, j)]] += prod
    bad = [k for k in range(659) if abs(jw[k] - jd[k]) > 1e-12 * (1 + abs(jd[k]))]
    assert bad == [], [(k, lu_pos[k], jw[k], jd[k]) for k in bad]
    zero_written = [k for k in range(659) if not jvs[k]]

    # --- Fun as written == full tendency minus the dropped term ------------
    nterms = sum(len(t) for t in tend)

    out = {
        "name": "RADM2 (KPP), as encoded in the reference's mgpu.cu",
        "species": species,
        "fixed": fixed,
        "nvar": 59,
        "nfix": 2,
        "nreact": 156,
        "rates": rates,
        "monomials": monomials,
        "tendency": tend,
        "fun_dropped_terms": dropped,
        "jac_nonzero": len(jpat),
        "lu": {"crow": crow, "diag": diag, "icol": icol, "nnz": 659},
        "jac_written_zero": zero_written,
        "wrf": parse_wrf_map(src),
        "provenance": {
            "species": "mgpu.cu:125-189", "rates": "mgpu.cu:763-918",
            "monomials": "mgpu.cu:1103-1258", "tendency": "mgpu.cu:1260-1385",
            "jacobian": "mgpu.cu:1685-2651", "lu_pattern": "mgpu.cu:2873-3090",
            "wrf": "mgpu.cu:30-44, 81-122, 513-646",
        },
    }
    dst = os.path.join(HERE, "radm2.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote %s: %d reactions, %d tendency terms, %d Jacobian nonzeros, "
          "%d LU positions (%d written as 0)" % (dst, len(rates), nterms, len(jpat),
                                                   len(lu_pos), len(zero_written)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/mgpu.cu")
