#!/usr/bin/env python3
"""Write the CPU oracle's mechanism tables (C header) from the mechanism JSON.

TEST INFRASTRUCTURE ONLY.  This is the oracle's own, independent reading of
``mech/radm2.json``: it emits plain data tables (rate monomials, tendency
terms, dropped terms, LU pattern) that ``radm2_oracle.c`` walks with loops,
plus the 156 rate laws written literally the way the reference writes them
(mgpu.cu:763-918, helpers mgpu.cu:708-750), with no shared subexpressions.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MECH = os.path.join(ROOT, "radm2-solver-multi-gpu_amd", "mech", "radm2.json")


def rate_to_c(e):
    """Rate-law notation -> C calling the oracle's literal helpers."""
    e = re.sub(r"\bJ\[(\d+)\]", r"J[\1]", e)
    e = re.sub(r"\bARR2\(([^,()]+),([^,()]+)\)", r"o_arr2(\1,\2, T)", e)
    e = re.sub(r"\bTROEE\(([^()]*)\)", r"o_troee(\1, T, M)", e)
    e = re.sub(r"\bTROE\(([^()]*)\)", r"o_troe(\1, T, M)", e)
    e = re.sub(r"\bTHERMAL_T2\(([^,()]+),([^,()]+)\)", r"o_thermal_t2(\1,\2, T)", e)
    e = e.replace("K46()", "o_k46(M, T)")
    return e


def main(dst=None):
    mech = json.load(open(MECH))
    dst = dst or os.path.join(HERE, "generated", "radm2_oracle_tables.h")
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    L = ["/* GENERATED by oracle/gen_oracle_tables.py -- test infrastructure. */",
         "#pragma once"]
    # monomials: factor code = species index (V) or 100 + fixed index (F); -1 ends
    L.append("static const int O_MONO[156][4] = {")
    for facs in mech["monomials"]:
        codes = [(i if k == "V" else 100 + i) for k, i in facs] + [-1] * (4 - len(facs))
        L.append("  {%s}," % ",".join(map(str, codes)))
    L.append("};")
    starts, rs, cs = [0], [], []
    for terms in mech["tendency"]:
        for r, c in terms:
            rs.append(r)
            cs.append(c)
        starts.append(len(rs))
    L.append("static const int O_TEND_START[60] = {%s};" % ",".join(map(str, starts)))
    L.append("static const int O_TEND_R[%d] = {%s};" % (len(rs), ",".join(map(str, rs))))
    L.append("static const double O_TEND_C[%d] = {%s};" % (len(cs), ",".join(
        (c if ("." in c or "e" in c.lower()) else c + ".0") for c in cs)))
    d = mech["fun_dropped_terms"]
    L.append("#define O_NDROP %d" % len(d))
    L.append("static const int O_DROP_I[%d] = {%s};" % (len(d), ",".join(str(x[0]) for x in d)))
    L.append("static const int O_DROP_R[%d] = {%s};" % (len(d), ",".join(str(x[1]) for x in d)))
    L.append("static const double O_DROP_C[%d] = {%s};" % (len(d), ",".join(x[2] + ".0" for x in d)))
    # photolysis: reactions whose rate law is a bare J[s]
    pj = {r: int(re.fullmatch(r"\(J\[(\d+)\]\)", e.strip()).group(1))
          for r, e in enumerate(mech["rates"]) if re.fullmatch(r"\(J\[(\d+)\]\)", e.strip())}
    L.append("#define O_NPHOTO %d" % len(pj))
    L.append("static const int O_PHOTO_R[%d] = {%s};" % (len(pj), ",".join(map(str, pj))))
    L.append("static const int O_PHOTO_J[%d] = {%s};" % (len(pj), ",".join(map(str, pj.values()))))
    lu = mech["lu"]
    L.append("static const int O_LU_CROW[60] = {%s};" % ",".join(map(str, lu["crow"])))
    L.append("static const int O_LU_DIAG[59] = {%s};" % ",".join(map(str, lu["diag"])))
    L.append("static const int O_LU_ICOL[659] = {%s};" % ",".join(map(str, lu["icol"])))
    w = mech["wrf"]
    L.append("static const int O_WRF_SRC[59] = {%s};" % ",".join(str(a) for a, _ in w["var_source"]))
    L.append("static const int O_WRF_FIELD[59] = {%s};" % ",".join(str(f) for _, f in w["var_source"]))
    for name, v in sorted(w["constants"].items()):
        L.append("#define O_%s (%r)" % (name.upper(), v))
    L.append("static void o_rconst_literal(double T, double M, const double *J, double RC_N2O5,"
             " double *k) {")
    for r, e in enumerate(mech["rates"]):
        L.append("  k[%d] = %s;" % (r, rate_to_c(e)))
    L.append("}")
    with open(dst, "w") as f:
        f.write("\n".join(L) + "\n")
    return dst


if __name__ == "__main__":
    print(main(*(sys.argv[1:2] or [None])))
