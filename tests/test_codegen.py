"""Invariants of the generated HIP code (codegen/gen_radm2.py), checked on
the text the generator writes: the hot-set tables place every rate
constant, state and LU entry exactly once, and the load-scheduled fused
Jacobian/decomposition never reads an LU entry before the statement that
stores it."""
import importlib.util
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "radm2-solver-multi-gpu_amd")
spec = importlib.util.spec_from_file_location("gen_radm2", os.path.join(PKG, "codegen", "gen_radm2.py"))
G = importlib.util.module_from_spec(spec)
spec.loader.exec_module(G)
MECH = json.load(open(os.path.join(PKG, "mech", "radm2.json")))


@pytest.fixture(scope="module")
def header(tmp_path_factory):
    p = tmp_path_factory.mktemp("gen") / "radm2_gen.h"
    G.generate(str(p))
    return p.read_text()


def _array(text, name):
    mo = re.search(r"%s\[(\d+)\] = \{([^}]*)\};" % re.escape(name), text)
    assert mo, name
    vals = [int(x) for x in mo.group(2).split(",")]
    assert len(vals) == int(mo.group(1))
    return vals


@pytest.mark.parametrize("budget", G.HOT_BUDGETS)
def test_hot_tables_partition_the_ranking(header, budget):
    S, R, P, L = budget
    ranked, acc = G.hot_ranking(MECH)
    sizes = (len(G.rate_forms(MECH)[0]), MECH["nvar"], MECH["lu"]["nnz"])
    assert len(ranked) == sum(sizes) and len(set(ranked)) == len(ranked)
    codes = [_array(header, "%s_%s" % (G.hot_name(*budget), n)) for n in ("K", "Y", "G")]
    lds, regs, split, hbm = [], [], [], []
    for a, arr in enumerate(codes):
        for i, c in enumerate(arr):
            if c >= 0:
                lds.append((c, (a, i)))
            elif c <= G.SPLIT_BASE:
                split.append((G.SPLIT_BASE - c, (a, i)))
            elif c <= G.HBM_BASE:
                hbm.append((G.HBM_BASE - c, (a, i)))
            else:
                assert c <= -2
                regs.append((-2 - c, (a, i)))
    # LDS slots 0..S-1, registers 0..R-1 and split entries 0..P*L-1 each hold
    # exactly one entry ...
    assert sorted(s for s, _ in lds) == list(range(S))
    assert sorted(r for r, _ in regs) == list(range(R))
    nsplit = min(P * L, len(ranked) - S - R)   # the tier may outsize the arrays
    assert sorted(j for j, _ in split) == list(range(nsplit))
    # ... the S most accessed in LDS, the next R in registers, the next P*L
    # split over the cell's lanes
    assert {e for _, e in lds} == set(ranked[:S])
    assert {e for _, e in regs} == set(ranked[S:S + R])
    assert {e for _, e in split} == set(ranked[S + R:S + R + P * L])
    # the rest in the workspace, numbered in ranking order
    n = S + R + nsplit
    assert [e for _, e in sorted(hbm)] == ranked[n:]
    # and nothing outside the tiers is accessed more than what is inside
    inside = min(acc[e] for e in ranked[:n]) if n else 0
    assert all(acc[e] <= inside for e in ranked[n:])


def _body(header, sig):
    start = header.index(sig)
    return header[start:header.index("\n}\n", start)]


def _formed(wide=False):
    t = G.RC_MAXTERMS_WIDE if wide else G.RC_MAXTERMS
    return (G.recomputed_l(MECH, maxmod=G.RC_MAXMOD, maxterms=t),
            G.recomputed_u(MECH, maxterms=t))


# (name, wide, constant suffix, load batch) of the two decomposition/solve
# pairs with formed entries
PAIRS = [("rc", False, "", G.RC_BATCH), ("rcw", True, "_WIDE", G.RC_BATCH_WIDE)]


@pytest.mark.parametrize("fn", ["jac_decomp_pf16", "jac_decomp_rc_pf%d" % G.RC_BATCH,
                                "jac_decomp_rcw_pf%d" % G.RC_BATCH_WIDE])
def test_lu_entries_are_read_after_they_are_stored(header, fn):
    body = _body(header, "int %s(" % fn)
    skipped = frozenset() if fn == "jac_decomp_pf16" else frozenset().union(
        *_formed(wide=fn.startswith("jac_decomp_rcw_")))
    stored = {}
    assert len(re.findall(r"RDP<\d+>\(g\)", body)) > 300   # the check is not vacuous
    for n, ln in enumerate(body.split("\n")):
        for mo in re.finditer(r"RDP<(\d+)>\(g\)", ln):
            q = int(mo.group(1))
            assert q in stored, "LU entry %d read before its store (line %d)" % (q, n)
        mo = re.match(r"\s*WR\(g, (\d+),", ln)
        if mo:
            q = int(mo.group(1))
            assert q not in stored, "LU entry %d stored twice" % q
            stored[q] = n
    assert sorted(stored) == sorted(set(range(MECH["lu"]["nnz"])) - skipped)


@pytest.mark.parametrize("pair,wide,suffix,batch", PAIRS)
def test_recomputed_l_entries_are_formed_as_the_decomposition_forms_them(header, pair, wide, suffix,
                                                                         batch):
    # the solves form these L entries instead of reading them back: each
    # must be an L position that no earlier elimination of its row updates,
    # never stored by the decomposition the kernels run, never read by their
    # solve, and formed there by the decomposition's own expression times
    # the same inverse pivot
    lu = MECH["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    rc, rc_u = _formed(wide)
    assert len(rc) == int(re.search(r"RECOMPUTED_L%s = (\d+);" % suffix, header).group(1)) > 50
    row_of = {kk: k for k in range(59) for kk in range(crow[k], crow[k + 1])}
    dec = _body(header, "int jac_decomp_pf16(")
    solve = _body(header, "void solve_%s_pf%d_reg(" % (pair, batch))
    read = {int(q) for q in re.findall(r"RDP<(\d+)>\(g\)", solve)}
    assert not (read & rc)
    # every other LU entry is read (the formed U entries are formed too)
    assert read == set(range(lu["nnz"])) - rc - rc_u

    def norm(e):
        return re.sub(r"V\(([kv])(\d+)_(?:e)?\d+\)", r"\1[\2]", e).strip()
    for kk in sorted(rc):
        k, j = row_of[kk], icol[kk]
        assert crow[k] <= kk < diag[k]
        assert not G.l_updates(MECH, k, j, G.lu_positions(MECH))
        mo = re.search(r"const double l%d = (.*) \* V\(p(\d+)_e%d\); "
                       r"x\[%d\] = fma\(-l%d, x\[%d\], x\[%d\]\);"
                       % (kk, kk, k, kk, j, k), solve)
        assert mo, kk
        assert int(mo.group(2)) == diag[j]
        md = re.search(r"double w%d_%d = (.*);" % (j, k), dec)
        assert md, kk
        assert norm(mo.group(1)) == norm(md.group(1)), kk
        # the decomposition scales the entry by the same inverse pivot and
        # applies no other update to it before
        upd = re.findall(r"\n\s*w%d_%d = ([^\n]*);" % (j, k), dec)
        assert upd[0] == "w%d_%d * V(u%d_%d)" % (j, k, diag[j], k), (kk, upd)


def test_workspace_and_hot_accesses_use_constant_indices(header):
    # the register tier needs compile-time indices: every generated access
    # to the rate constants, the state or the LU names its element literally
    for mo in re.finditer(r"\b(RD|WR)\((k|v|g), ([^,)]+)", header):
        assert re.fullmatch(r"\d+", mo.group(3).strip()), mo.group(0)
    # scheduled reads name theirs as a template argument
    n = 0
    for mo in re.finditer(r"\bRDP<([^>]+)>\((k|v|g)\)", header):
        assert re.fullmatch(r"\d+", mo.group(1).strip()), mo.group(0)
        n += 1
    assert n > 1000


def test_rate_storage_forms_rebuild_every_rate_constant():
    # each reaction's rate constant at a use is its stored value, a literal
    # or A times a stored exponential: the same expression as the rate law
    exprs, _ = G.rate_exprs(MECH)
    prims, forms = G.rate_forms(MECH)
    assert len(forms) == MECH["nreact"] and len(prims) < MECH["nreact"]
    kinds = set()
    for r, ex in enumerate(exprs):
        got = G.kexpr(forms, r, lambda p: prims[p])
        assert got == ex, (r, got, ex)
        kinds.add(forms[r][0])
        if forms[r][0] == "mul":
            assert re.fullmatch(r"rc_s\d+", prims[forms[r][2]])
    assert kinds == {"lit", "mul", "prim"}
    # the photolysis rate constants keep their reaction's index (the forcing
    # code writes them there)
    for r in range(len(G.photo_j(MECH))):
        assert forms[r] == ("prim", r)


def test_photolysis_map_matches_the_rate_laws():
    pj = G.photo_j(MECH)
    for r, s in enumerate(pj):
        assert MECH["rates"][r].strip() == "(J[%d])" % s
    assert sorted(pj) == list(range(21))


@pytest.mark.parametrize("pair,wide,suffix,batch", PAIRS)
def test_recomputed_u_entries_are_formed_as_the_decomposition_forms_them(header, pair, wide, suffix,
                                                                         batch):
    # the decomposition and the solves also form U entries: each an off-diagonal U
    # position that no elimination of its row updates, so that it is the
    # Jacobian expression itself; the decomposition neither stores nor reads
    # it (the rows that eliminate with it form it), the solve forms it in the
    # backward substitution
    lu = MECH["lu"]
    crow, diag, icol = lu["crow"], lu["diag"], lu["icol"]
    rc_l, rc_u = _formed(wide)
    assert len(rc_u) == int(re.search(r"RECOMPUTED_U%s = (\d+);" % suffix, header).group(1)) > 20
    row_of = {kk: k for k in range(59) for kk in range(crow[k], crow[k + 1])}
    dec0 = _body(header, "int jac_decomp_pf16(")
    dec = _body(header, "int jac_decomp_%s_pf%d(" % (pair, batch))
    solve = _body(header, "void solve_%s_pf%d_reg(" % (pair, batch))
    read = {int(q) for q in re.findall(r"RDP<(\d+)>\(g\)", solve)}
    assert read == set(range(lu["nnz"])) - rc_l - rc_u
    assert not ({int(q) for q in re.findall(r"RDP<(\d+)>\(g\)", dec)} & rc_u)

    def norm(e):
        return re.sub(r"V\(f?([kv])(\d+)_[u\d_]+\)", r"\1[\2]", e).strip()
    pos = G.lu_positions(MECH)
    for kk in sorted(rc_u):
        k, c = row_of[kk], icol[kk]
        assert diag[k] < kk < crow[k + 1]
        assert not any(pos.get((i, c), -1) > diag[i]
                       for i in (icol[q] for q in range(crow[k], diag[k])))
        md = re.search(r"double w%d_%d = (.*);" % (c, k), dec0)
        assert md, kk
        assert not re.search(r"\n\s*w%d_%d = " % (c, k), dec0), kk   # never updated
        expr = norm(md.group(1))
        mo = re.search(r"x\[%d\] = fma\(-\((.*)\), x\[%d\], x\[%d\]\);" % (k, c, k), solve)
        assert mo and norm(mo.group(1)) == expr, kk
        # every later row that eliminates with row k forms it the same way
        for r in range(k + 1, 59):
            if (r, k) in pos:
                mr = re.search(r"w%d_%d = fma\(-w%d_%d, (.*), w%d_%d\);" % (c, r, k, r, c, r), dec)
                assert mr and norm(mr.group(1)) == expr, (kk, r)
