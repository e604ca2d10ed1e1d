"""Host-side checks of the C ABI that need no GPU: the library loads, exports
every function include/radm2.h declares, and its KPP option validation
returns the reference's IERR codes (mgpu.cu:3715-3811) exactly as the
oracle's does."""
import ctypes
import itertools
import os
import re

import pytest

import oracle_ct as O

ROOT = O.ROOT
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def _declared():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(radm2_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import radm2_amd
    lib = radm2_amd.load_library()
    names = _declared()
    assert "radm2_integrate" in names and "radm2_rosenbrock" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    # and the Python mirror binds each of them
    from radm2_amd import mlp
    assert sorted(radm2_amd.EXPORTED_SYMBOLS + mlp.EXPORTED_SYMBOLS) == names


def test_schedule_and_placement_constants_match_the_python_mirror():
    # every RADM2_SCHED_* and RADM2_PLACE_* value of the header (an alias
    # resolved to its target) equals the Python constant of the same name
    import radm2_amd
    text = open(os.path.join(ROOT, "include", "radm2.h")).read()
    defs = dict(re.findall(r"^#define (RADM2_(?:SCHED|PLACE)_\w+) (\w+)", text, re.M))
    vals = {}
    for name, v in defs.items():
        while v in defs:
            v = defs[v]
        vals[name] = int(v)
    assert len(vals) == 10
    assert vals["RADM2_PLACE_LDS160_REG128"] == vals["RADM2_PLACE_LDS160_REG112"] == 4
    for name, v in vals.items():
        assert getattr(radm2_amd, name[len("RADM2_"):]) == v, name


def test_strerror_texts_follow_kpp():
    import radm2_amd
    assert radm2_amd.strerror(-5) == "Improper tolerance values"
    assert radm2_amd.strerror(-8) == "Matrix is repeatedly singular"
    assert radm2_amd.strerror(radm2_amd.E_ARG) == "invalid argument"


CASES = {
    "hmin": [0.0, 1e-3, -1.0],
    "hmax": [0.0, 100.0, -5.0],
    "hstart": [0.0, 6.0, -1e-3],
    "facmin": [0.0, 0.3, -0.1],
    "facmax": [0.0, 5.0, -1.0],
    "facrej": [0.0, 0.2, -1.0],
    "facsafe": [0.0, 0.8, -0.5],
    "max_steps": [0, 50, -1],
    "method": [0, 2, 1, 3, 6, -1],
    "sing_retries": [0, 5, -1],
}


def _expected(o, rtol, atol):
    """KPP's order of checks (mgpu.cu:3715-3811), written from the spec."""
    if o["max_steps"] < 0:
        return -1
    m = o["method"]
    if m != 0 and not (1 <= m <= 5):
        return -2
    if o["hmin"] < 0 or o["hmax"] < 0 or o["hstart"] < 0:
        return -3
    if o["facmin"] < 0 or o["facmax"] < 0 or o["facrej"] < 0 or o["facsafe"] < 0:
        return -4
    if atol <= 0 or rtol <= 10 * 2.0 ** -52 or rtol >= 1:
        return -5
    # KPP's five methods (the reference initialises only Ros3; the others
    # are implemented here); the singular-matrix retry count is not a KPP
    # option: a negative one is an argument error
    if o["sing_retries"] < 0:
        return -101
    return 0


def test_option_codes_match_oracle_and_kpp():
    import radm2_amd
    keys = list(CASES)
    n = 0
    for field in keys:
        for v in CASES[field]:
            for rtol, atol in ((1e-3, 1.0), (1e-16, 1.0), (1.0, 1.0), (1e-3, 0.0)):
                o = {k: CASES[k][0] for k in keys}
                o[field] = v
                po = radm2_amd.Opts(**o)
                oo = O.OptsIn(**o)
                got = radm2_amd.check_opts(600.0, rtol, atol, po)
                rc, _ = O.resolve(600.0, rtol, atol, oo)
                assert got == rc == _expected(o, rtol, atol), (o, rtol, atol, got, rc)
                n += 1
    assert n > 100
    assert radm2_amd.check_opts(600.0, 1e-3, 1.0, None) == 0


def test_integrate_defaults_resolve_like_integrate_call():
    # INTEGRATE(): Hstart = 0.01*dt, all else KPP defaults (mgpu.cu:3866-3873)
    rc, o = O.resolve(600.0, 1e-3, 1.0, None)
    assert rc == 0
    assert o.hstart == pytest.approx(6.0)
    assert (o.hmin, o.hmax, o.facmin, o.facmax, o.facrej, o.facsafe, o.max_steps) == \
        (0.0, 600.0, 0.2, 6.0, 0.1, 0.9, 100000)
    assert o.roundoff == 2.0 ** -52


def test_learned_part_argument_checks_need_no_device():
    # every call here is rejected before any launch (pointers are never read)
    from radm2_amd import mlp
    L = mlp.lib()
    P = ctypes.c_void_p(16)
    E = -101
    assert L.radm2_mlp_h0_select(None, 16, 16, P, 0.0, 600.0, 0.5, 0.0, P, None) == E     # hlo <= 0
    assert L.radm2_mlp_h0_select(None, 16, 16, P, 1.0, 0.5, 0.5, 0.0, P, None) == E       # hmax < hlo
    assert L.radm2_mlp_h0_select(None, 16, 1, P, 1e-9, 600.0, 0.5, 0.0, P, None) == E     # N < 2
    assert L.radm2_mlp_mse_loss(None, 16, 128, 2, 64, 65, P, P, None, P, P) == E           # live > width
    assert L.radm2_mlp_mse_loss(None, 16, 96, 2, 64, 59, P, P, None, P, P) == E            # heads*width > N
    assert L.radm2_mlp_surrogate_apply(None, 16, 16, 128, 58, P, P, P, 1.0, 1e3, 1e-3, 0.05, 0,
                                       P, P, None) == E                                     # width < 59
    assert L.radm2_mlp_surrogate_apply(None, 16, 16, 128, 64, P, P, P, 1.0, 1e3, 1e-3, 1.5, 0,
                                       P, P, None) == E                                     # audit > 1
    assert L.radm2_mlp_compact(None, 0, P, P, P) == E
    assert L.radm2_mlp_gather(None, 2, 1, 16, P, 4, P, 4, P) == E                          # elem size
    assert L.radm2_mlp_scatter(None, 8, 1, 2, P, 4, P, 16, P) == E                         # lds < m
    assert L.radm2_mlp_gather(None, 8, 1, 16, P, 0, None, 0, P) == 0                       # empty: no-op
    assert L.radm2_mlp_dlog_targets(None, 16, 16, P, P, 1.0, P, None, P) == E              # tmu w/o tsd
