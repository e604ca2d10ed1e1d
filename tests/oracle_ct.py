"""ctypes binding of the CPU oracle (oracle/liboracle_radm2.so) for tests.

The oracle is test infrastructure: only tests/, the smoke check and the
CPU-baseline leg of bench.py use it, always as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle_radm2.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


class OptsIn(ctypes.Structure):
    _fields_ = [("hmin", ctypes.c_double), ("hmax", ctypes.c_double),
                ("hstart", ctypes.c_double), ("facmin", ctypes.c_double),
                ("facmax", ctypes.c_double), ("facrej", ctypes.c_double),
                ("facsafe", ctypes.c_double), ("max_steps", ctypes.c_int32),
                ("method", ctypes.c_int32), ("sing_retries", ctypes.c_int32)]


class OptsOut(ctypes.Structure):
    _fields_ = [("hmin", ctypes.c_double), ("hmax", ctypes.c_double),
                ("hstart", ctypes.c_double), ("facmin", ctypes.c_double),
                ("facmax", ctypes.c_double), ("facrej", ctypes.c_double),
                ("facsafe", ctypes.c_double), ("roundoff", ctypes.c_double),
                ("max_steps", ctypes.c_int32), ("method", ctypes.c_int32),
                ("sing_retries", ctypes.c_int32)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load(path):
    """The oracle library at `path` with its functions' signatures bound."""
    L = ctypes.CDLL(path)
    L.oracle_rconst.argtypes = [ctypes.c_double, ctypes.c_double, _D, ctypes.c_double, _D]
    L.oracle_default_rc_n2o5.restype = ctypes.c_double
    L.oracle_fun.argtypes = [_D, ctypes.c_double, ctypes.c_double, _D, _D]
    L.oracle_jac.argtypes = [_D, ctypes.c_double, ctypes.c_double, _D, _D]
    L.oracle_jac_prep.argtypes = [_D, ctypes.c_double, ctypes.c_double, _D, ctypes.c_double, _D]
    L.oracle_decomp.argtypes = [_D, ctypes.c_int]
    L.oracle_decomp.restype = ctypes.c_int
    L.oracle_solve.argtypes = [_D, _D, ctypes.c_int]
    L.oracle_resolve.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.POINTER(OptsIn), ctypes.POINTER(OptsOut)]
    L.oracle_resolve.restype = ctypes.c_int
    L.oracle_ros3_cell.argtypes = [_D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.POINTER(OptsOut),
                                   ctypes.c_double, _I, _I, _I, _D]
    L.oracle_ros3_cell.restype = ctypes.c_int
    L.oracle_ros_cell.argtypes = ([ctypes.c_int, _D, _D, _D, _D]
                                  + L.oracle_ros3_cell.argtypes[2:])
    L.oracle_ros_cell.restype = ctypes.c_int
    L.oracle_method_coeffs.argtypes = [ctypes.c_int] + [_D] * 7
    L.oracle_method_coeffs.restype = None
    L.oracle_photo_at.argtypes = [_D, _D, _D, ctypes.c_double, ctypes.c_double]
    L.oracle_step_control.argtypes = [ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(OptsOut), _D, _D, _I, _I]
    L.oracle_step_control.restype = ctypes.c_int
    L.oracle_ros_attempt.argtypes = [ctypes.c_int, _D, _D] + [ctypes.c_double] * 5 + [_D, _D]
    L.oracle_ros_attempt.restype = ctypes.c_double
    L.oracle_ros_attempt_tab.argtypes = ([ctypes.c_int, _D, _D, _D, _D, _I, ctypes.c_double,
                                          _D, _D] + [ctypes.c_double] * 5 + [_D, _D])
    L.oracle_ros_attempt_tab.restype = ctypes.c_double
    L.oracle_photo_at.restype = None
    L.oracle_integrate.argtypes = [ctypes.c_int64, ctypes.c_int64, _D, _D, _D, _D, _D, _D, _D,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(OptsIn), _D, _I, _I, _I, _I, _D,
                                   ctypes.c_int]
    L.oracle_integrate.restype = ctypes.c_int
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = load(LIB)
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.float64:
        return a.ctypes.data_as(_D)
    assert a.dtype == np.int32
    return a.ctypes.data_as(_I)


def rconst(T, M, J, rc):
    k = np.zeros(156)
    lib().oracle_rconst(T, M, _p(np.ascontiguousarray(J, dtype=np.float64)), rc, _p(k))
    return k


def fun(V, h2o, m, k):
    out = np.zeros(59)
    lib().oracle_fun(_p(np.ascontiguousarray(V)), h2o, m, _p(np.ascontiguousarray(k)), _p(out))
    return out


def jac(V, h2o, m, k):
    out = np.zeros(659)
    lib().oracle_jac(_p(np.ascontiguousarray(V)), h2o, m, _p(np.ascontiguousarray(k)), _p(out))
    return out


def jac_prep(V, h2o, m, k, ghinv):
    out = np.zeros(659)
    lib().oracle_jac_prep(_p(np.ascontiguousarray(V)), h2o, m, _p(np.ascontiguousarray(k)),
                          ghinv, _p(out))
    return out


def decomp(G, ref_division=False):
    G = np.array(G, dtype=np.float64)
    rc = lib().oracle_decomp(_p(G), 1 if ref_division else 0)
    return rc, G


def solve(LU, b, ref_division=False):
    x = np.array(b, dtype=np.float64)
    lib().oracle_solve(_p(np.ascontiguousarray(LU)), _p(x), 1 if ref_division else 0)
    return x


def resolve(dt, rtol, atol, opts=None):
    out = OptsOut()
    rc = lib().oracle_resolve(dt, rtol, atol, ctypes.byref(opts) if opts else None, ctypes.byref(out))
    return rc, out


def method_coeffs(method):
    """The oracle's tables of KPP method `method` as numpy arrays."""
    arr = [np.zeros(15), np.zeros(15), np.zeros(6), np.zeros(6), np.zeros(6), np.zeros(6),
           np.zeros(2)]
    lib().oracle_method_coeffs(method, *[_p(a) for a in arr])
    return dict(zip(("A", "C", "M", "E", "alpha", "gamma", "g_elo"), arr),
                S=lib().oracle_method_stages(method))


def step_control(err_sum, elo, opts, H, T, rej_last, rej_more):
    """oracle_step_control on one attempt: (accepted, H, T, rej_last, rej_more)."""
    h, t = ctypes.c_double(H), ctypes.c_double(T)
    rl, rm = ctypes.c_int32(rej_last), ctypes.c_int32(rej_more)
    acc = lib().oracle_step_control(err_sum, elo, ctypes.byref(opts), ctypes.byref(h),
                                    ctypes.byref(t), ctypes.byref(rl), ctypes.byref(rm))
    return acc, h.value, t.value, rl.value, rm.value


def ros_attempt(method, y, k, h2o, m, H, rtol=1e-3, atol=1.0, table=None):
    """One autonomous attempt (oracle_ros_attempt, or oracle_ros_attempt_tab
    with `table` = method_coeffs()-style dict plus "newf"): (err_sum, Ynew,
    sum_j E_j K_j)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    k = np.ascontiguousarray(k, dtype=np.float64)
    yn, ye = np.zeros(59), np.zeros(59)
    if table is None:
        e = lib().oracle_ros_attempt(method, _p(y), _p(k), h2o, m, H, rtol, atol, _p(yn), _p(ye))
    else:
        c = {q: np.ascontiguousarray(table[q], dtype=np.float64) for q in "ACME"}
        nf = np.ascontiguousarray(table["newf"], dtype=np.int32)
        e = lib().oracle_ros_attempt_tab(int(table["S"]), _p(c["A"]), _p(c["C"]), _p(c["M"]),
                                         _p(c["E"]), _p(nf), float(table["g_elo"][0]), _p(y),
                                         _p(k), h2o, m, H, rtol, atol, _p(yn), _p(ye))
    return e, yn, ye


def integrate(cells, dt, rtol, atol, opts=None, h0=None, nthreads=0, n=None, use_rc=True,
              jrate_end=None, library=None):
    """Integrate make_cells()-style host arrays; returns (conc, counters dict, code).
    jrate_end: J rates at the end of the interval (non-autonomous), or None.
    library: an oracle library from load() (default: the test build)."""
    conc = np.array(cells["conc"], dtype=np.float64, order="C")
    ld = conc.shape[1]
    n = ld if n is None else n
    ns = np.zeros(ld, np.int32)
    na = np.zeros(ld, np.int32)
    nr = np.zeros(ld, np.int32)
    st = np.zeros(ld, np.int32)
    hx = np.zeros(ld)
    rc_arr = np.ascontiguousarray(cells["rc_n2o5"]) if use_rc else None
    h0a = None if h0 is None else np.ascontiguousarray(h0, dtype=np.float64)
    j1 = None if jrate_end is None else np.ascontiguousarray(jrate_end, dtype=np.float64)
    code = (library or lib()).oracle_integrate(
        n, ld, _p(conc), _p(np.ascontiguousarray(cells["temp"])),
        _p(np.ascontiguousarray(cells["cair"])), _p(np.ascontiguousarray(cells["h2o"])),
        _p(np.ascontiguousarray(cells["jrate"])), _p(j1), _p(rc_arr), dt, rtol, atol,
        ctypes.byref(opts) if opts else None, _p(h0a), _p(ns), _p(na), _p(nr), _p(st), _p(hx),
        nthreads)
    return conc, dict(nstep=ns, nacc=na, nrej=nr, status=st, hexit=hx), code


def wrf_convert(grid, phy, moist, ph, chem, rad, rc_fixed=False, ld=None):
    """oracle_wrf_convert: grid = (nx, ny, nz, ox, oy, oz, bx, by, bz), WRF
    arrays [field][ny][nz][nx] (float64); returns the solver inputs."""
    L = lib()
    L.oracle_wrf_convert.restype = None
    g = np.ascontiguousarray(grid, dtype=np.int32)
    n = int(g[6]) * int(g[7]) * int(g[8])
    ld = n if ld is None else ld
    out = dict(conc=np.zeros((59, ld)), temp=np.zeros(ld), cair=np.zeros(ld), h2o=np.zeros(ld),
               jrate=np.zeros((21, ld)), rc_n2o5=np.zeros(ld))
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (phy, moist, ph, chem, rad)]
    L.oracle_wrf_convert(_p(g), *[_p(a) for a in arrs], ctypes.c_int(1 if rc_fixed else 0),
                         ctypes.c_int64(ld), _p(out["conc"]), _p(out["temp"]), _p(out["cair"]),
                         _p(out["h2o"]), _p(out["jrate"]), _p(out["rc_n2o5"]))
    return out
