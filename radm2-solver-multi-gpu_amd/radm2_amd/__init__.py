"""Python host mirror of the RADM2 integrator's C ABI (include/radm2.h).

The solver itself is the HIP library ``libradm2_amd.so`` built next to this
file; this module only binds it with ctypes and moves torch device tensors'
pointers across.  There is no CPU path: if the library is missing, importing
the bindings raises.

Reference interface mirrored (Rizkyanno/RADM2-Solver-Multi-GPU, mgpu.cu):
  ``Rosenbrock(Tstart, Tend, AbsTol, RelTol, RPAR, IPAR)``  mgpu.cu:3664-3848
  ``INTEGRATE()`` default call                              mgpu.cu:3850-3889
  KPP return codes and messages (``ros_ErrorMsg``)          mgpu.cu:976-1005
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libradm2_amd.so")

NVAR = 59
NFIX = 2
NREACT = 156
NJ = 21
LU_NNZ = 659

OK = 0
E_ARG = -101
E_HIP = -102
E_UNSUPPORTED = -103
CELL_OK = 1
CELL_TOO_MANY_STEPS = -6
CELL_STEP_TOO_SMALL = -7
CELL_SINGULAR = -8
SCHED_GRID = 0
SCHED_PERSISTENT = 1
PLACE_PRIVATE = 0
PLACE_WORKSPACE = 1
PLACE_LDS160 = 2
PLACE_LDS312 = 3
PLACE_LDS160_REG112 = 4
PLACE_LDS160_REG128 = PLACE_LDS160_REG112   # former name
PLACE_SPLIT2 = 5
PLACE_SPLIT4 = 6

# reference species order (mgpu.cu:125-183)
SPECIES = (
    "SULF ORA1 ORA2 CO2 SO2 ETH O1D HC5 HC8 TOL XYL TPAN HONO H2O2 N2O5 HC3 CH4 PAA O3P "
    "HNO4 OP1 CSL PAN OL2 HNO3 CO ISO OLT OLI DCB GLY XNO2 KET MGLY TOLP XYLP OLTP OLN XO2 "
    "OL2P HC5P OP2 HCHO HC8P TCO3 O3 ONIT ALD OLIP KETP HO2 MO2 OH NO3 ACO3 HC3P ETHP NO NO2"
).split()
IND = {s: i for i, s in enumerate(SPECIES)}


class Radm2Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("radm2 error %d: %s" % (code, msg))
        self.code = code


class Opts(ctypes.Structure):
    """radm2_opts: RPAR[0..6], IPAR[2..3] of the reference's Rosenbrock(), and
    sing_retries (KPP's step halvings after a singular matrix; 0 = -8 at
    once, the reference's active code)."""
    _fields_ = [("hmin", ctypes.c_double), ("hmax", ctypes.c_double),
                ("hstart", ctypes.c_double), ("facmin", ctypes.c_double),
                ("facmax", ctypes.c_double), ("facrej", ctypes.c_double),
                ("facsafe", ctypes.c_double), ("max_steps", ctypes.c_int32),
                ("method", ctypes.c_int32), ("sing_retries", ctypes.c_int32)]


class Domain(ctypes.Structure):
    _fields_ = [("ncells", ctypes.c_int64), ("ld", ctypes.c_int64),
                ("conc", ctypes.c_void_p), ("temp", ctypes.c_void_p),
                ("cair", ctypes.c_void_p), ("h2o", ctypes.c_void_p),
                ("jrate", ctypes.c_void_p), ("rc_n2o5", ctypes.c_void_p),
                ("h0", ctypes.c_void_p), ("nstep", ctypes.c_void_p),
                ("nacc", ctypes.c_void_p), ("nrej", ctypes.c_void_p),
                ("status", ctypes.c_void_p), ("hexit", ctypes.c_void_p),
                ("hfirst", ctypes.c_void_p), ("dt", ctypes.c_double),
                ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
                ("opts", ctypes.POINTER(Opts)), ("jrate_end", ctypes.c_void_p)]


class WrfGrid(ctypes.Structure):
    """radm2_wrf_grid: input grid (nx, ny, nz), box offset and box size."""
    _fields_ = [(n, ctypes.c_int) for n in ("nx", "ny", "nz", "ox", "oy", "oz", "bx", "by", "bz")]


_EXPORTS = [
    ("radm2_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_void_p]),
    ("radm2_destroy", None, [ctypes.c_void_p]),
    ("radm2_set_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("radm2_set_kernel", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("radm2_integrate", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64] + [ctypes.c_void_p] * 6
     + [ctypes.c_double] * 3 + [ctypes.POINTER(Opts)] + [ctypes.c_void_p] * 6),
    ("radm2_integrate_noalloc", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64] + [ctypes.c_void_p] * 6
     + [ctypes.c_double] * 3 + [ctypes.POINTER(Opts)] + [ctypes.c_void_p] * 6),
    ("radm2_integrate_batched", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Domain)]),
    ("radm2_plan_create", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Domain), ctypes.POINTER(ctypes.c_void_p)]),
    ("radm2_plan_launch", ctypes.c_int, [ctypes.c_void_p]),
    ("radm2_plan_destroy", None, [ctypes.c_void_p]),
    ("radm2_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("radm2_last_elapsed_ms", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("radm2_enable_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("radm2_read_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    ("radm2_measure_fp64_peak", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    ("radm2_check_opts", ctypes.c_int,
     [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.POINTER(Opts)]),
    ("radm2_rosenbrock", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(Domain), ctypes.c_double, ctypes.c_double,
      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]),
    ("radm2_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("radm2_wrf_convert", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(WrfGrid)] + [ctypes.c_void_p] * 5
     + [ctypes.c_int, ctypes.c_int64] + [ctypes.c_void_p] * 6),
]
EXPORTED_SYMBOLS = tuple(n for n, _, _ in _EXPORTS)

_lib = None


def load_library(path=None):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError("libradm2_amd.so not built at %s (run __graft_entry__.build())" % p)
    lib = ctypes.CDLL(p)
    for name, res, args in _EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def strerror(code):
    return load_library().radm2_strerror(int(code)).decode()


def _check(code):
    if code != 0:
        raise Radm2Error(code, strerror(code))


def check_opts(dt, rtol, atol, opts=None):
    """KPP option validation: 0 or IERR -1..-5 (mgpu.cu:3715-3811)."""
    return load_library().radm2_check_opts(dt, rtol, atol, ctypes.byref(opts) if opts else None)


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _need(t, dtype_name, shape_tail=None):
    import torch
    if not t.is_cuda:
        raise ValueError("radm2: arrays must be device tensors")
    want = {"f64": torch.float64, "i32": torch.int32}[dtype_name]
    if t.dtype != want:
        raise ValueError("radm2: expected %s, got %s" % (want, t.dtype))
    if not t.is_contiguous():
        raise ValueError("radm2: arrays must be contiguous")
    return t


class CellBatch:
    """Device-resident SoA arrays of one domain (cell index fastest).

    conc [59][ld], temp/cair/h2o [ld], jrate [21][ld], optional rc_n2o5 and
    h0 [ld] and jrate_end [21][ld] (J at the end of the interval: a
    non-autonomous integration); outputs nstep/nacc/nrej/status (int32) and
    hexit/hfirst (f64).
    """

    def __init__(self, conc, temp, cair, h2o, jrate, rc_n2o5=None, h0=None,
                 outputs=True, hfirst=False, jrate_end=None):
        import torch
        self.conc = _need(conc, "f64")
        if conc.dim() != 2 or conc.shape[0] != NVAR:
            raise ValueError("radm2: conc must be [59][ld], got %s" % (tuple(conc.shape),))
        self.ld = conc.shape[1]
        self.ncells = self.ld
        self.temp = self._cellwise(temp, "temp")
        self.cair = self._cellwise(cair, "cair")
        self.h2o = self._cellwise(h2o, "h2o")
        self.jrate = self._species(jrate, NJ, "jrate")
        self.rc_n2o5 = None if rc_n2o5 is None else self._cellwise(rc_n2o5, "rc_n2o5")
        self.h0 = h0
        self.jrate_end = None if jrate_end is None else self._species(jrate_end, NJ, "jrate_end")
        dev = conc.device
        if outputs:
            self.nstep = torch.zeros(self.ld, dtype=torch.int32, device=dev)
            self.nacc = torch.zeros_like(self.nstep)
            self.nrej = torch.zeros_like(self.nstep)
            self.status = torch.zeros_like(self.nstep)
            self.hexit = torch.zeros(self.ld, dtype=torch.float64, device=dev)
        else:
            self.nstep = self.nacc = self.nrej = self.status = self.hexit = None
        self.hfirst = torch.zeros(self.ld, dtype=torch.float64, device=dev) if hfirst else None

    def _cellwise(self, t, name, kind="f64"):
        """A per-cell array: device, contiguous, shape [ld]."""
        _need(t, kind)
        if tuple(t.shape) != (self.ld,):
            raise ValueError("radm2: %s must have shape (%d,), got %s"
                             % (name, self.ld, tuple(t.shape)))
        if t.device != self.conc.device:
            raise ValueError("radm2: %s is on %s, conc on %s" % (name, t.device, self.conc.device))
        return t

    def _species(self, t, rows, name):
        _need(t, "f64")
        if tuple(t.shape) != (rows, self.ld):
            raise ValueError("radm2: %s must have shape (%d, %d), got %s"
                             % (name, rows, self.ld, tuple(t.shape)))
        return t

    @property
    def h0(self):
        """Per-cell proposed initial step (f64 [ld]) or None."""
        return self._h0

    @h0.setter
    def h0(self, t):
        self._h0 = None if t is None else self._cellwise(t, "h0")

    def domain(self, dt, rtol, atol, opts=None, ncells=None):
        d = Domain()
        d.ncells = self.ncells if ncells is None else ncells
        d.ld = self.ld
        d.conc, d.temp, d.cair, d.h2o, d.jrate = (_ptr(x) for x in (
            self.conc, self.temp, self.cair, self.h2o, self.jrate))
        d.rc_n2o5, d.h0 = _ptr(self.rc_n2o5), _ptr(self.h0)
        d.nstep, d.nacc, d.nrej, d.status, d.hexit, d.hfirst = (_ptr(x) for x in (
            self.nstep, self.nacc, self.nrej, self.status, self.hexit, self.hfirst))
        d.dt, d.rtol, d.atol = float(dt), float(rtol), float(atol)
        d.opts = ctypes.pointer(opts) if opts is not None else None
        d.jrate_end = _ptr(self.jrate_end)
        if d.ncells < 0 or d.ncells > self.ld:
            raise ValueError("radm2: ncells %d outside 0..%d" % (d.ncells, self.ld))
        # the struct holds raw device pointers: keep the arrays (and the
        # options it points to) alive as long as it is
        d._owner = (self, opts)
        return d


class Solver:
    """A context on one device (radm2_create / radm2_destroy).

    By default the context runs on a torch stream of its own, ordered with
    torch's current stream on every enqueue: the solver stream first waits
    for torch's current stream (inputs written by torch operations are
    complete), then torch's current stream waits for the solver stream
    (torch operations on the outputs see the results).  Both waits are
    device-side.  With an explicit ``stream`` (a hipStream_t handle) the
    ordering is the caller's job, as in the C ABI.
    """

    def __init__(self, device=0, stream=None):
        import torch
        self.lib = load_library()
        self.ctx = ctypes.c_void_p()
        self.device = torch.device("cuda", int(device))
        self._ts = None
        if stream is None:
            self._ts = torch.cuda.Stream(device=self.device)
            stream = self._ts.cuda_stream
        _check(self.lib.radm2_create(ctypes.byref(self.ctx), int(device),
                                     ctypes.c_void_p(stream)))

    def _before(self):
        if self._ts is not None:
            import torch
            self._ts.wait_stream(torch.cuda.current_stream(self.device))

    def _after(self):
        if self._ts is not None:
            import torch
            torch.cuda.current_stream(self.device).wait_stream(self._ts)

    def close(self):
        if self.ctx:
            self.lib.radm2_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_schedule(self, schedule=SCHED_PERSISTENT, block=0, refill_min=0):
        _check(self.lib.radm2_set_schedule(self.ctx, schedule, block, refill_min))

    def set_kernel(self, placement=-1, waves_per_simd=0):
        _check(self.lib.radm2_set_kernel(self.ctx, placement, waves_per_simd))

    def integrate(self, batch, dt, rtol=1e-3, atol=1.0, opts=None, ncells=None):
        """Asynchronous integration of one domain (radm2_integrate_batched)."""
        d = batch.domain(dt, rtol, atol, opts, ncells)
        self._before()
        code = self.lib.radm2_integrate_batched(self.ctx, 1, ctypes.byref(d))
        self._after()
        _check(code)

    def integrate_single(self, batch, dt, rtol=1e-3, atol=1.0, opts=None):
        """The single-domain entry point radm2_integrate (same semantics as
        integrate(), which goes through radm2_integrate_batched)."""
        self._before()
        code = self.lib.radm2_integrate(
            self.ctx, batch.ncells, batch.ld, _ptr(batch.conc), _ptr(batch.temp),
            _ptr(batch.cair), _ptr(batch.h2o), _ptr(batch.jrate), _ptr(batch.rc_n2o5),
            float(dt), float(rtol), float(atol), ctypes.pointer(opts) if opts is not None else None,
            _ptr(batch.h0), _ptr(batch.nstep), _ptr(batch.nacc), _ptr(batch.nrej),
            _ptr(batch.status), _ptr(batch.hexit))
        self._after()
        _check(code)

    def integrate_noalloc(self, batch, dt, rtol=1e-3, atol=1.0, opts=None):
        """radm2_integrate_noalloc: radm2_integrate's results without any
        device allocation (the library's per-device array holds the
        workspace; overlapping calls on one GPU race)."""
        self._before()
        code = self.lib.radm2_integrate_noalloc(
            self.ctx, batch.ncells, batch.ld, _ptr(batch.conc), _ptr(batch.temp),
            _ptr(batch.cair), _ptr(batch.h2o), _ptr(batch.jrate), _ptr(batch.rc_n2o5),
            float(dt), float(rtol), float(atol), ctypes.pointer(opts) if opts is not None else None,
            _ptr(batch.h0), _ptr(batch.nstep), _ptr(batch.nacc), _ptr(batch.nrej),
            _ptr(batch.status), _ptr(batch.hexit))
        self._after()
        _check(code)

    def integrate_batched(self, domains):
        arr = (Domain * len(domains))(*domains)
        self._before()
        code = self.lib.radm2_integrate_batched(self.ctx, len(domains), arr)
        self._after()
        _check(code)

    def plan(self, domains):
        arr = (Domain * len(domains))(*domains)
        p = ctypes.c_void_p()
        _check(self.lib.radm2_plan_create(self.ctx, len(domains), arr, ctypes.byref(p)))
        return Plan(self, p, domains)

    def sync(self):
        _check(self.lib.radm2_sync(self.ctx))

    def elapsed_ms(self):
        ms = ctypes.c_float()
        _check(self.lib.radm2_last_elapsed_ms(self.ctx, ctypes.byref(ms)))
        return ms.value

    def enable_stats(self, on=True):
        _check(self.lib.radm2_enable_stats(self.ctx, 1 if on else 0))

    def measure_fp64_peak(self):
        """Measured attainable FP64 FMA throughput, TFLOP/s."""
        t = ctypes.c_double()
        _check(self.lib.radm2_measure_fp64_peak(self.ctx, ctypes.byref(t)))
        return t.value

    def read_stats(self):
        out = (ctypes.c_uint64 * 4)()
        _check(self.lib.radm2_read_stats(self.ctx, out))
        return dict(attempts=out[0], accepted=out[1], rejected=out[2], refills=out[3])

    def rosenbrock(self, batch, Tstart, Tend, AbsTol, RelTol, RPAR, IPAR):
        """KPP convention (mgpu.cu:3664): returns IERR; RPAR/IPAR updated."""
        d = batch.domain(Tend - Tstart, 1e-3, 1.0)
        at = (ctypes.c_double * len(AbsTol))(*AbsTol)
        rt = (ctypes.c_double * len(RelTol))(*RelTol)
        rp = (ctypes.c_double * 20)(*RPAR)
        ip = (ctypes.c_int32 * 20)(*IPAR)
        self._before()
        code = self.lib.radm2_rosenbrock(self.ctx, ctypes.byref(d), Tstart, Tend, at, rt, rp, ip)
        self._after()
        RPAR[:] = list(rp)
        IPAR[:] = list(ip)
        return code


class Plan:
    """A pre-validated batch captured as a HIP graph (radm2_plan_*).  It
    keeps the domains it was built from, and through them their batches'
    arrays, alive: the graph holds their device addresses."""

    def __init__(self, solver, handle, domains=()):
        self.solver = solver
        self.handle = handle
        self._domains = list(domains)

    def launch(self):
        self.solver._before()
        code = self.solver.lib.radm2_plan_launch(self.handle)
        self.solver._after()
        _check(code)

    def close(self):
        if self.handle:
            self.solver.lib.radm2_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
