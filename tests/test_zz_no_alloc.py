"""radm2_integrate_noalloc against radm2_integrate through the C ABI.

Two calls in a row, the second on a smaller batch: each must give
radm2_integrate's results bit for bit (end states, step/accept/reject
counts, status, exit step), and neither may take device memory (the
device's free memory, read before and after, does not drop).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(cells):
    import radm2_amd
    from radm2_amd import synth
    dev = synth.to_device(cells)
    return radm2_amd.CellBatch(dev["conc"], dev["temp"], dev["cair"], dev["h2o"], dev["jrate"],
                               rc_n2o5=dev["rc_n2o5"])


def _call(lib, ctx, fn, b, dt, rtol, atol):
    import radm2_amd
    p = radm2_amd._ptr
    code = getattr(lib, fn)(ctx, b.ncells, b.ld, p(b.conc), p(b.temp), p(b.cair), p(b.h2o),
                            p(b.jrate), p(b.rc_n2o5), dt, rtol, atol, None, None, p(b.nstep),
                            p(b.nacc), p(b.nrej), p(b.status), p(b.hexit))
    assert code == 0, radm2_amd.strerror(code)
    assert lib.radm2_sync(ctx) == 0


def _out(b):
    return dict(conc=b.conc.cpu().numpy(), nstep=b.nstep.cpu().numpy(),
                nacc=b.nacc.cpu().numpy(), nrej=b.nrej.cpu().numpy(),
                status=b.status.cpu().numpy(), hexit=b.hexit.cpu().numpy())


def test_noalloc_matches_integrate_on_two_calls():
    import torch
    import radm2_amd
    from radm2_amd import synth
    lib = radm2_amd.load_library()
    ctx = ctypes.c_void_p()
    assert lib.radm2_create(ctypes.byref(ctx), 0, None) == 0
    ref_ctx = ctypes.c_void_p()
    assert lib.radm2_create(ctypes.byref(ref_ctx), 0, None) == 0
    try:
        # load the kernels' code objects onto the device before any
        # measurement (the runtime may do that at a kernel's first launch)
        warm = _batch(synth.make_cells(64, 60, 100.0, 0.05))
        _call(lib, ctx, "radm2_integrate_noalloc", warm, 60.0, 1e-3, 1.0)
        _call(lib, ref_ctx, "radm2_integrate", _batch(synth.make_cells(64, 60, 100.0, 0.05)),
              60.0, 1e-3, 1.0)
        for n, seed, dt in ((20000, 61, 600.0), (3000, 62, 300.0)):
            cells = synth.make_cells(n, seed, 100.0, 0.05)
            a, b = _batch(cells), _batch(cells)
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            _call(lib, ctx, "radm2_integrate_noalloc", a, dt, 1e-3, 1.0)
            free1 = torch.cuda.mem_get_info()[0]
            _call(lib, ref_ctx, "radm2_integrate", b, dt, 1e-3, 1.0)
            got, ref = _out(a), _out(b)
            print("noalloc: %d cells, free device memory %d -> %d B, mean steps %.1f"
                  % (n, free0, free1, got["nstep"].mean()))
            assert free1 >= free0
            assert (ref["status"] == 1).all()
            for k in ref:
                np.testing.assert_array_equal(got[k], ref[k], err_msg="%s (n=%d)" % (k, n))
    finally:
        lib.radm2_destroy(ctx)
        lib.radm2_destroy(ref_ctx)
