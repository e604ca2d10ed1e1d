"""WRF-Chem input adapter (the reference's dev_convert, mgpu.cu:513-646).

`convert()` turns WRF-ordered device arrays ([field][ny][nz][nx]) of one box
of grid cells into the solver's structure-of-arrays inputs with the
library's HIP kernel (include/radm2.h: radm2_wrf_convert); the result feeds
CellBatch directly.  The reference's own case is the 39 x 19 x 39 interior
of a 48 x 20 x 48 grid at offsets (ox, oy, oz) = (4, 4, 0).
"""
import ctypes

from . import WrfGrid, _check, _ptr, load_library

NCHEM = 44   # WRF chemistry fields the shuffle reads (highest index 43)
NRAD = 28    # WRF radical fields the shuffle reads (highest index 27)
NPH = 28     # photolysis fields of the reference's jv (the first 21 are used)

REFERENCE_GRID = dict(nx=48, ny=48, nz=20, ox=4, oy=4, oz=0, bx=39, by=39, bz=19)


def grid(nx, ny, nz, ox=0, oy=0, oz=0, bx=None, by=None, bz=None):
    return WrfGrid(nx, ny, nz, ox, oy, oz, nx - ox if bx is None else bx,
                   ny - oy if by is None else by, nz - oz if bz is None else bz)


def convert(g, phy, moist, ph, chem, rad, rc_fixed=False, ld=None):
    """Device tensors (float64, contiguous, [field][ny][nz][nx]) -> dict of
    conc [59][ld], temp, cair, h2o, rc_n2o5 [ld] and jrate [21][ld]."""
    import torch
    n = g.bx * g.by * g.bz
    ld = n if ld is None else ld
    dev = phy.device
    out = dict(conc=torch.empty((59, ld), dtype=torch.float64, device=dev),
               temp=torch.empty(ld, dtype=torch.float64, device=dev),
               cair=torch.empty(ld, dtype=torch.float64, device=dev),
               h2o=torch.empty(ld, dtype=torch.float64, device=dev),
               jrate=torch.empty((21, ld), dtype=torch.float64, device=dev),
               rc_n2o5=torch.empty(ld, dtype=torch.float64, device=dev))
    fs = g.nx * g.ny * g.nz
    for name, t, nf in (("phy", phy, 3), ("moist", moist, 2), ("ph", ph, 21), ("chem", chem, NCHEM),
                        ("rad", rad, NRAD)):
        if t.dtype != torch.float64 or not t.is_contiguous() or t.numel() < nf * fs:
            raise ValueError("radm2 wrf: %s must be contiguous float64 with >= %d fields" % (name, nf))
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _check(load_library().radm2_wrf_convert(
        st, ctypes.byref(g), _ptr(phy), _ptr(moist), _ptr(ph), _ptr(chem), _ptr(rad),
        1 if rc_fixed else 0, ld, _ptr(out["conc"]), _ptr(out["temp"]), _ptr(out["cair"]),
        _ptr(out["h2o"]), _ptr(out["jrate"]), _ptr(out["rc_n2o5"])))
    return out
