"""WRF-Chem input adapter on the GPU (radm2_wrf_convert) against the oracle's
dev_convert restatement, on the reference's own grid (39 x 19 x 39 cells of a
48 x 20 x 48 WRF grid with a 4-cell halo), and converted cells integrated
end to end."""
import numpy as np
import pytest

import oracle_ct as O

pytestmark = pytest.mark.gpu


def _wrf_inputs(seed, nx=48, ny=48, nz=20):
    rng = np.random.default_rng(seed)
    fs = (ny, nz, nx)
    T = rng.uniform(255, 305, fs)
    P = rng.uniform(5.0e4, 1.02e5, fs)
    rho = P / (287.05 * T)
    phy = np.stack([P, T, rho])
    moist = np.stack([np.zeros(fs), rng.uniform(1e-4, 2e-2, fs)])
    ph = rng.uniform(0, 1e-2, (28,) + fs) * 60.0 * rng.uniform(0, 1, fs)
    chem = np.exp(rng.uniform(np.log(1e-7), np.log(0.1), (44,) + fs))
    chem[:, rng.random(fs) < 0.05] *= -1.0            # a few negative mixing ratios
    rad = np.exp(rng.uniform(np.log(1e-12), np.log(1e-6), (28,) + fs))
    return phy, moist, ph, chem, rad


@pytest.mark.parametrize("rc_fixed", [False, True])
def test_wrf_convert_matches_oracle_on_reference_grid(rc_fixed):
    import torch
    from radm2_amd import wrf
    arrs = _wrf_inputs(5)
    g = wrf.grid(**wrf.REFERENCE_GRID)
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]
    out = wrf.convert(g, *dev, rc_fixed=rc_fixed)
    torch.cuda.synchronize()
    ref = O.wrf_convert(tuple(wrf.REFERENCE_GRID[k] for k in ("nx", "ny", "nz", "ox", "oy", "oz",
                                                                "bx", "by", "bz")),
                        *arrs, rc_fixed=rc_fixed)
    for k in ("conc", "temp", "cair", "h2o", "jrate"):
        np.testing.assert_array_equal(out[k].cpu().numpy(), ref[k], err_msg=k)
    # device vs host exp/pow: a few ulp
    np.testing.assert_allclose(out["rc_n2o5"].cpu().numpy(), ref["rc_n2o5"], rtol=1e-13)
    assert ref["conc"].shape == (59, 39 * 19 * 39)


def test_converted_cells_integrate_like_the_oracle(solver):
    import torch
    import radm2_amd
    from radm2_amd import wrf
    arrs = _wrf_inputs(6, nx=12, ny=12, nz=4)
    g = wrf.grid(12, 12, 4, 2, 2, 0, 8, 8, 4)          # 256 cells
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]
    c = wrf.convert(g, *dev, rc_fixed=True)
    b = radm2_amd.CellBatch(c["conc"], c["temp"], c["cair"], c["h2o"], c["jrate"],
                            rc_n2o5=c["rc_n2o5"])
    host = {k: v.cpu().numpy() for k, v in c.items()}
    solver.integrate(b, 600.0, 1e-3, 1.0)
    solver.sync()
    co, oo, code = O.integrate(host, 600.0, 1e-3, 1.0)
    assert code == 0
    st = b.status.cpu().numpy()
    assert (st == 1).all()
    cg = b.conc.cpu().numpy()
    big = np.abs(co) > 1e3
    rel = np.where(big, np.abs(cg - co) / np.maximum(np.abs(co), 1e-300), 0.0).max(axis=0)
    assert (rel <= 1e-6).all()
