"""Seeded synthetic RADM2 cells (the benchmark and test workloads).

The reference reads its cells from WRF-Chem dumps that it does not ship
(mgpu.cu:343-417) and converts them in dev_convert (mgpu.cu:539-584):
M from air density, H2O from qv, J = J_wrf/60, concentrations from ppmv.
This generator produces the same solver inputs directly, per cell:

  T ~ U(260, 300) K; P ~ U(500, 1013) hPa; M = P/(k_B T) (molec/cm3)
  x_H2O log-uniform 1e-3..2e-2, H2O = M x_H2O
  solar zenith angle ~ U(0, zmax) deg; J_i = J_i^clear * max(cos(SZA), 0)
    with zmax = 80 (4,096-cell case) or 100 (the 1M-cell and later cases;
    cos(SZA) <= 0 for SZA >= 90 deg, so 10% of cells are night cells with
    J = 0 -- the workload text says "~20%", which U(0,100) does not give)
  every variable species: M x with x log-uniform 1e-12..1e-7, except
    O3 U(20, 80) ppb, NOx log-uniform 0.1..20 ppb (NO share U(0.1, 0.5) of
    NOx), CO U(80, 300) ppb
  optional plume cells (a given fraction): NOx log-uniform 20..500 ppb and
    the emitted VOCs x10
  rc_n2o5 from the reference's own formula (mgpu.cu:546-552) evaluated at
    the cell's relative humidity.

The clear-sky photolysis table J^clear is this project's choice (the
reference takes J from WRF); it lives in ``mech/j_clear.json``.
"""
import json
import os

import numpy as np

from . import IND, NJ, NVAR

HERE = os.path.dirname(os.path.abspath(__file__))
J_CLEAR_PATH = os.path.join(os.path.dirname(HERE), "mech", "j_clear.json")
KB = 1.380649e-23
# emitted hydrocarbons and carbonyls scaled in plume cells
VOC = ("ETH", "HC3", "HC5", "HC8", "OL2", "OLT", "OLI", "ISO", "TOL", "XYL",
       "HCHO", "ALD", "KET")


def j_clear():
    with open(J_CLEAR_PATH) as f:
        d = json.load(f)
    v = np.array([d["rates"][k] for k in d["order"]], dtype=np.float64)
    assert v.shape == (NJ,)
    return v


def rc_n2o5_reference(T, P_pa, x_h2o):
    """mgpu.cu:546-552 as written (saturation vapour pressure, RH clamp,
    and the misplaced parenthesis that puts +300 inside the exponential)."""
    es = 1000.0 * 0.6112 * np.exp(17.67 * (T - 273.15) / (T - 29.65))
    qvs = es / (P_pa - es)
    qv = x_h2o * (2 * 1.0079 + 15.9994) / 28.97
    rh = np.clip(qv / qvs, 0.0, 1.0)
    return 1.0 / (3.6e4 * np.exp(-np.power(rh / 0.28, 2.8) + 300.0))


def make_cells(n, seed, zmax_deg=80.0, plume_frac=0.0):
    """Return a dict of host float64 arrays in solver layout:
    conc [59][n], temp/cair/h2o/rc_n2o5 [n], jrate [21][n], plus sza, plume."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(260.0, 300.0, n)
    P = rng.uniform(500.0, 1013.0, n) * 100.0
    M = P / (KB * T) * 1e-6
    xw = np.exp(rng.uniform(np.log(1e-3), np.log(2e-2), n))
    H2O = M * xw
    sza = rng.uniform(0.0, zmax_deg, n)
    mu = np.maximum(np.cos(np.deg2rad(sza)), 0.0)
    J = j_clear()[:, None] * mu[None, :]
    x = np.exp(rng.uniform(np.log(1e-12), np.log(1e-7), (NVAR, n)))
    x[IND["O3"]] = rng.uniform(20e-9, 80e-9, n)
    nox = np.exp(rng.uniform(np.log(0.1e-9), np.log(20e-9), n))
    fno = rng.uniform(0.1, 0.5, n)
    x[IND["CO"]] = rng.uniform(80e-9, 300e-9, n)
    plume = np.zeros(n, dtype=bool)
    if plume_frac > 0:
        plume = rng.uniform(0.0, 1.0, n) < plume_frac
        nox_p = np.exp(rng.uniform(np.log(20e-9), np.log(500e-9), n))
        nox = np.where(plume, nox_p, nox)
        for s in VOC:
            x[IND[s]] = np.where(plume, 10.0 * x[IND[s]], x[IND[s]])
    x[IND["NO"]] = nox * fno
    x[IND["NO2"]] = nox * (1.0 - fno)
    conc = M[None, :] * x
    return dict(conc=np.ascontiguousarray(conc), temp=T, cair=M, h2o=H2O,
                jrate=np.ascontiguousarray(J), rc_n2o5=rc_n2o5_reference(T, P, xw),
                sza=sza, plume=plume, pressure=P)


# The benchmark configurations (cells, seed, zenith range, plume fraction)
CONFIGS = {
    "c0_4k": dict(n=4096, seed=0, zmax_deg=80.0, plume_frac=0.0),
    "c1_1m": dict(n=1048576, seed=0, zmax_deg=100.0, plume_frac=0.05),
}
# the batched 64-domain workload: make_domains()


def make_domains(seed=4, ndom=64, lo=1000, hi=200000, dts=(60.0, 300.0, 600.0, 3600.0),
                 rtols=(1e-2, 1e-3, 1e-4)):
    """The batched workload: `ndom` domains with log-uniform sizes in
    [lo, hi] cells (2.4 million in expectation for the defaults; 3.3 million
    for seed 4), each with
    its own dt drawn from `dts` and rtol from `rtols` (atol 1), cell states
    as in the 1M-cell case (zenith U(0,100), 5% plume cells).  Returns a
    list of (cells, dt, rtol, atol)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), ndom)).astype(np.int64)
    out = []
    for d in range(ndom):
        cells = make_cells(int(sizes[d]), seed * 1000 + d, 100.0, 0.05)
        out.append((cells, float(rng.choice(dts)), float(rng.choice(rtols)), 1.0))
    return out


def to_device(cells, device="cuda"):
    """Copy make_cells() output to device tensors."""
    import torch
    out = {}
    for k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5"):
        out[k] = torch.from_numpy(np.ascontiguousarray(cells[k])).to(device)
    return out
