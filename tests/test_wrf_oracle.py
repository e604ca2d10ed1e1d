"""The oracle's WRF-Chem input adapter (dev_convert, mgpu.cu:513-646)
against closed forms, the reference's box/halo indexing and the species
names of its P_* field defines."""
import json
import os

import numpy as np
import pytest

import oracle_ct as O

MECH = json.load(open(os.path.join(O.ROOT, "radm2-solver-multi-gpu_amd", "mech", "radm2.json")))
NA, MWAIR, MWH2O = 6.022e23, 28.97, 2 * 1.0079 + 15.9994


def _fields(rng, nx, ny, nz, rho=None, T=None, P=None, qv=None):
    fs = (ny, nz, nx)
    phy = np.stack([np.full(fs, 9.0e4) if P is None else np.full(fs, P),
                    rng.uniform(260, 300, fs) if T is None else np.full(fs, T),
                    rng.uniform(0.6, 1.3, fs) if rho is None else np.full(fs, rho)])
    moist = np.stack([np.zeros(fs), rng.uniform(1e-3, 1.5e-2, fs) if qv is None else np.full(fs, qv)])
    ph = rng.uniform(0, 1, (28,) + fs)
    chem = rng.uniform(-1e-3, 0.1, (44,) + fs)
    rad = rng.uniform(-1e-9, 1e-6, (28,) + fs)
    return phy, moist, ph, chem, rad


def test_unit_conversions_closed_form():
    rng = np.random.default_rng(0)
    phy, moist, ph, chem, rad = _fields(rng, 2, 2, 1, rho=1.2, T=288.0, P=1.0e5, qv=0.01)
    o = O.wrf_convert((2, 2, 1, 0, 0, 0, 2, 2, 1), phy, moist, ph, chem, rad)
    M = 1.2e-3 / MWAIR * NA
    assert o["cair"] == pytest.approx(np.full(4, M), rel=1e-15)
    assert o["h2o"] == pytest.approx(np.full(4, 1e-3 / MWH2O * NA * 0.01 * 1.2), rel=1e-15)
    np.testing.assert_array_equal(o["temp"], np.full(4, 288.0))
    np.testing.assert_allclose(o["jrate"], ph[:21].reshape(21, -1) / 60.0, rtol=1e-15)
    # ppmv -> molec/cm3 with negatives clamped to 0, through the shuffle
    for s, (src, fld) in enumerate(MECH["wrf"]["var_source"]):
        x = (rad if src else chem)[fld].reshape(-1)
        np.testing.assert_allclose(o["conc"][s], 1e-6 * M * np.maximum(x, 0.0), rtol=1e-15)
        assert (o["conc"][s][x < 0] == 0).all()


def test_n2o5_hydrolysis_as_written_and_fixed():
    rng = np.random.default_rng(1)
    T, P = 288.0, 1.0e5
    es = 611.2 * np.exp(17.67 * (T - 273.15) / (T - 29.65))
    for rh_target in (0.0, 0.3, 0.8, 1.5):
        qv = rh_target * es / (P - es)
        phy, moist, ph, chem, rad = _fields(rng, 1, 1, 1, rho=1.1, T=T, P=P, qv=qv)
        rh = min(max(rh_target, 0.0), 1.0)
        written = O.wrf_convert((1, 1, 1, 0, 0, 0, 1, 1, 1), phy, moist, ph, chem, rad)["rc_n2o5"][0]
        fixed = O.wrf_convert((1, 1, 1, 0, 0, 0, 1, 1, 1), phy, moist, ph, chem, rad, True)["rc_n2o5"][0]
        # misplaced parenthesis: exp(300 - (rh/0.28)^2.8) in the denominator
        assert written == pytest.approx(1.0 / (3.6e4 * np.exp(300.0 - (rh / 0.28) ** 2.8)), rel=1e-9)
        assert written < 1e-115
        assert fixed == pytest.approx(1.0 / (3.6e4 * np.exp(-(rh / 0.28) ** 2.8) + 300.0), rel=1e-9)
    # dry air: the fixed rate is 1/(3.6e4 + 300)
    phy, moist, ph, chem, rad = _fields(rng, 1, 1, 1, qv=0.0)
    fixed = O.wrf_convert((1, 1, 1, 0, 0, 0, 1, 1, 1), phy, moist, ph, chem, rad, True)["rc_n2o5"][0]
    assert fixed == pytest.approx(1.0 / 36300.0, rel=1e-15)


def test_box_and_halo_indexing_follow_dev_convert():
    # the reference's grid: 48 x 20 x 48 inputs, 39 x 19 x 39 cells at a
    # 4-cell halo in x and y; output p = (j*bz + k)*bx + i reads input
    # (j+4, k, i+4): fields that encode their own input index check it
    nx, ny, nz = 48, 48, 20
    bx, by, bz = 39, 39, 19
    J, K, I = np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij")
    code = (J * 1000 + K) * 1000 + I
    rng = np.random.default_rng(2)
    phy, moist, ph, chem, rad = _fields(rng, nx, ny, nz)
    phy[1] = code.astype(np.float64)              # temperature carries the index
    o = O.wrf_convert((nx, ny, nz, 4, 4, 0, bx, by, bz), phy, moist, ph, chem, rad)
    j, k, i = np.meshgrid(np.arange(by), np.arange(bz), np.arange(bx), indexing="ij")
    want = ((j + 4) * 1000 + k) * 1000 + (i + 4)
    np.testing.assert_array_equal(o["temp"], want.reshape(-1).astype(np.float64))


def test_chemistry_shuffle_matches_field_names():
    # every WRF chemistry field with a P_* name that is a mechanism species
    # (P_HO is OH) feeds exactly that species
    species = MECH["species"]
    alias = {"HO": "OH"}
    src = MECH["wrf"]["var_source"]
    matched = 0
    for name, fld in MECH["wrf"]["chem_names"].items():
        sp = alias.get(name.upper(), name.upper())
        if sp in species:
            assert src[species.index(sp)] == [0, fld], (name, fld)
            matched += 1
    assert matched >= 40
    # every species has one source; one field is read twice, as written:
    # OL2P, a peroxy radical, takes OLI's chemistry field (mgpu.cu:605, 645)
    seen = {}
    for s, x in enumerate(src):
        seen.setdefault(tuple(x), []).append(species[s])
    dup = {k: v for k, v in seen.items() if len(v) > 1}
    assert dup == {(0, MECH["wrf"]["chem_names"]["oli"]): ["OLI", "OL2P"]}
