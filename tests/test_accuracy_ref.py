"""The accuracy reference the benchmark and the full-size GPU tests compare
with (tests/golden/accuracy_ref.npz, written by make_accuracy_ref.py): its
initial-state digests are those of the first cells of the 1,048,576-cell
workloads the generator makes, and re-running the CPU oracle at rtol 1e-10
on a few of those cells reproduces the stored end states and step counts
bit for bit."""
import hashlib
import os

import numpy as np
import pytest

import oracle_ct as O

REF = os.path.join(os.path.dirname(__file__), "golden", "accuracy_ref.npz")


@pytest.mark.parametrize("seed", [0, 2, 3])
def test_reference_is_the_oracle_on_the_workloads_first_cells(seed):
    from radm2_amd import synth
    d = np.load(REF)
    ref = d["s%d_conc" % seed]
    m = ref.shape[1]
    assert m >= 2048 and float(d["rtol"]) == 1e-10 and float(d["atol"]) == 1e-3
    full = synth.make_cells(1048576, seed, 100.0, 0.05)
    sub = {k: np.ascontiguousarray(full[k][..., :m])
           for k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
    del full
    assert hashlib.sha256(sub["conc"].tobytes()).hexdigest() == str(d["s%d_conc0_sha256" % seed])
    few = {k: np.ascontiguousarray(v[..., :2]) for k, v in sub.items()}
    co, oo, code = O.integrate(few, float(d["dt"]), 1e-10, 1e-3)
    assert code == 0
    np.testing.assert_array_equal(co, ref[:, :2])
    np.testing.assert_array_equal(oo["nstep"], d["s%d_nstep" % seed][:2])
