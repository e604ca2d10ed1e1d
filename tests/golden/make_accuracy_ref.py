#!/usr/bin/env python3
"""Accuracy references: the CPU oracle at rtol 1e-10, atol 1e-3 molec/cm3
on the first cells of the benchmark workloads (dt 600 s).

bench.py's accuracy leg measures every GPU path's end states against these
arrays (it reads them as data; the oracle itself does not run there).  The
workloads are the 1,048,576-cell sets of synth.make_cells with zenith
U(0, 100) deg and 5% plume cells, generated in full and sliced, so that the
cells are exactly the first cells the benchmark integrates:

  seed 0  the plain / persistent / learned-h0 timing workload, 4,096 cells
  seed 2  the learned-h0 held-out evaluation set, 2,048 cells
  seed 3  the surrogate evaluation set, 2,048 cells

Run from the repository root (about 35 minutes on 8 host cores):
    python tests/golden/make_accuracy_ref.py
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "radm2-solver-multi-gpu_amd"))

import oracle_ct as O  # noqa: E402
from radm2_amd import synth  # noqa: E402

SETS = ((0, 4096), (2, 2048), (3, 2048))
DT, RTOL, ATOL = 600.0, 1e-10, 1e-3


def main():
    out = dict(dt=DT, rtol=RTOL, atol=ATOL)
    for seed, n in SETS:
        full = synth.make_cells(1048576, seed, 100.0, 0.05)
        sub = {k: np.ascontiguousarray(full[k][..., :n])
               for k in ("conc", "temp", "cair", "h2o", "jrate", "rc_n2o5")}
        del full
        t = time.time()
        co, oo, code = O.integrate(sub, DT, RTOL, ATOL)
        assert code == 0 and (oo["status"] == 1).all()
        print("seed %d: %d cells in %.0f s, mean steps %.0f" % (seed, n, time.time() - t,
                                                              oo["nstep"].mean()), flush=True)
        out["s%d_conc" % seed] = co
        out["s%d_nstep" % seed] = oo["nstep"]
        # the initial states' digest: a reader checks that it integrates the same cells
        out["s%d_conc0_sha256" % seed] = np.array(hashlib.sha256(sub["conc"].tobytes()).hexdigest())
    dst = os.path.join(HERE, "accuracy_ref.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main()
