import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "radm2-solver-multi-gpu_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def solver():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import radm2_amd
    s = radm2_amd.Solver(0)
    yield s
    s.close()
