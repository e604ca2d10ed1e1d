"""bench.py on the CPU: more than one GPU is refused with a message (the
benchmark measures one GPU) before anything touches a device, and the
per-cell step histograms bin as a linear scan over the edges does."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_step_histograms_bin_like_a_linear_scan():
    import importlib.util
    import torch
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    B = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(B)
    for edges in (B.ATTEMPT_EDGES, B.REJECT_EDGES):
        x = torch.tensor([0, 1, 2, 24, 25, 174, 175, 176, 999, 1000, 5000] + list(edges), dtype=torch.int32)
        h = B.histogram(x, edges, torch)
        want = [0] * len(edges)
        for v in x.tolist():
            want[max(i for i, e in enumerate(edges) if e <= v)] += 1
        assert h == {"edges": list(edges), "counts": want}
