"""The parts of bench.py's contract that run without a GPU: the single-GPU
exit, and the algorithmic work its roofline object divides by."""
import importlib.util
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_kernel_flops_of_the_workload():
    """Four real circulants (4 N^3 FLOPs each) per (r, k3) slab of the fused
    kernel and one per slab of the axis-3 kernel, over nRc * N slabs; nRc is
    the number of (l, m, n) with |m|, |n| <= l < B."""
    m = _bench()
    B, N = 15, 30
    nrc = sum((2 * l + 1) ** 2 for l in range(B))
    assert nrc == 4495
    assert m.kernel_flops("k_drift_fused", B, N) == 4 * 4 * N ** 3 * nrc * N == 58_255_200_000
    assert 4 * m.kernel_flops("k_axis3_mfma", B, N) == m.kernel_flops("k_drift_fused", B, N)
    assert m.kernel_flops("k_other", B, N) == 0
