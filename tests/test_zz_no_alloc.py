"""hpup_plan_adjoint_stage_noalloc (the gradient stage with its partial sums
in the library's static device array) against hpup_plan_adjoint_stage, both
called through the C ABI on the same inputs: the stage output, the running
sum and the five parameter terms agree to the last bit (same kernels, same
fixed-order reductions), on two calls in a row, the second on a smaller
plan."""
import ctypes

import numpy as np
import pytest

import hpup

pytestmark = pytest.mark.gpu


def _rand(n, dtype, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.complex(torch.randn(n, generator=g, dtype=torch.float64),
                         torch.randn(n, generator=g, dtype=torch.float64)).to(dtype).cuda()


def _stage(fn, plan, ybar, y, scale, fbar_init):
    import torch
    out = torch.empty_like(ybar)
    fbar = _rand(plan.ncoef, plan.cdtype, 99)
    grad = torch.full((5,), 0.25, dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(plan._h, ctypes.c_void_p(ybar.data_ptr()), ctypes.c_void_p(y.data_ptr()), float(scale),
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(fbar.data_ptr()), 1 if fbar_init else 0,
            ctypes.c_void_p(grad.data_ptr()), stream)
    hpup._check(rc)
    torch.cuda.synchronize()
    return out, fbar, grad


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_stage_noalloc_matches(precision):
    import torch
    L = hpup.lib()
    for k, (B, N, init) in enumerate([(5, 12, False), (3, 8, True)]):  # second call: a smaller input
        print("noalloc stage", precision, B, N, flush=True)
        p = hpup.Plan(B, 2, N, 8.0, b=(0.3, 0.2, 0.0), G=np.array([[0.12, 0.03], [0.05, 0.18]]), dt=0.01,
                      method="rk4", precision=precision, operator=1)
        ybar, y = _rand(p.ncoef, p.cdtype, 10 + k), _rand(p.ncoef, p.cdtype, 20 + k)
        ref = _stage(L.hpup_plan_adjoint_stage, p, ybar, y, 0.3, init)
        got = _stage(L.hpup_plan_adjoint_stage_noalloc, p, ybar, y, 0.3, init)
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
        assert float(ref[2].abs().max()) > 0.25  # the five terms were added
