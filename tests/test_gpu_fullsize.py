"""The reduced-model workloads at their full sizes, checked through properties
that do not depend on a CPU reference (the oracle's dense derivative takes
hours there):

* reduced model, B = 15, 32^2 omega grid over [-4, 4]^2, matrix-Fisher
  s = (8, 4, 2) x N(0, I), b = (0.2, 0.2), G = diag(0.1, 0.1), dt = 0.005, RK4:
  total probability F^0_00(k = 0) is unchanged to 1e-12 per step (FP64), the
  step is linear (1e-12), the transforms reproduce the coefficients
  (forward(backward(F)) = F, 1e-12), and the FP32 plan stays within 1e-5 of
  the FP64 plan after 20 steps;
* gradient sweep, B = 10, 32^2 omega grid, 50 steps of 0.01, checkpoint every
  10 steps: the F0 gradient of J = Re<w, F_K> satisfies the dot-product
  identity Re<w, P^K x> = Re<grad, x> (1e-11), the b_1 gradient matches central
  differences (1e-6), and the sweep does not depend on the checkpoint spacing.
"""
import numpy as np
import pytest

import hpup
from hpup.adjoint import Gradient

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _rel(a, b):
    return float((a - b).abs().pow(2).sum().sqrt() / b.abs().pow(2).sum().sqrt())


def _initial(plan, B, N, L, s, seed):
    from hpup import density as D
    rng = np.random.default_rng(seed)
    fR, fx = D.matrix_fisher_gaussian(B, 2, N, L, s, D.random_rotation(rng), Sigma=np.eye(2))
    return plan.forward(D.grid_density(fR, fx, plan.rdtype))


def test_reduced_model_full_size():
    import torch
    B, N, L, dt, K = 15, 32, 8.0, 0.005, 20
    _print("case reduced full size B=15 N=32")
    kw = dict(b=(0.2, 0.2, 0.0), G=np.eye(2) * 0.1, dt=dt, method="RK4")
    p64 = hpup.Plan(B, 2, N, L, precision="fp64", **kw)
    F0 = _initial(p64, B, N, L, (8, 4, 2), 15)
    # transforms: the coefficients of a grid density built from coefficients
    assert _rel(p64.forward(p64.backward(F0)), F0) < 1e-12
    # mass per step and linearity
    F = F0.clone()
    m0 = float(F[p64.mass_index].real)
    for _ in range(K):
        p64.step(F, 1)
        assert abs(float(F[p64.mass_index].real) - m0) <= 1e-12 * abs(m0)
    g = torch.Generator(device="cpu").manual_seed(16)
    x = torch.complex(torch.randn(p64.ncoef, generator=g, dtype=torch.float64),
                      torch.randn(p64.ncoef, generator=g, dtype=torch.float64)).cuda()
    a = 0.7 - 0.3j
    lhs = (a * F0 + x).contiguous()
    p64.step(lhs, 1)
    P0, Px = F0.clone(), x.clone()
    p64.step(P0, 1)
    p64.step(Px, 1)
    assert _rel(lhs, a * P0 + Px) < 1e-12
    # FP32 plan from the same coefficients
    p32 = hpup.Plan(B, 2, N, L, precision="fp32", **kw)
    F32 = F0.to(torch.complex64)
    p32.step(F32, K)
    err = _rel(F32.to(torch.complex128), F)
    _print("  fp32 vs fp64 after %d steps: %.2e" % (K, err))
    assert err < 1e-5


def test_gradient_sweep_full_size():
    import torch
    B, N, L, dt, K = 10, 32, 8.0, 0.01, 50
    _print("case gradient sweep full size B=10 N=32")
    bt, G = (0.3, 0.2), np.diag([0.12, 0.18])
    gr = Gradient(B, N, L, b=bt, G=G, dt=dt, method="rk4")
    F0 = _initial(gr.fwd, B, N, L, (20, 10, 5), 17)
    g = torch.Generator(device="cpu").manual_seed(18)
    rnd = lambda: torch.complex(torch.randn(gr.fwd.ncoef, generator=g, dtype=torch.float64),
                                torch.randn(gr.fwd.ncoef, generator=g, dtype=torch.float64)).cuda()
    w, x = rnd(), rnd()
    FK, gF0, grads = gr.run(F0, K, lambda F: w, checkpoint=10)
    # J(F0) = Re<w, P^K F0> is linear in F0: its gradient is (P^K)^H w
    Px = x.clone()
    gr.fwd.step(Px, K)
    lhs = float(torch.vdot(w, Px).real)
    rhs = float(torch.vdot(gF0, x).real)
    assert abs(lhs - rhs) <= 1e-11 * float(w.abs().pow(2).sum().sqrt() * Px.abs().pow(2).sum().sqrt())
    # the damping gradient against central differences
    h = 1e-4

    def J(b1):
        p = hpup.Plan(B, 2, N, L, b=(b1, bt[1], 0.0), G=G, dt=dt, method="rk4", precision="fp64")
        F = F0.clone()
        p.step(F, K)
        return float(torch.vdot(w, F).real)
    fd = (J(bt[0] + h) - J(bt[0] - h)) / (2 * h)
    _print("  dJ/db1", grads["b"][0], fd)
    assert abs(grads["b"][0] - fd) <= 1e-6 * max(abs(fd), 1e-3 * abs(J(bt[0])))
    # checkpoint spacing changes nothing
    _, gF1, grads1 = gr.run(F0, K, lambda F: w, checkpoint=1)
    assert torch.equal(gF1, gF0)
    assert np.array_equal(grads1["b"], grads["b"]) and np.array_equal(grads1["G"], grads["G"])


def test_hybrid_full_size():
    """The large hybrid workload (B = 30, 60^3 rotation grid, 20^3 omega grid
    over [-6, 6]^3, 0.6/0.4 matrix-Fisher x Gaussian mixture, reflection reset
    with eps = 0.8 on a tilted wall, RK4 dt = 0.003), FP32 plan: the hybrid
    step is linear in F (every part is: drift, transforms and the reset for a
    fixed eps; 1e-5 relative in FP32), the post-reset grid it returns
    transforms to the coefficients it returns (1e-6), and the reset changes
    the total probability by dt times the grid quadrature of df."""
    import importlib.util
    import os
    import torch
    _print("hybrid full size")
    spec = importlib.util.spec_from_file_location(
        "bench_module", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py"))
    bm = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bm)
    p = bm.hybrid_plan("fp32")
    F1 = bm.hybrid_initial(p)
    assert len(p.reset_rows) > 0
    F2 = torch.roll(F1, 7)  # a second, different coefficient set
    f = p.empty_grid()
    H1 = p.hybrid_step(F1.clone(), f, 1)[0]
    H2 = p.hybrid_step(F2.clone(), f, 1)[0]
    H12 = p.hybrid_step(F1 + 2 * F2, f, 1)[0]
    torch.cuda.synchronize()
    assert _rel(H12, H1 + 2 * H2) < 1e-5
    # F_{k+1} is the forward transform of the grid the step returns
    F = F1.clone()
    p.hybrid_step(F, f, 1)
    assert _rel(p.forward(f), F) < 1e-6
    # mass: the drift conserves it, the reset moves it by dt sum w df dv
    m = p.mass_index
    g = p.backward(p.step(F1.clone(), 1))
    df = p.reset(g, mode=1)
    from hpup import density as D
    wgt = D.beta_weights(p.B)[1]
    wj = torch.tensor(wgt, device="cuda").view(1, 1, 2 * p.B, 1)
    dmass = float((df.double().view(-1, 2 * p.B, 2 * p.B, 2 * p.B) * wj).sum())
    F = F1.clone()
    p.hybrid_step(F, f, 1)
    torch.cuda.synchronize()
    got = float(F[m].real - F1[m].real)
    assert abs(got - p.model.dt * dmass) <= 1e-4 * abs(F1[m].real) + 1e-3 * abs(got)
