"""The 64-point MFMA drift kernel (reduced model, omega grids of 33 to 64
points per axis) against the oracle, including the reduced hybrid driver's
own grid N = 60 (pendulum_reduced_hybrid.m:58-59) and the edges N = 34
(two 16-wide tiles past 32) and N = 64 (no padding).

Checkers: the dense oracle (the reference's literal algorithm) on every
coefficient at B = 3, and the structured oracle (pinned against the dense one
on the CPU) for several steps at B = 8, where each persistent workgroup walks
several rotation rows (the FP64 kernel alternates its two LDS plane pairs
between rows).  A non-ramp X (the DFT of x + 0.3 sin(2 pi x / L), not a
coordinate ramp) runs the general-circulant GEMMs; the density-form gateway
the grid gravity.  The three-pass FMA kernels still serve 3-D grids above 32
points, checked against the structured integrator.

Tolerances: 1e-10 relative L2 (FP64), 2e-5 (FP32); total probability
F^0_00(k = 0) to 1e-12 per step (FP64)."""
import numpy as np
import pytest

import hpup
import oracle as O
from helpers import density, omega_grid, operators, rel_l2

pytestmark = pytest.mark.gpu

G2 = np.array([[0.1, 0.03], [0.02, 0.15]])
BT = np.array([0.2, 0.3])


def _print(*a):
    print(*a, flush=True)


def _case(B, N, L, seed):
    f = density(B, 2, N, L, (8, 4, 2), seed=seed, Sigma=np.eye(2) * 1.2)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    return O.fft_forward(B, f, d, w, 2, N), f, d, w


@pytest.mark.parametrize("N", [34, 60, 64])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
def test_plan_derivative_64(N, precision, tol):
    import torch
    _print("case derivative 64", N, precision)
    B, L = 3, 8.0
    Ffull, _, _, _ = _case(B, N, L, 80 + N)
    X, _, MR, u, cells = operators(B, 2, N, L)
    ref = O.derivative(B, 2, N, Ffull, X, None, MR, BT, G2, L, u, cells)
    p = hpup.Plan(B, 2, N, L, b=(BT[0], BT[1], 0.0), G=G2, dt=0.01, method="euler", precision=precision)
    F = p.upload_full(Ffull)
    dF = p.derivative(F)
    torch.cuda.synchronize()
    e = rel_l2(p.download_full(dF), ref)
    _print("  rel err %.2e" % e)
    assert e < tol


@pytest.mark.parametrize("method,nsteps", [("euler", 3), ("midpoint", 2), ("RK4", 2)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
def test_plan_steps_64(method, nsteps, precision, tol):
    """Several steps at B = 8 (680 rotation rows, more than the persistent
    grid's workgroups) against the structured integrator, and the total
    probability after them."""
    import torch
    _print("case steps 64", method, nsteps, precision)
    B, N, L, dt = 8, 60, 8.0, 0.004
    Ffull, _, _, _ = _case(B, N, L, 90)
    X, _, MR, u, cells = operators(B, 2, N, L)
    ref = Ffull
    for _ in range(nsteps):
        ref = O.integrate_structured(B, 2, N, ref, X, None, MR, BT, G2, dt, L, u, cells, method)
    p = hpup.Plan(B, 2, N, L, b=(BT[0], BT[1], 0.0), G=G2, dt=dt, method=method, precision=precision)
    F = p.upload_full(Ffull)
    m0 = float(F[p.mass_index].real)
    p.step(F, nsteps)
    torch.cuda.synchronize()
    e = rel_l2(p.download_full(F), ref)
    _print("  rel err %.2e" % e)
    assert e < tol
    if precision == "fp64":
        assert abs(float(F[p.mass_index].real) - m0) <= nsteps * 1e-12 * abs(m0)


@pytest.mark.parametrize("N", [8, 40])
def test_general_circulant_64(N):
    """X of a non-ramp function of the coordinate: the kernels' general
    circulant GEMMs (32- and 64-point tiles) through the reduced gateway."""
    _print("case general circulant", N)
    B, L, dt = 3, 8.0, 0.01
    Ffull, _, _, _ = _case(B, N, L, 100 + N)
    X, _, MR, u, cells = operators(B, 2, N, L)
    x = omega_grid(2, N, L)
    Xg = np.zeros_like(X)
    for p in range(2):
        Xg[..., p] = np.fft.fftn(x[p] + 0.3 * np.sin(2 * np.pi * x[p] / L))
    ref = O.integrate(B, 2, N, Ffull, Xg, None, MR, BT, G2, dt, L, u, cells, "RK2")
    out = hpup.pendulum_propagate_reduced(Ffull, Xg, MR, BT, G2, dt, L, u, cells, "RK2")
    e = rel_l2(out, ref)
    _print("  rel err %.2e" % e)
    assert e < 1e-10


def test_density_form_60():
    """Density-form gateway (gravity from the grid) at N = 60."""
    _print("case reduced den 60")
    B, N, L, dt = 3, 60, 3.4, 0.0025
    Ffull, f, d, w = _case(B, N, L, 110)
    X, _, _, u, _ = operators(B, 2, N, L)
    mR = O.grid_mR(B)
    ref = O.integrate_den(B, N, Ffull, f, X, mR, BT, G2, dt, L, u, d, w, "RK4")
    out = hpup.pendulum_propagate_reduced_den(Ffull, f, X, mR, BT, G2, dt, L, u, d, w, "RK4")
    e = rel_l2(out, ref)
    _print("  rel err %.2e" % e)
    assert e < 1e-10


def test_three_pass_kernels_3d_34():
    """3-D grids above 32 points keep the three-pass FMA kernels."""
    import torch
    _print("case 3-D N=34 three-pass")
    B, N, L, dt = 2, 34, 10.0, 0.004
    Jt = np.diag([1.0, 0.93539, 0.24935])
    f = density(B, 3, N, L, (10, 5, 1), seed=120, Sigma=np.eye(3) * 1.5)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Ffull = O.fft_forward(B, f, d, w, 3, N)
    X, _, MR, u, cells = operators(B, 3, N, L, Jt)
    bt, G = np.array([0.2, 0.15, 0.25]), np.diag([0.1, 0.2, 0.05])
    ref = O.integrate_structured(B, 3, N, Ffull, X, Jt, MR, bt, G, dt, L, u, cells, "RK2")
    p = hpup.Plan(B, 3, N, L, b=bt, G=G, Jt=Jt, dt=dt, method="RK2", precision="fp64")
    F = p.upload_full(Ffull)
    p.step(F, 1)
    torch.cuda.synchronize()
    e = rel_l2(p.download_full(F), ref)
    _print("  rel err %.2e" % e)
    assert e < 1e-10


def test_adjoint_and_split_need_32():
    """The adjoint operator and the 3xBF16 circulants are built for N <= 32."""
    import torch
    p = hpup.Plan(3, 2, 40, 8.0, precision="fp64", operator=1)
    with pytest.raises(hpup.HpupError):
        p.derivative(torch.zeros(p.ncoef, dtype=p.cdtype, device="cuda"))
    with pytest.raises(hpup.HpupError):
        q = hpup.Plan(3, 2, 40, 8.0, precision="bf16x3")
        q.derivative(torch.zeros(q.ncoef, dtype=q.cdtype, device="cuda"))
