"""The device-resident hybrid reset and hybrid step of the plan API against
the oracle.

* dim 2: the plan's reset equals the reference's no-noise reduced law as its
  drivers chain the gateways (getLambda_mex compact -> getOmega_mex 'old' ->
  getIndRule_mex -> pendulum_reduced_discrete_propagate(false, ...),
  pendulum_reduced_collision.m:221-289), restated by the oracle's gateway
  functions;
* dim 3: the plan's reset equals the oracle's 3-D law (refcpu_reset3.c), with
  and without a wall tilt and the 1/eps Jacobian, in FP64 and FP32;
* the hybrid step (pendulum_reduced_hybrid.m:224-234: integrator step,
  inverse transform, f += dt df, forward transform) equals the oracle's
  composition of the same four functions, in dim 2 and dim 3.
Tolerances: 1e-13 relative L2 for the FP64 reset (same geometry in the same
IEEE operations), 1e-6 FP32, 1e-10 for FP64 hybrid steps."""
import numpy as np
import pytest

import hpup
import oracle as O
from helpers import density, omega_grid, operators, rel_l2

pytestmark = pytest.mark.gpu

PAR = dict(d=0.12, h=0.2, r=0.025, theta_t=5 * np.pi / 180, lambda_max=100.0 * 0.1173)
EPS = 0.8


def _print(*a):
    print(*a, flush=True)


def _oracle_reduced_df(B, N, L, f):
    R = O.rotations(B)
    x = omega_grid(2, N, L)
    lam, indR, lists, PC = O.getLambda_compact(B, N, R, x, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"],
                                               PAR["lambda_max"])
    Om = O.getOmega(B, N, R, x, indR, EPS, PC, "old")
    ind, coeff = O.getIndRule(N, x, Om)
    return O.discrete_nonoise(B, N, f, lam, indR, lists, ind, coeff), indR


def _oracle_3d_df(B, N, L, f, tilt, jacobian):
    R = O.rotations(B)
    lam, PC, t = O.reset3_tables(B, R, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"], PAR["lambda_max"], tilt)
    return O.reset3_df(B, N, L, R, f, lam, PC, t, EPS, tilt, jacobian), lam


def _grid_tensor(p, f):
    import torch
    return torch.from_numpy(np.asfortranarray(f).ravel(order="F").copy()).to(p.rdtype).cuda()


def _grid_numpy(p, t):
    return t.double().cpu().numpy().reshape((2 * p.B,) * 3 + (p.N,) * p.dim, order="F")


@pytest.mark.parametrize("B,N,L", [(3, 8, 3.4), (4, 10, 4.0)])
def test_plan_reset_reduced_law(B, N, L):
    import torch
    _print("case reset reduced", B, N, L)
    f = density(B, 2, N, L, (6, 3, 1), seed=70 + B, Sigma=np.eye(2) * 0.6)
    df_ref, indR = _oracle_reduced_df(B, N, L, f)
    p = hpup.Plan(B, 2, N, L, precision="fp64")
    rows = p.set_reset(**PAR, epsilon=EPS)
    assert np.array_equal(rows, indR)
    ft = _grid_tensor(p, f)
    df = p.reset(ft, mode=1)
    torch.cuda.synchronize()
    g = _grid_numpy(p, df)
    assert np.count_nonzero(df_ref) > 0
    assert rel_l2(g, df_ref) < 1e-13
    # mode 0 in place: f + dt df
    dt = 0.01
    p.reset(ft, out=ft, scale=dt, mode=0)
    torch.cuda.synchronize()
    assert rel_l2(_grid_numpy(p, ft), f + dt * df_ref) < 1e-14


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
@pytest.mark.parametrize("tilt,jacobian", [(0.0, False), (0.3, True), (-0.2, False)])
def test_plan_reset_3d_law(tilt, jacobian, precision, tol):
    import torch
    _print("case reset 3d", tilt, jacobian, precision)
    B, N, L = 3, 8, 6.0
    f = density(B, 3, N, L, (6, 3, 1), seed=80, Sigma=np.eye(3) * 1.2)
    df_ref, lam = _oracle_3d_df(B, N, L, f, tilt, jacobian)
    p = hpup.Plan(B, 3, N, L, precision=precision)
    rows = p.set_reset(**PAR, epsilon=EPS, tilt=tilt, jacobian=jacobian)
    assert set(np.nonzero(lam)[0]) <= set(rows.tolist())
    df = p.reset(_grid_tensor(p, f), mode=1)
    torch.cuda.synchronize()
    assert np.count_nonzero(df_ref) > 0
    assert rel_l2(_grid_numpy(p, df), df_ref) < tol


def test_reset_rejects_bad_parameters():
    B, N, L = 2, 6, 4.0
    p2 = hpup.Plan(B, 2, N, L, precision="fp64")
    with pytest.raises(hpup.HpupError):
        p2.set_reset(**PAR, epsilon=EPS, tilt=0.1)  # the reduced model's wall is not tilted
    with pytest.raises(hpup.HpupError):
        p2.set_reset(**PAR, epsilon=0.0)
    with pytest.raises(hpup.HpupError):
        p2.reset(p2.empty_grid())  # no reset tables yet


@pytest.mark.parametrize("dim,method", [(2, "euler"), (3, "RK4"), (3, "euler")])
def test_hybrid_step_matches_oracle(dim, method):
    import torch
    _print("case hybrid step", dim, method)
    B, N, L, dt, nsteps = 3, 8, 6.0, 0.01, 2
    Jt = np.diag([1.0, 0.93539, 0.24935])
    Sigma = np.eye(dim) * 0.9
    f = density(B, dim, N, L, (8, 4, 2), seed=90 + dim, Sigma=Sigma)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    X, OJO, MR, u, cells = operators(B, dim, N, L, Jt if dim == 3 else None)
    bt = np.array([0.2, 0.15, 0.25])[:dim]
    G = np.diag([0.1, 0.2, 0.05])[:dim, :dim]
    tilt = 0.25 if dim == 3 else 0.0
    F = O.fft_forward(B, f, d, w, dim, N)
    p = hpup.Plan(B, dim, N, L, b=bt, G=G, Jt=Jt, dt=dt, method=method, precision="fp64")
    p.set_reset(**PAR, epsilon=EPS, tilt=tilt, jacobian=dim == 3)
    Fd = p.upload_full(F)
    fd = p.empty_grid()
    p.hybrid_step(Fd, fd, nsteps)
    torch.cuda.synchronize()
    for _ in range(nsteps):
        F = O.integrate(B, dim, N, F, X, OJO, MR, bt, G, dt, L, u, cells, method)
        g = O.fft_backward(B, F, d, dim, N)
        if dim == 2:
            df, _ = _oracle_reduced_df(B, N, L, g)
        else:
            df, _ = _oracle_3d_df(B, N, L, g, tilt, True)
        g = g + dt * df
        F = O.fft_forward(B, g, d, w, dim, N)
    assert rel_l2(_grid_numpy(p, fd), g) < 1e-10
    assert rel_l2(p.download_full(Fd), F) < 1e-10
