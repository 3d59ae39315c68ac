"""Gradients through hybrid steps in the restitution coefficient epsilon (no
counterpart in the reference, which only propagates forward).  Pinned by:

* the dot-product identity <R x, y> = <x, R^H y> of the reset's adjoint (dim 2
  and 3);
* the restitution derivative <g, (dR/deps) f> against central differences of
  the reset kernel in epsilon;
* the harness's dJ/deps after several hybrid steps against central
  differences of J through separately built plans, and through the oracle's
  CPU composition of the same four functions (the reference's path); the F0
  gradient against the exact change of the linear objective.

The density sits well inside a wide omega box (L = 16, unit widths), so the
box test of the pull-back, whose edge moves with epsilon, meets densities
below e^-18, and the differences are taken at a generic eps (0.77), where no
pulled-back lattice point sits on a cell face (the interpolant's kinks).
Tolerances: 1e-12 on the identity, 2e-6 relative on the differences
(h = 1e-6 for one reset, 1e-5 through hybrid steps)."""
import numpy as np
import pytest

import hpup
import oracle as O
from hpup.adjoint import HybridGradient
from helpers import density, omega_grid, operators

pytestmark = pytest.mark.gpu

PAR = dict(d=0.12, h=0.2, r=0.025, theta_t=5 * np.pi / 180, lambda_max=100.0 * 0.1173)
JT = np.diag([1.0, 0.93539, 0.24935])


def _print(*a):
    print(*a, flush=True)


def _grid(p, f):
    import torch
    return torch.from_numpy(np.asfortranarray(f).ravel(order="F").copy()).to(p.rdtype).cuda()


def _rand_grid(p, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(p.ngrid, generator=g, dtype=torch.float64).to(p.rdtype).cuda()


@pytest.mark.parametrize("dim,tilt,jac", [(2, 0.0, False), (3, 0.25, True), (3, -0.1, False)])
def test_reset_adjoint_identity(dim, tilt, jac):
    _print("reset adjoint identity", dim, tilt, jac)
    B, N, L = 3, 10, 6.0
    p = hpup.Plan(B, dim, N, L, precision="fp64")
    p.set_reset(**PAR, epsilon=0.8, tilt=tilt, jacobian=jac)
    x, y = _rand_grid(p, 1), _rand_grid(p, 2)
    Rx = p.reset(x, mode=1)
    RHy = p.reset_adjoint(y, scale=1.0) - y
    lhs, rhs = float((Rx * y).sum()), float((x * RHy).sum())
    scale = float(Rx.norm() * y.norm())
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)
    assert float(Rx.abs().max()) > 0


@pytest.mark.parametrize("dim,tilt,jac", [(2, 0.0, False), (3, 0.25, True), (3, -0.1, False)])
def test_restitution_derivative(dim, tilt, jac):
    import torch
    _print("restitution derivative", dim, tilt, jac)
    # a generic eps and grid spacing: at eps = 0.8 on a unit-spaced grid, pull-backs
    # of lattice points land exactly on cell faces, where the interpolant has kinks
    B, N, L, eps, h = 3, 14, 16.0, 0.77, 1e-6
    f = density(B, dim, N, L, (6, 3, 1), seed=21, Sigma=np.eye(dim))
    p = hpup.Plan(B, dim, N, L, precision="fp64")
    ft = _grid(p, f)
    # random weights under a Gaussian window: the pull-back's box edge, which
    # moves with eps, then meets weights and densities below 1e-13
    x = omega_grid(dim, N, L)
    win = np.exp(-0.125 * (x ** 2).sum(axis=0))
    g = _rand_grid(p, 3) * _grid(p, np.broadcast_to(win, f.shape))

    def val(e):
        p.set_reset(**PAR, epsilon=e, tilt=tilt, jacobian=jac)
        return float((g * p.reset(ft, mode=1)).sum())
    fd = (val(eps + h) - val(eps - h)) / (2 * h)
    p.set_reset(**PAR, epsilon=eps, tilt=tilt, jacobian=jac)
    acc = torch.zeros(1, dtype=torch.float64, device="cuda")
    p.reset_eps_grad(ft, g, acc)
    ana = float(acc.item())
    _print("   ", ana, fd)
    assert abs(ana - fd) <= 2e-6 * max(abs(fd), 1e-3), (ana, fd)


def _oracle_hybrid(B, N, L, F, ops, bt, G, dt, method, eps, tilt, K):
    X, OJO, MR, u, cells = ops
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    R = O.rotations(B)
    lam, PC, t = O.reset3_tables(B, R, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"], PAR["lambda_max"], tilt)
    for _ in range(K):
        F = O.integrate(B, 3, N, F, X, OJO, MR, bt, G, dt, L, u, cells, method)
        g = O.fft_backward(B, F, d, 3, N)
        g = g + dt * O.reset3_df(B, N, L, R, g, lam, PC, t, eps, tilt, True)
        F = O.fft_forward(B, g, d, w, 3, N)
    return F


def test_hybrid_restitution_gradient():
    import torch
    _print("hybrid restitution gradient")
    B, N, L, dt, K, method = 3, 10, 16.0, 0.01, 3, "RK4"
    eps, tilt, h = 0.77, 0.25, 1e-5
    bt, G = np.array([0.2, 0.15, 0.25]), np.diag([0.1, 0.2, 0.05])
    f0 = density(B, 3, N, L, (8, 4, 2), seed=31, Sigma=np.eye(3))
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Ffull = O.fft_forward(B, f0, d, w, 3, N)

    def grad_model(e):
        return HybridGradient(B, 3, N, L, dict(PAR, epsilon=e, tilt=tilt, jacobian=True), b=bt, G=G, Jt=JT, dt=dt,
                              method=method, precision="fp64")
    hg = grad_model(eps)
    F0 = hg.fwd.upload_full(Ffull)
    gen = torch.Generator(device="cpu").manual_seed(5)
    wv = torch.complex(torch.randn(hg.fwd.ncoef, generator=gen, dtype=torch.float64),
                       torch.randn(hg.fwd.ncoef, generator=gen, dtype=torch.float64)).cuda()
    FK, gF0, g = hg.run(F0, K, lambda F: wv, checkpoint=2)

    def J(e, Fs=None):
        m = grad_model(e)
        F = (F0 if Fs is None else Fs).clone()
        m.fwd.hybrid_step(F, m.fwd.empty_grid(), K)
        return float(torch.vdot(wv, F).real)
    fd = (J(eps + h) - J(eps - h)) / (2 * h)
    _print("   dJ/deps", g["epsilon"], fd)
    assert abs(g["epsilon"] - fd) <= 2e-6 * max(abs(fd), 1e-3)
    # the oracle's CPU hybrid steps (the reference's algorithm), differenced
    ops = operators(B, 3, N, L, JT)
    wfull = hg.fwd.download_full(wv)

    def Jo(e):
        return float(np.vdot(wfull, _oracle_hybrid(B, N, L, Ffull, ops, bt, G, dt, method, e, tilt, K)).real)
    fdo = (Jo(eps + h) - Jo(eps - h)) / (2 * h)
    _print("   oracle difference", fdo)
    assert abs(g["epsilon"] - fdo) <= 2e-6 * max(abs(fdo), 1e-3)
    # F0: J is linear in F0
    dF = torch.complex(torch.randn(hg.fwd.ncoef, generator=gen, dtype=torch.float64),
                       torch.randn(hg.fwd.ncoef, generator=gen, dtype=torch.float64)).cuda() * 1e-2
    dJ = J(eps, F0 + dF) - J(eps)
    assert abs(dJ - float(torch.vdot(gF0, dF).real)) <= 1e-9 * max(abs(dJ), 1e-12)
