"""The 3-D adjoint operator (SO(3) x R^3, pendulum.m's derivative) on the
plan API.  The reference has no adjoint; A^H is fixed by the dot-product
identity given the forward operator, whose parity with the oracle the other
GPU tests hold:

* <A x, y> = <x, A^H y> for random complex x, y, with the gyroscopic term
  (a diagonal inertia with distinct moments) and without it (isotropic
  inertia), on a small grid and on the 30-point benchmark grid;
* <P x, y> = <x, P^H y> for the step of every integrator.

Tolerances: FP64 1e-12 relative to |A x| |y|, FP32 2e-5."""
import numpy as np
import pytest

import hpup

pytestmark = pytest.mark.gpu

JT = np.diag([1.0, 0.93539, 0.24935])
BT = (0.2, 0.15, 0.25)
G3 = np.array([[0.1, 0.02, 0.0], [0.01, 0.2, 0.03], [0.0, 0.02, 0.05]])


def _print(*a):
    print(*a, flush=True)


def _rand(n, dtype, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    re = torch.randn(n, generator=g, dtype=torch.float64)
    im = torch.randn(n, generator=g, dtype=torch.float64)
    return torch.complex(re, im).to(dtype).cuda()


def _dot(x, y):
    import torch
    return complex(torch.vdot(x, y).item())


def _plan(B, N, L, Jt, precision="fp64", operator=0, dt=0.004, method="euler"):
    return hpup.Plan(B, 3, N, L, b=BT, G=G3, Jt=Jt, dt=dt, method=method, precision=precision, operator=operator)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
@pytest.mark.parametrize("B,N,L,gyro", [(4, 10, 8.0, True), (4, 10, 8.0, False), (3, 30, 10.0, True)])
def test_operator_adjoint_identity_3d(B, N, L, gyro, precision, tol):
    _print("3-D adjoint identity", B, N, gyro, precision)
    Jt = JT if gyro else np.eye(3)
    A, AH = _plan(B, N, L, Jt, precision), _plan(B, N, L, Jt, precision, operator=1)
    x, y = _rand(A.ncoef, A.cdtype, 11), _rand(A.ncoef, A.cdtype, 12)
    Ax, AHy = A.derivative(x), AH.derivative(y)
    lhs, rhs = _dot(Ax, y), _dot(x, AHy)
    scale = float(Ax.abs().pow(2).sum().sqrt() * y.abs().pow(2).sum().sqrt())
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs)
    # the operator is not self-adjoint, so the identity is not trivial
    if precision == "fp64":
        assert abs(_dot(A.derivative(y), x) - _dot(y, A.derivative(x))) > 1e6 * tol * scale


@pytest.mark.parametrize("method", ["euler", "midpoint", "RK2", "RK4"])
def test_step_adjoint_identity_3d(method):
    _print("3-D step adjoint identity", method)
    B, N, L = 4, 10, 8.0
    P, PH = _plan(B, N, L, JT, method=method), _plan(B, N, L, JT, method=method, operator=1)
    x, y = _rand(P.ncoef, P.cdtype, 13), _rand(P.ncoef, P.cdtype, 14)
    Px, PHy = P.step(x.clone(), 2), PH.step(y.clone(), 2)
    lhs, rhs = _dot(Px, y), _dot(x, PHy)
    scale = float(Px.abs().pow(2).sum().sqrt() * y.abs().pow(2).sum().sqrt())
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)


def test_parameter_gradients_stay_reduced():
    """The parameter-gradient stage is built for the reduced model only."""
    import torch
    AH = _plan(3, 6, 8.0, JT, operator=1)
    y = _rand(AH.ncoef, AH.cdtype, 15)
    g = torch.zeros(5, dtype=torch.float64, device="cuda")
    with pytest.raises(hpup.HpupError):
        AH.adjoint_stage(y, y.clone(), 0.1, g)
