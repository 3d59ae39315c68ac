"""The adjoint/gradient workload at its own size: reduced pendulum, B = 10,
32^2 omega grid over [-4, 4]^2, 50 RK4 steps of 0.01, checkpoint every 10
steps, and the log-likelihood of 256 seeded noisy measurements at the final
time (attitude noise matrix Fisher with concentration 20, angular-velocity
noise Gaussian with sigma 0.2; hpup.likelihood).  Gradients in the damping b
(both axes), the noise diagonal G_11, G_22 and the three initial matrix-Fisher
singular values, from the FP64 and the FP32 pipelines (torch autograd over
the hand-written adjoints), against central differences (h = 1e-6) of the
FP64 pipeline.  Tolerances: 1e-6 relative (FP64), 2e-3 (FP32)."""
import numpy as np
import pytest

import hpup
from hpup.adjoint import Gradient, mfg_density, propagate_function, transform_functions
from hpup.likelihood import MeasurementLikelihood, sample_mf_isotropic, sample_states

pytestmark = pytest.mark.gpu

B, N, L, K, DT, CP = 10, 32, 8.0, 50, 0.01, 10
S0 = (4.0, 2.0, 1.0)  # resolved at B = 10: the noisy measurements' likelihoods stay positive
B0, G0 = 0.2, (0.1, 0.1)


def _print(*a):
    print(*a, flush=True)


def _model(b, g11, g22, precision):
    return Gradient(B, N, L, b=(b, b), G=np.diag([g11, g22]), dt=DT, method="rk4", precision=precision)


def _final_density(th, model, U, grad=False):
    import torch
    Fwd, Bwd = transform_functions(model.fwd)
    rdt = model.fwd.rdtype
    s = th["s"]
    f0 = mfg_density(B, N, L, s, U, torch.zeros(2, dtype=torch.float64, device="cuda"),
                     torch.ones(2, dtype=torch.float64, device="cuda")).to(rdt)
    bt = torch.tensor(model.b, dtype=torch.float64, requires_grad=grad)
    Gt = torch.tensor(model.G, dtype=torch.float64, requires_grad=grad)
    FK = propagate_function().apply(Fwd.apply(f0), bt, Gt, model, K, CP)
    return Bwd.apply(FK), bt, Gt


@pytest.fixture(scope="module")
def setup():
    import torch
    from hpup import density as D
    rng = np.random.default_rng(4)
    U = D.random_rotation(rng)
    m = _model(B0, *G0, "fp64")
    th = {"s": torch.tensor(S0, dtype=torch.float64, device="cuda")}
    with torch.no_grad():
        fK, _, _ = _final_density(th, m, U)
    R, Om = sample_states(rng, m.fwd, fK, 256)
    Z = sample_mf_isotropic(rng, R, 20.0)
    y = Om + 0.2 * rng.standard_normal(Om.shape)
    return U, Z, y


def _loglik(th, b, g11, g22, U, Z, y, precision="fp64"):
    import torch
    m = _model(b, g11, g22, precision)
    with torch.no_grad():
        fK, _, _ = _final_density(th, m, U)
        p = MeasurementLikelihood(m.fwd, Z, y).probabilities(fK)
    if not bool(torch.isfinite(p).all()) or float(p.min()) <= 0:
        _print("   non-positive or non-finite probabilities: min %r, finite %r, grid finite %r"
               % (float(p.min()), bool(torch.isfinite(p).all()), bool(torch.isfinite(fK).all())))
    return float(p.log().sum())


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("fp32", 2e-3)])
def test_measurement_likelihood_gradients(setup, precision, tol):
    import torch
    U, Z, y = setup
    _print("likelihood gradients", precision)
    m = _model(B0, *G0, precision)
    s = torch.tensor(S0, dtype=torch.float64, device="cuda", requires_grad=True)
    fK, bt, Gt = _final_density({"s": s}, m, U, grad=True)
    ll = MeasurementLikelihood(m.fwd, Z, y)(fK)
    ll.backward()
    ana = {"b": float(bt.grad.sum()), "G11": float(Gt.grad[0, 0]), "G22": float(Gt.grad[1, 1])}
    for i in range(3):
        ana["s%d" % (i + 1)] = float(s.grad[i])
    h = 1e-6
    th0 = {"s": torch.tensor(S0, dtype=torch.float64, device="cuda")}

    def J(db=0.0, dg=(0.0, 0.0), ds=None):
        th = th0 if ds is None else {"s": th0["s"] + torch.tensor(ds, dtype=torch.float64, device="cuda")}
        return _loglik(th, B0 + db, G0[0] + dg[0], G0[1] + dg[1], U, Z, y)
    fd = {"b": (J(db=h) - J(db=-h)) / (2 * h),
          "G11": (J(dg=(h, 0)) - J(dg=(-h, 0))) / (2 * h),
          "G22": (J(dg=(0, h)) - J(dg=(0, -h))) / (2 * h)}
    for i in range(3):
        e = [0.0, 0.0, 0.0]
        e[i] = h
        fd["s%d" % (i + 1)] = (J(ds=e) - J(ds=[-v for v in e])) / (2 * h)
    for k in fd:
        _print("   %s analytic %.10g  difference %.10g" % (k, ana[k], fd[k]))
        assert abs(ana[k] - fd[k]) <= tol * max(abs(fd[k]), 1.0), (k, ana[k], fd[k])
    assert np.isfinite(float(ll)) and float(ll) > -1e6
