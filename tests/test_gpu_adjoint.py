"""Adjoint and parameter-derivative operators of the reduced model (dim 2),
and the gradient harness built on them (hpup.adjoint).  The reference has no
adjoint; these operators are pinned by identities that fix them given the
forward operator, whose parity with the oracle the other GPU tests hold:

* <A x, y> = <x, A^H y> for random complex x, y (fixes A^H);
* <P x, y> = <x, P^H y> for the step of every integrator;
* A(theta + e_t) x - A(theta) x = (dA/dtheta_t) x (A is linear in b and G);
* the harness's gradients (reverse mode through the Horner stages) against
  central differences of J(theta) through separately built forward plans
  and through the oracle's CPU propagation, and the F0 gradient against the
  exact change of the (linear) objective;
* the torch.autograd.Function binding against the harness.

Tolerances: FP64 1e-12 relative on the identities (roundoff of two fused
evaluations), 2e-7 on the central differences (h = 1e-4: the O(h^2) error of
a polynomial objective); FP32 identities 2e-5.
"""
import numpy as np
import pytest

import hpup
from hpup.adjoint import Gradient

pytestmark = pytest.mark.gpu

B, N, L = 5, 12, 8.0
BT = (0.3, 0.2)
G2 = np.array([[0.12, 0.03], [0.05, 0.18]])


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


def _plan(precision="fp64", operator=0, b=BT, G=G2, dt=0.01, method="euler"):
    return hpup.Plan(B, 2, N, L, b=tuple(b) + (0,), G=G, dt=dt, method=method, precision=precision,
                     operator=operator)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_operator_adjoint_identity(precision, tol):
    _print("adjoint identity", precision)
    A, AH = _plan(precision), _plan(precision, operator=1)
    x, y = _rand(A.ncoef, A.cdtype, 1), _rand(A.ncoef, A.cdtype, 2)
    Ax, AHy = A.derivative(x), AH.derivative(y)
    lhs, rhs = _dot(Ax, y), _dot(x, AHy)
    scale = float(Ax.abs().pow(2).sum().sqrt() * y.abs().pow(2).sum().sqrt())
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs)
    # A^H is not A: the operator is not self-adjoint
    assert abs(_dot(A.derivative(y), x) - _dot(y, A.derivative(x))) > 100 * tol * scale


@pytest.mark.parametrize("method", ["euler", "midpoint", "rk2", "rk4"])
def test_step_adjoint_identity(method):
    _print("step adjoint identity", method)
    P, PH = _plan(method=method, dt=0.02), _plan(method=method, dt=0.02, operator=1)
    x, y = _rand(P.ncoef, P.cdtype, 3), _rand(P.ncoef, P.cdtype, 4)
    Px, PHy = P.step(x.clone(), 2), PH.step(y.clone(), 2)
    lhs, rhs = _dot(Px, y), _dot(x, PHy)
    scale = float(Px.abs().pow(2).sum().sqrt() * y.abs().pow(2).sum().sqrt())
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)


@pytest.mark.parametrize("term", range(2, 8))
def test_parameter_operator_linearity(term):
    _print("parameter operator", term)
    b, G = list(BT), G2.copy()
    if term < 4:
        b[term - 2] += 1.0
    else:
        i, j = (term - 4) % 2, (term - 4) // 2
        G[i, j] += 1.0
    x = _rand(_plan().ncoef, _plan().cdtype, 5)
    diff = _plan(b=b, G=G).derivative(x) - _plan().derivative(x)
    dA = _plan(operator=term).derivative(x)
    assert float((diff - dA).abs().max()) <= 1e-12 * float(dA.abs().max())
    assert float(dA.abs().max()) > 0


def _objective(w, F0, b, G, K, method, dt):
    F = _plan(b=b, G=G, dt=dt, method=method).step(F0.clone(), K)
    return _dot(w, F).real


@pytest.mark.parametrize("method,checkpoint", [("rk4", 3), ("euler", None), ("midpoint", 2)])
def test_gradient_against_differences(method, checkpoint):
    _print("gradient", method, checkpoint)
    K, dt = 7, 0.02
    gr = Gradient(B, N, L, b=BT, G=G2, dt=dt, method=method)
    F0 = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 6)
    w = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 7)
    FK, gF0, g = gr.run(F0, K, lambda F: w, checkpoint=checkpoint)
    # the final state is the plain forward propagation
    ref = _plan(dt=dt, method=method).step(F0.clone(), K)
    assert float((FK - ref).abs().max()) <= 1e-13 * float(ref.abs().max())
    # F0: J is linear in F0
    d = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 8)
    J0 = _objective(w, F0, BT, G2, K, method, dt)
    dJ = _objective(w, F0 + d, BT, G2, K, method, dt) - J0
    assert abs(dJ - _dot(gF0, d).real) <= 1e-11 * abs(dJ)
    # parameters: central differences
    h = 1e-4
    for t in range(6):
        bp, bm, Gp, Gm = list(BT), list(BT), G2.copy(), G2.copy()
        if t < 2:
            bp[t] += h
            bm[t] -= h
            ana = g["b"][t]
        else:
            i, j = (t - 2) % 2, (t - 2) // 2
            Gp[i, j] += h
            Gm[i, j] -= h
            ana = g["G"][i, j]
        fd = (_objective(w, F0, bp, Gp, K, method, dt) - _objective(w, F0, bm, Gm, K, method, dt)) / (2 * h)
        _print("  term", t, ana, fd)
        assert abs(ana - fd) <= 2e-7 * max(abs(fd), 1e-3 * abs(J0)), (t, ana, fd)


def test_gradient_against_oracle_differences():
    """The GPU sweep's gradients of J = Re<w, F_K> in b and G against central
    differences of the oracle's CPU propagation (the reference's dense
    derivative and RK4, FP64) from a matrix-Fisher x Gaussian initial density,
    and the F0 gradient against the oracle's exact change of J along a
    direction (J is linear in F0)."""
    import oracle as O
    from helpers import density, operators, rel_l2
    _print("gradient vs oracle differences")
    Bo, No, Lo, K, dt = 3, 6, 8.0, 4, 0.02
    gr = Gradient(Bo, No, Lo, b=BT, G=G2, dt=dt, method="rk4")
    f = density(Bo, 2, No, Lo, (6, 3, 1), seed=12, Sigma=np.eye(2) * 0.8)
    d = O.wigner_table(Bo)
    _, _, _, wq = O.grid(Bo)
    Ffull = O.fft_forward(Bo, f, d, wq, 2, No)
    F0 = gr.fwd.upload_full(Ffull)
    wv = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 31)
    _, gF0, g = gr.run(F0, K, lambda F: wv, checkpoint=2)
    wfull = gr.fwd.download_full(wv)  # zero outside the compact entries
    X, _, MR, u, cells = operators(Bo, 2, No, Lo)
    # the model plan's operator is the oracle's reduced-model derivative
    ref = O.derivative(Bo, 2, No, Ffull, X, None, MR, np.asarray(BT, dtype=np.float64), G2, Lo, u, cells)
    assert rel_l2(gr.fwd.download_full(gr.fwd.derivative(F0)), ref) < 1e-10

    def J(b, G, F=Ffull):
        for _ in range(K):
            F = O.integrate(Bo, 2, No, F, X, None, MR, np.asarray(b, dtype=np.float64), G, dt, Lo, u, cells, "RK4")
        return float(np.real(np.vdot(wfull.ravel(order="F"), np.asarray(F).ravel(order="F"))))

    J0 = J(BT, G2)
    dirn = gr.fwd.download_full(_rand(gr.fwd.ncoef, gr.fwd.cdtype, 32))
    dJ = J(BT, G2, Ffull + dirn) - J0
    gfull = gr.fwd.download_full(gF0)
    assert abs(dJ - float(np.real(np.vdot(gfull.ravel(order="F"), dirn.ravel(order="F"))))) <= 1e-10 * abs(dJ)
    h = 1e-4
    for t in range(6):
        bp, bm, Gp, Gm = list(BT), list(BT), G2.copy(), G2.copy()
        if t < 2:
            bp[t] += h
            bm[t] -= h
            ana = g["b"][t]
        else:
            i, j = (t - 2) % 2, (t - 2) // 2
            Gp[i, j] += h
            Gm[i, j] -= h
            ana = g["G"][i, j]
        fd = (J(bp, Gp) - J(bm, Gm)) / (2 * h)
        _print("  term", t, ana, fd)
        assert abs(ana - fd) <= 2e-7 * max(abs(fd), 1e-3 * abs(J0)), (t, ana, fd)


def test_gradient_checkpoint_invariance():
    _print("gradient checkpoints")
    gr = Gradient(B, N, L, b=BT, G=G2, dt=0.02, method="rk4")
    F0 = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 9)
    w = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 10)
    _, g1, p1 = gr.run(F0, 6, lambda F: w)
    _, g2, p2 = gr.run(F0, 6, lambda F: w, checkpoint=2)
    assert float((g1 - g2).abs().max()) == 0.0
    assert np.array_equal(p1["b"], p2["b"]) and np.array_equal(p1["G"], p2["G"])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
def test_adjoint_stage_running_sum_modes(precision, tol):
    """hpup_plan_adjoint_stage: out = scale A^H ybar (the adjoint plan's
    derivative times scale); fbar += out, or fbar = ybar + out with fbar_init;
    the five parameter sums do not depend on the fbar mode (bit-identical)."""
    import torch
    _print("adjoint stage fbar modes", precision)
    AH = _plan(precision, operator=1)
    ybar, y = _rand(AH.ncoef, AH.cdtype, 21), _rand(AH.ncoef, AH.cdtype, 22)
    f0 = _rand(AH.ncoef, AH.cdtype, 23)
    scale = 0.37
    ref = AH.derivative(ybar) * scale
    g1 = torch.zeros(5, dtype=torch.float64, device="cuda")
    g2 = torch.zeros(5, dtype=torch.float64, device="cuda")
    fb = f0.clone()
    out = AH.adjoint_stage(ybar, y, scale, g1, fbar=fb)
    rel = lambda a, b: float((a - b).abs().pow(2).sum().sqrt() / b.abs().pow(2).sum().sqrt())
    assert rel(out, ref) < tol
    assert rel(fb, f0 + out) < tol
    fi = torch.full_like(f0, float("nan"))  # every element must be written
    out2 = AH.adjoint_stage(ybar, y, scale, g2, fbar=fi, fbar_init=True)
    assert torch.equal(out2, out)
    assert rel(fi, ybar + out) < tol
    assert torch.equal(g1, g2) and float(g1.abs().sum()) > 0


def test_operator_kind_errors():
    with pytest.raises(hpup.HpupError):
        _plan(operator=8)
    with pytest.raises(hpup.HpupError):
        hpup.Plan(4, 3, 8, L, precision="fp64", operator=2 + 3 + 9)  # past the last parameter operator
    big = hpup.Plan(3, 3, 40, L, precision="fp64", operator=1)  # the adjoint needs the fused path (N <= 32)
    with pytest.raises(hpup.HpupError):
        big.derivative(_rand(big.ncoef, big.cdtype, 1))


def test_autograd_function():
    """The torch.autograd.Function binding returns the harness's gradients."""
    import torch
    from hpup.adjoint import propagate_function
    _print("autograd function")
    gr = Gradient(B, N, L, b=BT, G=G2, dt=0.02, method="rk4")
    F0 = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 11)
    w = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 12)
    Fa = F0.clone().requires_grad_(True)
    b = torch.tensor(BT, dtype=torch.float64, requires_grad=True)
    G = torch.tensor(G2, dtype=torch.float64, requires_grad=True)
    FK = propagate_function().apply(Fa, b, G, gr, 5, 2)
    loss = (w.conj() * FK).real.sum()
    loss.backward()
    _, gF0, g = gr.run(F0, 5, lambda F: w, checkpoint=2)
    assert float((Fa.grad - gF0).abs().max()) <= 1e-12 * float(gF0.abs().max())
    assert np.allclose(b.grad.numpy(), g["b"], rtol=1e-12, atol=0)
    assert np.allclose(G.grad.numpy(), g["G"], rtol=1e-12, atol=0)


def _rand_real(n, dtype, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(n, generator=g, dtype=torch.float64).to(dtype).cuda()


@pytest.mark.parametrize("dim", [2, 3])
def test_transform_adjoint_identities(dim):
    """<B F, g> = Re<F, B^H g> and Re<A f, Fbar> = <f, A^H Fbar> (FP64)."""
    from hpup.adjoint import TransformAdjoint
    _print("transform adjoints", dim)
    p = hpup.Plan(4, dim, 4, L, precision="fp64")
    ta = TransformAdjoint(p)
    F = _rand(p.ncoef, p.cdtype, 13)
    g = _rand_real(p.ngrid, p.rdtype, 14)
    f = _rand_real(p.ngrid, p.rdtype, 15)
    Fb = _rand(p.ncoef, p.cdtype, 16)
    Bf = p.backward(F)
    lhs, rhs = float((Bf * g).sum()), _dot(F, ta.backward_adjoint(g)).real
    assert abs(lhs - rhs) <= 1e-12 * float(Bf.norm() * g.norm()), (lhs, rhs)
    AF = p.forward(f)
    lhs, rhs = _dot(AF, Fb).real, float((f * ta.forward_adjoint(Fb)).sum())
    assert abs(lhs - rhs) <= 1e-12 * float(AF.norm() * Fb.norm()), (lhs, rhs)


def test_grid_objective_gradient():
    """J = <B(F_K), g> on the grid density after K RK4 steps: the harness fed
    with B^H g against central differences in b and the exact F0 change."""
    from hpup.adjoint import TransformAdjoint
    _print("grid objective gradient")
    K, dt, h = 4, 0.02, 1e-4
    gr = Gradient(B, N, L, b=BT, G=G2, dt=dt, method="rk4")
    ta = TransformAdjoint(gr.fwd)
    F0 = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 17)
    g = _rand_real(gr.fwd.ngrid, gr.fwd.rdtype, 18)
    lam = ta.backward_adjoint(g)
    _, gF0, grads = gr.run(F0, K, lambda F: lam, checkpoint=2)

    def J(b, F=F0):
        p = _plan(b=b, dt=dt, method="rk4")
        return float((p.backward(p.step(F.clone(), K)) * g).sum())

    J0 = J(BT)
    for t in range(2):
        bp, bm = list(BT), list(BT)
        bp[t] += h
        bm[t] -= h
        fd = (J(bp) - J(bm)) / (2 * h)
        assert abs(grads["b"][t] - fd) <= 2e-7 * max(abs(fd), 1e-3 * abs(J0)), (t, grads["b"][t], fd)
    d = _rand(gr.fwd.ncoef, gr.fwd.cdtype, 19)
    dJ = J(BT, F0 + d) - J0
    assert abs(dJ - _dot(gF0, d).real) <= 1e-10 * abs(dJ)


def test_measurement_likelihood_gradient():
    """log-likelihood of synthetic (attitude, angular velocity) grid measurements
    after K RK4 steps, differentiated end to end with torch autograd through
    the density, both transforms and the propagation: gradients in the initial
    matrix-Fisher singular values, the Gaussian mean and width, the damping
    and the noise diagonal, against central differences of the FP64 pipeline."""
    import torch
    from hpup.adjoint import mfg_density, propagate_function, transform_functions
    _print("measurement likelihood gradient")
    K, dt, h = 4, 0.02, 1e-5
    U = np.eye(3)
    gr = Gradient(B, N, L, b=BT, G=G2, dt=dt, method="rk4")
    Fwd, Bwd = transform_functions(gr.fwd)
    Prop = propagate_function()
    theta0 = {"s": [4.0, 2.5, 1.0], "mu": [0.3, -0.2], "sig": [1.1, 0.9]}

    def pipeline(th, model=gr):
        # the propagation runs at the model's own b and G (its plans' operator)
        f0 = mfg_density(B, N, L, th["s"], U, th["mu"], th["sig"])
        F0 = Fwd.apply(f0)
        FK = Prop.apply(F0, torch.tensor(model.b, dtype=torch.float64, requires_grad=True),
                        torch.tensor(model.G, dtype=torch.float64), model, K, 2)
        return Bwd.apply(FK)

    t = {k: torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True) for k, v in theta0.items()}
    bt = torch.tensor(BT, dtype=torch.float64, requires_grad=True)
    Gt = torch.tensor(G2, dtype=torch.float64, requires_grad=True)
    f0 = mfg_density(B, N, L, t["s"], U, t["mu"], t["sig"])
    FK = Prop.apply(Fwd.apply(f0), bt, Gt, gr, K, 2)
    fK = Bwd.apply(FK)
    idx = torch.topk(fK.detach(), 12).indices  # synthetic measurements where the density is large
    loglik = torch.log(fK[idx]).sum()
    loglik.backward()

    def value(th, b=BT, G=G2):
        tt = {k: torch.tensor(v, dtype=torch.float64, device="cuda") for k, v in th.items()}
        same = tuple(b) == tuple(BT) and np.array_equal(G, G2)
        model = gr if same else Gradient(B, N, L, b=b, G=G, dt=dt, method="rk4")
        with torch.no_grad():
            return float(torch.log(pipeline(tt, model)[idx]).sum())

    for key in ("s", "mu", "sig"):
        for i in range(len(theta0[key])):
            tp = {k: list(v) for k, v in theta0.items()}
            tm = {k: list(v) for k, v in theta0.items()}
            tp[key][i] += h
            tm[key][i] -= h
            fd = (value(tp) - value(tm)) / (2 * h)
            ana = float(t[key].grad[i])
            _print("  ", key, i, ana, fd)
            assert abs(ana - fd) <= 1e-6 * max(abs(fd), 1.0), (key, i, ana, fd)
    for i in range(2):
        bp, bm = list(BT), list(BT)
        bp[i] += h
        bm[i] -= h
        fd = (value(theta0, bp) - value(theta0, bm)) / (2 * h)
        assert abs(float(bt.grad[i]) - fd) <= 1e-6 * max(abs(fd), 1.0), (i, float(bt.grad[i]), fd)
    for i in range(2):  # the noise diagonal
        Gp, Gm = G2.copy(), G2.copy()
        Gp[i, i] += h
        Gm[i, i] -= h
        fd = (value(theta0, BT, Gp) - value(theta0, BT, Gm)) / (2 * h)
        _print("   G", i, float(Gt.grad[i, i]), fd)
        assert abs(float(Gt.grad[i, i]) - fd) <= 1e-6 * max(abs(fd), 1.0), (i, float(Gt.grad[i, i]), fd)
