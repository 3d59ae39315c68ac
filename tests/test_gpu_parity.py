"""HIP path vs the FP64 oracle through the C ABI, at sizes the oracle runs in
seconds.  Tolerances: 1e-10 relative L2 for the FP64 path (the reference's
stated parity bar), 2e-5 for FP32."""
import numpy as np
import pytest

import hpup
import oracle as O
from helpers import density, operators, rel_l2, valid_mask

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("B,dim,N", [(4, 0, 1), (5, 2, 4), (3, 3, 4), (10, 0, 1), (15, 0, 1), (15, 2, 4), (17, 0, 1),
                                     (17, 2, 2), (30, 0, 1), (30, 2, 2)])
def test_fft_forward_backward(B, dim, N):
    _print("case fft", B, dim, N)
    L = 6.0
    f = density(B, dim, N, L, (6, 3, 1), seed=B)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F_ref = O.fft_forward(B, f, d, w, dim, N)
    F = hpup.fftSO3R_mex("forward", f, d, w)
    assert F.shape == F_ref.shape
    assert rel_l2(F, F_ref) < 1e-12
    f_ref = O.fft_backward(B, F_ref, d, dim, N)
    f2 = hpup.fftSO3R_mex("backward", F_ref, d)
    assert rel_l2(f2, f_ref) < 1e-12


def _reduced_case(B=4, N=8, L=8.0):
    f = density(B, 2, N, L, (8, 4, 2), seed=3, Sigma=np.eye(2) * 1.0)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F = O.fft_forward(B, f, d, w, 2, N)
    X, _, MR, u, cells = operators(B, 2, N, L)
    bt = np.array([0.2, 0.2])
    G = np.diag([0.1, 0.1])
    return F, X, MR, bt, G, u, cells, L


@pytest.mark.parametrize("method", ["euler", "midpoint", "RK2", "RK4"])
def test_propagate_reduced_fp64(method):
    _print("case reduced fp64", method)
    F, X, MR, bt, G, u, cells, L = _reduced_case()
    B, N, dt = F.shape[2], F.shape[3], 0.01
    ref = O.integrate(B, 2, N, F, X, None, MR, bt, G, dt, L, u, cells, method)
    out = hpup.pendulum_propagate_reduced(F, X, MR, bt, G, dt, L, u, cells, method)
    assert rel_l2(out, ref) < 1e-10
    # invalid-triangle entries are zero
    assert np.all(out[~np.broadcast_to(valid_mask(B, (N, N)), out.shape)] == 0)


def test_propagate_reduced_fp32():
    _print("case reduced fp32")
    F, X, MR, bt, G, u, cells, L = _reduced_case()
    B, N, dt = F.shape[2], F.shape[3], 0.01
    ref = O.integrate(B, 2, N, F, X, None, MR, bt, G, dt, L, u, cells, "euler")
    cells32 = {k: v.astype(np.float32) for k, v in cells.items()}
    out = hpup.pendulum_propagate_reduced32(F.astype(np.complex64), X.astype(np.complex64), MR.astype(np.complex64),
                                            bt.astype(np.float32), G.astype(np.float32), dt, L,
                                            u.astype(np.complex64), cells32, "euler")
    assert rel_l2(out, ref) < 2e-5


def _full_case(B=3, N=6, L=10.0):
    Sigma = np.array([[0.8, 0.1, 0.0], [0.1, 0.6, 0.05], [0.0, 0.05, 1.0]])
    f = density(B, 3, N, L, (10, 5, 1), seed=5, Sigma=Sigma)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F = O.fft_forward(B, f, d, w, 3, N)
    Jt = np.diag([1.0, 0.93539, 0.24935])
    X, OJO, MR, u, cells = operators(B, 3, N, L, Jt)
    bt = np.array([0.2, 0.2, 0.2])
    G = np.eye(3) * 0.1
    return F, X, OJO, MR, bt, G, u, cells, L


@pytest.mark.parametrize("method", ["euler", "RK4"])
def test_propagate_3d_fp64(method):
    _print("case 3d fp64", method)
    F, X, OJO, MR, bt, G, u, cells, L = _full_case()
    B, N, dt = F.shape[2], F.shape[3], 0.004
    ref = O.integrate(B, 3, N, F, X, OJO, MR, bt, G, dt, L, u, cells, method)
    out = hpup.pendulum_propagate(F, X, OJO, MR, bt, G, dt, L, u, cells, method)
    assert rel_l2(out, ref) < 1e-10


def test_derivative_3d_fp64_each_term():
    """dF itself (dt = 1 Euler minus F) against the oracle, including a case with
    only one term switched on at a time (damping, gyroscopic, gravity, noise)."""
    _print("case 3d derivative terms")
    F, X, OJO, MR, bt, G, u, cells, L = _full_case(B=4, N=4)
    B, N = F.shape[2], F.shape[3]
    zeros = np.zeros_like
    cases = {
        "all": (X, OJO, MR, bt, G),
        "kinematic": (X, zeros(OJO), zeros(MR), zeros(bt), zeros(G)),
        "damping": (X, zeros(OJO), zeros(MR), bt, zeros(G)),
        "gyro": (X, OJO, zeros(MR), zeros(bt), zeros(G)),
        "gravity": (X, zeros(OJO), MR, zeros(bt), zeros(G)),
        "noise": (X, zeros(OJO), zeros(MR), zeros(bt), np.array([[0.1, 0.02, 0.0], [0.02, 0.2, 0.01], [0.0, 0.01, 0.3]])),
    }
    for name, (Xc, Oc, Mc, bc, Gc) in cases.items():
        ref = O.derivative(B, 3, N, F, Xc, Oc, Mc, bc, Gc, L, u, cells)
        out = hpup.pendulum_propagate(F, Xc, Oc, Mc, bc, Gc, 1.0, L, u, cells, "euler") - F
        assert rel_l2(out, ref) < 1e-10, name


def test_propagate_3d_fp32():
    _print("case 3d fp32")
    F, X, OJO, MR, bt, G, u, cells, L = _full_case()
    B, N, dt = F.shape[2], F.shape[3], 0.004
    ref = O.integrate(B, 3, N, F, X, OJO, MR, bt, G, dt, L, u, cells, "euler")
    c = lambda a: a.astype(np.complex64)
    cells32 = {k: v.astype(np.float32) for k, v in cells.items()}
    out = hpup.pendulum_propagate32(c(F), c(X), c(OJO), c(MR), bt.astype(np.float32), G.astype(np.float32), dt, L,
                                    c(u), cells32, "euler")
    assert rel_l2(out, ref) < 2e-5


def test_bad_method_is_an_error():
    F, X, MR, bt, G, u, cells, L = _reduced_case(B=2, N=4)
    with pytest.raises(hpup.HpupError, match="method"):
        hpup.pendulum_propagate_reduced(F, X, MR, bt, G, 0.01, L, u, cells, "leapfrog")


@pytest.mark.parametrize("B,N,method", [(4, 8, "euler"), (4, 8, "RK4"), (3, 34, "RK2")])
def test_propagate_reduced_den(B, N, method):
    """Density-form gateway (gravity on the grid, pendulum_reduced_den.m); N = 34
    runs the 64-point drift kernel (more than 64 KB of LDS) and the large-grid
    transform kernels."""
    _print("case reduced den", B, N, method)
    L = 3.4
    f = density(B, 2, N, L, (8, 4, 2), seed=6, Sigma=np.eye(2) * 0.6)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F = O.fft_forward(B, f, d, w, 2, N)
    X, _, _, u, _ = operators(B, 2, N, L)
    mR = O.grid_mR(B)
    bt, G, dt = np.array([0.2, 0.25]), np.diag([0.1, 0.15]), 0.01
    ref = O.integrate_den(B, N, F, f, X, mR, bt, G, dt, L, u, d, w, method)
    out = hpup.pendulum_propagate_reduced_den(F, f, X, mR, bt, G, dt, L, u, d, w, method)
    assert rel_l2(out, ref) < 1e-10


def test_mixing_without_structure():
    """u_1 given a real part: the mixing coefficients lose the models'
    structure (u_1 imaginary, u_2 real, -i CG_1 real, -i CG_2 imaginary), the
    host detects it and the general complex mixing runs, against the oracle
    with the same u, at the 32-point and the 64-point tile."""
    for N in (8, 40):
        _print("case unstructured mixing", N)
        F, X, MR, bt, G, u, cells, L = _reduced_case(B=3, N=N)
        u2 = u.copy()
        u2[..., 0] = u2[..., 0] + 0.3 * np.abs(u2[..., 0])
        B, dt = F.shape[2], 0.01
        ref = O.integrate(B, 2, N, F, X, None, MR, bt, G, dt, L, u2, cells, "RK2")
        out = hpup.pendulum_propagate_reduced(F, X, MR, bt, G, dt, L, u2, cells, "RK2")
        assert rel_l2(out, ref) < 1e-10
        assert rel_l2(out, O.integrate(B, 2, N, F, X, None, MR, bt, G, dt, L, u, cells, "RK2")) > 1e-6


def test_reduced_gateway_large_omega_grid():
    """N = 34 on the reduced gateway: the 64-point MFMA drift kernel."""
    _print("case reduced N=34")
    F, X, MR, bt, G, u, cells, L = _reduced_case(B=2, N=34)
    B, N, dt = F.shape[2], F.shape[3], 0.01
    ref = O.integrate(B, 2, N, F, X, None, MR, bt, G, dt, L, u, cells, "RK2")
    out = hpup.pendulum_propagate_reduced(F, X, MR, bt, G, dt, L, u, cells, "RK2")
    assert rel_l2(out, ref) < 1e-10


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
def test_plan_steps_large_omega_grid(precision, tol):
    """A device-resident plan with N = 40 (the 64-point MFMA drift kernel)
    through two Horner-form RK4 steps, against the oracle's RK4."""
    import torch
    _print("case plan N=40", precision)
    B, N, L, dt = 3, 40, 8.0, 0.01
    f = density(B, 2, N, L, (8, 4, 2), seed=9, Sigma=np.eye(2) * 1.2)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Ffull = O.fft_forward(B, f, d, w, 2, N)
    X, _, MR, u, cells = operators(B, 2, N, L)
    bt, G = np.array([0.2, 0.3]), np.diag([0.1, 0.15])
    ref = Ffull
    for _ in range(2):
        ref = O.integrate(B, 2, N, ref, X, None, MR, bt, G, dt, L, u, cells, "RK4")
    p = hpup.Plan(B, 2, N, L, b=(0.2, 0.3, 0.0), G=G, dt=dt, method="RK4", precision=precision)
    F = p.upload_full(Ffull)
    p.step(F, 2)
    torch.cuda.synchronize()
    assert rel_l2(p.download_full(F), ref) < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-5), ("bf16x3", 2e-5)])
def test_plan_derivative_and_mass(precision, tol):
    """Device-resident plan: derivative against the oracle (plan built from the
    physical parameters, so it also checks the drivers' X/OJO/MR recipe), and
    total probability F^0_00(k=0) unchanged by a step."""
    import torch
    _print("case plan", precision)
    B, N, L = 4, 6, 10.0
    Jt = np.diag([1.0, 0.93539, 0.24935])
    Sigma = np.array([[0.8, 0.1, 0.0], [0.1, 0.6, 0.05], [0.0, 0.05, 1.0]])
    f = density(B, 3, N, L, (10, 5, 1), seed=7, Sigma=Sigma)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Ffull = O.fft_forward(B, f, d, w, 3, N)
    X, OJO, MR, u, cells = operators(B, 3, N, L, Jt)
    bt, G = np.array([0.2, 0.15, 0.25]), np.diag([0.1, 0.2, 0.05])
    ref = O.derivative(B, 3, N, Ffull, X, OJO, MR, bt, G, L, u, cells)
    p = hpup.Plan(B, 3, N, L, b=bt, G=G, Jt=Jt, dt=0.004, method="euler", precision=precision)
    F = p.upload_full(Ffull)
    dF = p.derivative(F)
    torch.cuda.synchronize()
    assert rel_l2(p.download_full(dF), ref) < tol
    m0 = float(F[p.mass_index].real)
    p.step(F, 3)
    torch.cuda.synchronize()
    mass_tol = 1e-12 if precision == "fp64" else 1e-6
    assert abs(float(F[p.mass_index].real) - m0) <= mass_tol * abs(m0)


@pytest.mark.parametrize("method,nsteps", [("euler", 1), ("euler", 2), ("euler", 3), ("midpoint", 2), ("RK2", 2),
                                            ("RK4", 1), ("RK4", 2)])
def test_plan_steps_vs_oracle(method, nsteps):
    """Several steps on the device-resident plan against the oracle
    integrator, at B = 5 (3-D), where the grid holds more workgroups than run
    at once: the fused stage reads neighbouring rotation rows of its input, so
    its stage buffers alternate (an odd Euler count ends with a copy)."""
    import torch
    _print("case plan steps", method, nsteps)
    B, N, L, dt = 5, 6, 10.0, 0.004
    Jt = np.diag([1.0, 0.93539, 0.24935])
    Sigma = np.array([[0.8, 0.1, 0.0], [0.1, 0.6, 0.05], [0.0, 0.05, 1.0]])
    f = density(B, 3, N, L, (10, 5, 1), seed=8, Sigma=Sigma)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Ffull = O.fft_forward(B, f, d, w, 3, N)
    X, OJO, MR, u, cells = operators(B, 3, N, L, Jt)
    bt, G = np.array([0.2, 0.15, 0.25]), np.diag([0.1, 0.2, 0.05])
    ref = Ffull
    for _ in range(nsteps):
        ref = O.integrate(B, 3, N, ref, X, OJO, MR, bt, G, dt, L, u, cells, method)
    p = hpup.Plan(B, 3, N, L, b=bt, G=G, Jt=Jt, dt=dt, method=method, precision="fp64")
    F = p.upload_full(Ffull)
    p.step(F, nsteps)
    torch.cuda.synchronize()
    assert rel_l2(p.download_full(F), ref) < 1e-10


def test_plan_transforms_roundtrip():
    """forward(backward(F)) == F for bandlimited coefficients (plan API, FP64)."""
    import torch
    _print("case plan transforms")
    B, N, L = 5, 8, 6.0
    f = density(B, 3, N, L, (6, 3, 1), seed=9)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Ffull = O.fft_forward(B, f, d, w, 3, N)
    p = hpup.Plan(B, 3, N, L, precision="fp64")
    F = p.upload_full(Ffull)
    g = p.backward(F)
    F2 = p.forward(g)
    torch.cuda.synchronize()
    assert rel_l2(p.download_full(F2), Ffull) < 1e-12
    f_ref = O.fft_backward(B, Ffull, d, 3, N)
    assert rel_l2(g.cpu().numpy(), f_ref.ravel(order="F")) < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_plan_transforms_b15(precision, tol):
    """Plan forward and backward at B = 15 (2B = 30: the prime-factor FFT form of
    the (alpha, gamma) stages) against the oracle's transforms."""
    import torch
    _print("case plan transforms B=15", precision)
    B, N, L = 15, 4, 6.0
    f = density(B, 2, N, L, (6, 3, 1), seed=11)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Ffull = O.fft_forward(B, f, d, w, 2, N)
    f_ref = O.fft_backward(B, Ffull, d, 2, N)
    p = hpup.Plan(B, 2, N, L, precision=precision)
    F = p.upload_full(Ffull)
    g = p.backward(F)
    torch.cuda.synchronize()
    assert rel_l2(g.cpu().numpy().astype(np.float64), f_ref.ravel(order="F")) < tol
    fdev = torch.from_numpy(np.asfortranarray(f).ravel(order="F").astype(np.float64 if precision == "fp64" else np.float32)).cuda()
    F2 = p.forward(fdev)
    torch.cuda.synchronize()
    mask = np.broadcast_to(valid_mask(B, (N, N)), Ffull.shape)
    assert rel_l2(p.download_full(F2)[mask], Ffull[mask]) < tol
