"""The HIP path against the oracle on the benchmark's own omega grids: N = 30
in 3-D (B = 15, 30^3 over [-5, 5]^3, the propagation benchmark) and N = 32 in
2-D (B = 15, 32^2 over [-4, 4]^2, the reduced model), plus N = 18, just past
one 16-wide MFMA tile.  These sizes run the MFMA kernels with their masked
edges (30 of 32 padded rows, a partial last column chunk of the axis-3
kernel) and the chirp-Toeplitz omega DFTs at the edge of their 96-entry
tables, where the small-grid tests cannot reach.

Checkers: the dense oracle (the reference's literal algorithm) on sampled
output omega ranges -- the first k1 line (through the k1 Nyquist bin) and
lines in the k3 = N/2 Nyquist plane -- and the structured oracle (pinned
against the dense one) on every coefficient.  Tolerances: 1e-10 relative L2
(FP64), 2e-5 (FP32), 1e-12 / 2e-6 for the transforms."""
import importlib.util
import os

import numpy as np
import pytest

import hpup
import oracle as O
from helpers import density, operators, rel_l2

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JT = np.diag([1.0, 0.93539, 0.24935])


def _print(*a):
    print(*a, flush=True)


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _ranges(dim, N):
    """Output omega ranges [x0, x1): the first k1 line, and a k1 line in the
    last axis' Nyquist plane (with the middle axis at its Nyquist bin too)."""
    nx = N ** dim
    mid = (N // 2) * N ** (dim - 1) + (N // 2) * N ** (dim - 2) if dim == 3 else (N // 2) * N
    return [(0, N), (mid, mid + N)] if mid + N <= nx else [(0, N)]


def _sel(A, dim, N, x0, x1):
    M = A.shape[0]
    return A.reshape((M, M, A.shape[2], N ** dim), order="F")[..., x0:x1]


# ------------------------------------------------------------ transforms ----

@pytest.mark.parametrize("B,dim,N", [(2, 3, 30), (3, 3, 30), (3, 2, 32), (4, 2, 32), (3, 3, 18)])
def test_gateway_transforms_on_metric_grids(B, dim, N):
    _print("case transforms", B, dim, N)
    L = 10.0
    f = density(B, dim, N, L, (10, 5, 1), seed=40 + B + N, Sigma=np.eye(dim) * 0.9)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F_ref = O.fft_forward(B, f, d, w, dim, N)
    F = hpup.fftSO3R_mex("forward", f, d, w)
    assert rel_l2(F, F_ref) < 1e-12
    f_ref = O.fft_backward(B, F_ref, d, dim, N)
    assert rel_l2(hpup.fftSO3R_mex("backward", F_ref, d), f_ref) < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
@pytest.mark.parametrize("B,dim,N", [(3, 3, 30), (4, 2, 32), (30, 3, 4), (30, 2, 6)])
def test_plan_transforms_on_metric_grids(B, dim, N, precision, tol):
    # B = 30 (2B = 60): the prime-factor 4 x 3 x 5 FFT kernels of the angular stages
    import torch
    _print("case plan transforms", B, dim, N, precision)
    L = 10.0
    f = density(B, dim, N, L, (10, 5, 1), seed=50 + N, Sigma=np.eye(dim) * 0.9)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Ffull = O.fft_forward(B, f, d, w, dim, N)
    f_ref = O.fft_backward(B, Ffull, d, dim, N)
    p = hpup.Plan(B, dim, N, L, precision=precision)
    g = p.backward(p.upload_full(Ffull))
    torch.cuda.synchronize()
    assert rel_l2(g.cpu().numpy().astype(np.float64), f_ref.ravel(order="F")) < tol
    fdev = torch.from_numpy(np.asfortranarray(f).ravel(order="F")).to(p.rdtype).cuda()
    F2 = p.download_full(p.forward(fdev))
    from helpers import valid_mask
    mask = np.broadcast_to(valid_mask(B, (N,) * dim), Ffull.shape)
    assert rel_l2(F2[mask], Ffull[mask]) < tol


# -------------------------------------------------- derivative and steps ----

def _model_case(B, dim, N, L, seed):
    """Random coefficients on the valid triangle (every omega frequency,
    Nyquist bins included, carries O(1) values, so sampled lines anywhere in
    the cube are meaningful) and the drivers' operators."""
    from helpers import valid_mask
    rng = np.random.default_rng(seed)
    M = 2 * B - 1
    shp = (M, M, B) + (N,) * dim
    F = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * valid_mask(B, (N,) * dim)
    F = np.asfortranarray(F)
    X, OJO, MR, u, cells = operators(B, dim, N, L, JT if dim == 3 else None)
    bt = np.array([0.2, 0.15, 0.25])[:dim]
    G = np.diag([0.1, 0.2, 0.05])[:dim, :dim]
    return F, X, OJO, MR, u, cells, bt, G


@pytest.mark.parametrize("B,dim,N", [(3, 3, 30), (4, 3, 18), (3, 2, 32), (5, 2, 30), (4, 2, 18)])
def test_derivative_on_metric_grids(B, dim, N):
    """Gateway derivative (dt = 1 Euler step minus F) in FP64 and FP32 against
    the structured oracle everywhere and the dense oracle on sampled lines."""
    _print("case derivative", B, dim, N)
    L = 10.0 if dim == 3 else 8.0
    F, X, OJO, MR, u, cells, bt, G = _model_case(B, dim, N, L, seed=60 + N)
    ref = O.derivative_structured(B, dim, N, F, X, JT if dim == 3 else None, MR, bt, G, L, u, cells)
    for x0, x1 in _ranges(dim, N):
        dense = O.derivative(B, dim, N, F, X, OJO, MR, bt, G, L, u, cells, x0, x1)
        assert rel_l2(_sel(ref, dim, N, x0, x1), _sel(dense, dim, N, x0, x1)) < 1e-13
    if dim == 3:
        out = hpup.pendulum_propagate(F, X, OJO, MR, bt, G, 1.0, L, u, cells, "euler") - F
    else:
        out = hpup.pendulum_propagate_reduced(F, X, MR, bt, G, 1.0, L, u, cells, "euler") - F
    assert rel_l2(out, ref) < 1e-10
    c = lambda a: a.astype(np.complex64)
    cells32 = {k: v.astype(np.float32) for k, v in cells.items()}
    if dim == 3:
        o32 = hpup.pendulum_propagate32(c(F), c(X), c(OJO), c(MR), bt.astype(np.float32), G.astype(np.float32), 1.0,
                                        L, c(u), cells32, "euler")
    else:
        o32 = hpup.pendulum_propagate_reduced32(c(F), c(X), c(MR), bt.astype(np.float32), G.astype(np.float32), 1.0,
                                                L, c(u), cells32, "euler")
    err = rel_l2(o32.astype(np.complex128) - F, ref)
    _print("  fp32 derivative rel err %.2e" % err)
    assert err < 2e-5


@pytest.mark.parametrize("B,dim,N,method", [(3, 3, 30, "RK4"), (4, 2, 32, "RK4"), (3, 3, 18, "midpoint"),
                                            (4, 2, 30, "RK2")])
def test_plan_steps_on_metric_grids(B, dim, N, method):
    """Device-resident plan (built from the physical parameters): two steps
    against the structured oracle's integrator, FP64 and FP32."""
    import torch
    _print("case plan steps", B, dim, N, method)
    L = 10.0 if dim == 3 else 8.0
    dt = 0.004
    F, X, OJO, MR, u, cells, bt, G = _model_case(B, dim, N, L, seed=70 + N)
    ref = F
    for _ in range(2):
        ref = O.integrate_structured(B, dim, N, ref, X, JT if dim == 3 else None, MR, bt, G, dt, L, u, cells, method)
    for precision, tol in (("fp64", 1e-10), ("fp32", 2e-5)):
        p = hpup.Plan(B, dim, N, L, b=tuple(bt) + (0.0,) * (3 - dim), G=G, Jt=JT, dt=dt, method=method,
                      precision=precision)
        Fd = p.upload_full(F)
        p.step(Fd, 2)
        torch.cuda.synchronize()
        e = rel_l2(p.download_full(Fd), ref)
        _print("  %s rel err %.2e" % (precision, e))
        assert e < tol


# ------------------------------------------ the benchmark's own workloads ----

def _bench_plan_derivative(precision):
    """The benchmark's FP64 initial coefficients (B = 15, 30^3), the plan's
    derivative in `precision`, and both in the MATLAB full layout."""
    import torch
    bm = _bench()
    cfg = dict(bm.CFG)
    p64 = bm.make_plan(cfg, "fp64", "RK4")
    F0 = bm.initial_coeffs(p64, cfg)
    if precision == "fp64":
        dF = p64.derivative(F0)
        plan = p64
    else:
        plan = bm.make_plan(cfg, precision, "RK4")
        dF = plan.derivative(F0.to(torch.complex64))
    torch.cuda.synchronize()
    Ffull = p64.download_full(F0)
    dFfull = plan.download_full(dF)
    del F0, dF
    torch.cuda.empty_cache()
    return cfg, Ffull, dFfull


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-5), ("bf16x3", 2e-5)])
def test_benchmark_workload_derivative(precision, tol):
    """B = 15, 30^3 omega grid over [-5, 5]^3, the benchmark's seeded
    matrix-Fisher x Gaussian coefficients: the plan's derivative against the
    dense oracle on two sampled k1 lines ((k2, k3) = (0, 0) through the k1
    Nyquist bin, and (2, -3)) and against the structured oracle on all 121 M coefficients."""
    _print("case benchmark workload derivative", precision)
    cfg, Ffull, dF = _bench_plan_derivative(precision)
    B, N, L = cfg["B"], cfg["N"], cfg["L"]
    X, OJO, MR, u, cells = operators(B, 3, N, L, JT)
    bt, G = np.array(cfg["b"]), np.eye(3) * cfg["G"]
    # the first k1 line, and the k1 line at (k2, k3) = (2, -3): the seeded
    # density's spectrum is at rounding level in the Nyquist planes
    mid = (N - 3) * N * N + 2 * N
    for x0, x1 in ((0, N), (mid, mid + N)):
        dense = O.derivative(B, 3, N, Ffull, X, OJO, MR, bt, G, L, u, cells, x0, x1)
        e = rel_l2(_sel(dF, 3, N, x0, x1), _sel(dense, 3, N, x0, x1))
        _print("  dense oracle, omega points [%d, %d): rel err %.2e" % (x0, x1, e))
        assert e < tol
        del dense
    ref = O.derivative_structured(B, 3, N, Ffull, X, JT, MR, bt, G, L, u, cells)
    e = rel_l2(dF, ref)
    _print("  structured oracle, all coefficients: rel err %.2e" % e)
    assert e < tol


def test_reduced_workload_derivative_and_step():
    """B = 15, 32^2 omega grid over [-4, 4]^2 (the reduced model's workload):
    FP64 and FP32 plan derivatives against the structured oracle everywhere
    and the dense oracle on sampled lines, and one RK4 step of 0.005."""
    import torch
    from hpup import density as D
    _print("case reduced workload")
    B, N, L, dt = 15, 32, 8.0, 0.005
    rng = np.random.default_rng(2)
    bt, G = np.array([0.2, 0.2]), np.eye(2) * 0.1
    p64 = hpup.Plan(B, 2, N, L, b=(0.2, 0.2, 0.0), G=G, dt=dt, method="RK4", precision="fp64")
    p32 = hpup.Plan(B, 2, N, L, b=(0.2, 0.2, 0.0), G=G, dt=dt, method="RK4", precision="fp32")
    fR, fx = D.matrix_fisher_gaussian(B, 2, N, L, (8, 4, 2), D.random_rotation(rng), Sigma=np.eye(2))
    F0 = p64.forward(D.grid_density(fR, fx))
    Ffull = p64.download_full(F0)
    X, _, MR, u, cells = operators(B, 2, N, L)
    ref = O.derivative_structured(B, 2, N, Ffull, X, None, MR, bt, G, L, u, cells)
    for x0, x1 in ((0, N), (3 * N, 4 * N)):
        dense = O.derivative(B, 2, N, Ffull, X, None, MR, bt, G, L, u, cells, x0, x1)
        assert rel_l2(_sel(ref, 2, N, x0, x1), _sel(dense, 2, N, x0, x1)) < 1e-13
    e64 = rel_l2(p64.download_full(p64.derivative(F0)), ref)
    e32 = rel_l2(p32.download_full(p32.derivative(F0.to(torch.complex64))), ref)
    step = O.integrate_structured(B, 2, N, Ffull, X, None, MR, bt, G, dt, L, u, cells, "RK4")
    F64, F32 = F0.clone(), F0.to(torch.complex64)
    p64.step(F64, 1)
    p32.step(F32, 1)
    torch.cuda.synchronize()
    s64 = rel_l2(p64.download_full(F64), step)
    s32 = rel_l2(p32.download_full(F32), step)
    _print("  derivative fp64 %.2e fp32 %.2e; RK4 step fp64 %.2e fp32 %.2e" % (e64, e32, s64, s32))
    assert e64 < 1e-10 and e32 < 2e-5 and s64 < 1e-10 and s32 < 2e-5
