"""get_stat on the GPU (one device pass over the grid density, then the host
fit) against the oracle's get_stat on the same grid (pendulum.m:547-594,
pendulum_reduced.m:485-532).  Tolerances: 1e-10 absolute on the FP64 paths
(the fit's own stopping rule is 1e-12 in the moment residual), 2e-5 for a
density stored in FP32.  U, V, EvR, EvRvR, ExvR and P are defined up to the
proper SVD's sign symmetry D (diag of +-1, det 1); they are compared after
aligning D, the other outputs directly."""
import numpy as np
import pytest

import hpup
import matfisher as MF
import oracle as O
from test_matfisher import _align, mfg_density

pytestmark = pytest.mark.gpu

INVARIANT = ("ER", "Ex", "Varx", "S", "Miu", "Sigma")


def _print(*a):
    print(*a, flush=True)


def _compare(st, ref, tol):
    for k in INVARIANT:
        assert np.abs(np.asarray(st[k]) - np.asarray(ref[k])).max() <= tol, k
    D = _align(st["U"], ref["U"])
    assert abs(np.linalg.det(D) - 1) < 1e-9
    assert np.abs(st["U"] @ D - ref["U"]).max() <= tol
    assert np.abs(st["V"] @ D - ref["V"]).max() <= tol
    assert np.abs(D @ st["EvR"] - ref["EvR"]).max() <= tol
    assert np.abs(D @ st["EvRvR"] @ D - ref["EvRvR"]).max() <= tol
    assert np.abs(st["ExvR"] @ D - ref["ExvR"]).max() <= tol
    assert np.abs(st["P"] @ D - ref["P"]).max() <= tol


def _case(dim):
    s0 = np.array([4.0, 2.5, 1.0])
    B, N, L = (6, 16, 9.0) if dim == 2 else (5, 10, 9.0)
    mu = np.array([0.3, -0.2, 0.1])[:dim]
    P0 = np.array([[0.2, -0.1, 0.05], [0.0, 0.15, -0.1], [0.05, 0.0, 0.1]])[:dim]
    Sc = np.array([[0.6, 0.1, 0.0], [0.1, 0.5, 0.05], [0.0, 0.05, 0.55]])[:dim, :dim]
    return mfg_density(B, dim, N, L, s0, mu, P0, Sc, seed=dim), B, N, L


@pytest.mark.parametrize("dim", [2, 3])
def test_get_stat_gateway(dim):
    _print("case get_stat gateway dim", dim)
    (f, R, x, w, _), B, N, L = _case(dim)
    ref = MF.get_stat(f, R, x, w)
    st = hpup.get_stat(f, R, x, w, return_fx=True)
    assert st["fit_ok"] and st["niter"] >= 0
    _compare(st, ref, 1e-10)
    assert np.abs(st["fx"] - ref["fx"]).max() <= 1e-12 * np.abs(ref["fx"]).max()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
def test_plan_get_stat(precision, tol):
    """Plan path: coefficients -> inverse transform on the device -> get_stat
    on the device grid, against the oracle's get_stat of the same grid."""
    import torch
    _print("case plan get_stat", precision)
    dim = 3
    (f, R, x, w, _), B, N, L = _case(dim)
    d = O.wigner_table(B)
    Ffull = O.fft_forward(B, f, d, w, dim, N)
    p = hpup.Plan(B, dim, N, L, precision=precision)
    F = p.upload_full(Ffull)
    g = p.backward(F)
    torch.cuda.synchronize()
    fg = g.double().cpu().numpy().reshape(f.shape, order="F")
    ref = MF.get_stat(fg, R, x, w)
    st = p.get_stat(g, return_fx=True)
    assert st["fit_ok"]
    _compare(st, ref, tol)
    ER, Ex, Varx, fx = p.stat_moments(g)
    assert np.abs(ER - ref["ER"]).max() <= tol and np.abs(Ex - ref["Ex"]).max() <= tol
    assert np.abs(Varx - ref["Varx"]).max() <= tol and np.abs(fx - ref["fx"]).max() <= tol * np.abs(ref["fx"]).max()


def test_get_stat_large_grid_partials():
    """A grid with many rotation tiles and omega chunks (B = 8, N = 24, dim 2:
    4 tiles x 9 chunks) so that the fixed-order partial reductions are
    exercised beyond one workgroup."""
    _print("case get_stat large grid")
    s0 = np.array([6.0, 3.0, -1.0])
    f, R, x, w, _ = mfg_density(8, 2, 24, 10.0, s0, np.array([0.1, 0.2]),
                                np.array([[0.1, 0.0, 0.2], [0.0, -0.1, 0.1]]), np.diag([0.5, 0.7]), seed=11)
    ref = MF.get_stat(f, R, x, w)
    st = hpup.get_stat(f, R, x, w)
    _compare(st, ref, 1e-10)
