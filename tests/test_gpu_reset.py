"""Hybrid reset gateways on the GPU against the oracle (reduced model, 2-D
omega), with the hybrid driver's parameters (pendulum_reduced_hybrid.m:28-49).
Indices must agree exactly (a top-nD slot may hold another source only where
two Gaussian densities tie to rounding); values to FP64 rounding."""
import numpy as np
import pytest

import hpup
import oracle as O
from helpers import density, omega_grid, rel_l2

pytestmark = pytest.mark.gpu

B, N, L = 3, 12, 3.4
PAR = dict(d=0.12, h=0.2, r=0.025, theta_t=5 * np.pi / 180, lambda_max=100.0 * 0.1177)
EPS = 0.8


@pytest.fixture(scope="module")
def setup():
    R = O.rotations(B)
    x = omega_grid(2, N, L)
    return R, x


def test_getLambda(setup):
    print("case getLambda", flush=True)
    R, x = setup
    lam, indR, PC = hpup.getLambda_mex(R, x, **PAR)
    lam_o, indR_o, PC_o = O.getLambda(B, N, R, x, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"], PAR["lambda_max"])
    assert len(indR) > 0
    np.testing.assert_array_equal(indR, indR_o)
    assert rel_l2(PC, PC_o) < 1e-14
    assert np.array_equal(lam == 0, lam_o == 0)
    assert rel_l2(lam, lam_o) < 1e-14
    lc, ic, lists, _ = hpup.getLambda_mex(R, x, compact=True, **PAR)
    lc_o, ic_o, lists_o, _ = O.getLambda_compact(B, N, R, x, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"],
                                                 PAR["lambda_max"])
    np.testing.assert_array_equal(ic, ic_o)
    assert rel_l2(lc, lc_o) < 1e-14
    assert all(np.array_equal(a, b) for a, b in zip(lists, lists_o))


@pytest.mark.parametrize("direction", ["new", "old"])
def test_getOmega_and_indrule(setup, direction):
    print("case getOmega", direction, flush=True)
    R, x = setup
    _, indR, PC = O.getLambda(B, N, R, x, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"], PAR["lambda_max"])
    Om = hpup.getOmega_mex(R, x, indR, EPS, PC, direction)
    Om_o = O.getOmega(B, N, R, x, indR, EPS, PC, direction)
    assert np.array_equal(np.isnan(Om), np.isnan(Om_o))
    ok = ~np.isnan(Om_o)
    assert np.abs(Om[ok] - Om_o[ok]).max() < 1e-13
    if direction == "old":
        ind, co = hpup.getIndRule_mex(x, Om_o)
        ind_o, co_o = O.getIndRule(N, x, Om_o)
        np.testing.assert_array_equal(ind, ind_o)
        assert np.array_equal(np.isnan(co), np.isnan(co_o))
        okc = ~np.isnan(co_o)
        assert np.abs(co[okc] - co_o[okc]).max() < 1e-13


def test_getFcL_and_noisy_propagate(setup):
    print("case getFcL", flush=True)
    R, x = setup
    lam, indR, PC = O.getLambda(B, N, R, x, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"], PAR["lambda_max"])
    Om = O.getOmega(B, N, R, x, indR, EPS, PC, "new")
    Gd = np.eye(2) * 0.05 ** 2 * 25
    nD = 10
    fcL, idx = hpup.getFcL_mex(x, Om, lam, Gd, nD)
    fcL_o, idx_o = O.getFcL(N, x, Om, lam, Gd, nD)
    # the normalizers are summed in a fixed order: a second call is bit-identical
    fcL2, idx2 = hpup.getFcL_mex(x, Om, lam, Gd, nD)
    assert np.array_equal(idx2, idx) and np.array_equal(fcL2.view(np.int64), fcL.view(np.int64))
    # identical top-nD selection; a slot may hold another source only where the
    # two candidates' Gaussian densities (numpy) tie to rounding, i.e. where
    # the last ulp of exp decides a stable sort
    nx, numR = N * N, len(indR)
    xl = np.asfortranarray(x).reshape(2, nx, order="F")
    Omf = Om.reshape(2, numR, nx, order="F")
    Gi = np.linalg.inv(Gd)

    def dens(k, dest, src):
        dv = xl[:, dest] - Omf[:, k, src]
        return np.exp(-0.5 * dv @ Gi @ dv)
    diff = np.argwhere(idx != idx_o)
    for t, k, d1, d2 in diff:
        dest = d1 + N * d2
        a, b = dens(k, dest, int(idx[t, k, d1, d2])), dens(k, dest, int(idx_o[t, k, d1, d2]))
        assert abs(a - b) <= 1e-13 * max(a, b), (t, k, dest, a, b)
    print("  top-nD slots holding a tied source: %d of %d" % (len(diff), idx.size), flush=True)
    # a tied slot moves a density from one source's normalizer to another's,
    # so every weight of those two sources (in that region point) changes;
    # all other weights agree to rounding
    touched = set()
    for t, k, d1, d2 in diff:
        touched.add((int(k), int(idx[t, k, d1, d2])))
        touched.add((int(k), int(idx_o[t, k, d1, d2])))
    kk = np.broadcast_to(np.arange(numR).reshape(1, numR, 1, 1), idx.shape)
    clean = (idx == idx_o) & ~np.vectorize(lambda k, s: (int(k), int(s)) in touched)(kk, idx_o)
    assert clean.sum() > 0.99 * idx.size
    assert np.abs(fcL[clean] - fcL_o[clean]).max() <= 1e-12 * np.abs(fcL_o).max()
    f = density(B, 2, N, L, (8, 4, 2), seed=4, Sigma=np.eye(2) * 0.5)
    df = hpup.pendulum_reduced_discrete_propagate(True, f, lam, indR, fcL_o, idx_o)
    df_o = O.discrete_noise(B, N, f, lam, indR, fcL_o, idx_o)
    assert rel_l2(df, df_o) < 1e-13


def test_nonoise_propagate(setup):
    print("case nonoise", flush=True)
    R, x = setup
    lc, indR, lists, PC = O.getLambda_compact(B, N, R, x, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"],
                                              PAR["lambda_max"])
    Om = O.getOmega(B, N, R, x, indR, EPS, PC, "old")
    ind, co = O.getIndRule(N, x, Om)
    f = density(B, 2, N, L, (8, 4, 2), seed=5, Sigma=np.eye(2) * 0.5)
    df = hpup.pendulum_reduced_discrete_propagate(False, f, lc, indR, lists, ind, co)
    df_o = O.discrete_nonoise(B, N, f, lc, indR, lists, ind, co)
    assert rel_l2(df, df_o) < 1e-13


def test_out_of_range_indices_are_rejected(setup):
    R, x = setup
    lam, indR, PC = O.getLambda(B, N, R, x, PAR["d"], PAR["h"], PAR["r"], PAR["theta_t"], PAR["lambda_max"])
    f = np.zeros((2 * B,) * 3 + (N, N), order="F")
    fcL = np.zeros((2, len(indR), N, N), order="F")
    idx = np.full((2, len(indR), N, N), N * N, dtype=np.int16, order="F")  # one past the end
    with pytest.raises(hpup.HpupError, match="out of range"):
        hpup.pendulum_reduced_discrete_propagate(True, f, lam, indR, fcL, idx)
