"""Matrix-Fisher statistics (get_stat's fit).  CPU tests:

* the oracle restatement (oracle/matfisher.py) pinned against brute-force
  Haar quadrature on the SO(3) grid: c(s) = E_Haar[exp(tr(diag(s) Q))], its
  gradient E[Q_ii e^...] and Hessian E[Q_ii Q_jj e^...], at generic s and at
  every repeated-|s| branch of the reference's Hessian system;
* the oracle's moment fit recovering known s, psvd invariants, and get_stat
  recovering the parameters of a matrix-Fisher-Gaussian density built with
  known (Miu, P, Sigma);
* the library's host-side fit (pdf_MF_normal, pdf_MF_M2S, psvd through the C
  ABI; no GPU involved) against the oracle.
"""
import numpy as np
import pytest

import hpup
import matfisher as MF
import oracle as O
from helpers import omega_grid, rot_random

# proper and degenerate singular-value sets; each repeated-|s| branch of
# pdf_MF_normal.m:80-205 is taken by at least one of them
S_CASES = [(0.0, 0.0, 0.0), (1.2, 0.5, 0.2), (3.0, 1.0, -0.5), (2.0, 2.0, 0.5), (1.0, 1.0, 1.0),
           (1.5, 1.5, -1.5), (2.0, 0.0, 0.0), (0.7, 0.7, 0.0), (1.0, 0.4, 0.4), (1.0, 0.4, -0.4),
           (2.5, -2.5, 0.3)]


def _haar(B=24):
    R = O.rotations(B).reshape(3, 3, -1)
    _, _, _, w = O.grid(B)
    n2 = 2 * B
    wt = np.broadcast_to(w[None, :, None], (n2, n2, n2)).reshape(-1)
    return R, wt


_HAAR = None


def brute(s):
    """c, grad c, Hessian c by quadrature on the B = 24 grid (unscaled)."""
    global _HAAR
    if _HAAR is None:
        _HAAR = _haar()
    R, wt = _HAAR
    diag = np.stack([R[0, 0], R[1, 1], R[2, 2]])
    e = np.exp(np.asarray(s) @ diag) * wt
    return e.sum(), diag @ e, (diag * e) @ diag.T


def _scaled(s, c, g, H):
    k = np.exp(-np.sum(s))
    one = np.ones(3)
    return k * c, k * (g - c), k * (H - np.outer(g, one) - np.outer(one, g) + c)


@pytest.mark.parametrize("s", S_CASES)
def test_oracle_normalizer_vs_haar(s):
    c, g, H = brute(s)
    for scaled in (False, True):
        cr, gr, Hr = _scaled(s, c, g, H) if scaled else (c, g, H)
        co, go, Ho = MF.mf_normal(s, scaled)
        if scaled and min(s[0] + s[1], s[0] + s[2], s[1] + s[2]) < 0:
            continue  # the scaled integrands assume s_i + s_j >= 0 (proper singular values)
        assert abs(co - cr) <= 1e-12 * cr, (s, scaled)
        assert np.abs(go - gr).max() <= 1e-11 * cr, (s, scaled)
        assert np.abs(Ho - Hr).max() <= 1e-10 * cr, (s, scaled)


@pytest.mark.parametrize("s", S_CASES + [(10.0, 5.0, -2.0), (40.0, 30.0, 20.0), (300.0, 200.0, 150.0)])
def test_library_normalizer_vs_oracle(s):
    for scaled in (False, True):
        if not scaled and max(abs(v) for v in s) > 20:
            continue
        if scaled and min(s[0] + s[1], s[0] + s[2], s[1] + s[2]) < 0:
            continue
        co, go, Ho = MF.mf_normal(s, scaled)
        c, g, H = hpup.pdf_MF_normal(s, scaled)
        assert abs(c - co) <= 1e-13 * co
        assert np.abs(g - go).max() <= 1e-12 * co
        assert np.abs(H - Ho).max() <= 1e-12 * co


def test_library_normalizer_off_proper_domain_vs_haar():
    """Off the proper domain (s_i + s_j < 0) the library's direct form still
    matches Haar quadrature."""
    s = (1.0, -1.0, 0.3)
    c, g, H = brute(s)
    cl, gl, Hl = hpup.pdf_MF_normal(s, False)
    assert abs(cl - c) <= 1e-12 * c and np.abs(gl - g).max() <= 1e-11 * c and np.abs(Hl - H).max() <= 1e-11 * c


@pytest.mark.parametrize("s", [(1.2, 0.5, 0.2), (3.0, 1.0, -0.5), (2.0, 2.0, 0.5), (0.4, 0.1, 0.05)])
def test_oracle_fit_recovers_s(s):
    c, g, _ = brute(s)
    d = g / c  # E[Q_ii] under the matrix Fisher density with F = diag(s)
    so, fval, it = MF.m2s(d)
    assert np.max(np.abs(fval)) <= 1e-12 and it < 100
    assert np.abs(so - s).max() <= 1e-9


# d inside the moment polytope (the convex hull of diag(Q), Q in SO(3)),
# from diffuse to concentrated
@pytest.mark.parametrize("d", [(0.6, 0.4, 0.1), (0.9, 0.85, 0.8), (0.2, 0.1, -0.05), (0.96, 0.95, 0.93),
                               (0.3, 0.3, 0.3), (0.0, 0.0, 0.0), (0.4, 0.3, -0.2)])
def test_library_fit_vs_oracle(d):
    so, _, _ = MF.m2s(d)
    s, fval, it = hpup.pdf_MF_M2S(d)
    assert it >= 0 and np.max(np.abs(fval)) <= 1e-12
    assert np.abs(s - so).max() <= 1e-9 * max(1.0, np.abs(so).max())


def test_library_fit_reports_non_convergence():
    """d outside the moment polytope (here d1 + d2 - d3 > 1) has no solution:
    the fit says so instead of reporting a NaN iterate as converged."""
    for d in ([1.2, 0.5, 0.1], [0.97, 0.96, -0.9]):
        s, fval, it = hpup.pdf_MF_M2S(d)
        assert it == -1


def _align(U, Uo):
    """D = diag(+-1), det 1, with U D ~ Uo (the proper SVD's symmetry)."""
    return np.diag(np.sign(np.sum(U * Uo, axis=0)))


@pytest.mark.parametrize("seed", range(6))
def test_psvd_vs_oracle(seed):
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((3, 3))
    if seed == 4:
        F[:, 2] = F[:, 0] + F[:, 1]  # rank 2
    if seed == 5:
        F = rot_random(rng) @ np.diag([0.9, 0.5, -0.2]) @ rot_random(rng).T
    U, S, V = hpup.psvd(F)
    Uo, So, Vo = MF.psvd(F)
    assert np.abs(U @ S @ V.T - F).max() <= 1e-14 * 10
    for M in (U, V):
        assert abs(np.linalg.det(M) - 1) <= 1e-13 and np.abs(M.T @ M - np.eye(3)).max() <= 1e-13
    s = np.diag(S)
    assert s[0] >= s[1] >= abs(s[2]) - 1e-15
    assert np.abs(S - So).max() <= 1e-14
    if seed != 4:  # distinct singular values: unique up to D
        D = _align(U, Uo)
        assert abs(np.linalg.det(D) - 1) < 1e-12
        assert np.abs(U @ D - Uo).max() <= 1e-12 and np.abs(V @ D - Vo).max() <= 1e-12


def mfg_density(B, dim, N, L, s0, mu, P0, Sc, seed=0):
    """Matrix-Fisher-Gaussian density on the grid:
    f(R, x) ~ exp(tr(F0^T R)) N(x; mu + P0 vR(R), Sc), F0 = U0 diag(s0) V0^T,
    vR from Q = U0^T R V0 (the get_stat definition).  Returns f, R, x, w, U0."""
    rng = np.random.default_rng(seed)
    U0, V0 = rot_random(rng), rot_random(rng)
    R = O.rotations(B)
    _, _, _, w = O.grid(B)
    Q = np.einsum("ca,cdijk,db->abijk", U0, R, V0)
    fR = np.exp(np.einsum("ii...->...", np.einsum("a,abijk->abijk", s0, Q)))
    vR = MF.vR_of(Q, s0)  # (3, n2, n2, n2)
    x = omega_grid(dim, N, L)
    m = mu.reshape((dim,) + (1,) * 3) + np.einsum("ai,ijkl->ajkl", P0, vR)  # (dim, n2, n2, n2)
    Si = np.linalg.inv(Sc)
    dlt = x.reshape((dim, 1, 1, 1) + x.shape[1:]) - m.reshape(m.shape + (1,) * dim)
    q = np.einsum("a...,ab,b...->...", dlt, Si, dlt)
    f = fR.reshape(fR.shape + (1,) * dim) * np.exp(-0.5 * q)
    wb = w.reshape((1, 2 * B, 1) + (1,) * dim)
    f /= np.sum(f * wb) * (L / N) ** dim
    return np.asfortranarray(f), R, x, w, U0


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_get_stat_recovers_mfg(dim):
    """get_stat of a matrix-Fisher-Gaussian density returns its parameters:
    S = diag(s0), Miu = mu, P = P0 (up to the proper-SVD sign symmetry), and
    Sigma = Sc + P0 (tr(S) I - S) P0^T (pendulum.m:582-584)."""
    s0 = np.array([3.0, 2.0, 1.0])
    B, N, L = (8, 20, 10.0) if dim == 2 else (7, 16, 10.0)
    mu = np.array([0.3, -0.2, 0.1])[:dim]
    P0 = np.array([[0.2, -0.1, 0.05], [0.0, 0.15, -0.1], [0.05, 0.0, 0.1]])[:dim]
    Sc = np.array([[0.5, 0.1, 0.0], [0.1, 0.4, 0.05], [0.0, 0.05, 0.45]])[:dim, :dim]
    f, R, x, w, U0 = mfg_density(B, dim, N, L, s0, mu, P0, Sc)
    st = MF.get_stat(f, R, x, w)
    tol = 1e-6  # grid-quadrature accuracy of this band and box
    assert np.abs(np.diag(st["S"]) - s0).max() <= tol
    assert np.abs(st["Miu"] - mu).max() <= tol
    D = _align(st["U"], U0)
    assert np.abs(st["P"] @ D - P0).max() <= tol
    Sig = Sc + P0 @ (np.sum(s0) * np.eye(3) - np.diag(s0)) @ P0.T
    assert np.abs(st["Sigma"] - Sig).max() <= tol
