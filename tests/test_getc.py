"""pendulum_plot_getc (getc.cpp:7-40): the oracle's restatement pinned by
closed forms.  A density f(R, x) = (1 + tr(A^T R)) h(x) has a rotation marginal
of degree one in R (bandlimited to l <= 1), so the transform and its point
evaluation reproduce 1 + tr(A^T R(alpha, beta, gamma)) at ANY Euler angles
(R = Rz(alpha) Ry(beta) Rz(gamma), the grid's convention, fftSO3R.m:38-49); and
at grid angles the point evaluation equals the inverse transform on the grid."""
import numpy as np
import pytest

import oracle as O
from helpers import density


def rzyz(a, b, g):
    def rz(t):
        return np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])

    rb = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    return rz(a) @ rb @ rz(g)


def degree_one_density(B, dim, N, L, seed):
    rng = np.random.default_rng(seed)
    A = 0.2 * rng.standard_normal((3, 3))
    R = O.rotations(B)
    g = 1.0 + np.einsum("ij,ijabc->abc", A, R)
    h = rng.uniform(0.5, 1.5, (N,) * dim)
    f = g.reshape(g.shape + (1,) * dim) * h.reshape((1, 1, 1) + h.shape)
    H = h.sum() * (L / N) ** dim
    return np.asfortranarray(f), A, H


def random_euler(rng, ns, na):
    e = np.empty((3, ns, na, 3))
    e[0] = rng.uniform(-np.pi, np.pi, (ns, na, 3))
    e[1] = rng.uniform(0, np.pi, (ns, na, 3))
    e[2] = rng.uniform(-np.pi, np.pi, (ns, na, 3))
    return e


def closed_form_c(e, A, H):
    ns, na = e.shape[1], e.shape[2]
    c = np.zeros((ns, 3))
    for k in range(3):
        for s in range(ns):
            for a in range(na):
                c[s, k] += H * (1.0 + np.sum(A * rzyz(*e[:, s, a, k])))
    return c


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_getc_degree_one_closed_form(dim):
    B, N, L = 3, 4, 6.0
    f, A, H = degree_one_density(B, dim, N, L, seed=dim)
    e = random_euler(np.random.default_rng(5), 4, 3)
    c = O.getc(f, e, L)
    ref = closed_form_c(e, A, H)
    assert np.abs(c - ref).max() < 1e-12 * np.abs(ref).max()


def test_oracle_getc_grid_points_match_inverse_transform():
    """At grid Euler angles the point evaluation is the inverse transform of the
    marginal's coefficients (a non-bandlimited matrix-Fisher marginal)."""
    B, N, L = 4, 4, 6.0
    f = density(B, 2, N, L, (6, 3, 1), seed=3)
    fR = f.reshape((2 * B) ** 3, -1, order="F").sum(axis=1) * (L / N) ** 2
    d = O.wigner_table(B)
    al, be, ga, w = O.grid(B)
    F = O.fft_forward(B, fR.reshape((2 * B,) * 3, order="F"), d, w, 0, 1)
    g = O.fft_backward(B, F, d, 0, 1)
    rng = np.random.default_rng(1)
    idx = rng.integers(0, 2 * B, (5, 3))
    e = np.zeros((3, 5, 1, 3))
    for s, (i, j, k) in enumerate(idx):
        for kk in range(3):
            e[:, s, 0, kk] = (al[i], be[j], ga[k])
    c = O.getc(f, e, L)
    for s, (i, j, k) in enumerate(idx):
        assert abs(c[s, 0] - g[i, j, k]) < 1e-12 * np.abs(g).max()
