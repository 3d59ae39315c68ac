"""The gyro random walk of gyro_bias.m (sum_ij G1_ij F u_i^T u_j^T) in the
oracle, pinned by the Casimir: with the left-invariant derivative matrices u_i
of getu.m, sum_i u_i u_i = -l (l+1) I on every l block, so an isotropic G1 = g I
multiplies block l by -g l (l+1); and the anisotropic term is bilinear in G1
(a sum of the three diagonal and the symmetric off-diagonal pieces)."""
import numpy as np

import oracle as O


def _rand_F(B, N, rng):
    M = 2 * B - 1
    return np.asfortranarray(rng.standard_normal((M, M, B, N, N, N)) + 1j * rng.standard_normal((M, M, B, N, N, N)))


def test_casimir():
    B = 5
    u = O.getu(B - 1)
    lmax = B - 1
    for l in range(B):
        ix = slice(lmax - l, lmax + l + 1)
        S = sum(u[ix, ix, l, i] @ u[ix, ix, l, i] for i in range(3))
        assert np.abs(S + l * (l + 1) * np.eye(2 * l + 1)).max() < 1e-12


def test_g1_isotropic_closed_form():
    B, N = 4, 2
    rng = np.random.default_rng(0)
    F = _rand_F(B, N, rng)
    u = O.getu(B - 1)
    T = O.gyrobias_g1_term(B, F, u, 0.3 * np.eye(3))
    lmax = B - 1
    for l in range(B):
        ix = slice(lmax - l, lmax + l + 1)
        assert np.abs(T[ix, ix, l] + 0.3 * l * (l + 1) * F[ix, ix, l]).max() < 1e-12 * np.abs(F).max() * (l + 1) ** 2


def test_g1_bilinear():
    B, N = 3, 2
    rng = np.random.default_rng(1)
    F = _rand_F(B, N, rng)
    u = O.getu(B - 1)
    G1 = rng.standard_normal((3, 3))
    T = O.gyrobias_g1_term(B, F, u, G1)
    parts = sum(O.gyrobias_g1_term(B, F, u, np.eye(3)[:, [i]] @ np.eye(3)[[j], :] * G1[i, j])
                for i in range(3) for j in range(3))
    assert np.abs(T - parts).max() < 1e-12 * np.abs(T).max()
