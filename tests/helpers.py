"""Shared test inputs: seeded matrix-Fisher x Gaussian densities on the
SO(3) x R^n grid (the drivers' initial condition, pendulum_reduced.m:184-194),
and the drivers' operator tensors from the oracle."""
import numpy as np

import oracle as O


def rot_random(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def omega_grid(dim, N, L):
    ax = -L / 2 + L / N * np.arange(N)
    mesh = np.meshgrid(*([ax] * dim), indexing="ij")
    return np.stack(mesh, axis=0)  # (dim, N, ..., N) MATLAB order


def density(B, dim, N, L, s, seed=0, Sigma=None, mu=None):
    """f(R, x) = exp(tr((U S)^T R)) N(x; mu, Sigma), unit mass by quadrature."""
    rng = np.random.default_rng(seed)
    U = rot_random(rng)
    R = O.rotations(B)
    _, _, _, w = O.grid(B)
    US = U @ np.diag(s)
    fR = np.exp(np.einsum("ij,ijabc->abc", US, R))
    if dim == 0:
        f = fR
        mass = np.sum(f * w[None, :, None])
        return np.asfortranarray(f / mass)
    x = omega_grid(dim, N, L)
    if Sigma is None:
        Sigma = np.eye(dim)
    if mu is None:
        mu = np.zeros(dim)
    Si = np.linalg.inv(Sigma)
    d = x - np.asarray(mu).reshape((dim,) + (1,) * dim)
    q = np.einsum("i...,ij,j...->...", d, Si, d)
    fx = np.exp(-0.5 * q)
    f = fR.reshape(fR.shape + (1,) * dim) * fx.reshape((1, 1, 1) + fx.shape)
    mass = np.sum(f * w[None, :, None].reshape((1, 2 * B, 1) + (1,) * dim)) * (L / N) ** dim
    return np.asfortranarray(f / mass)


def operators(B, dim, N, L, Jt=None):
    X = O.build_X(dim, N, L)
    OJO = O.build_OJO(N, L, Jt) if dim == 3 else None
    MR = O.build_MR(B, dim, Jt)
    u = O.getu(B - 1)
    cells = O.cg_table(B)
    return X, OJO, MR, u, cells


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def valid_mask(B, extra_shape=()):
    M = 2 * B - 1
    lmax = B - 1
    m = np.arange(M)[:, None, None] - lmax
    n = np.arange(M)[None, :, None] - lmax
    l = np.arange(B)[None, None, :]
    mask = (np.abs(m) <= l) & (np.abs(n) <= l)
    return mask.reshape(mask.shape + (1,) * len(extra_shape))
