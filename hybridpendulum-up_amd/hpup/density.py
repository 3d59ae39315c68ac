"""Initial densities on the SO(3) x R^n grid, built on the GPU.

The drivers' initial condition (pendulum.m:210-220, pendulum_reduced.m:184-194):
a matrix Fisher density exp(tr((U S)^T R)) on the rotation grid times a
Gaussian in omega.  Normalised to unit mass by the grid quadrature
sum_{ijk} w_j f (L/N)^dim (the same quadrature get_stat uses), instead of the
Bessel-integral constant c(S), so the discrete mass is exactly one.
"""
import math

import numpy as np


def random_rotation(rng):
    """Uniform random rotation from a normalised Gaussian quaternion."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def beta_weights(B):
    """fftSO3R.m:63-66."""
    j = np.arange(2 * B)
    beta = math.pi / (4 * B) * (2 * j + 1)
    t = np.arange(B)
    s = (np.sin(np.outer(beta, 2 * t + 1)) / (2 * t + 1)).sum(axis=1)
    return beta, 1.0 / (4 * B ** 3) * np.sin(beta) * s


def rotation_grid(B, device="cuda", dtype=None):
    """R(:,:,i,j,k) = Rz(alpha_i) Ry(beta_j) Rz(gamma_k) as a (3,3,(2B)^3) tensor, alpha fastest."""
    import torch
    dtype = dtype or torch.float64
    n2 = 2 * B
    idx = torch.arange(n2, device=device, dtype=dtype)
    a = (math.pi / B * idx).view(1, 1, n2)
    b = (math.pi / (4 * B) * (2 * idx + 1)).view(1, n2, 1)
    g = (math.pi / B * idx).view(n2, 1, 1)
    ca, sa, cb, sb, cg, sg = a.cos(), a.sin(), b.cos(), b.sin(), g.cos(), g.sin()
    shp = (n2, n2, n2)  # (gamma, beta, alpha) -> flattened alpha fastest
    R = torch.empty((3, 3) + shp, device=device, dtype=dtype)
    R[0, 0] = (ca * cb * cg - sa * sg).expand(shp)
    R[1, 0] = (sa * cb * cg + ca * sg).expand(shp)
    R[2, 0] = (-sb * cg).expand(shp)
    R[0, 1] = (-ca * cb * sg - sa * cg).expand(shp)
    R[1, 1] = (-sa * cb * sg + ca * cg).expand(shp)
    R[2, 1] = (sb * sg).expand(shp)
    R[0, 2] = (ca * sb).expand(shp)
    R[1, 2] = (sa * sb).expand(shp)
    R[2, 2] = cb.expand(shp)
    return R.reshape(3, 3, -1)


def matrix_fisher_gaussian(B, dim, N, L, s, U, Sigma=None, mu=None, weights=None, device="cuda"):
    """Returns (fR, fx): rotation factor ((2B)^3,) and omega factor (N^dim,) of a
    separable density, scaled so the product has unit grid mass.  A mixture
    of several components is not separable: build it with mixture_density
    (passing `weights` here is an error)."""
    import torch
    if weights is not None:
        raise ValueError("matrix_fisher_gaussian builds one separable component; use mixture_density for a "
                         "weighted mixture")
    R = rotation_grid(B, device=device)
    US = torch.as_tensor(np.asarray(U) @ np.diag(s), device=device, dtype=torch.float64)
    expo = (US.view(3, 3, 1) * R).sum(dim=(0, 1))
    fR = torch.exp(expo - expo.max())
    _, w = beta_weights(B)
    wj = torch.as_tensor(w, device=device).view(1, 2 * B, 1).expand(2 * B, 2 * B, 2 * B).reshape(-1)
    if dim == 0:
        fx = torch.ones(1, device=device, dtype=torch.float64)
        dv = 1.0
    else:
        ax = -L / 2 + L / N * torch.arange(N, device=device, dtype=torch.float64)
        grids = torch.meshgrid(*([ax] * dim), indexing="ij")
        # MATLAB order: first omega index fastest -> flatten with axis 0 fastest
        xs = torch.stack([gq.permute(*reversed(range(dim))).reshape(-1) for gq in grids], dim=0)
        S = np.eye(dim) if Sigma is None else np.asarray(Sigma, dtype=np.float64)
        m = np.zeros(dim) if mu is None else np.asarray(mu, dtype=np.float64)
        Si = torch.as_tensor(np.linalg.inv(S), device=device)
        d = xs - torch.as_tensor(m, device=device).view(dim, 1)
        q = (d * (Si @ d)).sum(dim=0)
        fx = torch.exp(-0.5 * q)
        dv = (L / N) ** dim
    mass = float((fR * wj).sum()) * float(fx.sum()) * dv
    return fR / mass, fx


def grid_density(fR, fx, dtype=None):
    """Flat grid f[R + nR*x] = fR[R] fx[x] (MATLAB order, rotation index fastest)."""
    import torch
    f = torch.outer(fx, fR).reshape(-1)
    return f if dtype is None else f.to(dtype)


def mixture_density(B, dim, N, L, components, weights, dtype=None, device="cuda"):
    """Grid density f = sum_k weights[k] f_k of matrix-Fisher x Gaussian
    components, each f_k of unit grid mass (so f has mass sum(weights)).
    components: sequence of dicts with keys s (3 proper singular values), U
    (3x3 rotation, F_k = U diag(s)) and optionally Sigma and mu of the
    Gaussian in omega.  Returned flat in MATLAB order (rotation index
    fastest), like grid_density."""
    import torch
    if len(components) != len(weights) or len(components) == 0:
        raise ValueError("one weight per component")
    f = None
    for comp, wk in zip(components, weights):
        fR, fx = matrix_fisher_gaussian(B, dim, N, L, comp["s"], comp["U"], Sigma=comp.get("Sigma"),
                                        mu=comp.get("mu"), device=device)
        if f is None:
            f = float(wk) * torch.outer(fx, fR).reshape(-1)
        else:
            f.add_(torch.outer(fx, fR).reshape(-1), alpha=float(wk))
        del fR, fx
    return f if dtype is None else f.to(dtype)
