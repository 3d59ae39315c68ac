"""Likelihood of noisy attitude and angular-velocity measurements under a
propagated grid density (no counterpart in the reference, which only
propagates forward).

A measurement (Z, y) of the state (R, Omega) has attitude noise
Z ~ MF(kappa R) (matrix Fisher with F = kappa R) and angular-velocity noise
y ~ N(Omega, sigma^2 I).  Its likelihood under a density f(R, Omega) is the
density convolved with the noise,

    p(Z, y) = int int MF(Z; kappa R) N(y; Omega, sigma^2 I) f(R, Omega) dR dOmega,

evaluated with the grid quadrature of the transforms (beta weights w_j on
SO(3), cells (L/N)^dim in omega) -- a separable linear functional of the grid
density per measurement:

    p_i = dv sum_x b_i(x) sum_R a_i(R) w_j f(R, x),
    a_i(R) = exp(kappa tr(Z_i^T R)) / c(kappa, kappa, kappa),  b_i(x) = N(y_i; x, sigma^2 I).

For P measurements that is two GEMMs over the grid (nx x nR times nR x P,
then a row contraction), so log-likelihoods and their gradients (torch
autograd) run on the device.  c is the matrix-Fisher normalizing constant
(pdf_MF_normal.m), which makes MF(Z; kappa R) a density under the Haar
measure of unit total mass.
"""
import math

import numpy as np

from . import pdf_MF_normal
from .density import beta_weights, rotation_grid


def _haar_sphere_axis(rng, n):
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _expm_so3(axis, angle):
    K = np.zeros((len(angle), 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -axis[:, 2], axis[:, 1], -axis[:, 0]
    K[:, 1, 0], K[:, 2, 0], K[:, 2, 1] = axis[:, 2], -axis[:, 1], axis[:, 0]
    s, c = np.sin(angle)[:, None, None], np.cos(angle)[:, None, None]
    return np.eye(3)[None] + s * K + (1 - c) * (K @ K)


def sample_mf_isotropic(rng, R, kappa):
    """Z_i ~ MF(kappa R_i) (n rotations R, shape (n, 3, 3)).  Z = R Q with
    Q ~ MF(kappa I): a uniform axis and a rotation angle t with density
    proportional to exp(2 kappa cos t) (1 - cos t) on [0, pi] (the matrix
    Fisher exponent tr(kappa Q) = kappa (1 + 2 cos t) times the Haar density
    of the angle), drawn by inverting its tabulated distribution function."""
    n = len(R)
    t = np.linspace(0.0, math.pi, 200001)
    logp = 2 * kappa * (np.cos(t) - 1) + np.log(np.maximum(1.0 - np.cos(t), 1e-300))
    p = np.exp(logp - logp.max())
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (p[1:] + p[:-1]) * np.diff(t))])
    cdf /= cdf[-1]
    ang = np.interp(rng.uniform(size=n), cdf, t)
    return np.einsum("nij,njk->nik", R, _expm_so3(_haar_sphere_axis(rng, n), ang))


def sample_states(rng, plan, f, n):
    """n states (R, Omega) drawn from the grid density f (a flat tensor of the
    plan's grid) with probabilities w_j f dv: grid rotations (n, 3, 3) and
    grid angular velocities (n, dim)."""
    import torch
    B, dim, N, L = plan.B, plan.dim, plan.N, plan.model.L
    n2 = 2 * B
    _, w = beta_weights(B)
    wj = torch.as_tensor(w, device=f.device, dtype=torch.float64).view(1, 1, n2, 1)
    p = (f.double().view(-1, n2, n2, n2) * wj).clamp_min(0).reshape(-1).cpu().numpy()
    idx = rng.choice(len(p), size=n, p=p / p.sum())
    nR = n2 ** 3
    iR, ix = idx % nR, idx // nR
    R = rotation_grid(B, device="cpu").permute(2, 0, 1).numpy()[iR]
    Om = np.stack([-L / 2 + L / N * ((ix // N ** a) % N) for a in range(dim)], axis=1)
    return R, Om


class MeasurementLikelihood:
    """log p(Z_i, y_i) summed over P measurements, as a differentiable torch
    function of the grid density (flat tensor of the plan's grid, MATLAB
    order, rotation index fastest)."""

    def __init__(self, plan, Z, y, kappa=20.0, sigma=0.2):
        import torch
        self.plan = plan
        B, dim, N, L = plan.B, plan.dim, plan.N, plan.model.L
        n2 = 2 * B
        Z = np.asarray(Z, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64).reshape(len(Z), dim)
        self.P = len(Z)
        dt = torch.float64
        R = rotation_grid(B, device="cuda", dtype=dt).reshape(9, -1)  # entry (a, b) at row 3 a + b, (9, nR)
        # tr(Z^T R) = sum_ab Z_ab R_ab
        Zf = torch.as_tensor(Z.reshape(self.P, 9), device="cuda", dtype=dt)  # same (a, b) order as R
        c, _, _ = pdf_MF_normal(np.array([kappa, kappa, kappa]), bool_scaled=True)
        # exp(kappa tr(Z^T R)) / c(kappa I), with the scaled normalizer c e^{-3 kappa}
        loga = kappa * (Zf @ R) - 3 * kappa - math.log(float(c))
        _, w = beta_weights(B)
        # grid rotation index a + 2B (j + 2B g): the weight varies with the middle (beta) index
        wj = torch.as_tensor(w, device="cuda", dtype=dt).view(1, n2, 1).expand(n2, n2, n2).reshape(1, -1)
        self.A = torch.exp(loga) * wj                       # (P, nR): a_i(R) w_j
        ax = -L / 2 + L / N * torch.arange(N, device="cuda", dtype=dt)
        grids = torch.meshgrid(*([ax] * dim), indexing="ij")
        x = torch.stack([g.permute(*reversed(range(dim))).reshape(-1) for g in grids], dim=1)  # (nx, dim)
        yt = torch.as_tensor(y, device="cuda", dtype=dt)
        d2 = ((yt.view(self.P, 1, dim) - x.view(1, -1, dim)) ** 2).sum(-1)
        self.Bm = torch.exp(-0.5 * d2 / sigma ** 2) / (2 * math.pi * sigma ** 2) ** (dim / 2)  # (P, nx)
        self.dv = (L / N) ** dim
        self.nR = n2 ** 3

    def probabilities(self, f):
        """p_i (P,) for the grid density f (torch, differentiable)."""
        fm = f.double().view(-1, self.nR)             # (nx, nR)
        M = fm @ self.A.t()                           # (nx, P): sum_R a_i(R) w_j f(R, x)
        return (self.Bm * M.t()).sum(dim=1) * self.dv

    def __call__(self, f):
        return self.probabilities(f).log().sum()
