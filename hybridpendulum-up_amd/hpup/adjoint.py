"""Gradients through the spectral propagator of the reduced pendulum (dim 2).

No counterpart in the reference, which only propagates forward.  The drift
operator A is linear and autonomous, and the plan evaluates one step of each
integrator in Horner form,

    y_0 = F_k,   y_j = F_k + c_j dt A y_{j-1}  (j = 1..s),   F_{k+1} = y_s,
    c = (1) Euler, (1/2, 1) midpoint and RK2, (1/4, 1/3, 1/2, 1) RK4,

so reverse mode runs through the same stages backwards: with ybar_s the
incoming adjoint,

    ybar_{j-1} = c_j dt A^H ybar_j,   Fbar_k = sum_{j=0..s} ybar_j,
    dJ/dtheta += c_j dt Re< ybar_j, (dA/dtheta) y_{j-1} >

(<x, y> = sum conj(x) y over the compact coefficients; a gradient g of a real
J means dJ = Re<g, dF>).  A^H is the adjoint plan (operator=1, the fused
kernels with transposed tables).  A is linear in the damping b and the noise
matrix G: dA/db_p = i K_p T_p (the damping term, applied by the operator=2+p
plan), and dA/dG_ij is the diagonal factor of the diffusion term, so its
products reduce elementwise.  The forward states are kept at checkpoints
every `checkpoint` steps and each step's stage values are recomputed from
them on the way back.
"""
import math

import numpy as np

from . import HpupError, Plan

HORNER = {"euler": [1.0], "midpoint": [0.5, 1.0], "rk2": [0.5, 1.0], "rk4": [0.25, 1.0 / 3.0, 0.5, 1.0]}


def omega_factors(N, L):
    """kc(k) (first derivative, Nyquist zeroed) and kd(k) (second derivative)
    of the omega-frequency index k, as the drift tables build them."""
    Bx = N // 2
    k = np.arange(N)
    kk = np.where(k < Bx, k, k - N)
    kp = np.where(k < Bx, k, np.where(k == Bx, 0, k - N))
    return 2 * math.pi * kp / L, -(2 * math.pi * kk / L) ** 2


class Gradient:
    """Forward propagation plus the adjoint sweep of the reduced model."""

    def __init__(self, B, N, L=1.0, b=(0, 0), G=None, dt=0.0, method="euler", precision="fp64"):
        import torch
        kw = dict(L=L, b=tuple(b) + (0,) * (3 - len(b)), G=G, dt=dt, method=method, precision=precision)
        self.method = method.lower()
        if self.method not in HORNER:
            raise HpupError("unknown integrator %r" % method)
        self.dt = dt
        self.N = N
        self.fwd = Plan(B, 2, N, **kw)
        self.adj = Plan(B, 2, N, operator=1, **kw)
        # dA/db_0, dA/db_1
        self.dAb = [Plan(B, 2, N, operator=2 + p, **kw) for p in range(2)]
        # dA/dG_ij as a factor of the (k2, k1) grid (k1 fastest): D(k) =
        # G_00 kd(k1) + G_11 kd(k2) - (G_10 + G_01) kc(k1) kc(k2)
        kc, kd = omega_factors(N, L)
        one = np.ones(N)
        f = {(0, 0): np.outer(one, kd), (1, 1): np.outer(kd, one), (1, 0): -np.outer(kc, kc)}
        f[(0, 1)] = f[(1, 0)]
        rdt = torch.float64 if precision == "fp64" else torch.float32
        self.dG = {ij: torch.from_numpy(v).to(rdt).cuda() for ij, v in f.items()}

    def _stages(self, F):
        """y_0 .. y_{s-1} of one step from F."""
        ys = [F]
        for c in HORNER[self.method][:-1]:
            ys.append(F + (c * self.dt) * self.fwd.derivative(ys[-1]))
        return ys

    def _step_back(self, F, lam, acc):
        """Adjoint of one step taken from F; returns Fbar, adds to acc (a
        device tensor of the six parameter gradients: no host sync here)."""
        import torch
        ys = self._stages(F)
        c = HORNER[self.method]
        Fbar = None
        ybar = lam
        N = self.N
        for j in range(len(c) - 1, -1, -1):
            Fbar = ybar.clone() if Fbar is None else Fbar + ybar
            s = c[j] * self.dt
            y = ys[j]
            terms = [torch.vdot(ybar, self.dAb[p].derivative(y)).real for p in range(2)]
            prod = (ybar.conj() * y).real.reshape(-1, N, N).sum(0)
            for ij in ((0, 0), (1, 0), (0, 1), (1, 1)):  # flat index 2 + i + 2 j
                terms.append((prod * self.dG[ij]).sum())
            acc += s * torch.stack(terms).to(acc.dtype)
            ybar = self.adj.derivative(ybar) * s
        return Fbar + ybar

    def run(self, F0, nsteps, dJ, checkpoint=None):
        """Propagates F0 by nsteps steps and back-propagates the gradient.

        dJ(F_K) returns the gradient of the real objective J with respect to
        the final coefficients (a tensor of the plan's shape and dtype).
        Returns (F_K, grad_F0, grads) with grads = {"b": (2,), "G": (2, 2)}.
        F0 is left unchanged.
        """
        import torch
        K = int(nsteps)
        cp = max(1, int(checkpoint or K or 1))
        checkpoints = [F0.clone()]
        F = F0.clone()
        for k in range(K):
            self.fwd.step(F, 1)
            if (k + 1) % cp == 0 and k + 1 < K:
                checkpoints.append(F.clone())
        FK = F
        lam = dJ(FK).clone()
        acc = torch.zeros(6, dtype=torch.float64, device=FK.device)
        for seg in range(len(checkpoints) - 1, -1, -1):
            k0 = seg * cp
            k1 = min(K, k0 + cp)
            states = [checkpoints[seg]]
            for _ in range(k0, k1 - 1):
                nxt = states[-1].clone()
                self.fwd.step(nxt, 1)
                states.append(nxt)
            for k in range(k1 - 1, k0 - 1, -1):
                lam = self._step_back(states[k - k0], lam, acc)
            del states
        acc = acc.cpu().numpy()
        G = np.zeros((2, 2))
        for i in range(2):
            for j in range(2):
                G[i, j] = acc[2 + i + 2 * j]
        return FK, lam, {"b": acc[:2].copy(), "G": G}


def propagate_function():
    """A torch.autograd.Function over the propagation:
    apply(F0, b, G, grad_model, nsteps, checkpoint) -> F_K, differentiable in F0
    (complex), b (2,) and G (2, 2) (float64 CPU tensors).  grad_model is a
    Gradient built at the same b and G (it defines the operator); the backward
    pass calls its adjoint sweep."""
    import torch

    class Propagate(torch.autograd.Function):
        @staticmethod
        def forward(ctx, F0, b, G, model, nsteps, checkpoint):
            ctx.model, ctx.nsteps, ctx.checkpoint = model, nsteps, checkpoint
            ctx.save_for_backward(F0)
            F = F0.detach().clone()
            model.fwd.step(F, nsteps)
            return F

        @staticmethod
        def backward(ctx, gout):
            (F0,) = ctx.saved_tensors
            # PyTorch's gradient of a real loss with respect to a complex
            # tensor is dL/dRe + i dL/dIm: the g of dJ = Re<g, dz> used here
            _, gF0, g = ctx.model.run(F0.detach(), ctx.nsteps, lambda F: gout.contiguous(),
                                      checkpoint=ctx.checkpoint)
            gb = torch.tensor(g["b"], dtype=torch.float64)
            gG = torch.tensor(g["G"], dtype=torch.float64)
            return gF0, gb, gG, None, None, None

    return Propagate
