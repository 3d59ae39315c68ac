"""Gradients through the spectral propagator of the reduced pendulum (dim 2).

No counterpart in the reference, which only propagates forward.  One step of
every integrator of the plan is a polynomial in the operator,

    F_{k+1} = P(dt A) F_k,   P(z) = sum_j c_j z^j
    (Euler 1 + z; midpoint and RK2 1 + z + z^2/2; RK4 up to z^4/24),

with real c_j, so the adjoint step is P(dt A^H) = P(dt A)^H, which the adjoint
plan (operator=1) applies with the same fused kernels.  A is linear in the
damping b and the noise matrix G, so dA/db_p and dA/dG_ij are operators of the
same form (operator=2+p, 2+dim+i+dim*j), and

    dJ/dtheta = sum_k Re< lam_{k+1}, dP/dtheta F_k >,
    dP/dtheta = sum_j c_j dt^j sum_{i<j} A^i (dA/dtheta) A^{j-1-i},
    lam_K = dJ/dF_K,  lam_k = P^H lam_{k+1}

(<x, y> = sum conj(x) y over the compact coefficients; a gradient g of a
real J means dJ = Re<g, dF>).  The forward states are kept at checkpoints
every `checkpoint` steps and recomputed one segment at a time on the way back.
"""
import numpy as np

from . import HpupError, Plan

POLY = {"euler": [1.0, 1.0], "midpoint": [1.0, 1.0, 0.5], "rk2": [1.0, 1.0, 0.5],
        "rk4": [1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0]}


def _dot(x, y):
    import torch
    return float(torch.vdot(x, y).real)


class Gradient:
    """Forward propagation plus the adjoint sweep of the reduced model."""

    def __init__(self, B, N, L=1.0, b=(0, 0), G=None, dt=0.0, method="euler", precision="fp64"):
        kw = dict(L=L, b=tuple(b) + (0,) * (3 - len(b)), G=G, dt=dt, method=method, precision=precision)
        self.dim = 2
        self.method = method.lower()
        if self.method not in POLY:
            raise HpupError("unknown integrator %r" % method)
        self.dt = dt
        self.fwd = Plan(B, 2, N, **kw)
        self.adj = Plan(B, 2, N, operator=1, **kw)
        # derivative operators, in the order b_0, b_1, G_00, G_10, G_01, G_11
        self.dA = [Plan(B, 2, N, operator=2 + t, **kw) for t in range(2 + 4)]

    def apply_step(self, F, nsteps=1):
        return self.fwd.step(F, nsteps)

    def apply_adjoint_step(self, F, nsteps=1):
        return self.adj.step(F, nsteps)

    def _param_terms(self, F, lam):
        """Re< lam, dP/dtheta F > for every parameter theta."""
        c = POLY[self.method]
        s = len(c) - 1
        dt = self.dt
        # y_q = A^q F and z_i = (A^H)^i lam, q, i < s
        y = [F]
        for _ in range(s - 1):
            y.append(self.fwd.derivative(y[-1]))
        z = [lam]
        for _ in range(s - 1):
            z.append(self.adj.derivative(z[-1]))
        # w_q = sum_{j > q} c_j dt^j z_{j-1-q}
        out = np.zeros(len(self.dA))
        for q in range(s):
            w = None
            for j in range(q + 1, s + 1):
                t = z[j - 1 - q] * (c[j] * dt ** j)
                w = t if w is None else w + t
            for t, op in enumerate(self.dA):
                out[t] += _dot(w, op.derivative(y[q]))
        return out

    def run(self, F0, nsteps, dJ, checkpoint=None):
        """Propagates F0 by nsteps steps and back-propagates the gradient.

        dJ(F_K) returns the gradient of the real objective J with respect to
        the final coefficients (a tensor of the plan's shape and dtype).
        Returns (F_K, grad_F0, grads) with grads = {"b": (2,), "G": (2, 2)}.
        F0 is left unchanged.
        """
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
        total = np.zeros(len(self.dA))
        for seg in range(len(checkpoints) - 1, -1, -1):
            k0 = seg * cp
            k1 = min(K, k0 + cp)
            states = [checkpoints[seg]]
            for _ in range(k0, k1 - 1):
                nxt = states[-1].clone()
                self.fwd.step(nxt, 1)
                states.append(nxt)
            for k in range(k1 - 1, k0 - 1, -1):
                total += self._param_terms(states[k - k0], lam)
                self.adj.step(lam, 1)
            del states
        G = np.zeros((2, 2))
        for i in range(2):
            for j in range(2):
                G[i, j] = total[2 + i + 2 * j]
        return FK, lam, {"b": total[:2].copy(), "G": G}
