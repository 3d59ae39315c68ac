"""Gradients through the spectral propagator (the reduced pendulum's
parameters, and the restitution coefficient of the hybrid step).

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
matrix G: dA/db_p = i K_p T_p (the damping term) and dA/dG_ij is the diagonal
factor of the diffusion term.  Each reverse stage is one fused kernel
(Plan.adjoint_stage) that applies A^H and reduces all five parameter terms;
the forward stage values come from the same fused stage kernel
(Plan.step_stages).  The forward states are kept at checkpoints every
`checkpoint` steps; on the way back each segment is recomputed from its
checkpoint once, keeping every stage value, and each step's reverse stages
run in one call (Plan.adjoint_step).
"""
import numpy as np

from . import HpupError, Plan

HORNER = {"euler": [1.0], "midpoint": [0.5, 1.0], "rk2": [0.5, 1.0], "rk4": [0.25, 1.0 / 3.0, 0.5, 1.0]}


class Gradient:
    """Forward propagation plus the adjoint sweep of the reduced model."""

    def __init__(self, B, N, L=1.0, b=(0, 0), G=None, dt=0.0, method="euler", precision="fp64"):
        kw = dict(L=L, b=tuple(b) + (0,) * (3 - len(b)), G=G, dt=dt, method=method, precision=precision)
        self.method = method.lower()
        if self.method not in HORNER:
            raise HpupError("unknown integrator %r" % method)
        self.dt = dt
        self.N = N
        self.b = np.array(tuple(b)[:2], dtype=np.float64)
        self.G = np.zeros((2, 2)) if G is None else np.asarray(G, dtype=np.float64).reshape(2, 2).copy()
        self.fwd = Plan(B, 2, N, **kw)
        self.adj = Plan(B, 2, N, operator=1, **kw)

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
        acc = torch.zeros(5, dtype=torch.float64, device=FK.device)
        ns = len(HORNER[self.method])
        for seg in range(len(checkpoints) - 1, -1, -1):
            k0 = seg * cp
            k1 = min(K, k0 + cp)
            # the segment's steps from its checkpoint, keeping every stage value
            # (the last step's result is the next checkpoint, not recomputed)
            stored = []
            F = checkpoints[seg]
            for k in range(k0, k1):
                ys = [torch.empty_like(F) for _ in range(ns - 1)]
                nxt = torch.empty_like(F) if k < k1 - 1 else None
                self.fwd.step_stages(F, ys, nxt)
                stored.append((F, ys))
                F = nxt
            for k in range(k1 - 1, k0 - 1, -1):
                Fk, ys = stored[k - k0]
                lam = self.adj.adjoint_step(Fk, ys, lam, acc)
                stored[k - k0] = None
            del stored
        acc = acc.cpu().numpy()
        G = np.array([[acc[2], acc[4]], [acc[4], acc[3]]])  # G_10 and G_01 enter A as their sum
        return FK, lam, {"b": acc[:2].copy(), "G": G}


def propagate_function():
    """A torch.autograd.Function over the propagation:
    apply(F0, b, G, grad_model, nsteps, checkpoint) -> F_K, differentiable in F0
    (complex), b (2,) and G (2, 2) (float64 CPU tensors).  grad_model is a
    Gradient built at the same b and G (it defines the operator; other values
    raise HpupError); the backward pass calls its adjoint sweep."""
    import torch

    class Propagate(torch.autograd.Function):
        @staticmethod
        def forward(ctx, F0, b, G, model, nsteps, checkpoint):
            # the operator is the one model's plans were built with: b and G
            # are the differentiation points, so they must be model's own
            bv = np.asarray(b.detach().cpu(), dtype=np.float64).reshape(-1)
            Gv = np.asarray(G.detach().cpu(), dtype=np.float64).reshape(2, 2)
            if not (np.array_equal(bv, model.b) and np.array_equal(Gv, model.G)):
                raise HpupError("b and G must equal the Gradient model's parameters (b=%s, G=%s); build a "
                                "Gradient at the new values" % (model.b.tolist(), model.G.tolist()))
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


class TransformAdjoint:
    """Adjoints of a plan's SO(3) x R^n transforms (no counterpart in the
    reference), for objectives on the grid density.  With A the forward
    transform (grid f -> coefficients F) and B the inverse (F -> f = Re(...)),
    stage by stage: the angular DFTs of A and B are each other's adjoints, the
    beta stages differ by the weights w_j (A) and 2l + 1 (B), and the omega
    DFTs by 1/nx, so

        B^H g    = ((2l + 1) / nx) A(g / w_j),
        A^H Fbar = w_j B(nx Fbar / (2l + 1)),

    adjoints for the real inner product on grids and Re<.,.> on coefficients
    (the same w_j as the plan's quadrature, 1/(4B^3) sin(beta_j) s_j)."""

    def __init__(self, plan):
        import torch
        from . import ridx
        from .density import beta_weights
        self.plan = plan
        B, n2, nx = plan.B, 2 * plan.B, plan.nx
        _, w = beta_weights(B)
        rdt = plan.rdtype
        # grid order a + 2B (j + 2B (g + 2B x)): view (nx, g, j, a)
        self.w = torch.tensor(w, dtype=rdt, device="cuda").view(1, 1, n2, 1)
        ell = np.zeros(plan.nrc)
        for l in range(B):
            for m in range(-l, l + 1):
                for n in range(-l, l + 1):
                    ell[ridx(B, l, m, n)] = 2 * l + 1
        self.lfac = torch.tensor(ell / nx, dtype=rdt, device="cuda").view(-1, 1)  # (2l + 1) / nx per row
        self.gshape = (nx, n2, n2, n2)

    def backward_adjoint(self, g):
        """B^H g for a real grid array g (the plan's grid size)."""
        F = self.plan.forward((g.view(self.gshape) / self.w).reshape(-1).contiguous())
        return (F.view(self.plan.nrc, -1) * self.lfac).reshape(-1)

    def forward_adjoint(self, Fbar):
        """A^H Fbar for coefficients Fbar (the plan's coefficient count)."""
        y = (Fbar.view(self.plan.nrc, -1) / self.lfac).reshape(-1).contiguous()
        return (self.plan.backward(y).view(self.gshape) * self.w).reshape(-1)


class HybridGradient:
    """Hybrid steps (pendulum_reduced_hybrid.m:224-234) and their adjoint
    sweep, for gradients in the restitution coefficient epsilon (no
    counterpart in the reference).  One step is

        F'_k = P F_k  (the plan's integrator),  f_k = B F'_k,
        g_k = f_k + dt R_eps f_k,               F_{k+1} = A g_k,

    with B, A the inverse and forward transforms and R_eps the reset's jump
    term, so with Fbar_{k+1} the incoming adjoint (dJ = Re<Fbar, dF>)

        gbar_k = A^H Fbar_{k+1},   dJ/deps += dt <gbar_k, (dR/deps) f_k>,
        fbar_k = gbar_k + dt R^H gbar_k,   Fbar_k = P^H B^H fbar_k,

    P^H the adjoint plan's step (operator=1, dim 2 or 3), R^H and dR/deps the
    plan's reset adjoint and restitution-derivative kernels.  The states F_k
    are kept every `checkpoint` steps and recomputed on the way back; f_k is
    recomputed from F_k (one step and one inverse transform)."""

    def __init__(self, B, dim, N, L, reset, b=(0, 0, 0), G=None, Jt=None, dt=0.0, method="euler",
                 precision="fp64"):
        kw = dict(L=L, b=b, G=G, Jt=Jt, dt=dt, method=method, precision=precision)
        self.dt = dt
        self.reset = dict(reset)
        self.fwd = Plan(B, dim, N, **kw)
        self.fwd.set_reset(**self.reset)
        self.adj = Plan(B, dim, N, operator=1, **kw)
        self.ta = TransformAdjoint(self.fwd)

    def _grid_before_reset(self, F):
        Fp = self.fwd.step(F.clone(), 1)
        return self.fwd.backward(Fp)

    def run(self, F0, nsteps, dJ, checkpoint=None):
        """Propagates F0 by nsteps hybrid steps and back-propagates dJ(F_K)
        (the gradient of a real objective in the final coefficients).
        Returns (F_K, grad_F0, {"epsilon": dJ/deps}); F0 is left unchanged."""
        import torch
        K = int(nsteps)
        cp = max(1, int(checkpoint or K or 1))
        checkpoints = [F0.clone()]
        F = F0.clone()
        f = self.fwd.empty_grid()
        for k in range(K):
            self.fwd.hybrid_step(F, f, 1)
            if (k + 1) % cp == 0 and k + 1 < K:
                checkpoints.append(F.clone())
        FK = F
        lam = dJ(FK).clone()
        acc = torch.zeros(1, dtype=torch.float64, device=FK.device)
        for seg in range(len(checkpoints) - 1, -1, -1):
            k0 = seg * cp
            k1 = min(K, k0 + cp)
            states = [checkpoints[seg]]
            for _ in range(k0, k1 - 1):
                nxt = states[-1].clone()
                self.fwd.hybrid_step(nxt, f, 1)
                states.append(nxt)
            for k in range(k1 - 1, k0 - 1, -1):
                fk = self._grid_before_reset(states[k - k0])
                gbar = self.ta.forward_adjoint(lam)
                self.fwd.reset_eps_grad(fk, gbar, acc, self.dt)
                del fk
                fbar = self.fwd.reset_adjoint(gbar, gbar, self.dt)
                lam = self.adj.step(self.ta.backward_adjoint(fbar).contiguous(), 1)
            del states
        return FK, lam, {"epsilon": float(acc.item())}


def transform_functions(plan):
    """torch.autograd.Function pair over the plan's transforms: Forward(f) -> F
    (backward through A^H) and Backward(F) -> f (backward through B^H)."""
    import torch
    ta = TransformAdjoint(plan)

    class Forward(torch.autograd.Function):
        @staticmethod
        def forward(ctx, f):
            return plan.forward(f.detach().contiguous())

        @staticmethod
        def backward(ctx, gF):
            return ta.forward_adjoint(gF.contiguous())

    class Backward(torch.autograd.Function):
        @staticmethod
        def forward(ctx, F):
            return plan.backward(F.detach().contiguous())

        @staticmethod
        def backward(ctx, gf):
            return ta.backward_adjoint(gf.contiguous())

    return Forward, Backward


def mfg_density(B, N, L, s, U, mu, sig):
    """Differentiable (torch, float64, on the GPU) matrix-Fisher x Gaussian grid
    density of the reduced model: f(R, x) proportional to
    exp(tr((U diag(s))^T R)) exp(-sum_p ((x_p - mu_p) / sig_p)^2 / 2), with unit
    mass under the grid quadrature (weights w_j, cells (L/N)^2).  s (3,), mu (2,)
    and sig (2,) may require gradients; U is a fixed 3 x 3 array."""
    import torch
    from .density import beta_weights, rotation_grid
    R = rotation_grid(B)
    Ut = torch.as_tensor(np.asarray(U, dtype=np.float64), device="cuda")
    expo = ((Ut * s.view(1, 3)).view(3, 3, 1) * R).sum(dim=(0, 1))
    fR = torch.exp(expo - expo.max().detach())
    _, w = beta_weights(B)
    wj = torch.as_tensor(w, device="cuda").view(1, 2 * B, 1).expand(2 * B, 2 * B, 2 * B).reshape(-1)
    ax = -L / 2 + L / N * torch.arange(N, device="cuda", dtype=torch.float64)
    x1 = ax.view(1, N).expand(N, N).reshape(-1)  # first omega index fastest
    x2 = ax.view(N, 1).expand(N, N).reshape(-1)
    q = ((x1 - mu[0]) / sig[0]) ** 2 + ((x2 - mu[1]) / sig[1]) ** 2
    fx = torch.exp(-0.5 * q)
    mass = (fR * wj).sum() * fx.sum() * (L / N) ** 2
    return torch.outer(fx, fR).reshape(-1) / mass
