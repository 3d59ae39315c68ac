"""oracle/matfisher.py -- matrix-Fisher statistics of get_stat restated in
numpy/scipy.  TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing that
ships); it is the checker of libhpup's get_stat, psvd, pdf_MF_normal and
pdf_MF_M2S, never a code path of the product.

Follows, step by step:
  * matrix Fisher/pdf_MF_normal.m:35-48 (c, scaled c_bar), :55-75 (first
    derivatives), :77-150 (unscaled Hessian system), :152-205 (scaled Hessian
    system), :212-268 (integrands);
  * matrix Fisher/pdf_MF_M2S.m:22-102 (Newton-Armijo with the three-point
    parabolic line search, including its MATLAB linear indexing of f_stack)
    and its local functions stack3/saturation/func (:104-159);
  * matrix Fisher/psvd.m:10-13;
  * propagation/pendulum.m:547-594 and pendulum_reduced.m:485-532 (get_stat).
MATLAB's `integral` (adaptive Gauss-Kronrod) is replaced by
scipy.integrate.quad (QUADPACK) at a tighter tolerance, and `besseli(.,.,1)`
by scipy.special.ive.
"""
import numpy as np
from scipy import integrate, special


def _quad(fn):
    return integrate.quad(fn, -1.0, 1.0, epsabs=1e-15, epsrel=1e-14, limit=400)[0]


def _J(u, a, b, scaled):
    # besseli(0, (a-b)(1-u)/2[, 1]) * besseli(0, (a+b)(1+u)/2[, 1])
    f = special.ive if scaled else special.iv
    return f(0, 0.5 * (a - b) * (1 - u)) * f(0, 0.5 * (a + b) * (1 + u))


def _cyc(i):
    # index = circshift([1 2 3], [0 4-i]) (0-based i): (i, j, k) cyclic
    return i, (i + 1) % 3, (i + 2) % 3


def mf_normal(s, scaled=True):
    """[c, dc, ddc] = pdf_MF_normal(s, scaled, 1, 1)."""
    s = np.asarray(s, dtype=np.float64).ravel()
    if not scaled:
        c = _quad(lambda u: 0.5 * np.exp(s[2] * u) * _J(u, s[0], s[1], False))
        dc = np.zeros(3)
        for i in range(3):
            _, j, k = _cyc(i)
            dc[i] = _quad(lambda u: 0.5 * _J(u, s[j], s[k], False) * u * np.exp(s[i] * u))
        return c, dc, _hessian_unscaled(s, c, dc)
    if s[0] >= s[1]:
        c = _quad(lambda u: 0.5 * np.exp((s[1] + s[2]) * (u - 1)) * _J(u, s[0], s[1], True))
    else:
        c = _quad(lambda u: 0.5 * np.exp((s[0] + s[2]) * (u - 1)) * _J(u, s[0], s[1], True))
    dc = np.zeros(3)
    for i in range(3):
        _, j, k = _cyc(i)
        si, sj, sk = s[i], s[j], s[k]
        dc[k] = _quad(lambda u: 0.5 * _J(u, si, sj, True) * (u - 1) * np.exp((min(si, sj) + sk) * (u - 1)))
    return c, dc, _hessian_scaled(s, c, dc)


def _others(i, j):
    return [k for k in range(3) if k not in (i, j)]


def _hessian_unscaled(s, c, dc):
    """pdf_MF_normal.m:80-150: second derivatives from c and dc by a 9x9 linear
    system, with the reference's case split on repeated |s_i|."""
    A = np.zeros((9, 9))
    b = np.zeros(9)
    a = np.abs(s)
    sg = np.sign
    for i in range(3):
        for j in range(3):
            r = 3 * i + j
            ks = _others(i, j)
            if i == j:
                k1, k2 = ks
                t1 = (-dc[i] * s[i] + dc[k1] * s[k1]) / (s[k1] ** 2 - s[i] ** 2) if a[i] != a[k1] else None
                t2 = (-dc[i] * s[i] + dc[k2] * s[k2]) / (s[k2] ** 2 - s[i] ** 2) if a[i] != a[k2] else None
                if a[i] != a[k1] and a[i] != a[k2]:
                    A[r, r] = 1
                    b[r] = c - t1 - t2
                elif a[i] != a[k1] and a[i] == a[k2] and s[i] != 0:
                    A[r, r] = 1.5
                    A[r, 3 * i + k2] = -0.5 * sg(s[i] * s[k2])
                    b[r] = c - t1 - dc[i] / 2 / s[i]
                elif a[i] != a[k1] and s[i] == s[k2] and s[i] == 0:
                    A[r, r] = 2
                    b[r] = c - t1
                elif a[i] == a[k1] and a[i] != a[k2] and s[i] != 0:
                    A[r, r] = 1.5
                    A[r, 3 * i + k1] = -0.5 * sg(s[i] * s[k1])
                    b[r] = c - dc[i] / 2 / s[i] - t2
                elif a[i] == a[k1] and a[i] == a[k2] and s[i] != 0:
                    A[r, r] = 2
                    A[r, 3 * i + k1] = -0.5 * sg(s[i] * s[k1])
                    A[r, 3 * i + k2] = -0.5 * sg(s[i] * s[k2])
                    b[r] = c - dc[i] / s[i]
                elif s[i] == s[k1] and a[i] != a[k2] and s[i] == 0:
                    A[r, r] = 2
                    b[r] = c - t2
                else:
                    A[r, r] = 3
                    b[r] = c
            else:
                k = ks[0]
                if a[i] != a[j]:
                    A[r, r] = 1
                    b[r] = dc[k] + (-dc[i] * s[j] + dc[j] * s[i]) / (s[j] ** 2 - s[i] ** 2)
                elif a[i] == a[j] and s[i] != 0:
                    A[r, r] = 1.5
                    A[r, 3 * i + i] = -0.5 * sg(s[i] * s[j])
                    b[r] = dc[k] - dc[j] / 2 / s[i]
                else:
                    A[r, r] = 2
                    b[r] = dc[k]
    x = np.linalg.solve(A, b)
    return x.reshape(3, 3).T


def _hessian_scaled(s, cb, db):
    """pdf_MF_normal.m:152-205: the same system for the scaled constant."""
    A = np.zeros((9, 9))
    b = np.zeros(9)
    a = np.abs(s)
    sg = np.sign

    def gen(i, k):  # the generic pair term of the diagonal equations
        return (-cb / (s[i] + s[k]) + db[i] * s[i] / (s[k] ** 2 - s[i] ** 2)
                - db[k] * s[k] / (s[k] ** 2 - s[i] ** 2))

    def rep(i, k, sig):  # the repeated-|s| pair term
        return -(0.5 - sig / 2 + 0.5 / s[i]) * cb - (0.5 / s[i] + 1 - sig / 2) * db[i] + sig * db[k] / 2

    for i in range(3):
        for j in range(3):
            r = 3 * i + j
            ks = _others(i, j)
            if i == j:
                k1, k2 = ks
                if a[i] != a[k1] and a[i] != a[k2]:
                    A[r, r] = 1
                    b[r] = -2 * db[i] + gen(i, k1) + gen(i, k2)
                elif a[i] != a[k1] and a[i] == a[k2] and s[i] != 0:
                    sig = sg(s[i] * s[k2])
                    A[r, r] = 1.5
                    A[r, 3 * i + k2] = -0.5 * sig
                    b[r] = -2 * db[i] + gen(i, k1) + rep(i, k2, sig)
                elif a[i] != a[k1] and s[i] == s[k2] and s[i] == 0:
                    A[r, r] = 2
                    b[r] = -2 * db[i] + gen(i, k1) - cb - 2 * db[i]
                elif a[i] == a[k1] and a[i] != a[k2] and s[i] != 0:
                    sig = sg(s[i] * s[k1])
                    A[r, r] = 1.5
                    A[r, 3 * i + k1] = -0.5 * sig
                    b[r] = -2 * db[i] + rep(i, k1, sig) + gen(i, k2)
                elif a[i] == a[k1] and a[i] == a[k2] and s[i] != 0:
                    sig1, sig2 = sg(s[i] * s[k1]), sg(s[i] * s[k2])
                    A[r, r] = 2
                    A[r, 3 * i + k1] = -0.5 * sig1
                    A[r, 3 * i + k2] = -0.5 * sig2
                    b[r] = -2 * db[i] + rep(i, k1, sig1) + rep(i, k2, sig2)
                elif s[i] == s[k1] and a[i] != a[k2] and s[i] == 0:
                    A[r, r] = 2
                    b[r] = -2 * db[i] - cb - 2 * db[i] + gen(i, k2)
                else:
                    A[r, r] = 3
                    b[r] = -2 * db[i] - cb - 2 * db[i] - cb - 2 * db[i]
            else:
                k = ks[0]
                if a[i] != a[j]:
                    A[r, r] = 1
                    b[r] = (-cb / (s[i] + s[j]) - (1 + s[j] / (s[j] ** 2 - s[i] ** 2)) * db[i]
                            - (1 - s[i] / (s[j] ** 2 - s[i] ** 2)) * db[j] + db[k])
                elif a[i] == a[j] and s[i] != 0:
                    sig = sg(s[i] * s[j])
                    A[r, r] = 1.5
                    A[r, 3 * i + i] = -0.5 * sig
                    b[r] = (-(0.5 - sig / 2 + 0.5 / s[i]) * cb - (1.5 - sig) * db[i] - (1.5 + 0.5 / s[i]) * db[j]
                            + db[k])
                else:
                    A[r, r] = 2
                    b[r] = -cb - 2 * db[i] - 2 * db[j] + db[k]
    x = np.linalg.solve(A, b)
    return x.reshape(3, 3).T


def _func(s, d, want_df):
    # pdf_MF_M2S.m func(), bool_scaled = 1
    cb, db, ddb = mf_normal(s, True)
    f = db / cb - (d - 1.0)
    if not want_df:
        return f
    df = ddb / cb - np.outer(db, db) / cb ** 2
    return f, df


def _stack3(X, x):
    Y = x.reshape(-1, 1) if X is None else np.hstack([X, x.reshape(-1, 1)])
    return Y[:, -3:]


def m2s(d, s0=(0.0, 0.0, 0.0)):
    """[s, FVAL, NITER] = pdf_MF_M2S(d, s0), pdf_MF_M2S.m:22-102.  The line
    search's f_stack(end), f_stack(2), f_stack(3) are MATLAB linear indices
    into the 3 x n stack (single components), kept as the reference has them."""
    d = np.asarray(d, dtype=np.float64).ravel()
    s = np.asarray(s0, dtype=np.float64).ravel().copy()
    eps, alpha, sigma0, sigma1, MAX_ITER = 1e-12, 0.05, 0.1, 0.5, 100
    nf, NITER, N_SUBITER = 2 * eps, 0, 0
    f_trial = None
    lin = lambda M, k: M.ravel(order="F")[k - 1]  # MATLAB M(k)
    while nf > eps and NITER < MAX_ITER:
        NITER += 1
        f, df = _func(s, d, True)
        s_dir = -np.linalg.solve(df, f)
        f_stack = _stack3(None, f)
        lam_stack = [0.0]
        lam = 1.0
        s_trial = s + lam * s_dir
        f_trial = _func(s_trial, d, False)
        f_stack = _stack3(f_stack, f_trial)
        lam_stack = (lam_stack + [lam])[-3:]
        N_SUBITER = 0
        nrm = np.linalg.norm(f)
        while abs(lin(f_stack, f_stack.size)) > (1 - alpha * lam) * nrm and N_SUBITER < MAX_ITER:
            N_SUBITER += 1
            if len(lam_stack) < 3:
                lam = sigma1
            else:
                lam2, lam3 = lam_stack[1], lam_stack[2]
                f2, f3 = abs(lin(f_stack, 2)), abs(lin(f_stack, 3))
                c2 = 1 / lam2 / lam3 / (lam2 - lam3) * (lam3 * (f2 - nrm) - lam2 * (f3 - nrm))
                if c2 > 0:
                    c1 = 1 / (lam2 - lam3) * (-lam3 * (f2 - nrm) / lam2 + lam2 * (f3 - nrm) / lam3)
                    lt = -c1 / 2 / c2
                    lam = min(max(lt, sigma0 * lam3), sigma1 * lam3)
                else:
                    lam = sigma1 * lam3
            s_trial = s + lam * s_dir
            f_trial = _func(s_trial, d, False)
            f_stack = _stack3(f_stack, f_trial)
            lam_stack = (lam_stack + [lam])[-3:]
        s = s_trial
        nf = np.max(np.abs(f_trial))
    return s, f_trial, NITER


def psvd(F):
    """[U, S, V] = psvd(F), psvd.m:10-13."""
    U, sv, Vh = np.linalg.svd(np.asarray(F, dtype=np.float64))
    V = Vh.T
    S = np.diag(sv) @ np.diag([1.0, 1.0, np.linalg.det(U @ V)])
    U = U @ np.diag([1.0, 1.0, np.linalg.det(U)])
    V = V @ np.diag([1.0, 1.0, np.linalg.det(V)])
    return U, S, V


def vR_of(Q, s):
    """vR = (s2 Q32 - s3 Q23, s3 Q13 - s1 Q31, s1 Q21 - s2 Q12) (pendulum.m:566-568)."""
    return np.stack([s[1] * Q[2, 1] - s[2] * Q[1, 2], s[2] * Q[0, 2] - s[0] * Q[2, 0],
                     s[0] * Q[1, 0] - s[1] * Q[0, 1]])


def get_stat(f, R, x, w):
    """get_stat(f, R, x, w), pendulum.m:547-594 (dim 3) / pendulum_reduced.m:485-532
    (dim 2).  f (2B,2B,2B,N,..), R (3,3,2B,2B,2B), x (dim,N,..), w (2B)."""
    f = np.asarray(f, dtype=np.float64)
    dim = f.ndim - 3
    N = x.shape[1]
    xi = (0,) * (dim - 1)
    L = x[(0, N - 1) + xi] + x[(0, 1) + xi] - 2 * x[(0, 0) + xi]
    dv = (L / N) ** dim
    xa = tuple(range(3, 3 + dim))
    fR = f.sum(axis=xa) * dv
    ER = np.einsum("abijk,ijk,j->ab", R, fR, w)
    fx = np.einsum("ijk...,j->...", f, w)
    xf, fxf = x.reshape(dim, -1, order="F"), fx.ravel(order="F")
    Ex = xf @ fxf * dv
    Varx = (xf * fxf) @ xf.T * dv - np.outer(Ex, Ex)
    U, D, V = psvd(ER)
    s, _, niter = m2s(np.diag(D), (0.0, 0.0, 0.0))
    S = np.diag(s)
    Q = np.einsum("ca,cdijk,db->abijk", U, R, V)
    vR = vR_of(Q, s)
    EvR = np.einsum("aijk,j,ijk->a", vR, w, fR)
    EvRvR = np.einsum("aijk,bijk,j,ijk->ab", vR, vR, w, fR)
    ExvR1 = np.einsum("bijk,j,ijk...->b...", vR, w, f)
    ExvR = xf @ ExvR1.reshape(3, -1, order="F").T * dv
    covxvR = ExvR - np.outer(Ex, EvR)
    covvRvR = EvRvR - np.outer(EvR, EvR)
    P = covxvR @ np.linalg.inv(covvRvR)
    Miu = Ex - P @ EvR
    Sigma = Varx - P @ covxvR.T + P @ (np.trace(S) * np.eye(3) - S) @ P.T
    return {"ER": ER, "Ex": Ex, "Varx": Varx, "EvR": EvR, "ExvR": ExvR, "EvRvR": EvRvR, "U": U, "S": S, "V": V,
            "P": P, "Miu": Miu, "Sigma": Sigma, "niter": niter, "fx": fx}
