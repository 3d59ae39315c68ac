"""ctypes binding of the FP64 CPU oracle (liboracle.so built from refcpu.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, the smoke entry and bench.py's
cpu_baseline leg to check or time the HIP path; never by the product.

Array conventions follow refcpu.h: numpy arrays in Fortran (MATLAB) order,
complex128 for complex data.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_S = ctypes.POINTER(ctypes.c_short)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.ref_clebsch_gordan.restype = ctypes.c_int
        L.ref_derivative.restype = ctypes.c_int
        L.ref_integrate.restype = ctypes.c_int
        L.ref_getLambda.restype = ctypes.c_int
        L.ref_getLambda_compact.restype = ctypes.c_int
        L.ref_derivative_structured.restype = ctypes.c_int
        L.ref_integrate_structured.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a, t=_D):
    return a.ctypes.data_as(t)


def _f64(shape, cplx=False):
    return np.zeros(shape, dtype=np.complex128 if cplx else np.float64, order="F")


def grid(B):
    al, be, ga, w = (np.zeros(2 * B) for _ in range(4))
    lib().ref_grid(B, _p(al), _p(be), _p(ga), _p(w))
    return al, be, ga, w


def rotations(B):
    R = _f64((3, 3, 2 * B, 2 * B, 2 * B))
    lib().ref_rotations(B, _p(R))
    return R


def wigner_d(beta, lmax):
    M = 2 * lmax + 1
    d = _f64((M, M, lmax + 1))
    lib().ref_wigner_d(ctypes.c_double(beta), lmax, _p(d))
    return d


def wigner_d_jacobi(beta, l, m, n):
    out = ctypes.c_double(0.0)
    lib().ref_wigner_d_jacobi(ctypes.c_double(beta), l, m, n, ctypes.byref(out))
    return out.value


def wigner_table(B):
    M = 2 * B - 1
    d = _f64((M, M, B, 2 * B))
    lib().ref_wigner_table(B, _p(d))
    return d


def clebsch_gordan(l1, l2):
    n = (2 * l1 + 1) * (2 * l2 + 1)
    C = _f64((n, n))
    lib().ref_clebsch_gordan(l1, l2, _p(C))
    return C


def cg_table(B):
    """cell CG{l1+1,l2+1} as a dict keyed (l1, l2), plus a pointer array."""
    cells = {}
    for l2 in range(B):
        for l1 in range(B):
            cells[(l1, l2)] = clebsch_gordan(l1, l2)
    return cells


def _cg_ptrs(cells, B):
    arr = (_D * (B * B))()
    for (l1, l2), C in cells.items():
        arr[l1 + B * l2] = _p(C)
    return arr


def getu(lmax):
    M = 2 * lmax + 1
    u = _f64((M, M, lmax + 1, 3), cplx=True)
    lib().ref_getu(lmax, _p(u))
    return u


def fft_forward(B, f, d, w, dim, N):
    M = 2 * B - 1
    f = np.asfortranarray(f, dtype=np.float64)
    F = _f64((M, M, B) + (N,) * dim, cplx=True)
    lib().ref_fftSO3R_forward(B, dim, N, _p(f), _p(np.asfortranarray(d)), _p(np.ascontiguousarray(w)), _p(F))
    return F


def fft_backward(B, F, d, dim, N):
    F = np.asfortranarray(F, dtype=np.complex128)
    f = _f64((2 * B,) * 3 + (N,) * dim)
    lib().ref_fftSO3R_backward(B, dim, N, _p(F), _p(np.asfortranarray(d)), _p(f))
    return f


def build_X(dim, N, L):
    X = _f64((N,) * dim + (dim,), cplx=True)
    lib().ref_build_X(dim, N, ctypes.c_double(L), _p(X))
    return X


def build_OJO(N, L, Jt):
    O = _f64((N, N, N, 3), cplx=True)
    lib().ref_build_OJO(N, ctypes.c_double(L), _p(np.asfortranarray(Jt, dtype=np.float64)), _p(O))
    return O


def build_MR(B, dim, Jt=None):
    M = 2 * B - 1
    MR = _f64((M, M, B, dim), cplx=True)
    J = np.asfortranarray(Jt if Jt is not None else np.eye(3), dtype=np.float64)
    lib().ref_build_MR(B, dim, _p(J), _p(MR))
    return MR


def derivative(B, dim, N, F, X, OJO, MR, b, G, L, u, cells, x0=0, x1=0):
    F = np.asfortranarray(F, dtype=np.complex128)
    dF = np.zeros_like(F, order="F")
    ptrs = _cg_ptrs(cells, B)
    O = _p(np.asfortranarray(OJO)) if OJO is not None else _D()
    lib().ref_derivative(B, dim, N, _p(F), _p(np.asfortranarray(X)), O, _p(np.asfortranarray(MR)),
                         _p(np.ascontiguousarray(b, dtype=np.float64)),
                         _p(np.asfortranarray(G, dtype=np.float64)), ctypes.c_double(L),
                         _p(np.asfortranarray(u)), ptrs, _p(dF), ctypes.c_long(x0), ctypes.c_long(x1))
    return dF


def integrate(B, dim, N, F, X, OJO, MR, b, G, dt, L, u, cells, method):
    F = np.asfortranarray(F, dtype=np.complex128)
    Fn = np.zeros_like(F, order="F")
    ptrs = _cg_ptrs(cells, B)
    O = _p(np.asfortranarray(OJO)) if OJO is not None else _D()
    rc = lib().ref_integrate(B, dim, N, _p(F), _p(np.asfortranarray(X)), O, _p(np.asfortranarray(MR)),
                             _p(np.ascontiguousarray(b, dtype=np.float64)),
                             _p(np.asfortranarray(G, dtype=np.float64)), ctypes.c_double(dt),
                             ctypes.c_double(L), _p(np.asfortranarray(u)), ptrs, method.encode(), _p(Fn))
    if rc != 0:
        raise ValueError("unknown method %r" % method)
    return Fn


def derivative_structured(B, dim, N, F, X, Jt, MR, b, G, L, u, cells):
    """derivative() on the operator's nonzero support (refcpu_structured.c);
    Jt (3-D) gives the gyroscopic term, None omits it."""
    F = np.asfortranarray(F, dtype=np.complex128)
    dF = np.zeros_like(F, order="F")
    J = _p(np.asfortranarray(Jt, dtype=np.float64)) if Jt is not None else _D()
    rc = lib().ref_derivative_structured(B, dim, N, _p(F), _p(np.asfortranarray(X)), J, _p(np.asfortranarray(MR)),
                                         _p(np.ascontiguousarray(b, dtype=np.float64)),
                                         _p(np.asfortranarray(G, dtype=np.float64)), ctypes.c_double(L),
                                         _p(np.asfortranarray(u)), _cg_ptrs(cells, B), _p(dF))
    if rc != 0:
        raise ValueError("operator lacks the structure the structured form needs")
    return dF


def integrate_structured(B, dim, N, F, X, Jt, MR, b, G, dt, L, u, cells, method):
    F = np.asfortranarray(F, dtype=np.complex128)
    Fn = np.zeros_like(F, order="F")
    J = _p(np.asfortranarray(Jt, dtype=np.float64)) if Jt is not None else _D()
    rc = lib().ref_integrate_structured(B, dim, N, _p(F), _p(np.asfortranarray(X)), J, _p(np.asfortranarray(MR)),
                                        _p(np.ascontiguousarray(b, dtype=np.float64)),
                                        _p(np.asfortranarray(G, dtype=np.float64)), ctypes.c_double(dt),
                                        ctypes.c_double(L), _p(np.asfortranarray(u)), _cg_ptrs(cells, B),
                                        method.encode(), _p(Fn))
    if rc == -2:
        raise ValueError("unknown method %r" % method)
    if rc != 0:
        raise ValueError("operator lacks the structure the structured form needs")
    return Fn


def reset3_tables(B, R, dwall, h, r, theta_t, lambda_max, tilt):
    """Per-rotation rate lambda (0 outside the collision region), contact
    point PC (3, nR) and tangent t (3, nR) of the 3-D reset."""
    nR = 8 * B ** 3
    lam = np.zeros(nR)
    PC = _f64((3, nR))
    t = _f64((3, nR))
    lib().ref_reset3_tables(B, _p(np.asfortranarray(R)), *(ctypes.c_double(v) for v in
                                                            (dwall, h, r, theta_t, lambda_max, tilt)),
                            _p(lam), _p(PC), _p(t))
    return lam, PC, t


def reset3_pullback(Rr, t, epsilon, Om):
    out = np.zeros(3)
    lib().ref_reset3_pullback(_p(np.asfortranarray(Rr, dtype=np.float64)), _p(np.ascontiguousarray(t, dtype=np.float64)),
                              ctypes.c_double(epsilon), _p(np.ascontiguousarray(Om, dtype=np.float64)), _p(out))
    return out


def reset3_df(B, N, L, R, f, lam, PC, t, epsilon, tilt, jacobian=True):
    f = np.asfortranarray(f, dtype=np.float64)
    df = np.zeros_like(f, order="F")
    lib().ref_reset3_df(B, N, ctypes.c_double(L), _p(np.asfortranarray(R)), _p(f), _p(np.ascontiguousarray(lam)),
                        _p(np.asfortranarray(PC)), _p(np.asfortranarray(t)), ctypes.c_double(epsilon),
                        ctypes.c_double(tilt), 1 if jacobian else 0, _p(df))
    return df


def getLambda(B, N, R, x, dwall, h, r, theta_t, lambda_max):
    nR, nx = 8 * B ** 3, N * N
    lam = np.zeros(nR * nx)
    indR = np.zeros(nR, dtype=np.int32)
    PC = _f64((3, nR))
    numR = lib().ref_getLambda(B, N, _p(np.asfortranarray(R)), _p(np.asfortranarray(x)),
                               *(ctypes.c_double(v) for v in (dwall, h, r, theta_t, lambda_max)),
                               _p(lam), _p(indR, _I), _p(PC))
    return lam[: numR * nx].reshape((numR, N, N), order="F"), indR[:numR].copy(), PC


def getLambda_compact(B, N, R, x, dwall, h, r, theta_t, lambda_max):
    nR, nx = 8 * B ** 3, N * N
    lam = np.zeros(nR)
    indR = np.zeros(nR, dtype=np.int32)
    indx = np.zeros(nR * nx, dtype=np.int32)
    numx = np.zeros(nR, dtype=np.int32)
    PC = _f64((3, nR))
    numR = lib().ref_getLambda_compact(B, N, _p(np.asfortranarray(R)), _p(np.asfortranarray(x)),
                                       *(ctypes.c_double(v) for v in (dwall, h, r, theta_t, lambda_max)),
                                       _p(lam), _p(indR, _I), _p(indx, _I), _p(numx, _I), _p(PC))
    lists = [indx[k * nx: k * nx + numx[k]].copy() for k in range(numR)]
    return lam[:numR].copy(), indR[:numR].copy(), lists, PC


def getOmega(B, N, R, x, indR, epsilon, PC, direction):
    numR = len(indR)
    Om = _f64((2, numR, N, N))
    lib().ref_getOmega(B, N, _p(np.asfortranarray(R)), _p(np.asfortranarray(x)),
                       _p(np.ascontiguousarray(indR, dtype=np.int32), _I), numR,
                       ctypes.c_double(epsilon), _p(np.asfortranarray(PC)),
                       1 if direction.lower() == "old" else 0, _p(Om))
    return Om


def getFcL(N, x, Omega, lam, Gd, nD):
    numR = Omega.shape[1]
    fcL = _f64((nD, numR, N, N))
    idx = np.zeros((nD, numR, N, N), dtype=np.int16, order="F")
    lib().ref_getFcL(N, _p(np.asfortranarray(x)), _p(np.asfortranarray(Omega)),
                     _p(np.asfortranarray(lam)), numR, _p(np.asfortranarray(Gd, dtype=np.float64)),
                     nD, _p(fcL), _p(idx, _S))
    return fcL, idx


def getIndRule(N, x, Omega):
    numR = Omega.shape[1]
    ind = np.zeros((4, numR, N, N), dtype=np.int32, order="F")
    coeff = _f64((4, numR, N, N))
    lib().ref_getIndRule(N, _p(np.asfortranarray(x)), _p(np.asfortranarray(Omega)), numR,
                         _p(ind, _I), _p(coeff))
    return ind, coeff


def discrete_noise(B, N, f, lam, indR, fcL, fcL_indx):
    df = np.zeros_like(np.asfortranarray(f))
    nD = fcL.shape[0]
    lib().ref_discrete_noise(B, N, _p(np.asfortranarray(f)), _p(np.asfortranarray(lam)),
                             _p(np.ascontiguousarray(indR, dtype=np.int32), _I), len(indR),
                             _p(np.asfortranarray(fcL)), _p(np.asfortranarray(fcL_indx, dtype=np.int16), _S),
                             nD, _p(df))
    return df


def discrete_nonoise(B, N, f, lam, indR, lambda_indx, ind_interp, coeff_interp):
    df = np.zeros_like(np.asfortranarray(f))
    numR = len(indR)
    keep = [np.ascontiguousarray(a, dtype=np.int32) for a in lambda_indx]
    ptrs = (_I * numR)(*[_p(a, _I) for a in keep])
    nums = np.array([len(a) for a in keep], dtype=np.int32)
    lib().ref_discrete_nonoise(B, N, _p(np.asfortranarray(f)), _p(np.ascontiguousarray(lam, dtype=np.float64)),
                               _p(np.ascontiguousarray(indR, dtype=np.int32), _I), numR, ptrs,
                               _p(nums, _I), _p(np.asfortranarray(ind_interp, dtype=np.int32), _I),
                               _p(np.asfortranarray(coeff_interp)), _p(df))
    return df


def stat_moments(B, dim, N, L, f, R, x, w):
    ER = np.zeros((3, 3), order="F")
    Ex = np.zeros(dim)
    Varx = np.zeros((dim, dim), order="F")
    fx = np.zeros((N,) * dim, order="F")
    lib().ref_stat_moments(B, dim, N, ctypes.c_double(L), _p(np.asfortranarray(f)),
                           _p(np.asfortranarray(R)), _p(np.asfortranarray(x)),
                           _p(np.ascontiguousarray(w)), _p(ER), _p(Ex), _p(Varx), _p(fx))
    return ER, Ex, Varx, fx


def derivative_den(B, N, F, f, X, mR, b, G, L, u, d, w):
    F = np.asfortranarray(F, dtype=np.complex128)
    dF = np.zeros_like(F, order="F")
    fp = _p(np.asfortranarray(f, dtype=np.float64)) if f is not None else _D()
    lib().ref_derivative_den(B, N, _p(F), fp, _p(np.asfortranarray(X)), _p(np.asfortranarray(mR)),
                             _p(np.ascontiguousarray(b, dtype=np.float64)), _p(np.asfortranarray(G, dtype=np.float64)),
                             ctypes.c_double(L), _p(np.asfortranarray(u)), _p(np.asfortranarray(d)),
                             _p(np.ascontiguousarray(w)), _p(dF))
    return dF


def integrate_den(B, N, F, f, X, mR, b, G, dt, L, u, d, w, method):
    F = np.asfortranarray(F, dtype=np.complex128)
    Fn = np.zeros_like(F, order="F")
    rc = lib().ref_integrate_den(B, N, _p(F), _p(np.asfortranarray(f, dtype=np.float64)), _p(np.asfortranarray(X)),
                                 _p(np.asfortranarray(mR)), _p(np.ascontiguousarray(b, dtype=np.float64)),
                                 _p(np.asfortranarray(G, dtype=np.float64)), ctypes.c_double(dt), ctypes.c_double(L),
                                 _p(np.asfortranarray(u)), _p(np.asfortranarray(d)), _p(np.ascontiguousarray(w)),
                                 method.encode(), _p(Fn))
    if rc != 0:
        raise ValueError("unknown method %r" % method)
    return Fn


def grid_mR(B):
    """mR = [R32; -R31] on the rotation grid, (2, 2B, 2B, 2B) (pendulum_reduced_hybrid.m:115)."""
    R = rotations(B)
    return np.asfortranarray(np.stack([R[2, 1], -R[2, 0]], axis=0))


def getc(f, e, L):
    """Plot colours of the attitude marginal, restating pendulum_plot_getc
    (propagation/mex/sources/pendulum_plot_getc/src/getc.cpp:7-40 get_color, 42-61
    get_margin, 145-175 ifftSO3_single with wigner_D 256-296): the rotation
    marginal (L/N)^(ndims-3) sum_x f, its SO(3) transform (fftSO3R.m with nx = 1,
    the refcpu forward), and sum_l (2l+1) sum_{m,n} Re(F^l_mn d^l_mn(beta)
    e^{-i m alpha} e^{-i n gamma}) at every Euler triple of e (3, ns, na, 3);
    c[s, k] = sum_a of the values.  Pure Python over the points: small cases."""
    f = np.asarray(f, dtype=np.float64)
    B = f.shape[0] // 2
    N = f.shape[3]
    dim = f.ndim - 3
    nR = (2 * B) ** 3
    fR = f.reshape(nR, -1, order="F").sum(axis=1) * (L / N) ** dim
    d = wigner_table(B)
    w = grid(B)[3]
    F = fft_forward(B, fR.reshape((2 * B,) * 3, order="F"), d, w, 0, 1)
    lmax = B - 1
    ms = np.arange(-lmax, lmax + 1)
    ns, na = e.shape[1], e.shape[2]
    c = np.zeros((ns, 3))
    for k in range(3):
        for s in range(ns):
            for a in range(na):
                al, be, ga = e[:, s, a, k]
                dd = wigner_d(be, lmax)
                ph = np.exp(-1j * ms[:, None] * al) * np.exp(-1j * ms[None, :] * ga)
                val = 0.0
                for l in range(B):
                    val += (2 * l + 1) * np.real(np.sum(F[:, :, l] * dd[:, :, l] * ph))
                c[s, k] += val
    return c


def gyrobias_g1_term(B, F, u, G1):
    """sum_ij G1_ij F u_i^T u_j^T on every l block (gyro_bias.m:273-284, the
    non-conjugate transposes of the MEX's CUBLAS_OP_T, integrate.cu:363-390)."""
    lmax = B - 1
    out = np.zeros_like(F)
    for l in range(B):
        ix = slice(lmax - l, lmax + l + 1)
        Fl = F[ix, ix, l]
        for i in range(3):
            for j in range(3):
                if G1[i, j] == 0:
                    continue
                T = np.einsum("mk...,nk->mn...", Fl, u[ix, ix, l, i])
                T = np.einsum("mk...,nk->mn...", T, u[ix, ix, l, j])
                out[ix, ix, l] += G1[i, j] * T
    return out


def gyrobias_step(B, N, F, X, dt, u, G1, G2):
    """One Euler step of gyro_bias.m's integrate() (gyro_bias.m:235-298): the
    kinematic term and the bias random walk are the refcpu derivative of the 3-D
    model with no damping, gyroscopic or gravity term on the bias box L = 2
    (gyro_bias.m:64), plus the gyro random walk above."""
    M = 2 * B - 1
    OJO = np.zeros((N, N, N, 3), dtype=np.complex128, order="F")
    MR = np.zeros((M, M, B, 3), dtype=np.complex128, order="F")
    dF = derivative(B, 3, N, F, X, OJO, MR, np.zeros(3), G2, 2.0, u, cg_table(B))
    dF = dF + gyrobias_g1_term(B, np.asarray(F), u, np.asarray(G1))
    return np.asfortranarray(F + dt * dF)
