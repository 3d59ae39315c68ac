"""hpup: host-side mirror of the reference's operator interface over libhpup.so.

Two layers, both over the C ABI declared in include/hpup.h:

* gateway functions with the reference MEX names and argument meaning
  (``fftSO3R_mex``, ``pendulum_propagate_reduced[32]``, ``pendulum_propagate[32]``,
  ``getLambda_mex``, ``getOmega_mex``, ``getFcL_mex``, ``getIndRule_mex``,
  ``pendulum_reduced_discrete_propagate``) taking numpy arrays in MATLAB
  (Fortran) order and returning the array the gateway returns;
* ``Plan``: device-resident tables and coefficients (torch tensors on the GPU),
  steps run back to back on the current HIP stream.

The library is required: every entry point raises if libhpup.so is missing or
no GPU is present.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libhpup.so")

HPUP_FP64, HPUP_FP32, HPUP_BF16X3 = 0, 1, 2
METHODS = {"euler": 0, "midpoint": 1, "rk2": 2, "rk4": 3}
PRECISIONS = {"fp64": HPUP_FP64, "fp32": HPUP_FP32, "bf16x3": HPUP_BF16X3}

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I16 = ctypes.POINTER(ctypes.c_int16)


class HpupError(RuntimeError):
    pass


class StatOut(ctypes.Structure):
    """hpup_stat: get_stat outputs, 3x3 blocks column-major (omega blocks with
    leading dimension dim)."""
    _fields_ = [(n, ctypes.c_double * k) for n, k in
                (("ER", 9), ("Ex", 3), ("Varx", 9), ("EvR", 3), ("ExvR", 9), ("EvRvR", 9),
                 ("U", 9), ("S", 9), ("V", 9), ("P", 9), ("Miu", 3), ("Sigma", 9))] + \
              [("niter", ctypes.c_int), ("fit_ok", ctypes.c_int)]

    def as_dict(self, dim):
        m = lambda n, r, c: np.array(getattr(self, n)[: r * c]).reshape((r, c), order="F")
        return {"ER": m("ER", 3, 3), "Ex": np.array(self.Ex[:dim]), "Varx": m("Varx", dim, dim),
                "EvR": np.array(self.EvR[:3]), "ExvR": m("ExvR", dim, 3), "EvRvR": m("EvRvR", 3, 3),
                "U": m("U", 3, 3), "S": m("S", 3, 3), "V": m("V", 3, 3), "P": m("P", dim, 3),
                "Miu": np.array(self.Miu[:dim]), "Sigma": m("Sigma", dim, dim),
                "niter": int(self.niter), "fit_ok": bool(self.fit_ok)}


class Model(ctypes.Structure):
    _fields_ = [("B", ctypes.c_int), ("dim", ctypes.c_int), ("N", ctypes.c_int), ("L", ctypes.c_double),
                ("b", ctypes.c_double * 3), ("G", ctypes.c_double * 9), ("Jt", ctypes.c_double * 9),
                ("dt", ctypes.c_double), ("method", ctypes.c_int), ("precision", ctypes.c_int)]


class ResetParams(ctypes.Structure):
    _fields_ = [("d", ctypes.c_double), ("h", ctypes.c_double), ("r", ctypes.c_double),
                ("theta_t", ctypes.c_double), ("lambda_max", ctypes.c_double), ("epsilon", ctypes.c_double),
                ("tilt", ctypes.c_double), ("jacobian", ctypes.c_int)]


_lib = None

# every symbol of include/hpup.h
EXPORTS = [
    "hpup_last_error", "hpup_version", "hpup_parse_method",
    "hpup_fftSO3R_forward", "hpup_fftSO3R_backward",
    "hpup_pendulum_propagate_reduced", "hpup_pendulum_propagate_reduced32",
    "hpup_pendulum_propagate", "hpup_pendulum_propagate32", "hpup_pendulum_propagate_reduced_den",
    "hpup_getLambda", "hpup_getLambda_compact", "hpup_getOmega", "hpup_getFcL", "hpup_getIndRule",
    "hpup_discrete_propagate_noise", "hpup_discrete_propagate_nonoise",
    "hpup_plan_create", "hpup_plan_destroy", "hpup_plan_num_coeffs", "hpup_plan_num_grid",
    "hpup_plan_forward", "hpup_plan_backward", "hpup_plan_step", "hpup_plan_derivative",
    "hpup_plan_full_to_compact", "hpup_plan_compact_to_full", "hpup_plan_stat_moments",
    "hpup_plan_profile", "hpup_plan_profile_name", "hpup_plan_profile_bytes",
    "hpup_get_stat", "hpup_psvd", "hpup_pdf_MF_normal", "hpup_pdf_MF_M2S", "hpup_plan_get_stat",
    "hpup_pendulum_plot_getc", "hpup_gyrobias_propagate", "hpup_plan_create_operator",
    "hpup_plan_adjoint_stage", "hpup_plan_axpy_stage",
    "hpup_plan_set_reset", "hpup_plan_reset_region", "hpup_plan_reset", "hpup_plan_hybrid_step",
    "hpup_plan_reset_adjoint", "hpup_plan_reset_eps_grad", "hpup_plan_adjoint_stage_noalloc",
    "hpup_plan_step_stages", "hpup_plan_adjoint_step",
]


def lib():
    """Loads libhpup.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        # torch bundles its own HIP runtime under the same soname as ROCm's
        # (libamdhip64.so.7); whichever loads first serves the whole process.
        # Loading torch first keeps one runtime that both torch and libhpup
        # initialise; the other order leaves torch without a device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise HpupError("libhpup.so is not built at %s (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.hpup_last_error.restype = ctypes.c_char_p
        L.hpup_plan_num_coeffs.restype = ctypes.c_int64
        L.hpup_plan_num_grid.restype = ctypes.c_int64
        L.hpup_plan_create.argtypes = [ctypes.POINTER(Model), ctypes.POINTER(ctypes.c_void_p)]
        L.hpup_plan_create_operator.argtypes = [ctypes.POINTER(Model), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        for name in ("hpup_plan_forward", "hpup_plan_backward", "hpup_plan_derivative",
                     "hpup_plan_full_to_compact", "hpup_plan_compact_to_full"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.hpup_plan_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.hpup_plan_axpy_stage.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                           ctypes.c_void_p, ctypes.c_void_p]
        L.hpup_plan_adjoint_stage.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_void_p]
        L.hpup_plan_adjoint_stage_noalloc.argtypes = L.hpup_plan_adjoint_stage.argtypes
        L.hpup_plan_step_stages.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.c_void_p, ctypes.c_void_p]
        L.hpup_plan_adjoint_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.hpup_plan_destroy.argtypes = [ctypes.c_void_p]
        L.hpup_plan_num_coeffs.argtypes = [ctypes.c_void_p]
        L.hpup_plan_num_grid.argtypes = [ctypes.c_void_p]
        L.hpup_plan_stat_moments.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _D, _D, _D, _D, ctypes.c_void_p]
        L.hpup_plan_profile.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int,
                                        ctypes.c_void_p]
        L.hpup_plan_profile_name.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.hpup_plan_profile_name.restype = ctypes.c_char_p
        L.hpup_plan_profile_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.hpup_plan_profile_bytes.restype = ctypes.c_double
        L.hpup_plan_get_stat.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(StatOut), _D,
                                         ctypes.c_void_p]
        L.hpup_get_stat.argtypes = [_D, _I64, ctypes.c_int, _D, _D, _D, ctypes.POINTER(StatOut), _D]
        L.hpup_psvd.argtypes = [_D, _D, _D, _D]
        L.hpup_pdf_MF_normal.argtypes = [_D, ctypes.c_int, _D, _D, _D]
        L.hpup_pdf_MF_M2S.argtypes = [_D, _D, _D, _D, ctypes.POINTER(ctypes.c_int)]
        L.hpup_plan_set_reset.argtypes = [ctypes.c_void_p, ctypes.POINTER(ResetParams)]
        L.hpup_plan_reset_region.argtypes = [ctypes.c_void_p, _I32, ctypes.c_int]
        L.hpup_plan_reset.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_void_p]
        L.hpup_plan_hybrid_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_void_p]
        L.hpup_plan_reset_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                              ctypes.c_void_p]
        L.hpup_plan_reset_eps_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                               ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def ridx(B, l, m, n):
    """Device compact rotation row of (l, m, n): m-major, then l = |m|..B-1,
    then n = -l..l (the same formula as the C library)."""
    lmax = B - 1
    sq = lambda k: k * (k + 1) * (2 * k + 1) // 6
    s = sq(lmax) - sq(-m) if m <= 0 else sq(lmax) + sq(m - 1)
    return (m + lmax) * B * B - s + l * l - m * m + (n + l)


def _check(rc):
    if rc != 0:
        raise HpupError("libhpup error %d: %s" % (rc, lib().hpup_last_error().decode()))


def _p(a, t=_D):
    return a.ctypes.data_as(t)


def _dims(a):
    return (ctypes.c_int64 * a.ndim)(*a.shape)


def _interleaved(a, dtype):
    """complex array -> contiguous Fortran-order buffer viewed as interleaved reals."""
    return np.asfortranarray(a, dtype=dtype)


# ------------------------------------------------------------ transforms --

def fftSO3R_mex(direction, *args):
    """F = fftSO3R_mex('forward', f, d, w) / f = fftSO3R_mex('backward', F, d)."""
    if direction.lower() == "forward":
        f, d, w = args
        f = np.asfortranarray(f, dtype=np.float64)
        B = f.shape[0] // 2
        F = np.zeros((2 * B - 1, 2 * B - 1, B) + f.shape[3:], dtype=np.complex128, order="F")
        _check(lib().hpup_fftSO3R_forward(_p(f), _dims(f), f.ndim, _p(np.asfortranarray(d, dtype=np.float64)),
                                          _p(np.ascontiguousarray(np.ravel(w), dtype=np.float64)), _p(F)))
        return F
    if direction.lower() == "backward":
        F, d = args
        F = _interleaved(F, np.complex128)
        B = F.shape[2]
        f = np.zeros((2 * B,) * 3 + F.shape[3:], dtype=np.float64, order="F")
        _check(lib().hpup_fftSO3R_backward(_p(F), _dims(F), F.ndim, _p(np.asfortranarray(d, dtype=np.float64)),
                                           _p(f)))
        return f
    raise ValueError("direction must be 'forward' or 'backward'")


def _cg_ptrs(CG, B, dtype, keep):
    t = _D if dtype == np.float64 else _F
    arr = (t * (B * B))()
    for key, C in CG.items():
        l1, l2 = key
        a = np.asfortranarray(C, dtype=dtype)
        keep.append(a)
        arr[l1 + B * l2] = a.ctypes.data_as(t)
    return arr


def _propagate(fn, dtype, F, X, OJO, MR, bt, G, dtt, L, u, CG, method):
    cdt = np.complex128 if dtype == np.float64 else np.complex64
    t = _D if dtype == np.float64 else _F
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    F = _interleaved(F, cdt)
    B = F.shape[2]
    keep = []
    cg = _cg_ptrs(CG, B, dtype, keep)
    out = np.zeros_like(F, order="F")
    args = [_p(F, t), _dims(F), _p(_interleaved(X, cdt), t)]
    if OJO is not None:
        args.append(_p(_interleaved(OJO, cdt), t))
    args += [_p(_interleaved(MR, cdt), t), _p(np.ascontiguousarray(np.ravel(bt), dtype=dtype), t),
             _p(np.asfortranarray(G, dtype=dtype), t), ct(dtt), ct(L), _p(_interleaved(u, cdt), t), cg,
             method.encode(), _p(out, t)]
    _check(fn(*args))
    return out


def pendulum_propagate_reduced(F, X, MR, bt, G, dtt, L, u, CG, method):
    return _propagate(lib().hpup_pendulum_propagate_reduced, np.float64, F, X, None, MR, bt, G, dtt, L, u, CG, method)


def pendulum_propagate_reduced32(F, X, MR, bt, G, dtt, L, u, CG, method):
    return _propagate(lib().hpup_pendulum_propagate_reduced32, np.float32, F, X, None, MR, bt, G, dtt, L, u, CG,
                      method)


def pendulum_propagate_reduced_den(F, f, X, mR, bt, G, dtt, L, u, d, w, method):
    """Fnew = pendulum_propagate_reduced_den(F, f, X, mR, bt, G, dtt, L, u, d, w, method)."""
    F = _interleaved(F, np.complex128)
    out = np.zeros_like(F, order="F")
    _check(lib().hpup_pendulum_propagate_reduced_den(
        _p(F), _dims(F), _p(np.asfortranarray(f, dtype=np.float64)), _p(_interleaved(X, np.complex128)),
        _p(np.asfortranarray(mR, dtype=np.float64)), _p(np.ascontiguousarray(np.ravel(bt), dtype=np.float64)),
        _p(np.asfortranarray(G, dtype=np.float64)), ctypes.c_double(dtt), ctypes.c_double(L),
        _p(_interleaved(u, np.complex128)), _p(np.asfortranarray(d, dtype=np.float64)),
        _p(np.ascontiguousarray(np.ravel(w), dtype=np.float64)), method.encode(), _p(out)))
    return out


def pendulum_propagate(F, X, OJO, MR, bt, G, dtt, L, u, CG, method):
    return _propagate(lib().hpup_pendulum_propagate, np.float64, F, X, OJO, MR, bt, G, dtt, L, u, CG, method)


def pendulum_propagate32(F, X, OJO, MR, bt, G, dtt, L, u, CG, method):
    return _propagate(lib().hpup_pendulum_propagate32, np.float32, F, X, OJO, MR, bt, G, dtt, L, u, CG, method)


# ------------------------------------------------------------ hybrid reset --
# Reduced model (2-D omega).  Index arrays are 0-based int32/int16 as the
# reference's C code produces them; x is the (2, N, N) omega grid.

def getLambda_mex(R, x, d, h, r, theta_t, lambda_max, compact=False):
    """[lambda, lambda_indR, PC] = getLambda_mex(..., false) or
    [lambda, lambda_indR, lambda_indx, PC] = getLambda_mex(..., true)."""
    R = np.asfortranarray(R, dtype=np.float64)
    x = np.asfortranarray(x, dtype=np.float64)
    B, N = R.shape[2] // 2, x.shape[1]
    nR, nx = 8 * B ** 3, N * N
    indR = np.zeros(nR, dtype=np.int32)
    PC = np.zeros((3, nR), order="F")
    numR = ctypes.c_int(0)
    vals = [ctypes.c_double(v) for v in (d, h, r, theta_t, lambda_max)]
    if not compact:
        lam = np.zeros(nR * nx)
        _check(lib().hpup_getLambda(_p(R), B, _p(x), N, *vals, _p(lam), _p(indR, _I32), ctypes.byref(numR), _p(PC)))
        k = numR.value
        return lam[: k * nx].reshape((k, N, N), order="F"), indR[:k].copy(), PC
    lam = np.zeros(nR)
    indx = np.zeros(nR * nx, dtype=np.int32)
    numx = np.zeros(nR, dtype=np.int32)
    _check(lib().hpup_getLambda_compact(_p(R), B, _p(x), N, *vals, _p(lam), _p(indR, _I32), ctypes.byref(numR),
                                        _p(indx, _I32), _p(numx, _I32), _p(PC)))
    k = numR.value
    return lam[:k].copy(), indR[:k].copy(), [indx[i * nx: i * nx + numx[i]].copy() for i in range(k)], PC


def getOmega_mex(R, x, indR, epsilon, PC, direction):
    R = np.asfortranarray(R, dtype=np.float64)
    x = np.asfortranarray(x, dtype=np.float64)
    B, N = R.shape[2] // 2, x.shape[1]
    indR = np.ascontiguousarray(indR, dtype=np.int32)
    Om = np.zeros((2, len(indR), N, N), order="F")
    _check(lib().hpup_getOmega(_p(R), B, _p(x), N, _p(indR, _I32), len(indR), ctypes.c_double(epsilon),
                               _p(np.asfortranarray(PC, dtype=np.float64)), direction.encode(), _p(Om)))
    return Om


def getFcL_mex(x, Omega, lam, Gd, nD):
    x = np.asfortranarray(x, dtype=np.float64)
    N, numR = x.shape[1], Omega.shape[1]
    fcL = np.zeros((nD, numR, N, N), order="F")
    idx = np.zeros((nD, numR, N, N), dtype=np.int16, order="F")
    _check(lib().hpup_getFcL(_p(x), N, _p(np.asfortranarray(Omega, dtype=np.float64)),
                             _p(np.asfortranarray(lam, dtype=np.float64)), numR,
                             _p(np.asfortranarray(Gd, dtype=np.float64)), int(nD), _p(fcL), _p(idx, _I16)))
    return fcL, idx


def getIndRule_mex(x, Omega):
    x = np.asfortranarray(x, dtype=np.float64)
    N, numR = x.shape[1], Omega.shape[1]
    ind = np.zeros((4, numR, N, N), dtype=np.int32, order="F")
    coeff = np.zeros((4, numR, N, N), order="F")
    _check(lib().hpup_getIndRule(_p(x), N, _p(np.asfortranarray(Omega, dtype=np.float64)), numR, _p(ind, _I32),
                                 _p(coeff)))
    return ind, coeff


def pendulum_reduced_discrete_propagate(noise, f, lam, indR, *args):
    """df = pendulum_reduced_discrete_propagate(true, f, lambda, indR, fcL, fcL_indx) or
    (false, f, lambda, indR, lambda_indx, ind_interp, coeff_interp)."""
    f = np.asfortranarray(f, dtype=np.float64)
    B, N = f.shape[0] // 2, f.shape[3]
    indR = np.ascontiguousarray(indR, dtype=np.int32)
    df = np.zeros_like(f, order="F")
    if noise:
        fcL, idx = args
        _check(lib().hpup_discrete_propagate_noise(_p(f), B, N, _p(np.asfortranarray(lam, dtype=np.float64)),
                                                   _p(indR, _I32), len(indR), _p(np.asfortranarray(fcL)),
                                                   _p(np.asfortranarray(idx, dtype=np.int16), _I16), fcL.shape[0],
                                                   _p(df)))
        return df
    lambda_indx, ind_interp, coeff_interp = args
    keep = [np.ascontiguousarray(a, dtype=np.int32) for a in lambda_indx]
    ptrs = (_I32 * max(len(keep), 1))(*[_p(a, _I32) for a in keep])
    nums = np.array([len(a) for a in keep] or [0], dtype=np.int32)
    _check(lib().hpup_discrete_propagate_nonoise(_p(f), B, N, _p(np.ascontiguousarray(lam, dtype=np.float64)),
                                                 _p(indR, _I32), len(indR), ptrs, _p(nums, _I32),
                                                 _p(np.asfortranarray(ind_interp, dtype=np.int32), _I32),
                                                 _p(np.asfortranarray(coeff_interp, dtype=np.float64)), _p(df)))
    return df


# ------------------------------------------------------------------ plan --

# ---------------------------------------------------- statistics, fit --

def get_stat(f, R, x, w, return_fx=False):
    """[ER, Ex, Varx, EvR, ExvR, EvRvR, U, S, V, P, Miu, Sigma(, fx)] = get_stat(f, R, x, w)
    (the drivers' local function, pendulum.m:547-594) as a dict; the grid
    sums run on the GPU."""
    f = np.asfortranarray(f, dtype=np.float64)
    dim = f.ndim - 3
    out = StatOut()
    fx = np.zeros(f.shape[3:], order="F")
    _check(lib().hpup_get_stat(_p(f), _dims(f), f.ndim, _p(np.asfortranarray(R, dtype=np.float64)),
                               _p(np.asfortranarray(x, dtype=np.float64)),
                               _p(np.ascontiguousarray(np.ravel(w), dtype=np.float64)), ctypes.byref(out), _p(fx)))
    r = out.as_dict(dim)
    if return_fx:
        r["fx"] = fx
    return r


def gyrobias_propagate(F, X, dt, u, G1, G2):
    """Fnew = gyrobias_propagate(F, X, dt, u, G1, G2) (gyrobias_propagate/kernel.cu:8-120): one
    Euler step of the attitude x gyro-bias model, FP64."""
    F = np.asfortranarray(F, dtype=np.complex128)
    out = np.zeros_like(F, order="F")
    _check(lib().hpup_gyrobias_propagate(_p(F), _dims(F), _p(_interleaved(X, np.complex128)), ctypes.c_double(dt),
                                         _p(_interleaved(u, np.complex128)),
                                         _p(np.asfortranarray(G1, dtype=np.float64)),
                                         _p(np.asfortranarray(G2, dtype=np.float64)), _p(out)))
    return out


def pendulum_plot_getc(f, e, L):
    """c = pendulum_plot_getc(f, e, L) (pendulum_plot_getc/src/mexFunction.cpp:5-33): the
    rotation marginal of the grid density f, SO(3)-transformed and evaluated at the Euler
    triples e (3, ns, na, 3); c (ns, 3) sums the na evaluations."""
    f = np.asfortranarray(f, dtype=np.float64)
    e = np.asfortranarray(e, dtype=np.float64)
    if e.ndim != 4:
        raise HpupError("e must be (3, ns, na, 3)")
    c = np.zeros((e.shape[1], 3), order="F")
    _check(lib().hpup_pendulum_plot_getc(_p(f), _dims(f), f.ndim, _p(e), _dims(e), e.ndim, ctypes.c_double(L), _p(c)))
    return c


def psvd(F):
    """[U, S, V] = psvd(F) (matrix Fisher/psvd.m)."""
    F = np.asfortranarray(F, dtype=np.float64)
    U, S, V = (np.zeros((3, 3), order="F") for _ in range(3))
    _check(lib().hpup_psvd(_p(F), _p(U), _p(S), _p(V)))
    return U, S, V


def pdf_MF_normal(s, bool_scaled=False):
    """[c, dc, ddc] = pdf_MF_normal(s, bool_scaled, 1, 1)."""
    s = np.ascontiguousarray(np.ravel(s), dtype=np.float64)
    c = np.zeros(1)
    dc = np.zeros(3)
    ddc = np.zeros((3, 3), order="F")
    _check(lib().hpup_pdf_MF_normal(_p(s), int(bool(bool_scaled)), _p(c), _p(dc), _p(ddc)))
    return float(c[0]), dc, ddc


def pdf_MF_M2S(d, s0=None):
    """[s, FVAL, NITER] = pdf_MF_M2S(d, s0); NITER = -1 if not converged."""
    d = np.ascontiguousarray(np.ravel(d), dtype=np.float64)
    s = np.zeros(3)
    fval = np.zeros(3)
    it = ctypes.c_int(0)
    s0p = _p(np.ascontiguousarray(np.ravel(s0), dtype=np.float64)) if s0 is not None else _D()
    _check(lib().hpup_pdf_MF_M2S(_p(d), s0p, _p(s), _p(fval), ctypes.byref(it)))
    return s, fval, it.value


class Plan:
    """Device-resident propagator.  Coefficients are torch tensors on cuda:0
    in the compact layout [r][k3][k2][k1] (complex128 for fp64, complex64
    otherwise); the grid is real (2B,2B,2B,N..) in MATLAB order, stored as a
    flat tensor.

    operator selects what the plan applies (no counterpart in the reference;
    used by the gradient harness in hpup.adjoint): 0 the model's operator A,
    1 its adjoint A^H (dim 2 or 3, N <= 32; step then applies the adjoint of
    the forward step), 2 + p the derivative dA/db_p, 2 + dim + i + dim j
    dA/dG_ij."""

    def __init__(self, B, dim, N=1, L=1.0, b=(0, 0, 0), G=None, Jt=None, dt=0.0, method="euler",
                 precision="fp64", operator=0):
        import torch
        if not torch.cuda.is_available():
            raise HpupError("Plan needs a GPU")
        m = Model()
        m.B, m.dim, m.N, m.L, m.dt = B, dim, N if dim else 1, L, dt
        m.method = METHODS[method.lower()]
        m.precision = PRECISIONS[precision.lower()]
        for i, v in enumerate(b):
            m.b[i] = v
        if G is not None:
            Gm = np.asarray(G, dtype=np.float64)
            for j in range(dim):
                for i in range(dim):
                    m.G[i + dim * j] = Gm[i, j]
        Jm = np.eye(3) if Jt is None else np.asarray(Jt, dtype=np.float64)
        for j in range(3):
            for i in range(3):
                m.Jt[i + 3 * j] = Jm[i, j]
        self.model = m
        self.precision = precision.lower()
        self.B, self.dim, self.N = B, dim, (N if dim else 1)
        self.nx = self.N ** dim
        self.nrc = B * (2 * B - 1) * (2 * B + 1) // 3
        h = ctypes.c_void_p()
        if operator:
            _check(lib().hpup_plan_create_operator(ctypes.byref(m), int(operator), ctypes.byref(h)))
        else:
            _check(lib().hpup_plan_create(ctypes.byref(m), ctypes.byref(h)))
        self._h = h
        self.operator = int(operator)
        self.ncoef = lib().hpup_plan_num_coeffs(h)
        self.ngrid = lib().hpup_plan_num_grid(h)
        self.cdtype = torch.complex128 if self.precision == "fp64" else torch.complex64
        self.rdtype = torch.float64 if self.precision == "fp64" else torch.float32

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.hpup_plan_destroy(h)
            self._h = None

    @staticmethod
    def _stream():
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _chk(self, t, n, dtype):
        import torch
        if not (t.is_cuda and t.is_contiguous() and t.numel() == n and t.dtype == dtype):
            raise HpupError("tensor must be a contiguous CUDA %s with %d elements" % (dtype, n))
        return ctypes.c_void_p(t.data_ptr())

    def empty_coeffs(self):
        import torch
        return torch.empty(self.ncoef, dtype=self.cdtype, device="cuda")

    def empty_grid(self):
        import torch
        return torch.empty(self.ngrid, dtype=self.rdtype, device="cuda")

    def forward(self, f, F=None):
        F = self.empty_coeffs() if F is None else F
        _check(lib().hpup_plan_forward(self._h, self._chk(f, self.ngrid, self.rdtype),
                                       self._chk(F, self.ncoef, self.cdtype), self._stream()))
        return F

    def backward(self, F, f=None):
        f = self.empty_grid() if f is None else f
        _check(lib().hpup_plan_backward(self._h, self._chk(F, self.ncoef, self.cdtype),
                                        self._chk(f, self.ngrid, self.rdtype), self._stream()))
        return f

    def step(self, F, nsteps=1):
        _check(lib().hpup_plan_step(self._h, self._chk(F, self.ncoef, self.cdtype), int(nsteps), self._stream()))
        return F

    def derivative(self, F, dF=None):
        dF = self.empty_coeffs() if dF is None else dF
        _check(lib().hpup_plan_derivative(self._h, self._chk(F, self.ncoef, self.cdtype),
                                          self._chk(dF, self.ncoef, self.cdtype), self._stream()))
        return dF

    def axpy_stage(self, x, F0, scale, out=None):
        """out = F0 + scale A x (one Horner stage of the plan's operator)."""
        out = self.empty_coeffs() if out is None else out
        _check(lib().hpup_plan_axpy_stage(self._h, self._chk(x, self.ncoef, self.cdtype),
                                          self._chk(F0, self.ncoef, self.cdtype), float(scale),
                                          self._chk(out, self.ncoef, self.cdtype), self._stream()))
        return out

    def adjoint_stage(self, ybar, y, scale, grad, out=None, fbar=None, fbar_init=False):
        """Adjoint plans (operator=1): out = scale A^H ybar, fbar += out (if
        given; fbar = ybar + out with fbar_init) and, on the device, grad
        (float64, 5) += scale Re<ybar, dA/dtheta y> for theta = b_1, b_2, G_00,
        G_11, G_01 (= G_10 term)."""
        import torch
        out = self.empty_coeffs() if out is None else out
        if not (grad.is_cuda and grad.dtype == torch.float64 and grad.numel() == 5 and grad.is_contiguous()):
            raise HpupError("grad must be a contiguous CUDA float64 tensor of 5 elements")
        _check(lib().hpup_plan_adjoint_stage(self._h, self._chk(ybar, self.ncoef, self.cdtype),
                                             self._chk(y, self.ncoef, self.cdtype), float(scale),
                                             self._chk(out, self.ncoef, self.cdtype),
                                             self._chk(fbar, self.ncoef, self.cdtype) if fbar is not None else None,
                                             1 if fbar_init else 0, ctypes.c_void_p(grad.data_ptr()),
                                             self._stream()))
        return out

    def step_stages(self, F, ys, F_next=None):
        """Forward model plans: ys (s - 1 coefficient tensors) receive the Horner
        stage values y_1..y_{s-1} of one step from F; F_next (optional) y_s."""
        arr = (ctypes.c_void_p * max(1, len(ys)))(*[self._chk(y, self.ncoef, self.cdtype) for y in ys])
        _check(lib().hpup_plan_step_stages(self._h, self._chk(F, self.ncoef, self.cdtype), arr,
                                           self._chk(F_next, self.ncoef, self.cdtype) if F_next is not None else None,
                                           self._stream()))
        return ys

    def adjoint_step(self, F, ys, lam, grad, Fbar=None):
        """Adjoint plans: the reverse of one step from its stage values (F, ys)
        and the incoming adjoint lam; returns Fbar, accumulates grad (device
        float64, 5)."""
        Fbar = self.empty_coeffs() if Fbar is None else Fbar
        arr = (ctypes.c_void_p * max(1, len(ys)))(*[self._chk(y, self.ncoef, self.cdtype) for y in ys])
        _check(lib().hpup_plan_adjoint_step(self._h, self._chk(F, self.ncoef, self.cdtype), arr,
                                            self._chk(lam, self.ncoef, self.cdtype),
                                            self._chk(Fbar, self.ncoef, self.cdtype),
                                            ctypes.c_void_p(grad.data_ptr()), self._stream()))
        return Fbar

    @property
    def mass_index(self):
        """Flat index of F^0_00(k = 0); the total probability is
        F[mass_index].real * (L/N)^dim."""
        return ridx(self.B, 0, 0, 0) * self.nx

    @property
    def nfull(self):
        return (2 * self.B - 1) ** 2 * self.B * self.nx

    def full_to_compact(self, Ffull, F=None):
        F = self.empty_coeffs() if F is None else F
        _check(lib().hpup_plan_full_to_compact(self._h, self._chk(Ffull, self.nfull, self.cdtype),
                                               self._chk(F, self.ncoef, self.cdtype), self._stream()))
        return F

    def compact_to_full(self, F, Ffull=None):
        import torch
        Ffull = torch.empty(self.nfull, dtype=self.cdtype, device="cuda") if Ffull is None else Ffull
        _check(lib().hpup_plan_compact_to_full(self._h, self._chk(F, self.ncoef, self.cdtype),
                                               self._chk(Ffull, self.nfull, self.cdtype), self._stream()))
        return Ffull

    def stat_moments(self, f):
        ER = np.zeros(9)
        Ex = np.zeros(max(self.dim, 1))
        Varx = np.zeros(max(self.dim, 1) ** 2)
        fx = np.zeros(self.nx)
        _check(lib().hpup_plan_stat_moments(self._h, self._chk(f, self.ngrid, self.rdtype), _p(ER), _p(Ex),
                                            _p(Varx), _p(fx), self._stream()))
        d = max(self.dim, 1)
        return (ER.reshape((3, 3), order="F"), Ex[: self.dim], Varx.reshape((d, d), order="F")[: self.dim, : self.dim],
                fx.reshape((self.N,) * self.dim, order="F") if self.dim else fx)

    def get_stat(self, f, return_fx=False):
        """Full get_stat of the device grid f on the plan's grids (dict)."""
        out = StatOut()
        fx = np.zeros(self.nx)
        _check(lib().hpup_plan_get_stat(self._h, self._chk(f, self.ngrid, self.rdtype), ctypes.byref(out), _p(fx),
                                        self._stream()))
        r = out.as_dict(self.dim)
        if return_fx:
            r["fx"] = fx.reshape((self.N,) * self.dim, order="F")
        return r

    def profile(self, F, nreps=3):
        """Average milliseconds per kernel of an Euler step (advances F by nreps
        steps), with each kernel's name and algorithmic bytes per launch."""
        ms = np.zeros(8)
        n = lib().hpup_plan_profile(self._h, self._chk(F, self.ncoef, self.cdtype), int(nreps), _p(ms), 8,
                                    self._stream())
        if n < 0:
            _check(n)
        return [(lib().hpup_plan_profile_name(self._h, i).decode(), float(ms[i]),
                 float(lib().hpup_plan_profile_bytes(self._h, i))) for i in range(n)]

    # hybrid collision reset (device grid)
    def set_reset(self, d, h, r, theta_t, lambda_max, epsilon, tilt=0.0, jacobian=False):
        """Builds the reset tables (getLambda_mex's d, h, r, theta_t,
        lambda_max; getOmega_mex's epsilon; the wall tilt and the 1/epsilon
        Jacobian of the 3-D law)."""
        prm = ResetParams(float(d), float(h), float(r), float(theta_t), float(lambda_max), float(epsilon),
                          float(tilt), 1 if jacobian else 0)
        _check(lib().hpup_plan_set_reset(self._h, ctypes.byref(prm)))
        n = lib().hpup_plan_reset_region(self._h, None, 0)
        if n < 0:
            _check(n)
        self.reset_rows = np.zeros(n, dtype=np.int32)
        if n:
            lib().hpup_plan_reset_region(self._h, _p(self.reset_rows, _I32), n)
        return self.reset_rows

    def reset(self, f, out=None, scale=1.0, mode=0):
        """mode 0: out = f + scale df(f) (out may be f); mode 1: out = scale df(f)."""
        out = self.empty_grid() if out is None else out
        _check(lib().hpup_plan_reset(self._h, self._chk(f, self.ngrid, self.rdtype),
                                     self._chk(out, self.ngrid, self.rdtype), float(scale), int(mode),
                                     self._stream()))
        return out

    def reset_adjoint(self, g, out=None, scale=1.0):
        """out = g + scale R^H g, R the reset's jump term (out may be g)."""
        out = self.empty_grid() if out is None else out
        _check(lib().hpup_plan_reset_adjoint(self._h, self._chk(g, self.ngrid, self.rdtype),
                                             self._chk(out, self.ngrid, self.rdtype), float(scale), self._stream()))
        return out

    def reset_eps_grad(self, f, g, acc, scale=1.0):
        """acc (CUDA float64, 1 element) += scale <g, dR/d epsilon f>."""
        import torch
        if not (acc.is_cuda and acc.dtype == torch.float64 and acc.numel() >= 1):
            raise HpupError("acc must be a CUDA float64 tensor")
        _check(lib().hpup_plan_reset_eps_grad(self._h, self._chk(f, self.ngrid, self.rdtype),
                                              self._chk(g, self.ngrid, self.rdtype), float(scale),
                                              ctypes.c_void_p(acc.data_ptr()), self._stream()))
        return acc

    def hybrid_step(self, F, f, nsteps=1):
        """nsteps hybrid steps: integrator step, inverse transform, reset with
        dt on the grid f, forward transform; f ends as the post-reset grid."""
        _check(lib().hpup_plan_hybrid_step(self._h, self._chk(F, self.ncoef, self.cdtype),
                                           self._chk(f, self.ngrid, self.rdtype), int(nsteps), self._stream()))
        return F, f

    # host conveniences (MATLAB full layout numpy arrays)
    def upload_full(self, Ffull_np):
        import torch
        a = np.asfortranarray(Ffull_np).ravel(order="F")
        t = torch.from_numpy(a.astype(np.complex128 if self.precision == "fp64" else np.complex64)).cuda()
        return self.full_to_compact(t)

    def download_full(self, F):
        full = self.compact_to_full(F).cpu().numpy()
        M = 2 * self.B - 1
        return full.reshape((M, M, self.B) + (self.N,) * self.dim, order="F")
