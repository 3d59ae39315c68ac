"""Pins the FP64 oracle (oracle/refcpu.c) against facts the mathematics fixes,
computed by independent routes: sympy's Wigner-d and Clebsch-Gordan, numpy's
FFT, quadrature exactness, the SO(3) product and derivative rules checked on
the grid by finite differences, the Taylor polynomial of the linear drift
for the integrators, scipy's Gaussian density and bilinear reproduction for
the reset's transition kernel and interpolation rule, and invariants of the
reset map.  The
reference ships no golden vectors for this path, so these are the anchors."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import density, omega_grid, operators, rel_l2


def _bandlimited(B, deg, rng, nx_shape=()):
    M, lmax = 2 * B - 1, B - 1
    F = rng.standard_normal((M, M, B) + nx_shape) + 1j * rng.standard_normal((M, M, B) + nx_shape)
    for l in range(B):
        for m in range(-lmax, lmax + 1):
            for n in range(-lmax, lmax + 1):
                if abs(m) > l or abs(n) > l or l > deg:
                    F[m + lmax, n + lmax, l] = 0
    return F


def _real_bandlimited(B, deg, seed):
    """a real grid function of degree <= deg and its coefficients"""
    rng = np.random.default_rng(seed)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    f = O.fft_backward(B, _bandlimited(B, deg, rng), d, 0, 1)
    return O.fft_forward(B, f, d, w, 0, 1), f, d, w


# ---------------------------------------------------------------- grid ----

@pytest.mark.parametrize("B", [2, 5, 10])
def test_weights_integrate_legendre_exactly(B):
    """sum_j w_j (2B)^2 P_l(cos beta_j) = delta_l0 for l < 2B (fftSO3R.m:63-66)."""
    from scipy.special import eval_legendre
    _, beta, _, w = O.grid(B)
    assert abs(w.sum() - 1.0 / (4 * B * B)) < 1e-15
    for l in range(2 * B):
        s = (2 * B) ** 2 * np.sum(w * eval_legendre(l, np.cos(beta)))
        assert abs(s - (1.0 if l == 0 else 0.0)) < 1e-13, l


def test_rotation_grid_is_zyz():
    B = 3
    R = O.rotations(B)
    a, b, g = math.pi / B * 2, math.pi / (4 * B) * 3, math.pi / B * 5
    rz = lambda t: np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1]])
    ry = lambda t: np.array([[math.cos(t), 0, math.sin(t)], [0, 1, 0], [-math.sin(t), 0, math.cos(t)]])
    assert np.abs(R[:, :, 2, 1, 5] - rz(a) @ ry(b) @ rz(g)).max() < 1e-15


# ------------------------------------------------------------ Wigner-d ----

def test_wigner_d_matches_sympy():
    from sympy import N as SN
    from sympy.physics.quantum.spin import Rotation
    lmax = 4
    for beta in (0.3, 1.7, 2.9):
        d = O.wigner_d(beta, lmax)
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                for n in range(-l, l + 1):
                    v = complex(SN(Rotation.d(l, m, n, beta).doit()))
                    assert abs(v.imag) < 1e-12
                    assert abs(d[m + lmax, n + lmax, l] - v.real) < 1e-13, (beta, l, m, n)


def test_wigner_d_recursion_matches_jacobi_form_high_degree():
    """The FP64 l-recursion against the closed Jacobi form (Wigner_d.m:7-9)
    evaluated in 50-digit arithmetic, up to degree 29 (B = 30)."""
    import mpmath as mp
    mp.mp.dps = 50
    lmax = 29
    beta = mp.mpf("1.1")
    d = O.wigner_d(1.1, lmax)

    def closed(l, m, n):
        k, a, b = l - n, n - m, n + m
        x = mp.cos(beta)
        P = mp.fsum(mp.binomial(k + a, k - s) * mp.binomial(k + b, s) * ((x - 1) / 2) ** s * ((x + 1) / 2) ** (k - s)
                    for s in range(k + 1))
        pre = mp.sqrt(mp.factorial(l + n) * mp.factorial(l - n) / (mp.factorial(l + m) * mp.factorial(l - m)))
        return pre * mp.sin(beta / 2) ** a * mp.cos(beta / 2) ** b * P

    for l in (15, 22, 29):
        for m in range(-l, l + 1, 5):
            for n in range(-l, l + 1, 6):
                assert abs(d[m + lmax, n + lmax, l] - float(closed(l, m, n))) < 1e-13, (l, m, n)
    # the FP64 explicit Jacobi sum agrees at moderate degree
    for l in (3, 7):
        for m in range(-l, l + 1):
            for n in range(-l, l + 1):
                assert abs(d[m + lmax, n + lmax, l] - O.wigner_d_jacobi(1.1, l, m, n)) < 1e-12


@pytest.mark.parametrize("B", [4, 8])
def test_wigner_d_orthogonal_under_quadrature(B):
    """sum_j w_j (2B)^2 (2l+1) d^l_mn d^l'_mn = delta_ll' for l, l' < B."""
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    lmax = B - 1
    for m, n in [(0, 0), (1, -1), (2, 1), (-3, 3)]:
        ls = [l for l in range(B) if l >= max(abs(m), abs(n))]
        for l in ls:
            for lp in ls:
                s = (2 * B) ** 2 * (2 * l + 1) * np.sum(w * d[m + lmax, n + lmax, l, :] * d[m + lmax, n + lmax, lp, :])
                assert abs(s - (1.0 if l == lp else 0.0)) < 1e-12


# ------------------------------------------------------- Clebsch-Gordan ----

def test_clebsch_gordan_matches_sympy_up_to_column_sign():
    from sympy import S
    from sympy.physics.quantum.cg import CG
    for l1, l2 in [(1, 1), (2, 1), (3, 2), (4, 1)]:
        C = O.clebsch_gordan(l1, l2)
        for l in range(abs(l1 - l2), l1 + l2 + 1):
            for m in range(-l, l + 1):
                col = l * l - (l2 - l1) ** 2 + l + m
                ours, ref = [], []
                for m1 in range(-l1, l1 + 1):
                    m2 = m - m1
                    if abs(m2) > l2:
                        continue
                    ours.append(C[(l1 + m1) * (2 * l2 + 1) + l2 + m2, col])
                    ref.append(float(CG(S(l1), S(m1), S(l2), S(m2), S(l), S(m)).doit()))
                ours, ref = np.array(ours), np.array(ref)
                sgn = np.sign(np.dot(ours, ref))
                assert np.abs(ours - sgn * ref).max() < 1e-12, (l1, l2, l, m)


def test_clebsch_gordan_is_orthogonal():
    for l1, l2 in [(3, 1), (5, 4), (9, 1)]:
        C = O.clebsch_gordan(l1, l2)
        assert np.abs(C.T @ C - np.eye(C.shape[0])).max() < 1e-11


def test_product_rule_with_gravity_moment():
    """The reference's CG sum (pendulum.m:474-502) equals the transform of the
    grid product mR_p f for f of degree <= B-2 (fftSO3_prod.m logic)."""
    B = 6
    lmax, M = B - 1, 2 * B - 1
    F, f, d, w = _real_bandlimited(B, B - 2, seed=1)
    MR = O.build_MR(B, 2)
    R = O.rotations(B)
    cells = O.cg_table(B)
    for p, mr in enumerate([R[2, 1], -R[2, 0]]):
        prod = O.fft_forward(B, f * mr, d, w, 0, 1)
        T = np.zeros((M, M, B), complex)
        for l in range(B):
            for l2 in range(B):
                for l1 in range(B):
                    if not (abs(l1 - l2) <= l <= l1 + l2):
                        continue
                    cl = (2 * l1 + 1) * (2 * l2 + 1) / (2 * l + 1)
                    col = l * l - (l1 - l2) ** 2
                    C = cells[(l1, l2)]
                    MRl = MR[lmax - l2:lmax + l2 + 1, lmax - l2:lmax + l2 + 1, l2, p]
                    Fl = F[lmax - l1:lmax + l1 + 1, lmax - l1:lmax + l1 + 1, l1]
                    for m in range(-l, l + 1):
                        Cm = C[:, col + m + l].reshape((2 * l1 + 1, 2 * l2 + 1)).T
                        for n in range(-l, l + 1):
                            Cn = C[:, col + n + l].reshape((2 * l1 + 1, 2 * l2 + 1)).T
                            T[m + lmax, n + lmax, l] += cl * np.sum(Fl * (Cm.T @ MRl @ Cn))
        assert rel_l2(T, prod) < 1e-13


# -------------------------------------------------------------------- u ----

def _euler_zyz(R):
    beta = math.acos(max(-1.0, min(1.0, R[2, 2])))
    alpha = math.atan2(R[1, 2], R[0, 2])
    gamma = math.atan2(R[2, 1], -R[2, 0])
    return alpha, beta, gamma


def _synth(F, B, R):
    """f(R) = sum (2l+1) F^l_mn e^{-i m alpha} d^l_mn(beta) e^{-i n gamma} (ifftSO3R.m)."""
    lmax = B - 1
    a, b, g = _euler_zyz(R)
    d = O.wigner_d(b, lmax)
    m = np.arange(-lmax, lmax + 1)
    ph = np.exp(-1j * m[:, None] * a) * np.exp(-1j * m[None, :] * g)
    return sum((2 * l + 1) * np.sum(F[:, :, l] * d[:, :, l] * ph) for l in range(B)).real


def test_getu_is_the_left_invariant_derivative():
    """F[d/dt f(R exp(t e_p^))] = F u_p^T per degree (fftSO3_deriv.m:12-26)."""
    B = 4
    lmax = B - 1
    F, f, d, w = _real_bandlimited(B, B - 1, seed=2)
    u = O.getu(lmax)
    R = O.rotations(B)
    h = 1e-5
    hat = [np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0.]]), np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0.]]),
           np.array([[0, -1, 0], [1, 0, 0], [0, 0, 0.]])]
    from scipy.linalg import expm
    for p in range(3):
        Ep, Em = expm(h * hat[p]), expm(-h * hat[p])
        df = np.zeros_like(f)
        for i in range(2 * B):
            for j in range(2 * B):
                for k in range(2 * B):
                    Rg = R[:, :, i, j, k]
                    df[i, j, k] = (_synth(F, B, Rg @ Ep) - _synth(F, B, Rg @ Em)) / (2 * h)
        dF = O.fft_forward(B, df, d, w, 0, 1)
        pred = np.zeros_like(F)
        for l in range(B):
            s = slice(lmax - l, lmax + l + 1)
            pred[s, s, l] = F[s, s, l] @ u[s, s, l, p].T
        assert rel_l2(dF, pred) < 1e-8, p


# ----------------------------------------------------------- transforms ----

@pytest.mark.parametrize("B,dim,N", [(4, 0, 1), (3, 2, 4), (3, 3, 4)])
def test_inverse_then_forward_is_identity(B, dim, N):
    rng = np.random.default_rng(B + dim)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    f = O.fft_backward(B, _bandlimited(B, B - 1, rng, (N,) * dim), d, dim, N)
    F = O.fft_forward(B, f, d, w, dim, N)          # a real-consistent bandlimited F
    F2 = O.fft_forward(B, O.fft_backward(B, F, d, dim, N), d, w, dim, N)
    assert rel_l2(F2, F) < 1e-13


def test_forward_of_matrix_fisher_has_unit_mass_and_mean():
    """F^0_00 is the mass; sum over the grid of R w fR equals the degree-1
    coefficients' content (E[R] from either side)."""
    B = 6
    f = density(B, 0, 1, 1.0, (6, 3, 1), seed=3)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F = O.fft_forward(B, f, d, w, 0, 1)
    lmax = B - 1
    assert abs(F[lmax, lmax, 0].real - 1.0) < 1e-13
    R = O.rotations(B)
    ER = np.einsum("ijabc,abc,b->ij", R, f, w)
    # the same expectation from the coefficients: E[R_ij] = sum f R_ij w = transform of R_ij . f at l = 0
    for i in range(3):
        for j in range(3):
            Fp = O.fft_forward(B, f * R[i, j], d, w, 0, 1)
            assert abs(Fp[lmax, lmax, 0].real - ER[i, j]) < 1e-13


def test_omega_axes_use_unshifted_numpy_fft():
    """The omega part of the transform is numpy's forward FFT, unshifted."""
    B, dim, N = 2, 2, 6
    rng = np.random.default_rng(4)
    f = rng.standard_normal((4, 4, 4, N, N))
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F = O.fft_forward(B, f, d, w, dim, N)
    F0 = O.fft_forward(B, np.fft.ifft2(np.fft.fft2(f, axes=(3, 4)), axes=(3, 4)).real, d, w, dim, N)
    G = np.fft.fft2(np.stack([O.fft_forward(B, f[..., a, b], d, w, 0, 1) for b in range(N) for a in range(N)],
                             axis=-1).reshape((3, 3, 2, N, N), order="F"), axes=(3, 4))
    assert rel_l2(F, G) < 1e-13 and rel_l2(F0, F) < 1e-13


def test_X_is_dft_of_coordinate_ramp():
    for dim, N, L in [(2, 8, 3.2), (3, 6, 10.0)]:
        X = O.build_X(dim, N, L)
        x = omega_grid(dim, N, L)
        for p in range(dim):
            assert rel_l2(X[..., p], np.fft.fftn(x[p])) < 1e-13


# ----------------------------------------------------------- derivative ----

def _F_of(B, dim, N, L, seed):
    f = density(B, dim, N, L, (5, 3, 1), seed=seed, Sigma=np.eye(dim) * 0.7)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    return O.fft_forward(B, f, d, w, dim, N)


def test_omega_terms_equal_pseudo_spectral_numpy():
    """Kinematic, damping and gyroscopic terms: the reference's dense circular
    convolutions (pendulum.m:400-466) equal the pseudo-spectral evaluation
    -sum_p [ifft(F) x_p]^ u_p^T - sum_p c(k_p) [ifft(F)(-b_p x_p + ojo_p)]^ with
    numpy's FFT and c = 2 pi i k / L (Nyquist 0)."""
    B, dim, N, L = 3, 3, 6, 4.0
    F = _F_of(B, dim, N, L, 5)
    Jt = np.diag([1.0, 0.8, 0.3])
    X = O.build_X(dim, N, L)
    OJO = O.build_OJO(N, L, Jt)
    MR0 = np.zeros((2 * B - 1, 2 * B - 1, B, dim), complex)
    u = O.getu(B - 1)
    b = np.array([0.3, 0.2, 0.1])
    dF = O.derivative(B, dim, N, F, X, OJO, MR0, b, np.zeros((3, 3)), L, u, O.cg_table(B))
    x = omega_grid(dim, N, L)
    ax = (3, 4, 5)
    fphys = np.fft.ifftn(F, axes=ax)
    k = np.fft.fftfreq(N, 1.0 / N)
    kc = 2 * np.pi * 1j * np.where(np.abs(k) == N // 2, 0, k) / L
    J = Jt
    v = x.reshape(3, -1)
    ojo = -np.linalg.inv(J) @ np.cross(v.T, (J @ v).T).T
    ojo = ojo.reshape((3,) + (N,) * 3)
    pred = np.zeros_like(F)
    lmax = B - 1
    for p in range(dim):
        Tp = np.fft.fftn(fphys * x[p], axes=ax)
        for l in range(B):
            s = slice(lmax - l, lmax + l + 1)
            pred[s, s, l] -= np.einsum("mj...,nj->mn...", Tp[s, s, l], u[s, s, l, p])
        drift = np.fft.fftn(fphys * (-b[p] * x[p] + ojo[p]), axes=ax)
        cp = kc.reshape([N if q == p else 1 for q in range(dim)])
        pred -= drift * cp
    assert rel_l2(dF, pred) < 1e-12


def test_omega_derivative_and_noise_scaling_on_a_gaussian():
    """c(k) and the diagonal noise factor act as d/dx and d^2/dx^2 on a smooth
    periodic-in-practice Gaussian (sign and 2 pi / L scaling)."""
    N, L = 64, 12.0
    xs = -L / 2 + L / N * np.arange(N)
    g = np.exp(-xs ** 2)
    G = np.fft.fft(g)
    k = np.fft.fftfreq(N, 1.0 / N)
    c = 2 * np.pi * 1j * np.where(k == -N // 2, 0, k) / L
    dg = np.fft.ifft(c * G).real
    assert np.abs(dg - (-2 * xs * g)).max() < 1e-12
    dd = -(2 * np.pi * k / L) ** 2
    d2 = np.fft.ifft(dd * G).real
    assert np.abs(d2 - (4 * xs ** 2 - 2) * g).max() < 1e-12


def test_derivative_conserves_mass_and_matches_gravity_product():
    B, dim, N, L = 4, 2, 6, 6.0
    F = _F_of(B, dim, N, L, 6)
    X = O.build_X(dim, N, L)
    MR = O.build_MR(B, dim)
    u = O.getu(B - 1)
    cells = O.cg_table(B)
    b = np.array([0.2, 0.3])
    G = np.array([[0.1, 0.02], [0.02, 0.2]])
    dF = O.derivative(B, dim, N, F, X, None, MR, b, G, L, u, cells)
    lmax = B - 1
    assert abs(dF[lmax, lmax, 0, 0, 0]) < 1e-14 * np.abs(dF).max()
    # gravity alone: -sum_p c(k_p) transform(mR_p f) for f of degree <= B-2
    rng = np.random.default_rng(8)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    Fb = O.fft_forward(B, O.fft_backward(B, _bandlimited(B, B - 2, rng, (N, N)), d, dim, N), d, w, dim, N)
    dG = O.derivative(B, dim, N, Fb, np.zeros_like(X), None, MR, np.zeros(2), np.zeros((2, 2)), L, u, cells)
    fb = O.fft_backward(B, Fb, d, dim, N)
    R = O.rotations(B)
    k = np.fft.fftfreq(N, 1.0 / N)
    kc = 2 * np.pi * 1j * np.where(np.abs(k) == N // 2, 0, k) / L
    pred = np.zeros_like(Fb)
    for p, mr in enumerate([R[2, 1], -R[2, 0]]):
        prod = O.fft_forward(B, fb * mr[..., None, None], d, w, dim, N)
        pred -= prod * kc.reshape([N if q == p else 1 for q in range(dim)])
    assert rel_l2(dG, pred) < 1e-12


# ---------------------------------------------------------- integrators ----

@pytest.mark.parametrize("dim,method,order", [(2, "euler", 1), (2, "midpoint", 2), (2, "RK2", 2),
                                              (2, "RK4", 4), (3, "RK4", 4)])
def test_integrators_equal_the_taylor_polynomial_of_the_drift(dim, method, order):
    """The drift F -> dF is linear and autonomous, so one step of an s-stage
    explicit Runge-Kutta method of order s <= 4 is exactly
    sum_{k<=s} (dt A)^k F / k!: the order conditions fix every coefficient of
    the stability polynomial, and a wrong Butcher weight or node leaves a
    dt^k term with the wrong coefficient.  A^k F comes from the oracle's
    derivative, which is pinned on its own."""
    B, N, L = 3, (6 if dim == 2 else 4), 6.0
    F = _F_of(B, dim, N, L, 7)
    X, OJO, MR, u, cells = operators(B, dim, N, L, np.diag([1.0, 0.8, 0.3]) if dim == 3 else None)
    b = np.array([0.3, 0.2, 0.1])[:dim]
    G = np.diag([0.1, 0.2, 0.15])[:dim, :dim] + 0.02 * (1 - np.eye(dim))
    dt = 0.05
    pred, term, last = F.copy(), F, None
    for k in range(1, order + 1):
        term = O.derivative(B, dim, N, term, X, OJO, MR, b, G, L, u, cells) * (dt / k)
        pred, last = pred + term, term
    Fn = O.integrate(B, dim, N, F, X, OJO, MR, b, G, dt, L, u, cells, method)
    assert rel_l2(Fn, pred) < 1e-13
    assert np.linalg.norm(last) > 1e-6 * np.linalg.norm(F)  # the highest term is resolved


def _den_case():
    B, N, L = 4, 6, 6.0
    F = _F_of(B, 2, N, L, 9)
    X, _, MR, u, cells = operators(B, 2, N, L)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    b, G = np.array([0.2, 0.3]), np.array([[0.1, 0.02], [0.02, 0.2]])
    A = lambda Y: O.derivative(B, 2, N, Y, X, None, MR, b, G, L, u, cells)
    D = lambda Y, f: O.derivative_den(B, N, Y, f, X, O.grid_mR(B), b, G, L, u, d, w)
    return B, N, L, F, X, u, d, w, b, G, A, D


def test_density_form_derivative_matches_the_spectral_derivative():
    """pendulum_reduced_den.m takes the gravity term from the grid density:
    with f the inverse transform of F it equals the spectral (Clebsch-Gordan)
    derivative exactly, because the grid product mR f has degree <= B and the
    forward transform projects it exactly; for any other grid f the derivative
    is that plus the gravity term of f - f_F alone, which for a grid function of
    degree <= B - 2 is the pinned spectral gravity term of its transform."""
    B, N, L, F, X, u, d, w, b, G, A, D = _den_case()
    fF = O.fft_backward(B, F, d, 2, N)
    aF = A(F)
    assert rel_l2(D(F, fF), aF) < 1e-12
    assert rel_l2(D(F, None), aF) < 1e-12
    rng = np.random.default_rng(23)
    g = O.fft_backward(B, O.fft_forward(B, O.fft_backward(B, _bandlimited(B, B - 2, rng, (N, N)), d, 2, N),
                                        d, w, 2, N), d, 2, N)
    grav = O.derivative(B, 2, N, O.fft_forward(B, g, d, w, 2, N), np.zeros_like(X), None, O.build_MR(B, 2),
                        np.zeros(2), np.zeros((2, 2)), L, u, O.cg_table(B))
    assert np.linalg.norm(grav) > 1e-3 * np.linalg.norm(aF)
    assert rel_l2(D(F, fF + g), aF + grav) < 1e-12


@pytest.mark.parametrize("method", ["euler", "midpoint", "RK2", "RK4"])
def test_density_form_integrators_use_the_grid_density_in_the_first_stage_only(method):
    """pendulum_reduced_den.m's integrators evaluate the first stage with the
    caller's grid density and every later stage from the stage coefficients.
    With k1 that first derivative, a = A F and A the (linear) spectral drift,
    composing the stages gives, for any f,
      euler     F + dt k1
      midpoint  F + dt a + dt^2/2 A k1
      RK2       F + dt/2 (k1 + a) + dt^2/2 A k1
      RK4       F + dt/6 (k1 + 5a) + dt^2/6 A (k1 + 2a) + dt^3/12 A^2 (k1 + a) + dt^4/24 A^3 k1
    which reduces to the Taylor polynomial of A when f is the inverse transform of F."""
    B, N, L, F, X, u, d, w, b, G, A, D = _den_case()
    dt = 0.05
    fF = O.fft_backward(B, F, d, 2, N)
    rng = np.random.default_rng(24)
    for f in (fF, fF + 0.3 * np.abs(fF).max() * rng.standard_normal(fF.shape)):
        k1, a = D(F, f), A(F)
        if method == "euler":
            pred = F + dt * k1
        elif method == "midpoint":
            pred = F + dt * a + dt ** 2 / 2 * A(k1)
        elif method == "RK2":
            pred = F + dt / 2 * (k1 + a) + dt ** 2 / 2 * A(k1)
        else:
            Ak1, Aa = A(k1), A(a)
            pred = (F + dt / 6 * (k1 + 5 * a) + dt ** 2 / 6 * (Ak1 + 2 * Aa)
                    + dt ** 3 / 12 * A(Ak1 + Aa) + dt ** 4 / 24 * A(A(Ak1)))
        out = O.integrate_den(B, N, F, f, X, O.grid_mR(B), b, G, dt, L, u, d, w, method)
        assert rel_l2(out, pred) < 1e-13


# ---------------------------------------------------------------- reset ----

def _reset_setup():
    B, N = 3, 10
    L = 3.4
    R = O.rotations(B)
    x = omega_grid(2, N, L)
    return B, N, L, R, x


def test_reset_map_involution_and_restitution():
    """old(new(omega)) = omega; the tangential component reverses with -eps
    (collision.tex:162-180, test.m:45-53)."""
    B, N, L, R, x = _reset_setup()
    lam, indR, PC = O.getLambda(B, N, R, x, 0.12, 0.2, 0.025, 5 * np.pi / 180, 100.0)
    assert len(indR) > 0
    eps = 0.8
    Om_new = O.getOmega(B, N, R, x, indR, eps, PC, "new")
    Rf = R.reshape(3, 3, -1, order="F")
    for k in range(0, len(indR), max(1, len(indR) // 7)):
        Rr = Rf[:, :, indR[k]]
        t = np.cross(Rr[:, 2], [1.0, 0, 0])
        t /= np.linalg.norm(t)
        for ix in range(0, N * N, 7):
            xv = x.reshape(2, -1, order="F")[:, ix]
            om = Rr @ np.array([xv[0], xv[1], 0.0])
            on = Rr @ np.array([*Om_new[:, k].reshape(2, -1, order="F")[:, ix], 0.0])
            # R^T omega+ keeps a zero third body component only up to the 2-D projection; compare in space
            om_plus = om - (1 + eps) * np.dot(om, t) * t
            assert abs(np.dot(om_plus, t) + eps * np.dot(om, t)) < 1e-12
            back = om_plus - (1 + eps) / eps * np.dot(om_plus, t) * t
            assert np.abs(back - om).max() < 1e-12
            assert np.abs((Rr.T @ om_plus)[:2] - (Rr.T @ on)[:2]).max() < 1e-12


def test_noisy_reset_conserves_mass_per_rotation():
    """Each source's outflow is renormalised over the destinations that keep it
    in their top-nD list (getFcL.cu:94-101), so per rotation the jump term
    integrates to minus the outflow of sources no destination kept."""
    B, N, L, R, x = _reset_setup()
    lam, indR, PC = O.getLambda(B, N, R, x, 0.12, 0.2, 0.025, 5 * np.pi / 180, 100.0)
    Om = O.getOmega(B, N, R, x, indR, 0.8, PC, "new")
    Gd = np.eye(2) * 0.05 ** 2 * 40
    fcL, idx = O.getFcL(N, x, Om, lam, Gd, 10)
    f = density(B, 2, N, L, (8, 4, 2), seed=11, Sigma=np.eye(2) * 0.5)
    df = O.discrete_noise(B, N, f, lam, indR, fcL, idx)
    fr = f.reshape(8 * B ** 3, N * N, order="F")
    s = df.reshape(8 * B ** 3, N * N, order="F").sum(axis=1)
    lam2 = lam.reshape(len(indR), N * N, order="F")
    idx2 = idx.reshape(10, len(indR), N * N, order="F")
    for k, iR in enumerate(indR):
        kept = np.zeros(N * N, bool)
        kept[np.unique(idx2[:, k, :])] = True
        lost = -np.sum(lam2[k, ~kept] * fr[iR, ~kept])
        assert abs(s[iR] - lost) < 1e-10 * max(1.0, np.abs(fr[iR]).max() * 100), (k, s[iR], lost)
    mask = np.ones(8 * B ** 3, bool)
    mask[indR] = False
    assert np.all(s[mask] == 0) and np.abs(df).max() > 0


def test_index_rule_is_bilinear_interpolation():
    """getIndRule's rule (getIndRule.cu:50-90) reproduces every bilinear
    function a + b x1 + c x2 + d x1 x2 at the query point, its four indices are
    the corners (x1, x2), (x1 + dx, x2), (x1, x2 + dx), (x1 + dx, x2 + dx) of the
    grid cell holding the query, its weights sum to one, and a NaN query gives
    the NaN rule."""
    _, N, L, _, x = _reset_setup()
    rng = np.random.default_rng(21)
    xl = np.asfortranarray(x).reshape(2, -1, order="F")
    dx = L / N
    numR = 4
    Om = np.asfortranarray(rng.uniform(xl.min(), xl.max() - 1e-9, size=(2, numR, N, N)))
    Om[:, 1, 2, 3] = np.nan
    ind, coeff = O.getIndRule(N, x, Om)
    Omf = Om.reshape(2, numR, -1, order="F")
    indf = ind.reshape(4, numR, -1, order="F")
    cf = coeff.reshape(4, numR, -1, order="F")
    g = lambda p: 0.3 + 1.1 * p[0] - 0.7 * p[1] + 0.45 * p[0] * p[1]
    for k in range(numR):
        for ix in range(N * N):
            q = Omf[:, k, ix]
            if np.isnan(q[0]):
                assert np.all(np.isnan(cf[:, k, ix]))
                continue
            pts = xl[:, indf[:, k, ix]]
            assert np.allclose(pts[:, 1] - pts[:, 0], [dx, 0], atol=1e-12)
            assert np.allclose(pts[:, 2] - pts[:, 0], [0, dx], atol=1e-12)
            assert np.allclose(pts[:, 3] - pts[:, 0], [dx, dx], atol=1e-12)
            assert np.all(pts[:, 0] <= q + 1e-12) and np.all(q <= pts[:, 3] + 1e-12)
            assert abs(cf[:, k, ix].sum() - 1) < 1e-13
            assert abs(np.dot(cf[:, k, ix], g(pts)) - g(q)) < 1e-12


def test_transition_kernel_keeps_the_densest_sources_and_conserves_outflow():
    """For each destination getFcL (getFcL.cu:11-166) keeps the nD sources of
    largest Gaussian density N(x_dest; Omega_src, Gd), evaluated here with
    scipy.stats, in descending order; each kept weight is that density over the
    sum of the source's kept densities, times lambda(source), so the kept
    weights of every source sum to lambda(source)."""
    from scipy.stats import multivariate_normal
    _, N, L, _, x = _reset_setup()
    rng = np.random.default_rng(22)
    xl = np.asfortranarray(x).reshape(2, -1, order="F")
    nx, numR, nD = N * N, 3, 8
    Om = np.asfortranarray(xl[:, None, :] * 0.9 + rng.uniform(-0.1, 0.1, size=(2, numR, nx))).reshape(
        2, numR, N, N, order="F")
    lam = np.asfortranarray(rng.uniform(1.0, 3.0, size=(numR, nx)))
    Gd = np.array([[0.09, 0.02], [0.02, 0.05]])
    fcL, idx = O.getFcL(N, x, Om, lam, Gd, nD)
    fcL = fcL.reshape(nD, numR, nx, order="F")
    idx = idx.reshape(nD, numR, nx, order="F").astype(np.int64)
    Omf = Om.reshape(2, numR, nx, order="F")
    for k in range(numR):
        dens = np.array([[multivariate_normal.pdf(xl[:, d], mean=Omf[:, k, s], cov=Gd) for s in range(nx)]
                         for d in range(nx)])  # (destination, source)
        ratio = {}
        for d in range(nx):
            top = np.argsort(-dens[d], kind="stable")[:nD]
            assert np.array_equal(idx[:, k, d], top)
            for t, s in enumerate(top):
                ratio.setdefault(s, []).append(fcL[t, k, d] / (lam[k, s] * dens[d, s]))
        for s, r in ratio.items():
            r = np.array(r)
            assert np.abs(r / r[0] - 1).max() < 1e-12  # one normalizer per source
            kept = fcL[:, k, :][idx[:, k, :] == s].sum()
            assert abs(kept - lam[k, s]) < 1e-12 * lam[k, s]
