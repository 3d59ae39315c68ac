"""More pins of the FP64 oracle, each against a fact fixed independently of
the oracle's own formula:

* the structured derivative (refcpu_structured.c) against the dense,
  literal one (refcpu.c), every term and every integrator, including a
  non-diagonal inertia and non-diagonal noise;
* the 3-D gravity term with Jt^{-1} against the transform of the grid
  product (the product rule, pendulum.m:166-193,470-506);
* getLambda's contact geometry (getLambda.cu:192-266): the contact point is
  the point of the end disc (radius r about h r3, normal r3) that reaches
  furthest toward the wall, it touches the wall plane x1 = d at theta0, the
  rate ramp is continuous, monotone and odd about theta0 between 0 and
  lambda_max, and the contact-velocity mask is the sign of (omega x PC).e1
  computed with numpy's cross product;
* the no-noise reset (propagate.cu:123-229) is exact on densities that are
  bilinear in omega on every rotation (bilinear interpolation reproduces them);
* the noisy reset (propagate.cu:6-121) applied to a point mass gives exactly
  the kept weights at exactly the destinations that list that source.
"""
import math

import numpy as np
import pytest

import oracle as O
from helpers import density, omega_grid, operators, rel_l2


def _case(B, dim, N, L, Jt, seed=3):
    f = density(B, dim, N, L, (8, 4, 2), seed=seed, Sigma=np.eye(dim) * 0.9)
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F = O.fft_forward(B, f, d, w, dim, N)
    X, OJO, MR, u, cells = operators(B, dim, N, L, Jt)
    bt = np.array([0.2, 0.3, 0.25])[:dim]
    G = np.array([[0.1, 0.02, 0.01], [0.02, 0.2, 0.03], [0.01, 0.03, 0.3]])[:dim, :dim]
    return F, X, OJO, MR, u, cells, bt, G


JT_DIAG = np.diag([1.0, 0.93539, 0.24935])
JT_FULL = np.array([[1.0, 0.1, 0.05], [0.1, 0.9, 0.02], [0.05, 0.02, 0.3]])


@pytest.mark.parametrize("B,dim,N,Jt", [(3, 2, 6, None), (4, 3, 6, JT_DIAG), (3, 3, 4, JT_FULL), (2, 3, 5, JT_DIAG)])
def test_structured_derivative_equals_dense(B, dim, N, Jt):
    """Each term on its own and all together, then all four integrators."""
    L = 8.0
    F, X, OJO, MR, u, cells, bt, G = _case(B, dim, N, L, Jt)
    z = np.zeros_like
    cases = {"all": (X, OJO, Jt, MR, bt, G), "kinematic": (X, None, None, z(MR), z(bt), z(G)),
             "damping": (z(X), None, None, z(MR), bt, z(G)), "gravity": (z(X), None, None, MR, z(bt), z(G)),
             "noise": (z(X), None, None, z(MR), z(bt), G)}
    if dim == 3:
        cases["gyro"] = (X, OJO, Jt, z(MR), z(bt), z(G))
    for name, (Xc, Oc, Jc, Mc, bc, Gc) in cases.items():
        if name == "damping":
            Xc = X  # damping is a circulant of X; the kinematic u term is then also on
        ref = O.derivative(B, dim, N, F, Xc, Oc if Oc is not None else (z(OJO) if dim == 3 else None), Mc, bc, Gc,
                           L, u, cells)
        out = O.derivative_structured(B, dim, N, F, Xc, Jc, Mc, bc, Gc, L, u, cells)
        assert np.isfinite(out).all()
        assert rel_l2(out, ref) < 1e-13, name
    for method in ("euler", "midpoint", "RK2", "RK4"):
        ref = O.integrate(B, dim, N, F, X, OJO, MR, bt, G, 0.01, L, u, cells, method)
        out = O.integrate_structured(B, dim, N, F, X, Jt, MR, bt, G, 0.01, L, u, cells, method)
        assert rel_l2(out, ref) < 1e-14, method


def test_structured_derivative_rejects_unstructured_operators():
    B, dim, N, L = 3, 2, 4, 6.0
    F, X, OJO, MR, u, cells, bt, G = _case(B, dim, N, L, None)
    Xb = X.copy()
    Xb[1, 1, 0] += 0.5  # off the k1 axis line
    with pytest.raises(ValueError):
        O.derivative_structured(B, dim, N, F, Xb, None, MR, bt, G, L, u, cells)
    Mb = MR.copy()
    Mb[B - 1, B - 1, 2, 0] = 0.1  # content on degree 2
    with pytest.raises(ValueError):
        O.derivative_structured(B, dim, N, F, X, None, Mb, bt, G, L, u, cells)


def test_gravity_term_3d_is_the_transform_of_the_moment_with_inverse_inertia():
    """For F of degree <= B-2, the 3-D gravity term alone equals
    -sum_p c(k_p) fftSO3R(mR_p f) with mR = Jt^{-1} [R32, -R31, 0] formed here
    from the rotation grid and numpy's inverse (product rule, product.tex:25-63)."""
    B, dim, N, L = 4, 3, 4, 6.0
    lmax, M = B - 1, 2 * B - 1
    rng = np.random.default_rng(31)
    Fb = rng.standard_normal((M, M, B, N, N, N)) + 1j * rng.standard_normal((M, M, B, N, N, N))
    for l in range(B):
        for m in range(-lmax, lmax + 1):
            for n in range(-lmax, lmax + 1):
                if abs(m) > l or abs(n) > l or l > B - 2:
                    Fb[m + lmax, n + lmax, l] = 0
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    fb = O.fft_backward(B, Fb, d, dim, N)
    Fb = O.fft_forward(B, fb, d, w, dim, N)
    MR = O.build_MR(B, 3, JT_FULL)
    X = O.build_X(3, N, L)
    u = O.getu(lmax)
    cells = O.cg_table(B)
    dG = O.derivative(B, dim, N, Fb, np.zeros_like(X), np.zeros_like(X), MR, np.zeros(3), np.zeros((3, 3)), L, u,
                      cells)
    R = O.rotations(B)
    Ji = np.linalg.inv(JT_FULL)
    mom = np.stack([R[2, 1], -R[2, 0], np.zeros_like(R[2, 0])])
    k = np.fft.fftfreq(N, 1.0 / N)
    kc = 2 * np.pi * 1j * np.where(np.abs(k) == N // 2, 0, k) / L
    pred = np.zeros_like(Fb)
    for p in range(3):
        mr = np.tensordot(Ji[p], mom, axes=1)
        prod = O.fft_forward(B, fb * mr[..., None, None, None], d, w, dim, N)
        pred -= prod * kc.reshape([N if q == p else 1 for q in range(dim)])
    assert rel_l2(dG, pred) < 1e-12


# ------------------------------------------------------------- getLambda ----

def _rot(psi, theta, phi):
    c, s = math.cos, math.sin
    Rx = np.array([[1, 0, 0], [0, c(psi), -s(psi)], [0, s(psi), c(psi)]])
    Ry = np.array([[c(theta), 0, s(theta)], [0, 1, 0], [-s(theta), 0, c(theta)]])
    Rz = np.array([[c(phi), -s(phi), 0], [s(phi), c(phi), 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


WALL = dict(dwall=0.12, h=0.2, r=0.025, theta_t=5 * np.pi / 180, lambda_max=100.0)


def test_getLambda_contact_point_geometry():
    """PC - h r3 is orthogonal to r3 with length r (a point of the end disc),
    PC.e1 is the largest e1 component on that circle (sampled densely with
    numpy), and at theta = theta0 the contact point lies on the wall x1 = d."""
    B, N, L = 3, 6, 3.4
    R = O.rotations(B)
    x = omega_grid(2, N, L)
    h, r, dwall = WALL["h"], WALL["r"], WALL["dwall"]
    _, _, PC = O.getLambda(B, N, R, x, **WALL)
    Rf = R.reshape(3, 3, -1, order="F")
    ang = np.linspace(0, 2 * np.pi, 20001)
    for i in range(0, Rf.shape[2], 5):
        r3 = Rf[:, 2, i]
        if abs(r3[0]) > 0.99:
            continue
        v = PC[:, i] - h * r3
        assert abs(np.dot(v, r3)) < 1e-14
        assert abs(np.linalg.norm(v) - r) < 1e-14
        a = np.cross(r3, [0.0, 0.0, 1.0]) if abs(r3[2]) < 0.9 else np.cross(r3, [0.0, 1.0, 0.0])
        a /= np.linalg.norm(a)
        b = np.cross(r3, a)
        circle = h * r3[:, None] + r * (np.cos(ang)[None] * a[:, None] + np.sin(ang)[None] * b[:, None])
        assert abs(circle[0].max() - PC[0, i]) < 1e-9
    # a rotation at theta0 touches the wall
    theta0 = math.asin(dwall / math.hypot(h, r)) - math.asin(r / math.hypot(h, r))
    Rs = np.zeros((3, 3, 8))
    for j in range(8):
        Rs[:, :, j] = _rot(0.3 * j, theta0, 0.7 * j)
    Rs = np.asfortranarray(Rs.reshape(3, 3, 2, 2, 2, order="F"))
    _, _, PCs = O.getLambda(1, N, Rs, x, **WALL)
    assert np.abs(PCs[0] - dwall).max() < 1e-14


def test_getLambda_rate_ramp_and_contact_mask():
    """On rotations with theta = asin(R13) swept across the ramp: no region
    membership below theta0 - theta_t, lambda_max above theta0 + theta_t,
    and in between a continuous, nondecreasing ramp from 0 to lambda_max,
    odd about (theta0, lambda_max/2) with zero slope at both ends.  lambda is
    zeroed exactly where (omega x PC).e1 < 0, omega = R [x; 0], by numpy."""
    h, r, dwall, tt, lmax = WALL["h"], WALL["r"], WALL["dwall"], WALL["theta_t"], WALL["lambda_max"]
    theta0 = math.asin(dwall / math.hypot(h, r)) - math.asin(r / math.hypot(h, r))
    B, N, L = 4, 6, 3.4  # 512 synthetic rotations
    n = 8 * B ** 3
    thetas = theta0 + np.linspace(-2 * tt, 2 * tt, n)
    rng = np.random.default_rng(41)
    Rs = np.zeros((3, 3, n))
    for i, th in enumerate(thetas):
        Rs[:, :, i] = _rot(rng.uniform(-np.pi, np.pi), th, rng.uniform(-np.pi, np.pi))
    R = np.asfortranarray(Rs.reshape(3, 3, 2 * B, 2 * B, 2 * B, order="F"))
    x = omega_grid(2, N, L)
    lam, indR, PC = O.getLambda(B, N, R, x, **WALL)
    lam_c, indR_c, lists, _ = O.getLambda_compact(B, N, R, x, **WALL)
    assert np.array_equal(indR, indR_c)
    th_in = thetas[indR]
    assert np.all(thetas[np.setdiff1d(np.arange(n), indR)] < theta0 - tt)
    assert np.all(th_in >= theta0 - tt)
    # the per-rotation rate is the compact lambda; the full one is it or 0
    above = th_in > theta0 + tt
    assert np.all(lam_c[above] == lmax)
    ramp = ~above
    tr, lr = th_in[ramp], lam_c[ramp]
    assert np.all(np.diff(lr) >= -1e-12) and lr.min() >= 0 and lr.max() <= lmax
    # continuity and zero end slopes: the increments are bounded by the
    # ramp's largest slope lambda_max pi / (4 theta_t) and vanish at the ends
    dth = np.diff(tr)
    slope = np.diff(lr) / dth
    assert slope.max() <= lmax * np.pi / (4 * tt) * (1 + 1e-9)
    assert lr[0] < 1e-3 * lmax and lmax - lr[-1] < 1e-3 * lmax
    # odd about theta0: lambda(theta0 + s) + lambda(theta0 - s) = lambda_max
    # (interpolated on the sampled ramp at symmetric points)
    s = np.linspace(0, 0.95 * tt, 40)
    up = np.interp(theta0 + s, tr, lr)
    dn = np.interp(theta0 - s, tr, lr)
    assert np.abs(up + dn - lmax).max() < 1e-3 * lmax
    assert abs(np.interp(theta0, tr, lr) - lmax / 2) < 1e-3 * lmax
    # contact mask
    Rf = R.reshape(3, 3, -1, order="F")
    xl = np.asfortranarray(x).reshape(2, -1, order="F")
    lam2 = lam.reshape(len(indR), -1, order="F")
    for k, iR in enumerate(indR):
        om = Rf[:, :, iR] @ np.vstack([xl, np.zeros(N * N)])
        vC = np.cross(om.T, PC[:, iR]).T[0]
        assert np.array_equal(lam2[k] == 0, vC < 0) or np.all(lam_c[k] == 0)
        assert np.all(lam2[k][vC >= 0] == lam_c[k])
        assert np.array_equal(lists[k], np.nonzero(vC >= 0)[0])


# ---------------------------------------------------------------- resets ----

def test_nonoise_reset_is_exact_for_bilinear_densities():
    """f(R, x) = a_R + b_R x1 + c_R x2 + d_R x1 x2: the no-noise inflow at a
    region rotation is lambda(R) f(R, Omega_old(x)) exactly (bilinear
    interpolation reproduces f), minus lambda f on the listed indices; NaN
    pullbacks give no inflow; rotations outside the region are untouched."""
    B, N, L = 3, 10, 3.4
    R = O.rotations(B)
    x = omega_grid(2, N, L)
    lam, indR, lists, PC = O.getLambda_compact(B, N, R, x, **WALL)
    Om = O.getOmega(B, N, R, x, indR, 0.8, PC, "old")
    ind, coeff = O.getIndRule(N, x, Om)
    nR, nx = 8 * B ** 3, N * N
    rng = np.random.default_rng(51)
    a, b, c, dd = (rng.uniform(0.5, 1.5, nR) for _ in range(4))
    xl = np.asfortranarray(x).reshape(2, -1, order="F")
    fb = lambda q1, q2: a[:, None] + b[:, None] * q1 + c[:, None] * q2 + dd[:, None] * q1 * q2
    f = np.asfortranarray(fb(xl[0][None], xl[1][None]).reshape((2 * B,) * 3 + (N, N), order="F"))
    df = O.discrete_nonoise(B, N, f, lam, indR, lists, ind, coeff)
    dfr = df.reshape(nR, nx, order="F")
    fr = f.reshape(nR, nx, order="F")
    Omf = Om.reshape(2, len(indR), nx, order="F")
    checked = 0
    for k, iR in enumerate(indR):
        q = Omf[:, k, :]
        ok = ~np.isnan(q[0])
        pred = np.zeros(nx)
        pred[ok] = lam[k] * (a[iR] + b[iR] * q[0, ok] + c[iR] * q[1, ok] + dd[iR] * q[0, ok] * q[1, ok])
        pred[lists[k]] -= lam[k] * fr[iR, lists[k]]
        assert np.abs(dfr[iR] - pred).max() < 1e-12 * max(1.0, np.abs(pred).max())
        checked += int(ok.sum())
    assert checked > 0
    out = np.ones(nR, bool)
    out[indR] = False
    assert np.all(dfr[out] == 0)


def test_noisy_reset_moves_a_point_mass_to_exactly_the_listed_destinations():
    """With f = 1 at one (rotation, source) grid point and 0 elsewhere, the
    noisy inflow at every destination is the weight of that source in the
    destination's kept list (0 where it is not kept), and the outflow is
    -lambda at the source itself: a wrong gather index or direction fails."""
    B, N, L = 3, 8, 3.4
    R = O.rotations(B)
    x = omega_grid(2, N, L)
    lam, indR, PC = O.getLambda(B, N, R, x, **WALL)
    Om = O.getOmega(B, N, R, x, indR, 0.8, PC, "new")
    nD = 6
    fcL, idx = O.getFcL(N, x, Om, lam, np.eye(2) * 0.2 ** 2, nD)
    nR, nx, numR = 8 * B ** 3, N * N, len(indR)
    fcLr = fcL.reshape(nD, numR, nx, order="F")
    idxr = idx.reshape(nD, numR, nx, order="F").astype(np.int64)
    lamr = lam.reshape(numR, nx, order="F")
    for k in (0, numR // 2, numR - 1):
        iR = indR[k]
        for src in (3, nx // 2 + 1):
            f = np.zeros(nR * nx)
            f[iR + nR * src] = 1.0
            f = np.asfortranarray(f.reshape((2 * B,) * 3 + (N, N), order="F"))
            df = O.discrete_noise(B, N, f, lam, indR, fcL, idx).reshape(nR, nx, order="F")
            pred = np.zeros(nx)
            for dest in range(nx):
                hit = np.nonzero(idxr[:, k, dest] == src)[0]
                if hit.size:
                    pred[dest] = fcLr[hit[0], k, dest]
            pred[src] -= lamr[k, src]
            assert np.array_equal(np.nonzero(df[iR])[0], np.nonzero(pred)[0])
            assert np.abs(df[iR] - pred).max() < 1e-15 * max(1.0, np.abs(pred).max())
            other = np.ones(nR, bool)
            other[iR] = False
            assert np.all(df[other] == 0)
