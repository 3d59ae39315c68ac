/*
 * refcpu_reset3.c -- the hybrid collision reset on a 3-D angular-velocity
 * grid (FP64).  TEST INFRASTRUCTURE ONLY: the oracle of the 3-D reset; only
 * tests/ load it.
 *
 * The reference builds the reset for the reduced model only (2-D omega,
 * propagation/discrete/mex/sources/pendulum_reduced_collision).  Its
 * derivation (notes/10 collision/collision.tex:100-190) holds for a full
 * body angular velocity Omega in R^3; this restates the reference's reset
 * with Omega in R^3 and the wall normal e_w = (cos tau, 0, sin tau) (tilt
 * tau about e2; tau = 0 is the reference's wall x1 = d):
 *   theta = asin(r3 . e_w)                       (getLambda.cu:192-206, r3 = R e3)
 *   lambda(R) = 0 below theta0 - theta_t, lambda_max above theta0 + theta_t,
 *               lambda_max/2 sin(pi/(2 theta_t)(theta - theta0)) + lambda_max/2
 *               between                           (getLambda.cu:218-232)
 *   PC = (h - r tan theta) r3 + r sec theta e_w   (getLambda.cu:208-216)
 *   t = r3 x e_w / |r3 x e_w|, omega = R Omega    (getOmega.cu:64-97)
 *   omega- = omega - ((1+eps)/eps)(omega.t) t,  Omega- = R^T omega-
 *                                                  (getOmega.cu:99-137, the 'old' map g)
 *   vC(Omega) = (R Omega x PC) . e_w              (getLambda.cu:234-266)
 *             = sigma (omega . t),  sigma = (PC x e_w) . t = (h - r tan theta) |r3 x e_w|
 *   vC(Omega-) = -vC(Omega)/eps (the law of restitution), so both contact tests
 *   are decided by the sign of sigma (omega . t): outflow where it is > 0,
 *   inflow where it is < 0.  On the switching plane omega . t = 0 the map is
 *   the identity and nothing moves (neither term: with J = 1/eps, counting
 *   both would create mass (1/eps - 1) lambda f on the lattice points of that
 *   plane).  Evaluated this way the tests do not depend on eps, even where
 *   omega . t is at rounding level
 *   df(R, Omega) = J lambda(R) f(R, Omega-) [vC(Omega-) > 0, Omega- in the box]
 *                 - lambda(R) f(R, Omega) [vC(Omega) > 0]
 *                                                  (propagate.cu:123-229, collision.tex:160-190)
 * with f(R, Omega-) by trilinear interpolation of the grid values extended by
 * zero outside the omega grid (the 3-D form of getIndRule.cu:50-90's bilinear
 * rule; where the reference drops a pull-back outside [x_0, x_{N-1}] whole,
 * the zero extension lets the inflow fade to zero over the outermost cell,
 * so it stays continuous in eps) and
 * J = 1/eps (|det| of the map g, which conserves probability) or J = 1 (the
 * reference's reduced law omits it).
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* per-rotation constants: lambda (0 outside the region), PC (3), t (3) */
void ref_reset3_tables(int B, const double* R, double dwall, double h, double r, double theta_t, double lambda_max,
                       double tilt, double* lam, double* PC, double* t)
{
    long nR = 8L * B * B * B;
    double ew[3] = {cos(tilt), 0.0, sin(tilt)};
    double theta0 = asin(dwall / sqrt(h * h + r * r)) - asin(r / sqrt(h * h + r * r));
    for (long i = 0; i < nR; i++) {
        const double* Rr = R + 9 * i;
        double r3[3] = {Rr[6], Rr[7], Rr[8]};
        double th = asin(r3[0] * ew[0] + r3[1] * ew[1] + r3[2] * ew[2]);
        double l;
        if (th < theta0 - theta_t) l = 0.0;
        else if (th > theta0 + theta_t) l = lambda_max;
        else l = lambda_max / 2 * sin(M_PI / (2 * theta_t) * (th - theta0)) + lambda_max / 2;
        lam[i] = l;
        double lr3 = h - r * tan(th);
        for (int a = 0; a < 3; a++) PC[3 * i + a] = lr3 * r3[a] + r / cos(th) * ew[a];
        double c[3] = {r3[1] * ew[2] - r3[2] * ew[1], r3[2] * ew[0] - r3[0] * ew[2], r3[0] * ew[1] - r3[1] * ew[0]};
        double nc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int a = 0; a < 3; a++) t[3 * i + a] = nc > 0 ? c[a] / nc : 0.0;
    }
}

/* sigma = (PC x e_w) . t: vC(omega) = sigma (omega . t) */
static double contact_sigma(const double* P, const double* ew, const double* t)
{
    double c[3] = {P[1] * ew[2] - P[2] * ew[1], P[2] * ew[0] - P[0] * ew[2], P[0] * ew[1] - P[1] * ew[0]};
    return c[0] * t[0] + c[1] * t[1] + c[2] * t[2];
}

/* Omega- = g(R, Omega) for one rotation (body frame, 3 components) */
void ref_reset3_pullback(const double* Rr, const double* t, double epsilon, const double* Om, double* Om_old)
{
    double w[3];
    for (int a = 0; a < 3; a++) w[a] = Rr[a] * Om[0] + Rr[a + 3] * Om[1] + Rr[a + 6] * Om[2];
    double fac = (1.0 + epsilon) / epsilon * (w[0] * t[0] + w[1] * t[1] + w[2] * t[2]);
    for (int a = 0; a < 3; a++) w[a] -= fac * t[a];
    for (int a = 0; a < 3; a++) Om_old[a] = Rr[3 * a] * w[0] + Rr[3 * a + 1] * w[1] + Rr[3 * a + 2] * w[2];
}

/* df on the grid f (nR, N^3), rotation index fastest. */
void ref_reset3_df(int B, int N, double L, const double* R, const double* f, const double* lam, const double* PC,
                   const double* t, double epsilon, double tilt, int jacobian, double* df)
{
    long nR = 8L * B * B * B, nx = (long)N * N * N;
    double ew[3] = {cos(tilt), 0.0, sin(tilt)};
    double hx = L / N, x0 = -L / 2;
    double J = jacobian ? 1.0 / epsilon : 1.0;
    memset(df, 0, sizeof(double) * nR * nx);
#pragma omp parallel for schedule(static)
    for (long iR = 0; iR < nR; iR++) {
        if (lam[iR] == 0.0) continue;
        const double* Rr = R + 9 * iR;
        const double* P = PC + 3 * iR;
        const double* tt = t + 3 * iR;
        const double sigma = contact_sigma(P, ew, tt);
        for (long ix = 0; ix < nx; ix++) {
            long i[3] = {ix % N, (ix / N) % N, ix / ((long)N * N)};
            double Om[3], w[3], Oo[3];
            for (int a = 0; a < 3; a++) Om[a] = x0 + hx * i[a];
            for (int a = 0; a < 3; a++) w[a] = Rr[a] * Om[0] + Rr[a + 3] * Om[1] + Rr[a + 6] * Om[2];
            const double vq = sigma * (w[0] * tt[0] + w[1] * tt[1] + w[2] * tt[2]);
            double out = vq > 0.0 ? -lam[iR] * f[iR + nR * ix] : 0.0;
            ref_reset3_pullback(Rr, tt, epsilon, Om, Oo);
            long c[3];
            double s[3];
            int support = 1;
            for (int a = 0; a < 3; a++) {
                c[a] = (long)floor((Oo[a] - x0) / hx);  /* cell [x_c, x_{c+1}], c in [-1, N-1] */
                if (!(c[a] >= -1 && c[a] <= N - 1)) support = 0;
                s[a] = (Oo[a] - (x0 + hx * c[a])) / hx;
            }
            if (support && vq < 0.0) {
                double v = 0.0;
                for (int q = 0; q < 8; q++) {
                    int b0 = q & 1, b1 = (q >> 1) & 1, b2 = (q >> 2) & 1;
                    long j0 = c[0] + b0, j1 = c[1] + b1, j2 = c[2] + b2;
                    if (j0 < 0 || j0 >= N || j1 < 0 || j1 >= N || j2 < 0 || j2 >= N) continue;  /* zero outside */
                    double wt = (b0 ? s[0] : 1 - s[0]) * (b1 ? s[1] : 1 - s[1]) * (b2 ? s[2] : 1 - s[2]);
                    v += wt * f[iR + nR * (j0 + N * (j1 + (long)N * j2))];
                }
                out += J * lam[iR] * v;
            }
            df[iR + nR * ix] = out;
        }
    }
}
