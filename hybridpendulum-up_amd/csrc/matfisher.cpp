// Matrix-Fisher statistics on the host (get_stat's fit: pendulum.m:547-594,
// psvd.m:10-13, pdf_MF_M2S.m:22-102, pdf_MF_normal.m:35-75).
//
// Design: the reference evaluates the scaled normalizer c(s) and its three
// first derivatives as separate adaptive integrals (one per cyclic relabelling
// of s) and obtains the Hessian from a 9x9 linear system with a case split on
// repeated singular values.  Here all ten quantities c, dc/ds_i,
// d2c/ds_i ds_j come from ONE adaptive Gauss-Kronrod pass over a vector
// integrand: the Kunze representation differentiated under the integral
// sign, using I0' = I1 and I1' = I0 - I1/x.  No case split is needed, and the
// fit minimises the convex log c(s) - d.s by Newton's method (the Hessian of
// log c is the covariance of diag(Q), positive definite), which has the same
// root as the reference's Newton-Armijo iteration on grad log c - d.
#include "matfisher.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace hpup {
namespace mf {

static const double kPi = 3.14159265358979323846;

// ------------------------------------------------------------- Bessel I --
// Power series for |x| <= 25 (all terms positive, no cancellation); the
// Hankel asymptotic expansion beyond, truncated where its terms fall below
// 1e-17 (their minimum there is about e^{-2|x|}).

double i0e(double x) {
    x = std::fabs(x);
    if (x <= 25.0) {
        const double q = 0.25 * x * x;
        double term = 1.0, sum = 1.0;
        for (int k = 1; k < 400; k++) {
            term *= q / ((double)k * k);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        return sum * std::exp(-x);
    }
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 100; k++) {
        const double r = (2.0 * k - 1) * (2.0 * k - 1) / (8.0 * k * x);
        if (r >= 1.0) break;
        term *= r;
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum / std::sqrt(2.0 * kPi * x);
}

double i1e(double x) {
    const double ax = std::fabs(x), sg = x < 0 ? -1.0 : 1.0;
    if (ax <= 25.0) {
        const double q = 0.25 * ax * ax;
        double term = 0.5 * ax, sum = term;
        for (int k = 1; k < 400; k++) {
            term *= q / ((double)k * (k + 1));
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        return sg * sum * std::exp(-ax);
    }
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 100; k++) {
        // term ratio ((2k-1)^2 - 4 nu^2) / (8 k x), nu = 1
        const double r = ((2.0 * k - 1) * (2.0 * k - 1) - 4.0) / (8.0 * k * ax);
        if (std::fabs(r) >= 1.0) break;
        term *= r;
        sum += term;
        if (std::fabs(term) < 1e-17 * std::fabs(sum)) break;
    }
    return sg * sum / std::sqrt(2.0 * kPi * ax);
}

// e^{-|x|} I0''(x) = e^{-|x|} (I0(x) - I1(x)/x), 1/2 at x = 0
static double i0ppe(double x, double a0, double a1) {
    if (std::fabs(x) < 1e-8) return 0.5 * a0;
    return a0 - a1 / x;
}

// ------------------------------------------------------ the normalizer --

namespace {

constexpr int kNV = 10;  // C, D1..D3, H11, H22, H33, H12, H13, H23
using Vec = double[kNV];

struct Kunze {
    double s1, s2, s3, K;
    void eval(double u, double* out) const {
        const double a = 0.5 * (s1 - s2) * (1 - u), b = 0.5 * (s1 + s2) * (1 + u);
        const double e = 0.5 * std::exp(std::fabs(a) + std::fabs(b) + s3 * u - K);
        const double A0 = i0e(a), A1 = i1e(a), A2 = i0ppe(a, A0, A1);
        const double B0 = i0e(b), B1 = i1e(b), B2 = i0ppe(b, B0, B1);
        const double p = 0.5 * (1 - u), q = 0.5 * (1 + u);
        const double P = A0 * B0;
        const double t1 = p * A1 * B0, t2 = q * A0 * B1;
        const double P1 = t1 + t2, P2 = -t1 + t2;
        const double paa = p * p * A2 * B0, pbb = q * q * A0 * B2, pab = 2 * p * q * A1 * B1;
        out[0] = e * P;
        out[1] = e * P1;
        out[2] = e * P2;
        out[3] = e * u * P;
        out[4] = e * (paa + pab + pbb);
        out[5] = e * (paa - pab + pbb);
        out[6] = e * u * u * P;
        out[7] = e * (pbb - paa);
        out[8] = e * u * P1;
        out[9] = e * u * P2;
    }
};

// Gauss-Kronrod 7/15 abscissae and weights (nodes in descending order, the
// Gauss nodes are the odd-indexed ones).
const double kXgk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                        0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                        0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                        0.207784955007898467600689403773245, 0.0};
const double kWgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                        0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                        0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                        0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
const double kWg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                       0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

struct Piece {
    double a, b, err;
    Vec val;
};

void gk15(const Kunze& f, double a, double b, Piece& pc) {
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    Vec k, g, v1, v2;
    f.eval(c, v1);
    for (int i = 0; i < kNV; i++) {
        k[i] = kWgk[7] * v1[i];
        g[i] = kWg[3] * v1[i];
    }
    for (int j = 0; j < 7; j++) {
        f.eval(c - h * kXgk[j], v1);
        f.eval(c + h * kXgk[j], v2);
        for (int i = 0; i < kNV; i++) {
            k[i] += kWgk[j] * (v1[i] + v2[i]);
            if (j & 1) g[i] += kWg[j >> 1] * (v1[i] + v2[i]);
        }
    }
    pc.a = a;
    pc.b = b;
    pc.err = 0.0;
    for (int i = 0; i < kNV; i++) {
        pc.val[i] = k[i] * h;
        pc.err = std::max(pc.err, std::fabs((k[i] - g[i]) * h));
    }
}

}  // namespace

Normalizer normalizer(const double s[3]) {
    Kunze f{s[0], s[1], s[2], 0.0};
    f.K = std::max(std::fabs(s[0] + s[1]) + s[2], std::fabs(s[0] - s[1]) - s[2]);
    // global adaptive bisection of the piece with the largest error estimate
    std::vector<Piece> pcs(1);
    gk15(f, -1.0, 1.0, pcs[0]);
    for (int it = 0; it < 2000; it++) {
        double tot_err = 0.0, C = 0.0;
        size_t worst = 0;
        for (size_t i = 0; i < pcs.size(); i++) {
            tot_err += pcs[i].err;
            C += pcs[i].val[0];
            if (pcs[i].err > pcs[worst].err) worst = i;
        }
        if (tot_err <= 1e-14 * std::fabs(C)) break;
        const double a = pcs[worst].a, b = pcs[worst].b, m = 0.5 * (a + b);
        Piece left, right;
        gk15(f, a, m, left);
        gk15(f, m, b, right);
        pcs[worst] = left;
        pcs.push_back(right);
    }
    Vec tot = {0};
    for (const Piece& p : pcs)
        for (int i = 0; i < kNV; i++) tot[i] += p.val[i];
    Normalizer n;
    n.K = f.K;
    n.C = tot[0];
    for (int i = 0; i < 3; i++) n.D[i] = tot[1 + i];
    const int hmap[3][3] = {{4, 7, 8}, {7, 5, 9}, {8, 9, 6}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) n.H[i + 3 * j] = tot[hmap[i][j]];
    return n;
}

// --------------------------------------------------------- small LA ----

// Solves A x = b (3x3 column-major) by Gaussian elimination with partial
// pivoting; false if singular.
static bool solve3(const double Ain[9], const double bin[3], double x[3]) {
    double A[3][4];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) A[i][j] = Ain[i + 3 * j];
        A[i][3] = bin[i];
    }
    for (int c = 0; c < 3; c++) {
        int p = c;
        for (int r = c + 1; r < 3; r++)
            if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
        if (A[p][c] == 0.0 || !std::isfinite(A[p][c])) return false;
        if (p != c)
            for (int j = 0; j < 4; j++) std::swap(A[p][j], A[c][j]);
        for (int r = c + 1; r < 3; r++) {
            const double f = A[r][c] / A[c][c];
            for (int j = c; j < 4; j++) A[r][j] -= f * A[c][j];
        }
    }
    for (int i = 2; i >= 0; i--) {
        double s = A[i][3];
        for (int j = i + 1; j < 3; j++) s -= A[i][j] * x[j];
        x[i] = s / A[i][i];
    }
    return true;
}

static double det3(const double M[9]) {
    return M[0] * (M[4] * M[8] - M[7] * M[5]) - M[3] * (M[1] * M[8] - M[7] * M[2]) +
           M[6] * (M[1] * M[5] - M[4] * M[2]);
}

// ------------------------------------------------------------- the fit --

int fit_s(const double d[3], const double s0[3], double s[3], double fval[3], double tol, int max_iter) {
    auto psi = [&](const Normalizer& n, const double* x) {
        return std::log(n.C) + n.K - (d[0] * x[0] + d[1] * x[1] + d[2] * x[2]);
    };
    for (int i = 0; i < 3; i++) s[i] = s0[i];
    Normalizer n = normalizer(s);
    for (int it = 0;; it++) {
        double r[3], g[3], rmax = 0.0;
        bool finite = std::isfinite(n.C) && n.C > 0;
        for (int i = 0; i < 3; i++) {
            g[i] = n.D[i] / n.C;
            r[i] = g[i] - d[i];
            fval[i] = r[i];
            finite = finite && std::isfinite(r[i]);
            rmax = std::max(rmax, std::fabs(r[i]));
        }
        if (!finite) return -1;
        if (rmax <= tol) return it;
        if (it >= max_iter) return -1;
        double Hm[9], step[3], mr[3] = {-r[0], -r[1], -r[2]};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Hm[i + 3 * j] = n.H[i + 3 * j] / n.C - g[i] * g[j];
        if (!solve3(Hm, mr, step)) return -1;
        const double slope = r[0] * step[0] + r[1] * step[1] + r[2] * step[2];
        const double p0 = psi(n, s);
        double t = 1.0, x[3];
        Normalizer nt;
        for (int ls = 0; ls < 60; ls++, t *= 0.5) {
            for (int i = 0; i < 3; i++) x[i] = s[i] + t * step[i];
            nt = normalizer(x);
            // Armijo on the convex objective, with a rounding allowance so that
            // the final quadratically convergent steps are never refused
            if (psi(nt, x) <= p0 + 1e-4 * t * slope + 1e-13 * (1.0 + std::fabs(p0))) break;
        }
        for (int i = 0; i < 3; i++) s[i] = x[i];
        n = nt;
    }
}

void svd3(const double F[9], double U[9], double sig[3], double V[9]) {
    double A[9], W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::memcpy(A, F, sizeof(A));
    const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (const auto& pq : pr) {
            double* ap = A + 3 * pq[0];
            double* aq = A + 3 * pq[1];
            double* vp = W + 3 * pq[0];
            double* vq = W + 3 * pq[1];
            const double al = ap[0] * ap[0] + ap[1] * ap[1] + ap[2] * ap[2];
            const double be = aq[0] * aq[0] + aq[1] * aq[1] + aq[2] * aq[2];
            const double ga = ap[0] * aq[0] + ap[1] * aq[1] + ap[2] * aq[2];
            if (al == 0.0 || be == 0.0) continue;
            const double rel = std::fabs(ga) / std::sqrt(al * be);
            if (rel <= 1e-16) continue;
            off = std::max(off, rel);
            const double ze = (be - al) / (2.0 * ga);
            const double t = (ze >= 0 ? 1.0 : -1.0) / (std::fabs(ze) + std::sqrt(1.0 + ze * ze));
            const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
            for (int i = 0; i < 3; i++) {
                const double x = ap[i], y = aq[i];
                ap[i] = c * x - sn * y;
                aq[i] = sn * x + c * y;
                const double vx = vp[i], vy = vq[i];
                vp[i] = c * vx - sn * vy;
                vq[i] = sn * vx + c * vy;
            }
        }
        if (off <= 1e-15) break;
    }
    double nrm[3];
    for (int j = 0; j < 3; j++) nrm[j] = std::sqrt(A[3 * j] * A[3 * j] + A[3 * j + 1] * A[3 * j + 1] + A[3 * j + 2] * A[3 * j + 2]);
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int b) { return nrm[a] > nrm[b]; });
    for (int j = 0; j < 3; j++) {
        sig[j] = nrm[ord[j]];
        for (int i = 0; i < 3; i++) {
            V[i + 3 * j] = W[i + 3 * ord[j]];
            U[i + 3 * j] = sig[j] > 0 ? A[i + 3 * ord[j]] / sig[j] : 0.0;
        }
    }
    // complete U where singular values vanish (relative to the largest)
    const double tiny = 1e-300 + 1e-15 * sig[0];
    auto cross = [](const double* a, const double* b, double* c) {
        c[0] = a[1] * b[2] - a[2] * b[1];
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
    };
    if (sig[0] <= tiny) {
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::memcpy(U, I, sizeof(I));
    } else if (sig[1] <= tiny) {
        // any unit vector orthogonal to u1, then u3 = u1 x u2
        const double* u1 = U;
        double e[3] = {0, 0, 0};
        int k = 0;
        for (int i = 1; i < 3; i++)
            if (std::fabs(u1[i]) < std::fabs(u1[k])) k = i;
        e[k] = 1.0;
        double u2[3];
        cross(u1, e, u2);
        const double nn = std::sqrt(u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
        for (int i = 0; i < 3; i++) U[3 + i] = u2[i] / nn;
        cross(U, U + 3, U + 6);
    } else if (sig[2] <= tiny) {
        cross(U, U + 3, U + 6);
    }
}

void psvd(const double F[9], double U[9], double s[3], double V[9]) {
    svd3(F, U, s, V);
    const double du = det3(U), dvv = det3(V);
    s[2] *= (du < 0 ? -1.0 : 1.0) * (dvv < 0 ? -1.0 : 1.0);
    if (du < 0)
        for (int i = 0; i < 3; i++) U[6 + i] = -U[6 + i];
    if (dvv < 0)
        for (int i = 0; i < 3; i++) V[6 + i] = -V[6 + i];
}

// ------------------------------------------------------------ get_stat --

void finish_stat(int dim, int64_t nR, int64_t nx, int n2, const double* w, const double* fR, const double* gR,
                 const double* fx, const double* xg, double dv,
                 const std::function<void(int64_t, double*)>& Rof, bool fit, Stat& o) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    std::memset(&o, 0, sizeof(o));
    // E[R] = sum_R R w_j fR dv                                (pendulum.m:552-553)
    double R[9];
    for (int64_t iR = 0; iR < nR; iR++) {
        Rof(iR, R);
        const double wt = fR[iR] * w[(iR / n2) % n2] * dv;
        for (int q = 0; q < 9; q++) o.ER[q] += R[q] * wt;
    }
    // E[x], Var[x] from the omega marginal                    (:555-558)
    double S2[9] = {0};
    for (int64_t ix = 0; ix < nx; ix++)
        for (int a = 0; a < dim; a++) {
            o.Ex[a] += xg[dim * ix + a] * fx[ix];
            for (int b = 0; b < dim; b++) S2[a + dim * b] += xg[dim * ix + a] * xg[dim * ix + b] * fx[ix];
        }
    for (int a = 0; a < dim; a++) o.Ex[a] *= dv;
    for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) o.Varx[a + dim * b] = S2[a + dim * b] * dv - o.Ex[a] * o.Ex[b];
    if (!fit) return;
    // proper SVD and the matrix-Fisher fit                    (:560-563)
    double sv[3];
    psvd(o.ER, o.U, sv, o.V);
    double s[3], fv[3];
    const double zero[3] = {0, 0, 0};
    o.niter = fit_s(sv, zero, s, fv);
    for (int i = 0; i < 3; i++) o.S[4 * i] = s[i];
    // vR(R) = (s2 Q32 - s3 Q23, s3 Q13 - s1 Q31, s1 Q21 - s2 Q12), Q = U^T R V  (:565-568)
    double tmp[9], Q[9];
    for (int64_t iR = 0; iR < nR; iR++) {
        Rof(iR, R);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) tmp[i + 3 * j] = R[i] * o.V[3 * j] + R[i + 3] * o.V[1 + 3 * j] + R[i + 6] * o.V[2 + 3 * j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                Q[i + 3 * j] = o.U[3 * i] * tmp[3 * j] + o.U[1 + 3 * i] * tmp[1 + 3 * j] + o.U[2 + 3 * i] * tmp[2 + 3 * j];
        const double vR[3] = {s[1] * Q[2 + 3 * 1] - s[2] * Q[1 + 3 * 2], s[2] * Q[0 + 3 * 2] - s[0] * Q[2 + 3 * 0],
                              s[0] * Q[1 + 3 * 0] - s[1] * Q[0 + 3 * 1]};
        const double wj = w[(iR / n2) % n2];
        const double wf = wj * fR[iR] * dv;
        for (int i = 0; i < 3; i++) {
            o.EvR[i] += vR[i] * wf;
            for (int j = 0; j < 3; j++) o.EvRvR[i + 3 * j] += vR[i] * vR[j] * wf;
            for (int a = 0; a < dim; a++) o.ExvR[a + dim * i] += vR[i] * wj * gR[(int64_t)a * nR + iR] * dv;
        }
    }
    // MFG moment matching                                      (:577-584)
    if (dim > 0) {
        double cxv[9], cvv[9];
        for (int a = 0; a < dim; a++)
            for (int i = 0; i < 3; i++) cxv[a + dim * i] = o.ExvR[a + dim * i] - o.Ex[a] * o.EvR[i];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) cvv[i + 3 * j] = o.EvRvR[i + 3 * j] - o.EvR[i] * o.EvR[j];
        // P cvv = cxv  <=>  cvv^T P^T = cxv^T  (rows of P)
        double cvvT[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) cvvT[i + 3 * j] = cvv[j + 3 * i];
        bool ok = true;
        for (int a = 0; a < dim; a++) {
            const double rhs[3] = {cxv[a], cxv[a + dim], cxv[a + 2 * dim]};
            double row[3];
            ok = ok && solve3(cvvT, rhs, row);
            for (int i = 0; i < 3; i++) o.P[a + dim * i] = ok ? row[i] : nan;
        }
        for (int a = 0; a < dim; a++) {
            o.Miu[a] = o.Ex[a];
            for (int i = 0; i < 3; i++) o.Miu[a] -= o.P[a + dim * i] * o.EvR[i];
        }
        const double trS = s[0] + s[1] + s[2];
        for (int a = 0; a < dim; a++)
            for (int b = 0; b < dim; b++) {
                double v = o.Varx[a + dim * b];
                for (int i = 0; i < 3; i++) v -= o.P[a + dim * i] * cxv[b + dim * i];
                for (int i = 0; i < 3; i++) v += o.P[a + dim * i] * (trS - s[i]) * o.P[b + dim * i];
                o.Sigma[a + dim * b] = v;
            }
    }
    // a failed fit leaves NaN, as the reference's catch branch does (:586-594)
    bool finite = true;
    auto chk = [&](const double* a, int n) {
        for (int i = 0; i < n; i++) finite = finite && std::isfinite(a[i]);
    };
    chk(o.S, 9);
    chk(o.EvR, 3);
    chk(o.EvRvR, 9);
    chk(o.ExvR, 3 * dim);
    chk(o.P, 3 * dim);
    chk(o.Miu, dim);
    chk(o.Sigma, dim * dim);
    o.fit_ok = finite ? 1 : 0;
    if (!finite) {
        for (double* a : {o.S, o.EvRvR, o.ExvR, o.P, o.Sigma})
            for (int i = 0; i < 9; i++) a[i] = nan;
        for (double* a : {o.EvR, o.Miu})
            for (int i = 0; i < 3; i++) a[i] = nan;
    }
}

}  // namespace mf
}  // namespace hpup
