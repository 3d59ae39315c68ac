/*
 * refcpu.c -- FP64 CPU restatement of fdcl-gwu/hybridPendulum-UP's spectral
 * Fokker-Planck path on SO(3) x R^n.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle.  Only tests/, the smoke entry
 * and bench.py's cpu_baseline leg load it; the product path never does.
 *
 * Every function restates one reference routine and cites it.  The derivative
 * is the reference's own dense algorithm (literal circular convolutions over
 * the whole omega-frequency grid and the full (l1,l2) Clebsch-Gordan sweep),
 * so it is both the parity oracle and the timed CPU baseline.
 */
#include "refcpu.h"

#include <complex.h>
#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

typedef double _Complex cplx;

#define PI 3.14159265358979323846

/* ---------------------------------------------------------------- grid ---- */

/* fftSO3R.m:27-29 (grid) and fftSO3R.m:63-66 (beta quadrature weights). */
void ref_grid(int B, double* alpha, double* beta, double* gamma, double* w)
{
    for (int j = 0; j < 2 * B; j++) {
        if (alpha) alpha[j] = PI / B * j;
        if (gamma) gamma[j] = PI / B * j;
        double bj = PI / (4.0 * B) * (2 * j + 1);
        if (beta) beta[j] = bj;
        if (w) {
            double s = 0.0;
            for (int t = 0; t < B; t++) s += sin((2 * t + 1) * bj) / (2 * t + 1);
            w[j] = 1.0 / (4.0 * B * B * B) * sin(bj) * s;
        }
    }
}

/* fftSO3R.m:38-49: R(:,:,i,j,k) = Rz(alpha_i) Ry(beta_j) Rz(gamma_k), column-major 3x3. */
void ref_rotations(int B, double* R)
{
    int n = 2 * B;
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                double a = PI / B * i, b = PI / (4.0 * B) * (2 * j + 1), g = PI / B * k;
                double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
                double Ra[9] = {ca, sa, 0, -sa, ca, 0, 0, 0, 1};
                double Rb[9] = {cb, 0, -sb, 0, 1, 0, sb, 0, cb};
                double Rg[9] = {cg, sg, 0, -sg, cg, 0, 0, 0, 1};
                double T[9], O[9];
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) {
                        double s = 0;
                        for (int q = 0; q < 3; q++) s += Ra[r + 3 * q] * Rb[q + 3 * c];
                        T[r + 3 * c] = s;
                    }
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) {
                        double s = 0;
                        for (int q = 0; q < 3; q++) s += T[r + 3 * q] * Rg[q + 3 * c];
                        O[r + 3 * c] = s;
                    }
                memcpy(R + 9 * (i + n * (j + n * k)), O, sizeof(O));
            }
}

/* ------------------------------------------------------------ Wigner-d ---- */

static double fact(int n)
{
    double r = 1.0;
    for (int i = 2; i <= n; i++) r *= i;
    return r;
}

#define DIDX(m, n, l) (((m) + lmax) + M * (((n) + lmax) + M * (l)))

/* Wigner_d.m:50-64 (symm): fill the rest of degree l from the computed wedge. */
static void wigner_symm(double* d, int l, int lmax)
{
    int M = 2 * lmax + 1;
    for (int m = 1; m <= l; m++)
        for (int n = -m + 1; n <= m; n++) d[DIDX(m, n, l)] = d[DIDX(-n, -m, l)];
    for (int m = -l; m <= l - 1; m++)
        for (int n = m + 1; n <= l; n++)
            d[DIDX(m, n, l)] = d[DIDX(n, m, l)] * (((m - n) & 1) ? -1.0 : 1.0);
}

/* Wigner_d.m:10-45: three-term recursion in l, seeds l = 0,1, then symmetries. */
void ref_wigner_d(double beta, int lmax, double* d)
{
    int M = 2 * lmax + 1;
    memset(d, 0, sizeof(double) * M * M * (lmax + 1));
    double cb = cos(beta), cb2 = cos(beta / 2), sb2 = sin(beta / 2);
    d[DIDX(0, 0, 0)] = 1.0;
    if (lmax < 1) return;
    d[DIDX(-1, -1, 1)] = cb2 * cb2;
    d[DIDX(0, -1, 1)] = -sqrt(2.0) * cb2 * sb2;
    d[DIDX(0, 0, 1)] = cb;
    d[DIDX(1, -1, 1)] = sb2 * sb2;
    wigner_symm(d, 1, lmax);
    for (int l = 2; l <= lmax; l++) {
        for (int n = -l; n <= 0; n++)
            for (int m = n; m <= -n; m++) {
                double v;
                if (n == -l) {
                    v = sqrt(fact(2 * l) / fact(l + m) / fact(l - m)) * pow(cb2, l - m) *
                        pow(-sb2, l + m);
                } else {
                    double a = l * (2.0 * l - 1) / sqrt((double)(l * l - m * m) * (l * l - n * n));
                    v = a * (cb - (double)m * n / (l - 1) / l) * d[DIDX(m, n, l - 1)];
                    if (n != -l + 1)
                        v -= sqrt((double)((l - 1) * (l - 1) - m * m) * ((l - 1) * (l - 1) - n * n)) /
                             sqrt((double)(l * l - m * m) * (l * l - n * n)) * l / (l - 1.0) *
                             d[DIDX(m, n, l - 2)];
                }
                d[DIDX(m, n, l)] = v;
            }
        wigner_symm(d, l, lmax);
    }
}

/* Wigner_d.m:7-9: closed Jacobi form (independent path, used as a check). */
void ref_wigner_d_jacobi(double beta, int l, int m, int n, double* out)
{
    /* d = sqrt((l+n)!(l-n)!/((l+m)!(l-m)!)) sin(b/2)^(n-m) cos(b/2)^(n+m) P_{l-n}^{(n-m,n+m)}(cos b) */
    int k = l - n, a = n - m, bb = n + m;
    double x = cos(beta);
    /* Jacobi polynomial by its explicit sum (valid for negative a,b as a polynomial). */
    double P = 0.0;
    for (int s = 0; s <= k; s++) {
        /* binom(k+a, k-s) * binom(k+bb, s) * ((x-1)/2)^s ((x+1)/2)^(k-s) with generalized binomials */
        double c1 = 1.0, c2 = 1.0;
        for (int t = 0; t < k - s; t++) c1 *= (double)(k + a - t) / (t + 1);
        for (int t = 0; t < s; t++) c2 *= (double)(k + bb - t) / (t + 1);
        P += c1 * c2 * pow((x - 1) / 2, s) * pow((x + 1) / 2, k - s);
    }
    double pre = sqrt(fact(l + n) * fact(l - n) / fact(l + m) / fact(l - m));
    *out = pre * pow(sin(beta / 2), a) * pow(cos(beta / 2), bb) * P;
}

/* fftSO3R.m:69-73: d(:,:,:,j) = Wigner_d(beta_j, lmax). */
void ref_wigner_table(int B, double* d)
{
    int lmax = B - 1, M = 2 * B - 1;
    double* beta = (double*)malloc(sizeof(double) * 2 * B);
    ref_grid(B, NULL, beta, NULL, NULL);
    for (int j = 0; j < 2 * B; j++) ref_wigner_d(beta[j], lmax, d + (size_t)j * M * M * B);
    free(beta);
}

/* ------------------------------------------------------ Clebsch-Gordan ---- */

/* Dense LU with partial pivoting (MATLAB linsolve on a square matrix). */
static void lu_solve(int n, double* A, double* b)
{
    int* piv = (int*)malloc(sizeof(int) * n);
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = fabs(A[k + n * k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i + n * k]) > mx) { mx = fabs(A[i + n * k]); p = i; }
        piv[k] = p;
        if (p != k) {
            for (int j = 0; j < n; j++) { double t = A[k + n * j]; A[k + n * j] = A[p + n * j]; A[p + n * j] = t; }
            double t = b[k]; b[k] = b[p]; b[p] = t;
        }
        double dkk = A[k + n * k];
        if (dkk == 0.0) continue;
        for (int i = k + 1; i < n; i++) {
            double f = A[i + n * k] / dkk;
            A[i + n * k] = f;
            for (int j = k + 1; j < n; j++) A[i + n * j] -= f * A[k + n * j];
            b[i] -= f * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= A[i + n * j] * b[j];
        b[i] = s / A[i + n * i];
    }
    free(piv);
}

/* clebsch_gordan.m:1-41: null vector of (J^2 - l(l+1)) restricted to m1+m2 = m,
 * last (largest m1) component fixed positive, unit norm. */
int ref_clebsch_gordan(int l1, int l2, double* CG)
{
    int dim = (2 * l1 + 1) * (2 * l2 + 1);
    memset(CG, 0, sizeof(double) * dim * dim);
    int nmax = 2 * (l1 < l2 ? l1 : l2) + 1;
    double* A = (double*)malloc(sizeof(double) * nmax * nmax);
    double* Ab = (double*)malloc(sizeof(double) * nmax * nmax);
    double* rhs = (double*)malloc(sizeof(double) * nmax);
    int* m1 = (int*)malloc(sizeof(int) * nmax);
    for (int l = abs(l1 - l2); l <= l1 + l2; l++) {
        for (int m = -l; m <= l; m++) {
            int lo = (m - l1 - l2 + abs(l1 - l2 + m)) / 2;
            int hi = (m + l1 + l2 - abs(l1 - l2 - m)) / 2;
            int nz = hi - lo + 1;
            for (int k = 0; k < nz; k++) m1[k] = lo + k;
            double* C = rhs;
            if (nz == 1) {
                C[0] = 1.0;
            } else {
                memset(A, 0, sizeof(double) * nz * nz);
                for (int k = 0; k < nz; k++) {
                    int a1 = m1[k], a2 = m - m1[k];
                    A[k + nz * k] = l1 * (l1 + 1.0) + l2 * (l2 + 1.0) + 2.0 * a1 * a2 - l * (l + 1.0);
                    if (k > 0) {
                        int p1 = m1[k - 1], p2 = m - m1[k - 1];
                        double v = sqrt(l1 * (l1 + 1.0) - (double)p1 * a1) * sqrt(l2 * (l2 + 1.0) - (double)p2 * a2);
                        A[(k - 1) + nz * k] = v;
                        A[k + nz * (k - 1)] = v;
                    }
                }
                int n1 = nz - 1;
                for (int j = 0; j < n1; j++)
                    for (int i = 0; i < n1; i++) Ab[i + n1 * j] = A[i + nz * j];
                for (int i = 0; i < n1; i++) C[i] = 0.0;
                C[n1 - 1] = -A[(nz - 2) + nz * (nz - 1)];
                lu_solve(n1, Ab, C);
                double s = 0.0;
                for (int i = 0; i < n1; i++) s += C[i] * C[i];
                double Cn = sqrt(1.0 / (1.0 + s));
                for (int i = 0; i < n1; i++) C[i] *= Cn;
                C[n1] = Cn;
            }
            int col = l * l - (l2 - l1) * (l2 - l1) + l + m;
            for (int k = 0; k < nz; k++) {
                int row = (l1 + m1[k]) * (2 * l2 + 1) + l2 + (m - m1[k]);
                CG[row + (size_t)dim * col] = C[k];
            }
        }
    }
    free(A); free(Ab); free(rhs); free(m1);
    return dim;
}

/* ------------------------------------------------------------------ u ---- */

/* getu.m:176-207 (complex branch): u_1, u_2 bidiagonal, u_3 diagonal. */
void ref_getu(int lmax, double* u)
{
    int M = 2 * lmax + 1, B = lmax + 1;
    cplx* U = (cplx*)u;
    memset(u, 0, sizeof(double) * 2 * M * M * B * 3);
#define UIDX(m, n, l, p) (((m) + lmax) + M * (((n) + lmax) + M * ((l) + B * (p))))
    for (int l = 0; l <= lmax; l++) {
        for (int n = -l; n <= l; n++) {
            if (n > -l) U[UIDX(n - 1, n, l, 0)] = -0.5 * I * sqrt((double)(l + n) * (l - n + 1));
            if (n < l) U[UIDX(n + 1, n, l, 0)] = -0.5 * I * sqrt((double)(l - n) * (l + n + 1));
        }
        for (int n = -l; n <= l; n++) {
            if (n > -l) U[UIDX(n - 1, n, l, 1)] = 0.5 * sqrt((double)(l + n) * (l - n + 1));
            if (n < l) U[UIDX(n + 1, n, l, 1)] = -0.5 * sqrt((double)(l - n) * (l + n + 1));
        }
        for (int m = -l; m <= l; m++) U[UIDX(m, m, l, 2)] = -I * (double)m;
    }
#undef UIDX
}

/* ------------------------------------------------------------ DFT utils ---- */

/* In-place naive DFT of `cnt` interleaved vectors of length n with element
 * stride `st`, vectors enumerated by (outer, inner) with inner stride 1 and
 * outer stride n*st.  sign=-1 forward, +1 inverse (unnormalized). */
static void dft_axis(cplx* a, long inner, long n, long outer, int sign)
{
    cplx* tw = (cplx*)malloc(sizeof(cplx) * n);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * n);
    for (long k = 0; k < n; k++) tw[k] = cexp(sign * 2.0 * PI * I * (double)k / (double)n);
    for (long o = 0; o < outer; o++)
        for (long i = 0; i < inner; i++) {
            cplx* base = a + o * n * inner + i;
            for (long k = 0; k < n; k++) {
                cplx s = 0;
                for (long j = 0; j < n; j++) s += base[j * inner] * tw[(j * k) % n];
                tmp[k] = s;
            }
            for (long k = 0; k < n; k++) base[k * inner] = tmp[k];
        }
    free(tw);
    free(tmp);
}

static long ipow(long a, int e)
{
    long r = 1;
    while (e-- > 0) r *= a;
    return r;
}

/* ------------------------------------------------------- SO(3)xR^n FFT ---- */

/* fftSO3R.m:90-106: FFT over (alpha, gamma, omega axes), fftshift+flip on
 * alpha and gamma (frequency -m lands at index m+lmax), then the beta
 * quadrature sum_j w_j d^l_mn(beta_j).  Entries with max(|m|,|n|) > l are 0. */
void ref_fftSO3R_forward(int B, int dim, int N, const double* f, const double* d,
                         const double* w, double* F)
{
    long n2 = 2 * B, M = 2 * B - 1, lmax = B - 1;
    long nx = ipow(N, dim);
    long nR = n2 * n2 * n2;
    cplx* G = (cplx*)malloc(sizeof(cplx) * nR * nx);
    for (long i = 0; i < nR * nx; i++) G[i] = f[i];
    dft_axis(G, 1, n2, n2 * n2 * nx, -1);             /* alpha */
    dft_axis(G, n2 * n2, n2, nx, -1);                 /* gamma */
    for (int p = 0; p < dim; p++)                     /* omega axes */
        dft_axis(G, nR * ipow(N, p), N, ipow(N, dim - 1 - p), -1);
    cplx* Fo = (cplx*)F;
    for (long x = 0; x < nx; x++)
        for (long l = 0; l <= lmax; l++)
            for (long n = -lmax; n <= lmax; n++)
                for (long m = -lmax; m <= lmax; m++) {
                    long o = (m + lmax) + M * ((n + lmax) + M * (l + B * x));
                    if (labs(m) > l || labs(n) > l) { Fo[o] = 0; continue; }
                    long ia = (-m + n2) % n2, ig = (-n + n2) % n2;
                    cplx s = 0;
                    for (long j = 0; j < n2; j++)
                        s += w[j] * d[(m + lmax) + M * ((n + lmax) + M * (l + B * j))] *
                             G[ia + n2 * (j + n2 * (ig + n2 * x))];
                    Fo[o] = s;
                }
    free(G);
}

/* ifftSO3R.m:16-37 and pendulum_reduced.m:258-278: per (m,n,j) the sum over
 * l >= max(|m|,|n|) of (2l+1) F d, ifftshift+flip, inverse omega DFT (1/nx),
 * unnormalized inverse DFT over (alpha, gamma), real part. */
void ref_fftSO3R_backward(int B, int dim, int N, const double* F, const double* d, double* f)
{
    long n2 = 2 * B, M = 2 * B - 1, lmax = B - 1;
    long nx = ipow(N, dim);
    long nR = n2 * n2 * n2;
    const cplx* Fi = (const cplx*)F;
    cplx* H = (cplx*)calloc(nR * nx, sizeof(cplx));
    for (long x = 0; x < nx; x++)
        for (long ig = 0; ig < n2; ig++)
            for (long j = 0; j < n2; j++)
                for (long ia = 0; ia < n2; ia++) {
                    if (ia == B || ig == B) continue; /* Nyquist of alpha / gamma */
                    long m = ia <= lmax ? -ia : n2 - ia;
                    long n = ig <= lmax ? -ig : n2 - ig;
                    long lmin = labs(m) > labs(n) ? labs(m) : labs(n);
                    cplx s = 0;
                    for (long l = lmin; l <= lmax; l++)
                        s += (2.0 * l + 1.0) * Fi[(m + lmax) + M * ((n + lmax) + M * (l + B * x))] *
                             d[(m + lmax) + M * ((n + lmax) + M * (l + B * j))];
                    H[ia + n2 * (j + n2 * (ig + n2 * x))] = s;
                }
    dft_axis(H, 1, n2, n2 * n2 * nx, +1);
    dft_axis(H, n2 * n2, n2, nx, +1);
    for (int p = 0; p < dim; p++) dft_axis(H, nR * ipow(N, p), N, ipow(N, dim - 1 - p), +1);
    double sc = 1.0 / (double)nx;
    for (long i = 0; i < nR * nx; i++) f[i] = creal(H[i]) * sc;
    free(H);
}

/* ----------------------------------------------------- operator tensors ---- */

/* pendulum.m:90-99,141-143 (pendulum_reduced.m:89-95,137-139): X_p = fftn(x_p)
 * with x_p = -L/2 + L/N * i_p.  X layout (N,..,N,dim) complex. */
void ref_build_X(int dim, int N, double L, double* X)
{
    long nx = ipow(N, dim);
    cplx* Xc = (cplx*)X;
    for (int p = 0; p < dim; p++) {
        cplx* a = Xc + nx * p;
        for (long x = 0; x < nx; x++) {
            long ip = (x / ipow(N, p)) % N;
            a[x] = -L / 2 + L / N * ip;
        }
        for (int q = 0; q < dim; q++) dft_axis(a, ipow(N, q), N, ipow(N, dim - 1 - q), -1);
    }
}

/* pendulum.m:145-158: ojo = Jt^{-1} (-(x cross Jt x)), OJO_p = fftn(ojo_p). */
void ref_build_OJO(int N, double L, const double* Jt, double* OJO)
{
    long nx = ipow(N, 3);
    cplx* O = (cplx*)OJO;
    /* inverse of Jt (3x3 column-major) */
    double a = Jt[0], b = Jt[3], c = Jt[6], dd = Jt[1], e = Jt[4], ff = Jt[7], g = Jt[2], h = Jt[5], k = Jt[8];
    double det = a * (e * k - ff * h) - b * (dd * k - ff * g) + c * (dd * h - e * g);
    double Ji[9] = {(e * k - ff * h) / det, -(dd * k - ff * g) / det, (dd * h - e * g) / det,
                    -(b * k - c * h) / det, (a * k - c * g) / det, -(a * h - b * g) / det,
                    (b * ff - c * e) / det, -(a * ff - c * dd) / det, (a * e - b * dd) / det};
    for (long x = 0; x < nx; x++) {
        double v[3];
        for (int p = 0; p < 3; p++) v[p] = -L / 2 + L / N * ((x / ipow(N, p)) % N);
        double Jv[3], cr[3], o[3];
        for (int r = 0; r < 3; r++) Jv[r] = Jt[r] * v[0] + Jt[r + 3] * v[1] + Jt[r + 6] * v[2];
        cr[0] = v[1] * Jv[2] - v[2] * Jv[1];
        cr[1] = v[2] * Jv[0] - v[0] * Jv[2];
        cr[2] = v[0] * Jv[1] - v[1] * Jv[0];
        for (int r = 0; r < 3; r++) o[r] = -(Ji[r] * cr[0] + Ji[r + 3] * cr[1] + Ji[r + 6] * cr[2]);
        for (int p = 0; p < 3; p++) O[x + nx * p] = o[p];
    }
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) dft_axis(O + nx * p, ipow(N, q), N, ipow(N, 2 - q), -1);
}

/* pendulum_reduced.m:141-166 (reduced: mR = [R32; -R31]) and pendulum.m:166-192
 * (3-D: mR = Jt^{-1} [R32; -R31; 0]); MR_p = SO(3) transform of mR_p. */
void ref_build_MR(int B, int dim, const double* Jt, double* MR)
{
    long n2 = 2 * B, nR = n2 * n2 * n2, M = 2 * B - 1;
    long nrot = M * M * B;
    double* R = (double*)malloc(sizeof(double) * 9 * nR);
    double* d = (double*)malloc(sizeof(double) * M * M * B * n2);
    double* w = (double*)malloc(sizeof(double) * n2);
    double* mr = (double*)malloc(sizeof(double) * nR);
    ref_rotations(B, R);
    ref_wigner_table(B, d);
    ref_grid(B, NULL, NULL, NULL, w);
    double Ji[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (dim == 3 && Jt) {
        double a = Jt[0], b = Jt[3], c = Jt[6], dd = Jt[1], e = Jt[4], ff = Jt[7], g = Jt[2], h = Jt[5], k = Jt[8];
        double det = a * (e * k - ff * h) - b * (dd * k - ff * g) + c * (dd * h - e * g);
        double t[9] = {(e * k - ff * h) / det, -(dd * k - ff * g) / det, (dd * h - e * g) / det,
                       -(b * k - c * h) / det, (a * k - c * g) / det, -(a * h - b * g) / det,
                       (b * ff - c * e) / det, -(a * ff - c * dd) / det, (a * e - b * dd) / det};
        memcpy(Ji, t, sizeof(t));
    }
    for (int p = 0; p < dim; p++) {
        for (long iR = 0; iR < nR; iR++) {
            double r32 = R[9 * iR + 2 + 3 * 1], r31 = R[9 * iR + 2];
            double v[3] = {r32, -r31, 0.0};
            mr[iR] = (dim == 3) ? Ji[p] * v[0] + Ji[p + 3] * v[1] + Ji[p + 6] * v[2] : v[p];
        }
        ref_fftSO3R_forward(B, 0, 1, mr, d, w, MR + 2 * nrot * p);
    }
    free(R); free(d); free(w); free(mr);
}

/* ----------------------------------------------------------- derivative ---- */

/* deriv_x (pendulum_reduced.m:468-482): 2*pi*i*k/L, Nyquist bin -> 0. */
static cplx deriv_x(long k, long Bx, double L)
{
    if (k < Bx) return 2.0 * PI * I * (double)k / L;
    if (k == Bx) return 0.0;
    return 2.0 * PI * I * (double)(k - 2 * Bx) / L;
}

/* pendulum_reduced.m:362-465 (dim 2) and pendulum.m:388-527 (dim 3): the
 * reference's derivative() with literal dense circular convolutions over the
 * omega-frequency grid and the full Clebsch-Gordan sweep.  Computes dF only
 * for output omega indices x in [x0, x1) (x1 <= 0 means all). */
int ref_derivative(int B, int dim, int N, const double* F, const double* X, const double* OJO,
                   const double* MR, const double* b, const double* G, double L,
                   const double* u, const double* const* CG, double* dF, long x0, long x1)
{
    long lmax = B - 1, M = 2 * B - 1, nrot = M * M * B;
    long nx = ipow(N, dim), Bx = N / 2;
    if (x1 <= 0 || x1 > nx) x1 = nx;
    const cplx* Fc = (const cplx*)F;
    const cplx* Xc = (const cplx*)X;
    const cplx* Oc = (const cplx*)OJO;
    const cplx* MRc = (const cplx*)MR;
    const cplx* Uc = (const cplx*)u;
    cplx* dFc = (cplx*)dF;
    for (long i = x0 * nrot; i < x1 * nrot; i++) dFc[i] = 0;

    long nxs = x1 - x0;
    /* gravity: Clebsch-Gordan product (pendulum.m:470-506).  As in the
     * reference, K = CG_m.' * MR^{l2} * CG_n is formed once per
     * (l, l2, l1, m, n) and applied to every omega point.  CG = NULL skips
     * the term (the density-form variant evaluates it on the grid). */
    cplx* tG = (cplx*)calloc((size_t)nrot * nxs * dim, sizeof(cplx));
#pragma omp parallel if (CG != NULL)
    if (CG) {
        cplx* KM = (cplx*)malloc(sizeof(cplx) * M * M);
        cplx* K = (cplx*)malloc(sizeof(cplx) * M * M);
#pragma omp for schedule(dynamic, 1)
        for (long lm = 0; lm < B * B; lm++) {
            long l = (long)sqrt((double)lm);
            while (l * l > lm) l--;
            while ((l + 1) * (l + 1) <= lm) l++;
            long m = lm - l * l - l;
            for (long l2 = 0; l2 <= lmax; l2++)
                for (long l1 = 0; l1 <= lmax; l1++) {
                    if (!(l >= labs(l1 - l2) && l <= l1 + l2)) continue;
                    double cl = (2.0 * l1 + 1) * (2.0 * l2 + 1) / (2.0 * l + 1);
                    long col = l * l - (l1 - l2) * (l1 - l2);
                    const double* C = CG[l1 + B * l2];
                    long d1 = 2 * l1 + 1, d2 = 2 * l2 + 1, dimc = d1 * d2;
                    const double* Cm = C + dimc * (col + m + l);
                    for (int p = 0; p < dim; p++) {
                        const cplx* MRp = MRc + nrot * p;
                        for (long i1 = 0; i1 < d1; i1++)
                            for (long j2 = 0; j2 < d2; j2++) {
                                cplx s = 0;
                                for (long i2 = 0; i2 < d2; i2++)
                                    s += Cm[i2 + d2 * i1] * MRp[(i2 - l2 + lmax) + M * ((j2 - l2 + lmax) + M * l2)];
                                KM[i1 + d1 * j2] = s;
                            }
                        for (long n = -l; n <= l; n++) {
                            const double* Cn = C + dimc * (col + n + l);
                            for (long j1 = 0; j1 < d1; j1++)
                                for (long i1 = 0; i1 < d1; i1++) {
                                    cplx s = 0;
                                    for (long j2 = 0; j2 < d2; j2++) s += KM[i1 + d1 * j2] * Cn[j2 + d2 * j1];
                                    K[i1 + d1 * j1] = s;
                                }
                            long ro = (m + lmax) + M * ((n + lmax) + M * l);
                            for (long xs = 0; xs < nxs; xs++) {
                                const cplx* Fx = Fc + nrot * (x0 + xs);
                                cplx acc = 0;
                                for (long j1 = 0; j1 < d1; j1++)
                                    for (long i1 = 0; i1 < d1; i1++)
                                        acc += Fx[(i1 - l1 + lmax) + M * ((j1 - l1 + lmax) + M * l1)] * K[i1 + d1 * j1];
                                tG[(size_t)nrot * (xs + nxs * p) + ro] += cl * acc;
                            }
                        }
                    }
                }
        }
        free(KM); free(K);
    }

#pragma omp parallel
    {
        cplx* tX = (cplx*)malloc(sizeof(cplx) * nrot * 3);
        cplx* tB = (cplx*)malloc(sizeof(cplx) * nrot * 3);
        cplx* tO = (cplx*)malloc(sizeof(cplx) * nrot * 3);
#pragma omp for schedule(dynamic, 1)
        for (long x = x0; x < x1; x++) {
            long kx[3] = {x % N, (x / N) % N, dim == 3 ? (x / (N * N)) % N : 0};
            memset(tX, 0, sizeof(cplx) * nrot * 3);
            memset(tB, 0, sizeof(cplx) * nrot * 3);
            memset(tO, 0, sizeof(cplx) * nrot * 3);
            /* dense circular convolutions (pendulum.m:400-414, 429-446, 449-466):
             * temp(k) = sum_a X[(k-a) mod N] F(a) / nx */
            for (long a = 0; a < nx; a++) {
                long ka[3] = {a % N, (a / N) % N, dim == 3 ? (a / (N * N)) % N : 0};
                long q = 0, st = 1;
                for (int p = 0; p < dim; p++) { q += ((kx[p] - ka[p] + N) % N) * st; st *= N; }
                const cplx* Fa = Fc + nrot * a;
                for (int p = 0; p < dim; p++) {
                    cplx xv = Xc[q + nx * p] / (double)nx;
                    cplx bv = -b[p] * Xc[q + nx * p] / (double)nx;
                    cplx ov = (dim == 3 && Oc) ? Oc[q + nx * p] / (double)nx : 0;
                    cplx* tx = tX + nrot * p;
                    cplx* tb = tB + nrot * p;
                    cplx* to = tO + nrot * p;
                    for (long r = 0; r < nrot; r++) {
                        cplx fv = Fa[r];
                        tx[r] += xv * fv;
                        tb[r] += bv * fv;
                        to[r] += ov * fv;
                    }
                }
            }
            cplx* dFx = dFc + nrot * x;
            const cplx* Fx = Fc + nrot * x;
            /* Omega-hat: dF^l -= temp_p^l * u_p^l.' (pendulum.m:416-422) */
            for (int p = 0; p < dim; p++)
                for (long l = 0; l <= lmax; l++)
                    for (long n = -l; n <= l; n++)
                        for (long m = -l; m <= l; m++) {
                            cplx s = 0;
                            for (long j = -l; j <= l; j++)
                                s += tX[nrot * p + (m + lmax) + M * ((j + lmax) + M * l)] *
                                     Uc[(n + lmax) + M * ((j + lmax) + M * (l + B * p))];
                            dFx[(m + lmax) + M * ((n + lmax) + M * l)] -= s;
                        }
            /* damping, gyroscopic and gravity terms times deriv_x(k_p)
             * (pendulum.m:442-443, 462-463, 504-506) */
            for (int p = 0; p < dim; p++) {
                cplx c = deriv_x(kx[p], Bx, L);
                const cplx* g = tG + (size_t)nrot * ((x - x0) + nxs * p);
                for (long r = 0; r < nrot; r++) dFx[r] -= tB[nrot * p + r] * c;
                if (dim == 3 && Oc)
                    for (long r = 0; r < nrot; r++) dFx[r] -= tO[nrot * p + r] * c;
                for (long r = 0; r < nrot; r++) dFx[r] -= g[r] * c;
            }
            /* noise (pendulum.m:510-523) */
            for (int i = 0; i < dim; i++)
                for (int j = 0; j < dim; j++) {
                    double c;
                    if (i == j) {
                        long k = kx[i];
                        double kk = (double)(k < Bx ? k : k - 2 * Bx);
                        c = -4.0 * PI * PI * kk * kk / (L * L);
                    } else {
                        long ki = kx[i], kj = kx[j];
                        double ci = 2 * PI * (double)(ki < Bx ? ki : (ki == Bx ? 0 : ki - 2 * Bx)) / L;
                        double cj = 2 * PI * (double)(kj < Bx ? kj : (kj == Bx ? 0 : kj - 2 * Bx)) / L;
                        c = -ci * cj;
                    }
                    double gij = G[i + dim * j];
                    for (long r = 0; r < nrot; r++) dFx[r] += gij * Fx[r] * c;
                }
        }
        free(tX); free(tB); free(tO);
    }
    free(tG);
    return 0;
}

/* ------------------------------------------------------------ integrate ---- */

/* pendulum_reduced.m:320-359 / pendulum.m:346-385: euler, midpoint, RK2, RK4. */
int ref_integrate(int B, int dim, int N, const double* F, const double* X, const double* OJO,
                  const double* MR, const double* b, const double* G, double dt, double L,
                  const double* u, const double* const* CG, const char* method, double* Fnew)
{
    long nrot = (long)(2 * B - 1) * (2 * B - 1) * B;
    long n = 2 * nrot * ipow(N, dim);
    int meth;
    if (!strcasecmp(method, "euler")) meth = 0;
    else if (!strcasecmp(method, "midpoint")) meth = 1;
    else if (!strcasecmp(method, "RK2")) meth = 2;
    else if (!strcasecmp(method, "RK4")) meth = 3;
    else return -1;
    double* dF1 = (double*)malloc(sizeof(double) * n);
    ref_derivative(B, dim, N, F, X, OJO, MR, b, G, L, u, CG, dF1, 0, 0);
    if (meth == 0) {
        for (long i = 0; i < n; i++) Fnew[i] = F[i] + dt * dF1[i];
        free(dF1);
        return 0;
    }
    double* F2 = (double*)malloc(sizeof(double) * n);
    double* dF2 = (double*)malloc(sizeof(double) * n);
    if (meth == 2) {
        for (long i = 0; i < n; i++) F2[i] = F[i] + dF1[i] * dt;
        ref_derivative(B, dim, N, F2, X, OJO, MR, b, G, L, u, CG, dF2, 0, 0);
        for (long i = 0; i < n; i++) Fnew[i] = F[i] + dt / 2 * (dF1[i] + dF2[i]);
        free(dF1); free(F2); free(dF2);
        return 0;
    }
    for (long i = 0; i < n; i++) F2[i] = F[i] + dF1[i] * dt / 2;
    ref_derivative(B, dim, N, F2, X, OJO, MR, b, G, L, u, CG, dF2, 0, 0);
    if (meth == 1) {
        for (long i = 0; i < n; i++) Fnew[i] = F[i] + dt * dF2[i];
        free(dF1); free(F2); free(dF2);
        return 0;
    }
    double* dF3 = (double*)malloc(sizeof(double) * n);
    double* dF4 = (double*)malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) F2[i] = F[i] + dF2[i] * dt / 2;
    ref_derivative(B, dim, N, F2, X, OJO, MR, b, G, L, u, CG, dF3, 0, 0);
    for (long i = 0; i < n; i++) F2[i] = F[i] + dF3[i] * dt;
    ref_derivative(B, dim, N, F2, X, OJO, MR, b, G, L, u, CG, dF4, 0, 0);
    for (long i = 0; i < n; i++)
        Fnew[i] = F[i] + 1.0 / 6.0 * dt * (dF1[i] + 2 * dF2[i] + 2 * dF3[i] + dF4[i]);
    free(dF1); free(F2); free(dF2); free(dF3); free(dF4);
    return 0;
}

/* --------------------------------------------------------- hybrid reset ---- */

/* getLambda.cu:6-92,192-266 (gateway mexLambda.cpp, compact = false):
 * theta = asin(R13); collision category against theta0 +- theta_t; contact
 * point PC; lambda per collision-region R, zeroed where the contact velocity
 * (omega x PC).e1 < 0.  lambda_full is (numR, nx) with numR fastest. */
int ref_getLambda(int B, int N, const double* R, const double* x, double dwall, double h, double r,
                  double theta_t, double lambda_max, double* lambda_full, int* indR, double* PC)
{
    long nR = 8L * B * B * B, nx = (long)N * N;
    double theta0 = asin(dwall / sqrt(h * h + r * r)) - asin(r / sqrt(h * h + r * r));
    double* theta = (double*)malloc(sizeof(double) * nR);
    char* cat = (char*)malloc(nR);
    for (long i = 0; i < nR; i++) {
        theta[i] = asin(R[9 * i + 6]);
        cat[i] = theta[i] < theta0 - theta_t ? 0 : (theta[i] > theta0 + theta_t ? 2 : 1);
    }
    int numR = 0;
    for (long i = 0; i < nR; i++)
        if (cat[i] != 0) indR[numR++] = (int)i;
    for (long i = 0; i < nR; i++) {
        double lr3 = h - r * tan(theta[i]);
        PC[3 * i] = lr3 * R[9 * i + 6] + r / cos(theta[i]);
        PC[3 * i + 1] = lr3 * R[9 * i + 7];
        PC[3 * i + 2] = lr3 * R[9 * i + 8];
    }
    for (int k = 0; k < numR; k++) {
        long iR = indR[k];
        double lam;
        if (cat[iR] == 2) lam = lambda_max;
        else lam = lambda_max / 2 * sin(PI / (2 * theta_t) * (theta[iR] - theta0)) + lambda_max / 2;
        const double* Rr = R + 9 * iR;
        for (long ix = 0; ix < nx; ix++) {
            double o1 = Rr[1] * x[2 * ix] + Rr[4] * x[2 * ix + 1];
            double o2 = Rr[2] * x[2 * ix] + Rr[5] * x[2 * ix + 1];
            double vC = o1 * PC[3 * iR + 2] - o2 * PC[3 * iR + 1];
            lambda_full[k + (long)numR * ix] = vC < 0.0 ? 0.0 : lam;
        }
    }
    free(theta);
    free(cat);
    return numR;
}

/* getLambda.cu:94-190 (compact = true): lambda per region rotation and, per
 * region, the ascending omega indices whose contact velocity is >= 0. */
int ref_getLambda_compact(int B, int N, const double* R, const double* x, double dwall, double h, double r,
                          double theta_t, double lambda_max, double* lambda, int* indR, int* indx, int* numx,
                          double* PC)
{
    long nR = 8L * B * B * B, nx = (long)N * N;
    double* full = (double*)malloc(sizeof(double) * nR * nx);
    int numR = ref_getLambda(B, N, R, x, dwall, h, r, theta_t, lambda_max, full, indR, PC);
    double theta0 = asin(dwall / sqrt(h * h + r * r)) - asin(r / sqrt(h * h + r * r));
    for (int k = 0; k < numR; k++) {
        long iR = indR[k];
        double th = asin(R[9 * iR + 6]);
        lambda[k] = th > theta0 + theta_t ? lambda_max
                                          : lambda_max / 2 * sin(PI / (2 * theta_t) * (th - theta0)) + lambda_max / 2;
        const double* Rr = R + 9 * iR;
        int c = 0;
        for (long ix = 0; ix < nx; ix++) {
            double o1 = Rr[1] * x[2 * ix] + Rr[4] * x[2 * ix + 1];
            double o2 = Rr[2] * x[2 * ix] + Rr[5] * x[2 * ix + 1];
            double vC = o1 * PC[3 * iR + 2] - o2 * PC[3 * iR + 1];
            if (!(vC < 0.0)) indx[(long)k * nx + c++] = (int)ix;
        }
        numx[k] = c;
    }
    free(full);
    return numR;
}

/* getOmega.cu:64-137: tangent t = r3 x e1 / |.|; 'new': omega+ = omega -
 * (1+eps)(omega.t)t; 'old': (1+eps)/eps, NaN when the contact velocity sign
 * is invalid or the result leaves the grid box.  Omega is (2, numR, nx). */
void ref_getOmega(int B, int N, const double* R, const double* x, const int* indR, int numR,
                  double epsilon, const double* PC, int old_dir, double* Omega)
{
    long nx = (long)N * N;
    double xlo = x[0] + 1e-10, xhi = x[2 * (N - 1)] - 1e-10;
    for (int k = 0; k < numR; k++) {
        const double* Rr = R + 9L * indR[k];
        double nt = sqrt(Rr[7] * Rr[7] + Rr[8] * Rr[8]);
        double t0 = Rr[8] / nt, t1 = -Rr[7] / nt;
        for (long ix = 0; ix < nx; ix++) {
            double o[3];
            o[0] = Rr[0] * x[2 * ix] + Rr[3] * x[2 * ix + 1];
            o[1] = Rr[1] * x[2 * ix] + Rr[4] * x[2 * ix + 1];
            o[2] = Rr[2] * x[2 * ix] + Rr[5] * x[2 * ix + 1];
            double fac = old_dir ? (1.0 + epsilon) / epsilon : (1.0 + epsilon);
            double ot = fac * (o[1] * t0 + o[2] * t1);
            o[1] -= ot * t0;
            o[2] -= ot * t1;
            double* out = Omega + 2 * (k + (long)numR * ix);
            double a0 = Rr[0] * o[0] + Rr[1] * o[1] + Rr[2] * o[2];
            double a1 = Rr[3] * o[0] + Rr[4] * o[1] + Rr[5] * o[2];
            if (old_dir) {
                const double* P = PC + 3L * indR[k];
                double vC = o[1] * P[2] - o[2] * P[1];
                if (vC < 0.0 || a0 < xlo || a0 > xhi || a1 < xlo || a1 > xhi) {
                    out[0] = NAN; out[1] = NAN;
                    continue;
                }
            }
            out[0] = a0;
            out[1] = a1;
        }
    }
}

typedef struct { double v; int i; } validx;
static int cmp_desc_stable(const void* a, const void* b)
{
    const validx* p = (const validx*)a;
    const validx* q = (const validx*)b;
    if (p->v > q->v) return -1;
    if (p->v < q->v) return 1;
    return p->i - q->i;
}

/* getFcL.cu:11-166: for each (R, destination) the Gaussian transition density
 * from every source Omega+; keep the nD largest (stable descending sort),
 * normalize per source so each source's outflow integrates to one, then
 * multiply by lambda(source) dx^2.  fcL, fcL_indx are (nD, numR, nx). */
void ref_getFcL(int N, const double* x, const double* Omega, const double* lambda, int numR,
                const double* Gd, int nD, double* fcL, short* fcL_indx)
{
    long nx = (long)N * N;
    double det = Gd[0] * Gd[3] - Gd[2] * Gd[1];
    double cn = 1.0 / (2 * PI * sqrt(det));
    double iG[4] = {Gd[3] / det, -Gd[2] / det, -Gd[1] / det, Gd[0] / det};
    double dx2 = (x[2] - x[0]) * (x[2] - x[0]);
    validx* buf = (validx*)malloc(sizeof(validx) * nx);
    double* norm = (double*)malloc(sizeof(double) * nx);
    for (int k = 0; k < numR; k++) {
        for (long s = 0; s < nx; s++) norm[s] = 0.0;
        for (long ixd = 0; ixd < nx; ixd++) {
            for (long s = 0; s < nx; s++) {
                const double* Om = Omega + 2 * (k + (long)numR * s);
                double d0 = x[2 * ixd] - Om[0], d1 = x[2 * ixd + 1] - Om[1];
                double q = iG[0] * d0 * d0 + (iG[1] + iG[2]) * d0 * d1 + iG[3] * d1 * d1;
                buf[s].v = exp(-0.5 * q) * cn;
                buf[s].i = (int)s;
            }
            qsort(buf, nx, sizeof(validx), cmp_desc_stable);
            for (int t = 0; t < nD; t++) {
                long o = t + (long)nD * (k + (long)numR * ixd);
                fcL[o] = buf[t].v;
                fcL_indx[o] = (short)buf[t].i;
            }
        }
        /* per-source normalization, summed in destination order */
        for (long ixd = 0; ixd < nx; ixd++)
            for (int t = 0; t < nD; t++) {
                long o = t + (long)nD * (k + (long)numR * ixd);
                norm[fcL_indx[o]] += fcL[o];
            }
        for (long ixd = 0; ixd < nx; ixd++)
            for (int t = 0; t < nD; t++) {
                long o = t + (long)nD * (k + (long)numR * ixd);
                fcL[o] = fcL[o] / (norm[fcL_indx[o]] * dx2);
            }
    }
    for (int k = 0; k < numR; k++)
        for (long ixd = 0; ixd < nx; ixd++)
            for (int t = 0; t < nD; t++) {
                long o = t + (long)nD * (k + (long)numR * ixd);
                fcL[o] *= lambda[k + (long)numR * fcL_indx[o]] * dx2;
            }
    free(buf);
    free(norm);
}

/* getIndRule.cu:50-90: bilinear interpolation rule for the pulled-back omega. */
void ref_getIndRule(int N, const double* x, const double* Omega, int numR, int* ind_interp,
                    double* coeff_interp)
{
    long nx = (long)N * N;
    double dx = x[2] - x[0], dx2 = dx * dx;
    double L = x[2 * (N - 1)] - x[0] + dx;
    for (int k = 0; k < numR; k++)
        for (long ix = 0; ix < nx; ix++) {
            long o = k + (long)numR * ix;
            const double* Om = Omega + 2 * o;
            int* ii = ind_interp + 4 * o;
            double* cc = coeff_interp + 4 * o;
            if (isnan(Om[0])) {
                for (int i = 0; i < 4; i++) { ii[i] = 0; cc[i] = NAN; }
                continue;
            }
            int x1 = (int)((Om[0] / L + 0.5) * N);
            int x2 = (int)((Om[1] / L + 0.5) * N);
            ii[0] = x1 + x2 * N;
            ii[1] = ii[0] + 1;
            ii[2] = ii[0] + N;
            ii[3] = ii[2] + 1;
            double g10 = x[2 * x1], g11 = x[2 * (x1 + 1)];
            double g20 = x[2 * x2 * N + 1], g21 = x[2 * (x2 + 1) * N + 1];
            double a = Om[0] - g10, bq = g11 - Om[0], c = Om[1] - g20, e = g21 - Om[1];
            cc[0] = bq * e / dx2;
            cc[1] = a * e / dx2;
            cc[2] = bq * c / dx2;
            cc[3] = a * c / dx2;
        }
}

/* propagate.cu:6-121,181-220: fin = sum_k f(R, src_k) fcL_k; df = fin - lambda f
 * on the collision region, 0 elsewhere.  f is the grid (nR, nx). */
void ref_discrete_noise(int B, int N, const double* f, const double* lambda, const int* indR,
                        int numR, const double* fcL, const short* fcL_indx, int nD, double* df)
{
    long nR = 8L * B * B * B, nx = (long)N * N;
    memset(df, 0, sizeof(double) * nR * nx);
    for (int k = 0; k < numR; k++)
        for (long ix = 0; ix < nx; ix++) {
            double s = 0.0;
            for (int t = 0; t < nD; t++) {
                long o = t + (long)nD * (k + (long)numR * ix);
                s += f[indR[k] + nR * fcL_indx[o]] * fcL[o];
            }
            long idf = indR[k] + nR * ix;
            df[idf] = s - f[idf] * lambda[k + (long)numR * ix];
        }
}

/* propagate.cu:123-229,189-211,222-229: no-noise law, bilinear pullback times
 * lambda(R), minus lambda f on the listed contact-region omega indices. */
void ref_discrete_nonoise(int B, int N, const double* f, const double* lambda, const int* indR,
                          int numR, const int* const* lambda_indx, const int* lambda_numx,
                          const int* ind_interp, const double* coeff_interp, double* df)
{
    long nR = 8L * B * B * B, nx = (long)N * N;
    memset(df, 0, sizeof(double) * nR * nx);
    for (int k = 0; k < numR; k++)
        for (long ix = 0; ix < nx; ix++) {
            long o = 4 * (k + (long)numR * ix);
            long idf = indR[k] + nR * ix;
            if (isnan(coeff_interp[o])) { df[idf] = 0; continue; }
            double s = 0.0;
            for (int i = 0; i < 4; i++) s += f[indR[k] + nR * ind_interp[o + i]] * coeff_interp[o + i];
            df[idf] = s * lambda[k];
        }
    for (int k = 0; k < numR; k++)
        for (int t = 0; t < lambda_numx[k]; t++) {
            long idf = indR[k] + nR * lambda_indx[k][t];
            df[idf] -= f[idf] * lambda[k];
        }
}

/* ----------------------------------------------------------- statistics ---- */

/* get_stat (pendulum_reduced.m:485-496, pendulum.m:547-558): rotation
 * marginal fR, E[R], omega marginal fx, E[x], Var[x]. */
void ref_stat_moments(int B, int dim, int N, double L, const double* f, const double* R,
                      const double* x, const double* w, double* ER, double* Ex, double* Varx,
                      double* fx)
{
    long n2 = 2 * B, nR = n2 * n2 * n2, nx = ipow(N, dim);
    double dv = pow(L / N, dim);
    for (int i = 0; i < 9; i++) ER[i] = 0;
    for (long iR = 0; iR < nR; iR++) {
        double s = 0.0;
        for (long ix = 0; ix < nx; ix++) s += f[iR + nR * ix];
        s *= dv;
        double wj = w[(iR / n2) % n2];
        for (int i = 0; i < 9; i++) ER[i] += R[9 * iR + i] * s * wj;
    }
    for (long ix = 0; ix < nx; ix++) {
        double s = 0.0;
        for (long iR = 0; iR < nR; iR++) s += f[iR + nR * ix] * w[(iR / n2) % n2];
        fx[ix] = s;
    }
    for (int p = 0; p < dim; p++) {
        double s = 0.0;
        for (long ix = 0; ix < nx; ix++) s += x[dim * ix + p] * fx[ix];
        Ex[p] = s * dv;
    }
    for (int p = 0; p < dim; p++)
        for (int q = 0; q < dim; q++) {
            double s = 0.0;
            for (long ix = 0; ix < nx; ix++) s += x[dim * ix + p] * x[dim * ix + q] * fx[ix];
            Varx[p + dim * q] = s * dv - Ex[p] * Ex[q];
        }
}

/* ------------------------------------------------- density-form variant ---- */

/* pendulum_reduced_den.m:397-533: the reduced derivative with the gravity
 * moment evaluated pseudo-spectrally, -sum_p c(k_p) fftSO3R(f mR_p), on the
 * grid density f; f = NULL reconstructs it from F by the inverse transform
 * (:453-476; the reference keeps the complex ifftn result there, its
 * imaginary part is rounding for a real-consistent F, this restatement keeps
 * the real part).  The other terms are ref_derivative's with MR = 0. */
int ref_derivative_den(int B, int N, const double* F, const double* f, const double* X, const double* mR,
                       const double* b, const double* G, double L, const double* u, const double* d,
                       const double* w, double* dF)
{
    long M = 2 * B - 1, nrot = M * M * B, nx = (long)N * N, n2 = 2 * B, nR = n2 * n2 * n2, Bx = N / 2;
    double* MR0 = (double*)calloc((size_t)2 * nrot * 2, sizeof(double));
    ref_derivative(B, 2, N, F, X, NULL, MR0, b, G, L, u, NULL, dF, 0, 0);
    double* fg = NULL;
    if (!f) {
        fg = (double*)malloc(sizeof(double) * nR * nx);
        ref_fftSO3R_backward(B, 2, N, F, d, fg);
        f = fg;
    }
    double* prod = (double*)malloc(sizeof(double) * nR * nx);
    double* FM = (double*)malloc(sizeof(double) * 2 * nrot * nx);
    cplx* dFc = (cplx*)dF;
    const cplx* FMc = (const cplx*)FM;
    for (int p = 0; p < 2; p++) {
        for (long x = 0; x < nx; x++)
            for (long iR = 0; iR < nR; iR++) prod[iR + nR * x] = f[iR + nR * x] * mR[p + 2 * iR];
        ref_fftSO3R_forward(B, 2, N, prod, d, w, FM);
        for (long x = 0; x < nx; x++) {
            long kp = p == 0 ? x % N : x / N;
            cplx c = deriv_x(kp, Bx, L);
            for (long r = 0; r < nrot; r++) dFc[r + nrot * x] -= FMc[r + nrot * x] * c;
        }
    }
    free(prod); free(FM); free(fg); free(MR0);
    return 0;
}

/* pendulum_reduced_den.m:359-395: the first stage uses the given f, the RK
 * substeps reconstruct it from their coefficients. */
int ref_integrate_den(int B, int N, const double* F, const double* f, const double* X, const double* mR,
                      const double* b, const double* G, double dt, double L, const double* u, const double* d,
                      const double* w, const char* method, double* Fnew)
{
    long nrot = (long)(2 * B - 1) * (2 * B - 1) * B;
    long n = 2 * nrot * (long)N * N;
    int meth;
    if (!strcasecmp(method, "euler")) meth = 0;
    else if (!strcasecmp(method, "midpoint")) meth = 1;
    else if (!strcasecmp(method, "RK2")) meth = 2;
    else if (!strcasecmp(method, "RK4")) meth = 3;
    else return -1;
    double* k1 = (double*)malloc(sizeof(double) * n);
    ref_derivative_den(B, N, F, f, X, mR, b, G, L, u, d, w, k1);
    if (meth == 0) {
        for (long i = 0; i < n; i++) Fnew[i] = F[i] + dt * k1[i];
        free(k1);
        return 0;
    }
    double* F2 = (double*)malloc(sizeof(double) * n);
    double* k2 = (double*)malloc(sizeof(double) * n);
    if (meth == 2) {
        for (long i = 0; i < n; i++) F2[i] = F[i] + k1[i] * dt;
        ref_derivative_den(B, N, F2, NULL, X, mR, b, G, L, u, d, w, k2);
        for (long i = 0; i < n; i++) Fnew[i] = F[i] + dt / 2 * (k1[i] + k2[i]);
        free(k1); free(F2); free(k2);
        return 0;
    }
    for (long i = 0; i < n; i++) F2[i] = F[i] + k1[i] * dt / 2;
    ref_derivative_den(B, N, F2, NULL, X, mR, b, G, L, u, d, w, k2);
    if (meth == 1) {
        for (long i = 0; i < n; i++) Fnew[i] = F[i] + dt * k2[i];
        free(k1); free(F2); free(k2);
        return 0;
    }
    double* k3 = (double*)malloc(sizeof(double) * n);
    double* k4 = (double*)malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) F2[i] = F[i] + k2[i] * dt / 2;
    ref_derivative_den(B, N, F2, NULL, X, mR, b, G, L, u, d, w, k3);
    for (long i = 0; i < n; i++) F2[i] = F[i] + k3[i] * dt;
    ref_derivative_den(B, N, F2, NULL, X, mR, b, G, L, u, d, w, k4);
    for (long i = 0; i < n; i++) Fnew[i] = F[i] + 1.0 / 6.0 * dt * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
    free(k1); free(F2); free(k2); free(k3); free(k4);
    return 0;
}
