/*
 * refcpu_structured.c -- the reference's derivative on the operator's exact
 * nonzero support (FP64, OpenMP).  TEST INFRASTRUCTURE ONLY: a second CPU
 * form of the oracle, used by tests/ as a fast checker at large grids (after
 * it is checked against the dense ref_derivative at small ones) and by
 * bench.py's CPU leg as the optimized-CPU comparison.  The product never
 * loads it.
 *
 * Same operator as ref_derivative (pendulum_reduced.m:362-465,
 * pendulum.m:388-527), written on its structure instead of literally:
 *  - X_p = fftn(x_p) is nonzero only on the k_p axis line
 *    (pendulum.m:141-143), so (F (*) X_p)/nx is a 1-D circular convolution
 *    T_p along omega axis p with kernel X_p[j e_p]/nx;
 *  - ojo = -Jt^{-1}(x cross Jt x) (pendulum.m:145-158) is a quadratic form,
 *    ojo_p = sum_{i<=j} P^p_ij x_i x_j, and fftn(x_i x_j) = X_i (*) X_j / nx,
 *    so (F (*) OJO_p)/nx = sum_{i<=j} P^p_ij T_i T_j F, exactly;
 *  - MR_p lives on degree l2 = 1 (R31, R32 are degree-one functions), so the
 *    Clebsch-Gordan product (pendulum.m:470-506) couples l to l1 = l-1, l, l+1
 *    only, with the reference's coefficient cl (C_m^T MR C_n)[i1, j1] kept
 *    where it is nonzero;
 *  - u_1, u_2 are bidiagonal and u_3 diagonal (getu.m:133-163): only their
 *    nonzero entries are applied.
 * Layouts are refcpu.h's (MATLAB full layout, interleaved complex).
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#define PI 3.14159265358979323846

static long ipw(long a, int e)
{
    long r = 1;
    while (e-- > 0) r *= a;
    return r;
}

/* natural FFT order index -> signed frequency */
static double kfreq(long k, long Bx) { return (double)(k < Bx ? k : k - 2 * Bx); }

/* out = T_p in: circular convolution along omega axis p with kernel kap
 * (N complex, interleaved), on the valid rows max(|m|,|n|) <= l (the others
 * are left as they are).  One job is one omega line and one degree l: its
 * (2l+1)^2 rows stay in cache across the N taps. */
static void circ_axis(const double* in, double* out, const double* kap, int B, int N, int dim, int p)
{
    long lmax = B - 1, M = 2 * B - 1, nrot = M * M * B;
    long Np = ipw(N, p), nouter = ipw(N, dim - 1 - p), nlines = Np * nouter;
#pragma omp parallel
    {
        double* acc = (double*)malloc(sizeof(double) * 2 * M * M);
#pragma omp for schedule(dynamic, 1)
        for (long job = 0; job < nlines * B; job++) {
            long l = lmax - job % B, line = job / B;  /* large blocks first */
            long ii = line % Np, oo = line / Np, w = 2 * l + 1;
            long r0 = (lmax - l) + M * (lmax - l) + M * M * l;  /* row of (m, n) = (-l, -l) */
            for (long k = 0; k < N; k++) {
                memset(acc, 0, sizeof(double) * 2 * w * w);
                for (long a = 0; a < N; a++) {
                    long j = (k - a + N) % N;
                    double kr = kap[2 * j], ki = kap[2 * j + 1];
                    if (kr == 0.0 && ki == 0.0) continue;
                    const double* s = in + 2 * (r0 + nrot * (ii + Np * (a + (long)N * oo)));
                    for (long c = 0; c < w; c++) {
                        const double* sc = s + 2 * M * c;
                        double* ac = acc + 2 * w * c;
                        for (long r = 0; r < w; r++) {
                            double sr = sc[2 * r], si = sc[2 * r + 1];
                            ac[2 * r] += kr * sr - ki * si;
                            ac[2 * r + 1] += kr * si + ki * sr;
                        }
                    }
                }
                double* o = out + 2 * (r0 + nrot * (ii + Np * (k + (long)N * oo)));
                for (long c = 0; c < w; c++) memcpy(o + 2 * M * c, acc + 2 * w * c, sizeof(double) * 2 * w);
            }
        }
        free(acc);
    }
}

typedef struct { long src; double re, im; } term;

/* Returns 0, or -1 when X is not supported on the axis lines or MR has
 * content off degree 1 (relative 1e-12): the structure this form relies on. */
int ref_derivative_structured(int B, int dim, int N, const double* F, const double* X, const double* Jt,
                              const double* MR, const double* b, const double* G, double L, const double* u,
                              const double* const* CG, double* dF)
{
    long lmax = B - 1, M = 2 * B - 1, nrot = M * M * B, nx = ipw(N, dim), Bx = N / 2;
    long nc = nrot * nx;

    /* circulant kernels kap_p[j] = X_p[j e_p] / nx, and the support check */
    double kap[3][2 * 64];
    if (N > 64) return -1;
    for (int p = 0; p < dim; p++) {
        const double* Xp = X + 2 * nx * p;
        double mx = 0, off = 0;
        for (long x = 0; x < nx; x++) {
            double a = hypot(Xp[2 * x], Xp[2 * x + 1]);
            long kk[3] = {x % N, (x / N) % N, (x / (N * N)) % N};
            int online = 1;
            for (int q = 0; q < dim; q++)
                if (q != p && kk[q] != 0) online = 0;
            if (online) mx = a > mx ? a : mx;
            else off = a > off ? a : off;
        }
        if (off > 1e-12 * mx) return -1;
        for (long j = 0; j < N; j++) {
            long x = j * ipw(N, p);
            kap[p][2 * j] = Xp[2 * x] / (double)nx;
            kap[p][2 * j + 1] = Xp[2 * x + 1] / (double)nx;
        }
    }
    /* degree-one support of MR */
    double mrmax = 0, mroff = 0;
    for (int p = 0; p < dim; p++)
        for (long l2 = 0; l2 <= lmax; l2++)
            for (long i = 0; i < M * M; i++) {
                long o = 2 * (i + M * M * l2 + nrot * p);
                double a = hypot(MR[o], MR[o + 1]);
                if (l2 == 1) mrmax = a > mrmax ? a : mrmax;
                else mroff = a > mroff ? a : mroff;
            }
    if (mroff > 1e-12 * (mrmax > 0 ? mrmax : 1.0)) return -1;

    /* gyroscopic quadratic forms: P[p][i][j] (i <= j) coefficient of x_i x_j
     * in ojo_p = -(Jt^{-1} (x cross Jt x))_p */
    double P[3][3][3];
    memset(P, 0, sizeof(P));
    int gyro = (dim == 3 && Jt != NULL);
    if (gyro) {
        double a = Jt[0], bb = Jt[3], c = Jt[6], d = Jt[1], e = Jt[4], f = Jt[7], g = Jt[2], h = Jt[5], k = Jt[8];
        double det = a * (e * k - f * h) - bb * (d * k - f * g) + c * (d * h - e * g);
        double Ji[9] = {(e * k - f * h) / det, -(d * k - f * g) / det, (d * h - e * g) / det,
                        -(bb * k - c * h) / det, (a * k - c * g) / det, -(a * h - bb * g) / det,
                        (bb * f - c * e) / det, -(a * f - c * d) / det, (a * e - bb * d) / det};
        double o[3][3][3];  /* o[i][j][p]: ojo_p at x = e_i + e_j (i <= j), i == j: at e_i */
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) {
                double v[3] = {0, 0, 0};
                v[i] += 1;
                if (j != i) v[j] += 1;
                double Jv[3], cr[3];
                for (int r = 0; r < 3; r++) Jv[r] = Jt[r] * v[0] + Jt[r + 3] * v[1] + Jt[r + 6] * v[2];
                cr[0] = v[1] * Jv[2] - v[2] * Jv[1];
                cr[1] = v[2] * Jv[0] - v[0] * Jv[2];
                cr[2] = v[0] * Jv[1] - v[1] * Jv[0];
                for (int r = 0; r < 3; r++) o[i][j][r] = -(Ji[r] * cr[0] + Ji[r + 3] * cr[1] + Ji[r + 6] * cr[2]);
            }
        for (int p = 0; p < 3; p++)
            for (int i = 0; i < 3; i++)
                for (int j = i; j < 3; j++)
                    P[p][i][j] = (i == j) ? o[i][i][p] : o[i][j][p] - o[i][i][p] - o[j][j][p];
    }

    /* Clebsch-Gordan terms per (row, p): coefficient cl (C_m^T MR C_n)[i1,j1]
     * with its source row, kept where nonzero */
    long* cg_start = (long*)calloc((size_t)nrot * dim, sizeof(long));
    long* cg_end = (long*)calloc((size_t)nrot * dim, sizeof(long));
    long cap = 64 * nrot * dim + 64, nterm = 0;
    term* terms = (term*)malloc(sizeof(term) * cap);
    double* K = (double*)malloc(sizeof(double) * 2 * M * M);
    for (int p = 0; p < dim; p++)
        for (long l = 0; l <= lmax; l++)
            for (long n = -l; n <= l; n++)
                for (long m = -l; m <= l; m++) {
                    long row = (m + lmax) + M * ((n + lmax) + M * l);
                    cg_start[row + nrot * p] = nterm;
                    if (!CG || B < 2) continue;
                    for (long l1 = l - 1; l1 <= l + 1; l1++) {
                        long l2 = 1, d1 = 2 * l1 + 1, d2 = 3, dimc = d1 * d2;
                        if (l1 < 0 || l1 > lmax || l < labs(l1 - l2) || l > l1 + l2) continue;
                        long col = l * l - (l1 - l2) * (l1 - l2);
                        double cl = (2.0 * l1 + 1) * (2.0 * l2 + 1) / (2.0 * l + 1);
                        const double* C = CG[l1 + B * l2];
                        const double* Cm = C + dimc * (col + m + l);
                        const double* Cn = C + dimc * (col + n + l);
                        memset(K, 0, sizeof(double) * 2 * d1 * d1);
                        for (long i2 = 0; i2 < d2; i2++)
                            for (long j2 = 0; j2 < d2; j2++) {
                                long o = 2 * ((i2 - l2 + lmax) + M * ((j2 - l2 + lmax) + M * l2) + nrot * p);
                                double mr = MR[o], mi = MR[o + 1];
                                if (mr == 0.0 && mi == 0.0) continue;
                                for (long i1 = 0; i1 < d1; i1++) {
                                    double cm = Cm[i2 + d2 * i1];
                                    if (cm == 0.0) continue;
                                    for (long j1 = 0; j1 < d1; j1++) {
                                        double cn = Cn[j2 + d2 * j1];
                                        if (cn == 0.0) continue;
                                        K[2 * (i1 + d1 * j1)] += cm * mr * cn;
                                        K[2 * (i1 + d1 * j1) + 1] += cm * mi * cn;
                                    }
                                }
                            }
                        for (long j1 = 0; j1 < d1; j1++)
                            for (long i1 = 0; i1 < d1; i1++) {
                                double kr = K[2 * (i1 + d1 * j1)], ki = K[2 * (i1 + d1 * j1) + 1];
                                if (kr == 0.0 && ki == 0.0) continue;
                                if (nterm == cap) {
                                    cap *= 2;
                                    terms = (term*)realloc(terms, sizeof(term) * cap);
                                }
                                terms[nterm].src = (i1 - l1 + lmax) + M * ((j1 - l1 + lmax) + M * l1);
                                terms[nterm].re = cl * kr;
                                terms[nterm].im = cl * ki;
                                nterm++;
                            }
                    }
                    cg_end[row + nrot * p] = nterm;
                }
    free(K);

    /* nonzero entries of u_p[n, j] per (l, n): dF^l[m, n] -= sum_j T^l[m, j] u^l[n, j] */
    long* u_start = (long*)calloc((size_t)M * B * dim, sizeof(long));
    long* u_end = (long*)calloc((size_t)M * B * dim, sizeof(long));
    term* uterm = (term*)malloc(sizeof(term) * (size_t)(nrot * dim * 3 + 1));
    long nu = 0;
    for (int p = 0; p < dim; p++)
        for (long l = 0; l <= lmax; l++)
            for (long n = -l; n <= l; n++) {
                long key = (n + lmax) + M * l + M * B * p;
                u_start[key] = nu;
                for (long j = -l; j <= l; j++) {
                    long o = 2 * ((n + lmax) + M * ((j + lmax) + M * (l + B * p)));
                    if (u[o] == 0.0 && u[o + 1] == 0.0) continue;
                    uterm[nu].src = j + lmax;
                    uterm[nu].re = u[o];
                    uterm[nu].im = u[o + 1];
                    nu++;
                }
                u_end[key] = nu;
            }

    /* entries outside the valid triangle max(|m|,|n|) <= l stay zero */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 2 * nc; i++) dF[i] = 0.0;

    /* T_p F for every axis */
    double* TF = (double*)calloc((size_t)2 * nc * dim, sizeof(double));
    for (int p = 0; p < dim; p++) circ_axis(F, TF + 2 * nc * p, kap[p], B, N, dim, p);

    /* pointwise part: kinematic (u), damping, gravity, diffusion */
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nx; x++) {
        long kx[3] = {x % N, (x / N) % N, dim == 3 ? (x / (N * N)) % N : 0};
        double cim[3], dd[3][3];
        for (int p = 0; p < dim; p++) cim[p] = kx[p] == Bx ? 0.0 : 2 * PI * kfreq(kx[p], Bx) / L;  /* c_p = i cim */
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) {
                if (i == j) {
                    double kk = kfreq(kx[i], Bx);
                    dd[i][j] = -4.0 * PI * PI * kk * kk / (L * L);
                } else {
                    dd[i][j] = -cim[i] * cim[j];
                }
            }
        double diff = 0;
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) diff += G[i + dim * j] * dd[i][j];
        const double* Fx = F + 2 * nrot * x;
        double* dFx = dF + 2 * nrot * x;
        for (long l = 0; l <= lmax; l++)
            for (long n = -l; n <= l; n++)
                for (long m = -l; m <= l; m++) {
                    long row = (m + lmax) + M * ((n + lmax) + M * l);
                    double ar = diff * Fx[2 * row], ai = diff * Fx[2 * row + 1];
                    for (int p = 0; p < dim; p++) {
                        const double* T = TF + 2 * (nc * p + nrot * x);
                        long key = (n + lmax) + M * l + M * B * p;
                        for (long t = u_start[key]; t < u_end[key]; t++) {
                            long s = (m + lmax) + M * (uterm[t].src + M * l);
                            ar -= T[2 * s] * uterm[t].re - T[2 * s + 1] * uterm[t].im;
                            ai -= T[2 * s] * uterm[t].im + T[2 * s + 1] * uterm[t].re;
                        }
                        /* W = b_p T_p F - CG_p[F]; dF += c_p W with c_p = i cim_p */
                        double wr = b[p] * T[2 * row], wi = b[p] * T[2 * row + 1];
                        for (long t = cg_start[row + nrot * p]; t < cg_end[row + nrot * p]; t++) {
                            double sr = Fx[2 * terms[t].src], si = Fx[2 * terms[t].src + 1];
                            wr -= sr * terms[t].re - si * terms[t].im;
                            wi -= sr * terms[t].im + si * terms[t].re;
                        }
                        ar -= cim[p] * wi;
                        ai += cim[p] * wr;
                    }
                    dFx[2 * row] = ar;
                    dFx[2 * row + 1] = ai;
                }
    }

    /* gyroscopic: dF -= c_p P^p_ij T_i T_j F */
    if (gyro) {
        double* tmp = (double*)calloc((size_t)2 * nc, sizeof(double));
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) {
                if (P[0][i][j] == 0.0 && P[1][i][j] == 0.0 && P[2][i][j] == 0.0) continue;
                circ_axis(TF + 2 * nc * j, tmp, kap[i], B, N, dim, i);
#pragma omp parallel for schedule(static)
                for (long x = 0; x < nx; x++) {
                    long kx[3] = {x % N, (x / N) % N, (x / (N * N)) % N};
                    double s = 0;  /* sum_p P^p_ij cim_p; the term is -i s T_iT_j F */
                    for (int p = 0; p < 3; p++) s += P[p][i][j] * (kx[p] == Bx ? 0.0 : 2 * PI * kfreq(kx[p], Bx) / L);
                    const double* t = tmp + 2 * nrot * x;
                    double* o = dF + 2 * nrot * x;
                    for (long r = 0; r < nrot; r++) {
                        o[2 * r] += s * t[2 * r + 1];
                        o[2 * r + 1] -= s * t[2 * r];
                    }
                }
            }
        free(tmp);
    }
    free(TF); free(terms); free(cg_start); free(cg_end); free(uterm); free(u_start); free(u_end);
    return 0;
}

/* ref_integrate (pendulum.m:346-385) on the structured derivative. */
int ref_integrate_structured(int B, int dim, int N, const double* F, const double* X, const double* Jt,
                             const double* MR, const double* b, const double* G, double dt, double L,
                             const double* u, const double* const* CG, const char* method, double* Fnew)
{
    long nrot = (long)(2 * B - 1) * (2 * B - 1) * B;
    long n = 2 * nrot * ipw(N, dim);
    int meth;
    if (!strcasecmp(method, "euler")) meth = 0;
    else if (!strcasecmp(method, "midpoint")) meth = 1;
    else if (!strcasecmp(method, "RK2")) meth = 2;
    else if (!strcasecmp(method, "RK4")) meth = 3;
    else return -2;
    int nst = meth == 0 ? 1 : (meth == 3 ? 4 : 2);
    double* k[4] = {NULL, NULL, NULL, NULL};
    double* Fs = (double*)malloc(sizeof(double) * n);
    int rc = 0;
    /* stage inputs: RK2 F + dt k1; midpoint F + dt/2 k1; RK4 F + dt/2 k1, F + dt/2 k2, F + dt k3 */
    for (int s = 0; s < nst && rc == 0; s++) {
        k[s] = (double*)malloc(sizeof(double) * n);
        const double* in = F;
        if (s > 0) {
            double c = (meth == 2) ? dt : ((meth == 3 && s == 3) ? dt : dt / 2);
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; i++) Fs[i] = F[i] + c * k[s - 1][i];
            in = Fs;
        }
        rc = ref_derivative_structured(B, dim, N, in, X, Jt, MR, b, G, L, u, CG, k[s]);
    }
    if (rc == 0) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            double v;
            if (meth == 0) v = F[i] + dt * k[0][i];
            else if (meth == 1) v = F[i] + dt * k[1][i];
            else if (meth == 2) v = F[i] + dt / 2 * (k[0][i] + k[1][i]);
            else v = F[i] + dt / 6.0 * (k[0][i] + 2 * k[1][i] + 2 * k[2][i] + k[3][i]);
            Fnew[i] = v;
        }
    }
    for (int s = 0; s < 4; s++) free(k[s]);
    free(Fs);
    return rc;
}
