// Plot colours of the attitude marginal (pendulum_plot_getc: getc.cpp:7-40,
// 42-61, 145-175): the rotation marginal of the grid density, its SO(3)
// transform, and the inverse transform evaluated at arbitrary Euler angles,
//   f(alpha, beta, gamma) = sum_l (2l+1) sum_{m,n} Re(F^l_mn d^l_mn(beta) e^{-i m alpha} e^{-i n gamma}).
//
// One thread per Euler triple.  d^l_mn(beta) is generated per (m, n) by the
// three-term recursion in l of Wigner_d.m (seed at l = max(|m|,|n|) on the
// wedge n <= 0, |m| <= -n), and each wedge pair serves its up to four partners
// through d_{m,n} = d_{-n,-m} = (-1)^{m-n} d_{n,m} = (-1)^{m-n} d_{-m,-n}.
// The coefficients are read at wave-uniform indices through the scalar cache.
#include <algorithm>

#include "drift.h"
#include "kernels_common.h"
#include "getc.h"

namespace hpup {

// fR[iR] = scale * sum_x f[iR + nR x]   (getc.cpp:42-61)
__global__ __launch_bounds__(256) void k_rot_marginal(const double* __restrict__ f, int64_t nR, int64_t nx,
                                                      double scale, double* __restrict__ fR) {
    const int64_t iR = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iR >= nR) return;
    double s = 0.0;
    for (int64_t x = 0; x < nx; x++) s += f[iR + nR * x];
    fR[iR] = s * scale;
}

typedef __attribute__((address_space(4))) const double* kdp;

// F: compact SO(3) coefficients (m-major), sqb[l][m + l] = sqrt((2l)! / ((l+m)! (l-m)!))
// (row stride 2B - 1); e: Euler triples (alpha, beta, gamma)
__global__ __launch_bounds__(256) void k_point_eval(const double2* __restrict__ F, int B,
                                                    const double* __restrict__ sqb, const double* __restrict__ e,
                                                    int64_t P, double* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t pc = p < P ? p : P - 1;
    const kdp Fk = (kdp)(const double*)F;
    const kdp sq = (kdp)sqb;
    const int lmax = B - 1, M = 2 * B - 1;
    const double al = e[3 * pc], be = e[3 * pc + 1], ga = e[3 * pc + 2];
    const double cb = cos(be), cb2 = cos(0.5 * be), sb2 = sin(0.5 * be);
    double acc = 0.0;
    for (int np = -lmax; np <= 0; np++) {
        const int l0 = -np;
        for (int mp = np; mp <= -np; mp++) {
            // the wedge pair's distinct partners (m, n) and their signs
            const double sg = ((mp - np) & 1) ? -1.0 : 1.0;
            int pm[4], pn[4];
            double ps[4];
            int cnt = 0;
            auto add = [&](int m, int n, double s) {
                bool seen = false;
                for (int q = 0; q < cnt; q++) seen = seen || (pm[q] == m && pn[q] == n);
                if (!seen) {
                    pm[cnt] = m;
                    pn[cnt] = n;
                    ps[cnt] = s;
                    cnt++;
                }
            };
            add(mp, np, 1.0);
            add(-np, -mp, 1.0);
            add(np, mp, sg);
            add(-mp, -np, sg);
            double phr[4], phi[4];  // s e^{-i (m alpha + n gamma)}
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (q < cnt) {
                    double sn, cs;
                    sincos(pm[q] * al + pn[q] * ga, &sn, &cs);
                    phr[q] = ps[q] * cs;
                    phi[q] = -ps[q] * sn;
                } else {
                    phr[q] = phi[q] = 0.0;
                }
            }
            // d^{l0}_{mp,np}: the wedge seed of Wigner_d.m
            double dl = sq[l0 * M + mp + l0] * pow(cb2, (double)(l0 - mp)) * pow(-sb2, (double)(l0 + mp));
            double dm1 = 0.0;
            const double m = mp, n = np;
            for (int l = l0; l <= lmax; l++) {
                if (l > l0) {
                    double dn;
                    if (l == 1) {
                        dn = cb;  // d^1_00 (l0 = 0)
                    } else {
                        const double L = l, r = sqrt((L * L - m * m) * (L * L - n * n));
                        dn = L * (2 * L - 1) / r * (cb - m * n / ((L - 1) * L)) * dl;
                        if (l >= l0 + 2)
                            dn -= sqrt(((L - 1) * (L - 1) - m * m) * ((L - 1) * (L - 1) - n * n)) / r * L / (L - 1) * dm1;
                    }
                    dm1 = dl;
                    dl = dn;
                }
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (q < cnt) {
                        const int64_t i = 2 * d_ridx(B, l, pm[q], pn[q]);
                        s += Fk[i] * phr[q] - Fk[i + 1] * phi[q];
                    }
                }
                acc += (2 * l + 1) * dl * s;
            }
        }
    }
    if (p < P) out[p] = acc;
}

void launch_rot_marginal(const double* f, int64_t nR, int64_t nx, double scale, double* fR, hipStream_t st) {
    hipLaunchKernelGGL(k_rot_marginal, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, st, f, nR, nx, scale, fR);
}

void launch_point_eval(const void* F, int B, const double* sqb, const double* e, int64_t P, double* out,
                       hipStream_t st) {
    if (P <= 0) return;
    hipLaunchKernelGGL(k_point_eval, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, (const double2*)F, B, sqb,
                       e, P, out);
}

}  // namespace hpup
