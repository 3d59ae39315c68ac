// Host-side FP64 tables: grid, weights, Wigner-d, Clebsch-Gordan, and the
// structured drift operator (circulant kernels, gyroscopic coefficients,
// Clebsch-Gordan and Lie-derivative term lists) that the HIP kernels apply.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

#include "hpup_internal.h"

namespace hpup {

using cd = std::complex<double>;
static const double kPi = 3.14159265358979323846;

// fftSO3R.m:28 (beta grid) and fftSO3R.m:63-66 (quadrature weights).
void host_grid(int B, std::vector<double>& beta, std::vector<double>& w) {
    beta.resize(2 * B);
    w.resize(2 * B);
    for (int j = 0; j < 2 * B; j++) {
        double bj = kPi / (4.0 * B) * (2 * j + 1);
        beta[j] = bj;
        double s = 0.0;
        for (int t = 0; t < B; t++) s += std::sin((2 * t + 1) * bj) / (2 * t + 1);
        w[j] = 1.0 / (4.0 * B * B * B) * std::sin(bj) * s;
    }
}

// Wigner_d.m:10-64: seeds at l = 0,1, the three-term recursion in l over the
// wedge n <= 0, n <= m <= -n, then the two symmetry passes (symm).
void host_wigner_d(double beta, int lmax, double* d) {
    const int M = 2 * lmax + 1;
    auto at = [&](int m, int n, int l) -> double& { return d[(m + lmax) + M * ((n + lmax) + M * l)]; };
    std::memset(d, 0, sizeof(double) * M * M * (lmax + 1));
    const double cb = std::cos(beta), c2 = std::cos(beta / 2), s2 = std::sin(beta / 2);
    auto symm = [&](int l) {
        for (int m = 1; m <= l; m++)
            for (int n = -m + 1; n <= m; n++) at(m, n, l) = at(-n, -m, l);
        for (int m = -l; m <= l - 1; m++)
            for (int n = m + 1; n <= l; n++) at(m, n, l) = at(n, m, l) * (((m - n) & 1) ? -1.0 : 1.0);
    };
    at(0, 0, 0) = 1.0;
    if (lmax < 1) return;
    at(-1, -1, 1) = c2 * c2;
    at(0, -1, 1) = -std::sqrt(2.0) * c2 * s2;
    at(0, 0, 1) = cb;
    at(1, -1, 1) = s2 * s2;
    symm(1);
    for (int l = 2; l <= lmax; l++) {
        for (int n = -l; n <= 0; n++) {
            for (int m = n; m <= -n; m++) {
                double v;
                if (n == -l) {
                    // sqrt((2l)!/((l+m)!(l-m)!)) computed as a product to stay finite
                    double lf = 0.0;
                    for (int t = 2; t <= 2 * l; t++) lf += std::log((double)t);
                    for (int t = 2; t <= l + m; t++) lf -= std::log((double)t);
                    for (int t = 2; t <= l - m; t++) lf -= std::log((double)t);
                    v = std::exp(0.5 * lf) * std::pow(c2, l - m) * std::pow(-s2, l + m);
                } else {
                    const double den = std::sqrt((double)(l * l - m * m) * (double)(l * l - n * n));
                    v = l * (2.0 * l - 1) / den * (cb - (double)m * n / (l - 1) / l) * at(m, n, l - 1);
                    if (n != -l + 1)
                        v -= std::sqrt((double)((l - 1) * (l - 1) - m * m) * (double)((l - 1) * (l - 1) - n * n)) /
                             den * l / (l - 1.0) * at(m, n, l - 2);
                }
                at(m, n, l) = v;
            }
        }
        symm(l);
    }
}

// clebsch_gordan.m:1-41: per (l, m) the null vector of the tridiagonal
// J^2 - l(l+1) block on m1 + m2 = m, last (largest m1) component positive,
// unit norm; solved by LU with partial pivoting as MATLAB's linsolve does.
// Adjoint of a model operator (for the gradient harness; no counterpart in
// the reference).  With A = sum of (row mixing) o (omega circulant) terms,
// (K T)^H = T^H K^H and the row mixing acts on the rotation index while T_p
// acts on omega, so each factor is transposed on its own: mixing r <- src
// with c becomes src <- r with conj(c) (negated for the Clebsch-Gordan terms,
// which enter as -i c F), T_p's kernel becomes conj(kappa_p(-d)), and the
// damping i b_p c_p T_p F becomes T_p^H(-i b_p c_p F), applied by the
// kernel's adj path.  In 3-D the axis-3 term (e3 + i b3 kc3) T3 F becomes
// T3^H [(e3^* - i b3 kc3) F] (the kernel's scaled axis-3 pass).
DriftSpec adjoint_spec(const DriftSpec& s) {
    HPUP_REQUIRE((s.dim == 2 || s.dim == 3) && !s.adj, HPUP_ERR_UNSUPPORTED,
                 "adjoint operator: omega dimension 2 or 3");
    DriftSpec a = s;
    a.adj = true;
    const int N = s.N, dim = s.dim;
    const int64_t nrc = nrc_of(s.B);
    // 3-D: the diagonal u_3 term conjugates; the gyroscopic terms a_p c_q T_r
    // become a_p c_q^* T_r^H = -a_p c_q T_r^H (c_q = i kc(k_q) is imaginary and
    // commutes with T_r, r != q)
    for (int64_t r = 0; r < (int64_t)a.e3.size() / 2; r++) a.e3[2 * r + 1] = -a.e3[2 * r + 1];
    for (int p = 0; p < 3; p++) a.gyro[p] = -s.gyro[p];
    for (int p = 0; p < dim; p++)
        for (int j = 0; j < N; j++) {
            const int jm = (N - j) % N;
            a.kappa[2 * (N * p + j)] = s.kappa[2 * (N * p + jm)];
            a.kappa[2 * (N * p + j) + 1] = -s.kappa[2 * (N * p + jm) + 1];
        }
    // Clebsch-Gordan terms
    std::vector<std::vector<std::pair<int32_t, std::vector<cd>>>> cg(nrc);
    for (int64_t r = 0; r < nrc; r++)
        for (int t = 0; t < s.TG; t++) {
            const int32_t src = s.cg_src[r * s.TG + t];
            if (src < 0) break;
            std::vector<cd> c(dim);
            for (int p = 0; p < dim; p++) {
                const size_t k = ((r * s.TG + t) * dim + p) * 2;
                c[p] = -std::conj(cd(s.cg_coef[k], s.cg_coef[k + 1]));
            }
            cg[src].push_back({(int32_t)r, c});
        }
    int TG = 0;
    for (auto& v : cg) TG = std::max<int>(TG, (int)v.size());
    a.TG = TG;
    a.cg_src.assign(nrc * std::max(TG, 1), -1);
    a.cg_coef.assign(nrc * std::max(TG, 1) * dim * 2, 0.0);
    for (int64_t r = 0; r < nrc; r++)
        for (size_t t = 0; t < cg[r].size(); t++) {
            a.cg_src[r * TG + t] = cg[r][t].first;
            for (int p = 0; p < dim; p++) {
                a.cg_coef[((r * TG + t) * dim + p) * 2] = cg[r][t].second[p].real();
                a.cg_coef[((r * TG + t) * dim + p) * 2 + 1] = cg[r][t].second[p].imag();
            }
        }
    // u_1, u_2 terms
    std::vector<std::vector<std::pair<int32_t, cd>>> ut(2 * nrc);
    for (int i = 0; i < 2; i++)
        for (int64_t r = 0; r < nrc; r++)
            for (int t = 0; t < s.TU; t++) {
                const int64_t k = (i * nrc + r) * s.TU + t;
                const int32_t src = s.u_src[k];
                if (src < 0) break;
                ut[i * nrc + src].push_back({(int32_t)r, std::conj(cd(s.u_coef[2 * k], s.u_coef[2 * k + 1]))});
            }
    int TU = 0;
    for (auto& v : ut) TU = std::max<int>(TU, (int)v.size());
    HPUP_REQUIRE(TU <= 4, HPUP_ERR_UNSUPPORTED, "adjoint u terms not banded");
    a.TU = TU;
    a.u_src.assign(2 * nrc * std::max(TU, 1), -1);
    a.u_coef.assign(2 * nrc * std::max(TU, 1) * 2, 0.0);
    for (int64_t i = 0; i < 2 * nrc; i++)
        for (size_t t = 0; t < ut[i].size(); t++) {
            a.u_src[i * TU + t] = ut[i][t].first;
            a.u_coef[(i * TU + t) * 2] = ut[i][t].second.real();
            a.u_coef[(i * TU + t) * 2 + 1] = ut[i][t].second.imag();
        }
    return a;
}

DriftSpec parameter_spec(const DriftSpec& s, int term) {
    const int dim = s.dim;
    HPUP_REQUIRE(term >= 2 && term < 2 + dim + dim * dim, HPUP_ERR_ARG, "parameter operator index out of range");
    DriftSpec a = s;
    const int64_t nrc = nrc_of(s.B);
    a.adj = false;
    a.TG = 0;
    a.cg_src.assign(nrc, -1);
    a.cg_coef.assign(nrc * dim * 2, 0.0);
    a.TU = 0;
    a.u_src.assign(2 * nrc, -1);
    a.u_coef.assign(2 * nrc * 2, 0.0);
    std::fill(a.e3.begin(), a.e3.end(), 0.0);
    for (int p = 0; p < 3; p++) a.b[p] = a.gyro[p] = 0.0;
    a.has_gyro = false;
    for (int i = 0; i < 9; i++) a.G[i] = 0.0;
    if (term < 2 + dim) a.b[term - 2] = 1.0;
    else a.G[term - 2 - dim] = 1.0;
    return a;
}

void host_clebsch_gordan(int l1, int l2, std::vector<double>& CG) {
    const int dim = (2 * l1 + 1) * (2 * l2 + 1);
    CG.assign((size_t)dim * dim, 0.0);
    for (int l = std::abs(l1 - l2); l <= l1 + l2; l++) {
        for (int m = -l; m <= l; m++) {
            const int lo = (m - l1 - l2 + std::abs(l1 - l2 + m)) / 2;
            const int hi = (m + l1 + l2 - std::abs(l1 - l2 - m)) / 2;
            const int nz = hi - lo + 1;
            std::vector<double> C(nz);
            if (nz == 1) {
                C[0] = 1.0;
            } else {
                std::vector<double> A((size_t)nz * nz, 0.0);
                for (int k = 0; k < nz; k++) {
                    const int a1 = lo + k, a2 = m - a1;
                    A[k + nz * k] = l1 * (l1 + 1.0) + l2 * (l2 + 1.0) + 2.0 * a1 * a2 - l * (l + 1.0);
                    if (k > 0) {
                        const int p1 = a1 - 1, p2 = m - p1;
                        const double v = std::sqrt(l1 * (l1 + 1.0) - (double)p1 * a1) *
                                         std::sqrt(l2 * (l2 + 1.0) - (double)p2 * a2);
                        A[(k - 1) + nz * k] = v;
                        A[k + nz * (k - 1)] = v;
                    }
                }
                const int n1 = nz - 1;
                std::vector<double> Ab((size_t)n1 * n1);
                for (int j = 0; j < n1; j++)
                    for (int i = 0; i < n1; i++) Ab[i + n1 * j] = A[i + nz * j];
                std::vector<double> x(n1, 0.0);
                x[n1 - 1] = -A[(nz - 2) + nz * (nz - 1)];
                for (int k = 0; k < n1; k++) {  // LU, partial pivoting
                    int p = k;
                    double mx = std::fabs(Ab[k + n1 * k]);
                    for (int i = k + 1; i < n1; i++)
                        if (std::fabs(Ab[i + n1 * k]) > mx) { mx = std::fabs(Ab[i + n1 * k]); p = i; }
                    if (p != k) {
                        for (int j = 0; j < n1; j++) std::swap(Ab[k + n1 * j], Ab[p + n1 * j]);
                        std::swap(x[k], x[p]);
                    }
                    const double dkk = Ab[k + n1 * k];
                    if (dkk == 0.0) continue;
                    for (int i = k + 1; i < n1; i++) {
                        const double f = Ab[i + n1 * k] / dkk;
                        Ab[i + n1 * k] = f;
                        for (int j = k + 1; j < n1; j++) Ab[i + n1 * j] -= f * Ab[k + n1 * j];
                        x[i] -= f * x[k];
                    }
                }
                for (int i = n1 - 1; i >= 0; i--) {
                    double s = x[i];
                    for (int j = i + 1; j < n1; j++) s -= Ab[i + n1 * j] * x[j];
                    x[i] = s / Ab[i + n1 * i];
                }
                double s = 0.0;
                for (double v : x) s += v * v;
                const double Cn = std::sqrt(1.0 / (1.0 + s));
                for (int i = 0; i < n1; i++) C[i] = x[i] * Cn;
                C[n1] = Cn;
            }
            const int col = l * l - (l2 - l1) * (l2 - l1) + l + m;
            for (int k = 0; k < nz; k++) {
                const int a1 = lo + k;
                const int row = (l1 + a1) * (2 * l2 + 1) + l2 + (m - a1);
                CG[row + (size_t)dim * col] = C[k];
            }
        }
    }
}

// ------------------------------------------------------ operator analysis --

namespace {

double max_abs(const cd* a, int64_t n) {
    double m = 0.0;
    for (int64_t i = 0; i < n; i++) m = std::max(m, std::abs(a[i]));
    return m;
}

}  // namespace

// The reference's drift operator, in the structure the kernels apply:
//  * X_p (pendulum.m:141-143) is the DFT of a coordinate ramp, nonzero only on
//    the k_p axis line, so (F (*) X_p)/nx is a 1-D circulant along axis p with
//    kernel kappa_p[j] = X_p[j e_p]/nx.
//  * OJO_p (pendulum.m:145-162) for a diagonal Jt is the DFT of a_p x_q x_r,
//    so (F (*) OJO_p)/nx = a_p T_q T_r F.
//  * MR_p (pendulum.m:166-192) is supported on degree l2 = 1, so the
//    Clebsch-Gordan product (pendulum.m:470-506) couples l1 in {l-1, l, l+1}.
//  * u_1, u_2 (getu.m:180-201) couple n -> n +- 1; u_3 is diagonal.
DriftSpec build_drift_spec_from_arrays(int B, int dim, int N, double L, const double* Xr,
                                       const double* OJOr, const double* MRr, const double* bt,
                                       const double* G, const double* ur,
                                       const double* const* CG, double tol) {
    HPUP_REQUIRE(B >= 1 && (dim == 2 || dim == 3) && N >= 2 && N % 2 == 0, HPUP_ERR_ARG,
                 "drift: need B >= 1, dim in {2,3}, even N >= 2");
    DriftSpec s;
    s.B = B; s.dim = dim; s.N = N; s.L = L;
    const int lmax = B - 1, M = 2 * B - 1;
    const int64_t nx = ipow64(N, dim), nrc = nrc_of(B), nrot = (int64_t)M * M * B;
    for (int p = 0; p < dim; p++) s.b[p] = bt[p];
    for (int i = 0; i < dim * dim; i++) s.G[i] = G[i];
    const cd* X = reinterpret_cast<const cd*>(Xr);
    const cd* MR = reinterpret_cast<const cd*>(MRr);
    const cd* u = reinterpret_cast<const cd*>(ur);

    // ---- X: circulant kernels
    s.kappa.assign(2 * N * dim, 0.0);
    for (int p = 0; p < dim; p++) {
        const cd* Xp = X + nx * p;
        const double mx = max_abs(Xp, nx);
        const int64_t st = ipow64(N, p);
        double off = 0.0;
        for (int64_t x = 0; x < nx; x++) {
            int64_t kp = (x / st) % N;
            if (x != kp * st) off = std::max(off, std::abs(Xp[x]));
        }
        HPUP_REQUIRE(off <= 0.1 * tol * mx + 1e-300, HPUP_ERR_UNSUPPORTED,
                     "X_" + std::to_string(p + 1) + " is not supported on its k_p axis line (not the DFT of a coordinate ramp)");
        for (int j = 0; j < N; j++) {
            cd k = Xp[j * st] / (double)nx;
            s.kappa[2 * (N * p + j)] = k.real();
            s.kappa[2 * (N * p + j) + 1] = k.imag();
        }
    }
    auto kap = [&](int p, int j) { return cd(s.kappa[2 * (N * p + j)], s.kappa[2 * (N * p + j) + 1]); };
    // constant real part on every axis -> the kernels' real-circulant path
    s.ramp = true;
    for (int p = 0; p < dim; p++) {
        double mx = 0.0, dev = 0.0;
        for (int j = 0; j < N; j++) mx = std::max(mx, std::abs(kap(p, j)));
        for (int j = 0; j < N; j++) dev = std::max(dev, std::abs(kap(p, j).real() - kap(p, 0).real()));
        s.c0[p] = kap(p, 0).real();
        if (!(dev <= 1e-3 * tol * mx)) s.ramp = false;
    }

    // ---- OJO: separable gyroscopic coefficients
    if (dim == 3 && OJOr) {
        const cd* O = reinterpret_cast<const cd*>(OJOr);
        for (int p = 0; p < 3; p++) {
            const cd* Op = O + nx * p;
            const int q = (p + 1) % 3, r = (p + 2) % 3;
            const double mx = max_abs(Op, nx);
            if (mx == 0.0) { s.gyro[p] = 0.0; continue; }
            // least-squares fit of Op/nx = a * kappa_q kappa_r on the k_p = 0 plane
            cd num = 0.0;
            double den = 0.0;
            for (int64_t x = 0; x < nx; x++) {
                int k[3] = {(int)(x % N), (int)((x / N) % N), (int)(x / (N * N))};
                if (k[p] != 0) continue;
                cd P = kap(q, k[q]) * kap(r, k[r]);
                num += std::conj(P) * (Op[x] / (double)nx);
                den += std::norm(P);
            }
            cd a = den > 0 ? num / den : 0.0;
            double res = 0.0, nrm = 0.0;
            for (int64_t x = 0; x < nx; x++) {
                int k[3] = {(int)(x % N), (int)((x / N) % N), (int)(x / (N * N))};
                cd pred = (k[p] == 0) ? a * kap(q, k[q]) * kap(r, k[r]) * (double)nx : cd(0.0);
                res = std::max(res, std::abs(Op[x] - pred));
                nrm = std::max(nrm, std::abs(Op[x]));
            }
            HPUP_REQUIRE(res <= tol * nrm, HPUP_ERR_UNSUPPORTED,
                         "OJO_" + std::to_string(p + 1) +
                             " is not a_p * x_q * x_r (gyroscopic term of a diagonal inertia); residual " +
                             std::to_string(res / nrm));
            HPUP_REQUIRE(std::abs(a.imag()) <= tol * std::abs(a) + 1e-300, HPUP_ERR_UNSUPPORTED,
                         "OJO gyroscopic coefficient is not real");
            s.gyro[p] = a.real();
            if (a.real() != 0.0) s.has_gyro = true;
        }
    }

    // ---- MR: only the degree-1 block may be nonzero
    double mr_max = 0.0, mr_off = 0.0;
    for (int p = 0; p < dim; p++)
        for (int l = 0; l <= lmax; l++)
            for (int n = -lmax; n <= lmax; n++)
                for (int m = -lmax; m <= lmax; m++) {
                    double v = std::abs(MR[(m + lmax) + M * ((n + lmax) + M * (l + (int64_t)B * p))]);
                    if (l == 1 && std::abs(m) <= 1 && std::abs(n) <= 1) mr_max = std::max(mr_max, v);
                    else mr_off = std::max(mr_off, v);
                }
    HPUP_REQUIRE(mr_off <= tol * mr_max + 1e-14, HPUP_ERR_UNSUPPORTED,
                 "MR has content outside degree l2 = 1 (gravity moment must be linear in R)");
    auto mr1 = [&](int p, int m2, int n2) {
        if (B < 2) return cd(0.0);
        return MR[(m2 + lmax) + M * ((n2 + lmax) + M * (1 + (int64_t)B * p))];
    };

    // ---- Clebsch-Gordan gravity terms (pendulum.m:474-502 restricted to l2 = 1)
    struct Term { int64_t src; cd c[3]; };
    std::vector<std::vector<Term>> cgt(nrc);
    int TG = 0;
    if (B >= 2 && mr_max > 0.0) {
        for (int l = 0; l <= lmax; l++)
            for (int l1 = std::max(0, l - 1); l1 <= std::min(lmax, l + 1); l1++) {
                if (!(l >= std::abs(l1 - 1) && l <= l1 + 1)) continue;
                const double* C = CG[l1 + (int64_t)B * 1];
                HPUP_REQUIRE(C != nullptr, HPUP_ERR_ARG, "CG cell {l1," + std::to_string(1) + "} missing");
                const int dimc = (2 * l1 + 1) * 3;
                const double cl = (2.0 * l1 + 1) * 3.0 / (2.0 * l + 1);
                const int col = l * l - (l1 - 1) * (l1 - 1);
                // C_m[i2, i1] = C[i1*3 + i2, col + m + l]
                auto cgv = [&](int m, int m1, int m2) -> double {
                    if (std::abs(m1) > l1 || std::abs(m2) > 1) return 0.0;
                    return C[(int64_t)(l1 + m1) * 3 + (1 + m2) + (int64_t)dimc * (col + m + l)];
                };
                for (int n = -l; n <= l; n++)
                    for (int m = -l; m <= l; m++) {
                        const int64_t r = ridx_B(B, l, m, n);
                        for (int m2 = -1; m2 <= 1; m2++)
                            for (int n2 = -1; n2 <= 1; n2++) {
                                const int m1 = m - m2, n1 = n - n2;
                                if (std::abs(m1) > l1 || std::abs(n1) > l1) continue;
                                const double cc = cl * cgv(m, m1, m2) * cgv(n, n1, n2);
                                if (cc == 0.0) continue;
                                Term t{ridx_B(B, l1, m1, n1), {0.0, 0.0, 0.0}};
                                double mag = 0.0;
                                for (int p = 0; p < dim; p++) {
                                    t.c[p] = cc * mr1(p, m2, n2);
                                    mag = std::max(mag, std::abs(t.c[p]));
                                }
                                if (mag <= 1e-13 * mr_max) continue;
                                bool merged = false;
                                for (auto& e : cgt[r])
                                    if (e.src == t.src) {
                                        for (int p = 0; p < dim; p++) e.c[p] += t.c[p];
                                        merged = true;
                                    }
                                if (!merged) cgt[r].push_back(t);
                            }
                        TG = std::max<int>(TG, (int)cgt[r].size());
                    }
            }
    }
    s.TG = TG;
    s.cg_src.assign(nrc * std::max(TG, 1), -1);
    s.cg_coef.assign(nrc * std::max(TG, 1) * dim * 2, 0.0);
    for (int64_t r = 0; r < nrc; r++)
        for (size_t t = 0; t < cgt[r].size(); t++) {
            s.cg_src[r * TG + t] = (int32_t)cgt[r][t].src;
            for (int p = 0; p < dim; p++) {
                s.cg_coef[((r * TG + t) * dim + p) * 2] = cgt[r][t].c[p].real();
                s.cg_coef[((r * TG + t) * dim + p) * 2 + 1] = cgt[r][t].c[p].imag();
            }
        }

    // ---- u: Lie-derivative terms  out_p[l,m,n] = -sum_j u_p^l[n,j] F[l,m,j]
    auto uval = [&](int p, int l, int a, int bb) {
        return u[(a + lmax) + M * ((bb + lmax) + M * (l + (int64_t)B * p))];
    };
    double umax = 0.0;
    for (int64_t i = 0; i < nrot * 3; i++) umax = std::max(umax, std::abs(u[i]));
    const double utol = 1e-4 * tol * std::max(umax, 1.0);
    int TU = 0;
    std::vector<std::vector<std::pair<int64_t, cd>>> ut(2 * nrc);
    for (int p = 0; p < 2; p++)
        for (int l = 0; l <= lmax; l++)
            for (int n = -l; n <= l; n++)
                for (int m = -l; m <= l; m++) {
                    auto& v = ut[p * nrc + ridx_B(B, l, m, n)];
                    for (int j = -l; j <= l; j++) {
                        cd c = uval(p, l, n, j);
                        if (std::abs(c) > utol) v.push_back({ridx_B(B, l, m, j), -c});
                    }
                    TU = std::max<int>(TU, (int)v.size());
                }
    HPUP_REQUIRE(TU <= 4, HPUP_ERR_UNSUPPORTED, "u_1/u_2 are not banded (more than 4 terms per row)");
    s.TU = TU;
    s.u_src.assign(2 * nrc * std::max(TU, 1), -1);
    s.u_coef.assign(2 * nrc * std::max(TU, 1) * 2, 0.0);
    for (int64_t i = 0; i < 2 * nrc; i++)
        for (size_t t = 0; t < ut[i].size(); t++) {
            s.u_src[i * TU + t] = (int32_t)ut[i][t].first;
            s.u_coef[(i * TU + t) * 2] = ut[i][t].second.real();
            s.u_coef[(i * TU + t) * 2 + 1] = ut[i][t].second.imag();
        }
    s.e3.assign(nrc * 2, 0.0);
    if (dim == 3) {
        for (int l = 0; l <= lmax; l++)
            for (int n = -l; n <= l; n++)
                for (int j = -l; j <= l; j++) {
                    cd c = uval(2, l, n, j);
                    if (j != n) {
                        HPUP_REQUIRE(std::abs(c) <= utol, HPUP_ERR_UNSUPPORTED, "u_3 is not diagonal");
                        continue;
                    }
                    for (int m = -l; m <= l; m++) {
                        s.e3[2 * ridx_B(B, l, m, n)] = -c.real();
                        s.e3[2 * ridx_B(B, l, m, n) + 1] = -c.imag();
                    }
                }
    }
    return s;
}

// ------------------------------------------------ operator from the model --

namespace {

// naive in-place DFT along one axis (sign -1 forward), used for small host tables
void dft_axis(std::vector<cd>& a, int64_t inner, int64_t n, int64_t outer, int sign) {
    std::vector<cd> tw(n), tmp(n);
    for (int64_t k = 0; k < n; k++) tw[k] = std::polar(1.0, sign * 2.0 * kPi * (double)k / (double)n);
    for (int64_t o = 0; o < outer; o++)
        for (int64_t i = 0; i < inner; i++) {
            cd* base = a.data() + o * n * inner + i;
            for (int64_t k = 0; k < n; k++) {
                cd s = 0.0;
                for (int64_t j = 0; j < n; j++) s += base[j * inner] * tw[(j * k) % n];
                tmp[k] = s;
            }
            for (int64_t k = 0; k < n; k++) base[k * inner] = tmp[k];
        }
}

void inv3(const double* J, double* Ji) {
    const double a = J[0], b = J[3], c = J[6], d = J[1], e = J[4], f = J[7], g = J[2], h = J[5], k = J[8];
    const double det = a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g);
    HPUP_REQUIRE(det != 0.0, HPUP_ERR_ARG, "Jt is singular");
    const double t[9] = {(e * k - f * h) / det, -(d * k - f * g) / det, (d * h - e * g) / det,
                         -(b * k - c * h) / det, (a * k - c * g) / det, -(a * h - b * g) / det,
                         (b * f - c * e) / det, -(a * f - c * d) / det, (a * e - b * d) / det};
    std::memcpy(Ji, t, sizeof(t));
}

}  // namespace

// The drivers' recipe (pendulum_reduced.m:109-166,169-182; pendulum.m:113-208):
// X from the coordinate ramps, OJO from Jt, MR as the SO(3) transform of
// Jt^{-1}[R32; -R31; 0], u = getu(lmax), CG{l1,1} -- then the same analysis.
DriftSpec build_drift_spec_from_model(const hpup_model& md) {
    const int B = md.B, dim = md.dim, N = md.N, lmax = B - 1, M = 2 * B - 1;
    HPUP_REQUIRE(B >= 1 && (dim == 2 || dim == 3) && N >= 2 && N % 2 == 0, HPUP_ERR_ARG,
                 "model: need B >= 1, dim in {2,3}, even N");
    const int64_t nx = ipow64(N, dim), n2 = 2 * B, nR = n2 * n2 * n2;
    const double L = md.L;
    // X
    std::vector<cd> X(nx * dim);
    for (int p = 0; p < dim; p++) {
        for (int64_t x = 0; x < nx; x++) X[nx * p + x] = -L / 2 + L / N * ((x / ipow64(N, p)) % N);
        std::vector<cd> a(X.begin() + nx * p, X.begin() + nx * (p + 1));
        for (int q = 0; q < dim; q++) dft_axis(a, ipow64(N, q), N, ipow64(N, dim - 1 - q), -1);
        std::copy(a.begin(), a.end(), X.begin() + nx * p);
    }
    double Ji[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::vector<cd> OJO;
    if (dim == 3) {
        inv3(md.Jt, Ji);
        OJO.assign(nx * 3, 0.0);
        for (int64_t x = 0; x < nx; x++) {
            double v[3], Jv[3], cr[3];
            for (int p = 0; p < 3; p++) v[p] = -L / 2 + L / N * ((x / ipow64(N, p)) % N);
            for (int r = 0; r < 3; r++) Jv[r] = md.Jt[r] * v[0] + md.Jt[r + 3] * v[1] + md.Jt[r + 6] * v[2];
            cr[0] = v[1] * Jv[2] - v[2] * Jv[1];
            cr[1] = v[2] * Jv[0] - v[0] * Jv[2];
            cr[2] = v[0] * Jv[1] - v[1] * Jv[0];
            for (int r = 0; r < 3; r++) OJO[x + nx * r] = -(Ji[r] * cr[0] + Ji[r + 3] * cr[1] + Ji[r + 6] * cr[2]);
        }
        for (int p = 0; p < 3; p++) {
            std::vector<cd> a(OJO.begin() + nx * p, OJO.begin() + nx * (p + 1));
            for (int q = 0; q < 3; q++) dft_axis(a, ipow64(N, q), N, ipow64(N, 2 - q), -1);
            std::copy(a.begin(), a.end(), OJO.begin() + nx * p);
        }
    }
    // MR: SO(3) transform (nx = 1) of mR_p on the rotation grid
    std::vector<double> beta, w;
    host_grid(B, beta, w);
    std::vector<double> d((size_t)M * M * B * n2);
    for (int j = 0; j < n2; j++) host_wigner_d(beta[j], lmax, d.data() + (size_t)j * M * M * B);
    std::vector<cd> MR((size_t)M * M * B * dim, 0.0);
    for (int p = 0; p < dim; p++) {
        std::vector<cd> g(nR);
        for (int64_t k = 0; k < n2; k++)
            for (int64_t j = 0; j < n2; j++)
                for (int64_t i = 0; i < n2; i++) {
                    const double b_ = beta[j], gm = kPi / B * k;
                    const double r32 = std::sin(b_) * std::sin(gm), r31 = -std::sin(b_) * std::cos(gm);
                    const double v[3] = {r32, -r31, 0.0};
                    g[i + n2 * (j + n2 * k)] = (dim == 3) ? Ji[p] * v[0] + Ji[p + 3] * v[1] + Ji[p + 6] * v[2] : v[p];
                }
        dft_axis(g, 1, n2, n2 * n2, -1);
        dft_axis(g, n2 * n2, n2, 1, -1);
        for (int l = 0; l <= lmax; l++)
            for (int n = -l; n <= l; n++)
                for (int m = -l; m <= l; m++) {
                    cd sacc = 0.0;
                    const int64_t ia = (-m + n2) % n2, ig = (-n + n2) % n2;
                    for (int64_t j = 0; j < n2; j++)
                        sacc += w[j] * d[(m + lmax) + M * ((n + lmax) + M * (l + (int64_t)B * j))] *
                                g[ia + n2 * (j + n2 * ig)];
                    MR[(m + lmax) + M * ((n + lmax) + M * (l + (int64_t)B * p))] = sacc;
                }
    }
    // u (getu.m complex branch)
    std::vector<cd> u((size_t)M * M * B * 3, 0.0);
    auto U = [&](int a, int bb, int l, int p) -> cd& { return u[(a + lmax) + M * ((bb + lmax) + M * (l + (int64_t)B * p))]; };
    const cd I1(0.0, 1.0);
    for (int l = 0; l <= lmax; l++) {
        for (int n = -l; n <= l; n++) {
            if (n > -l) U(n - 1, n, l, 0) = -0.5 * I1 * std::sqrt((double)(l + n) * (l - n + 1));
            if (n < l) U(n + 1, n, l, 0) = -0.5 * I1 * std::sqrt((double)(l - n) * (l + n + 1));
            if (n > -l) U(n - 1, n, l, 1) = 0.5 * std::sqrt((double)(l + n) * (l - n + 1));
            if (n < l) U(n + 1, n, l, 1) = -0.5 * std::sqrt((double)(l - n) * (l + n + 1));
            U(n, n, l, 2) = -I1 * (double)n;
        }
    }
    // CG cells {l1, 1}
    std::vector<std::vector<double>> cells(B);
    std::vector<const double*> ptr((size_t)B * B, nullptr);
    if (B >= 2)
        for (int l1 = 0; l1 < B; l1++) {
            host_clebsch_gordan(l1, 1, cells[l1]);
            ptr[l1 + (size_t)B * 1] = cells[l1].data();
        }
    double Gd[9];
    for (int i = 0; i < dim * dim; i++) Gd[i] = md.G[i];
    return build_drift_spec_from_arrays(B, dim, N, L, reinterpret_cast<const double*>(X.data()),
                                        dim == 3 ? reinterpret_cast<const double*>(OJO.data()) : nullptr,
                                        reinterpret_cast<const double*>(MR.data()), md.b, Gd,
                                        reinterpret_cast<const double*>(u.data()), ptr.data());
}

}  // namespace hpup
