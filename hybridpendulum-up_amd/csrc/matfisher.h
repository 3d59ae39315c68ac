// Host-side matrix-Fisher statistics: the normalizing constant of the matrix
// Fisher density on SO(3) with its derivatives, the proper SVD, and the
// moment -> proper-singular-value fit.  All FP64; these are 3x3 problems
// solved once per output time step, so they stay on the CPU.
#pragma once

#include <cstdint>
#include <functional>

#include "hpup.h"

namespace hpup {
namespace mf {

// Exponentially scaled modified Bessel functions e^{-|x|} I0(x), e^{-|x|} I1(x).
double i0e(double x);
double i1e(double x);

// c(s) = int_SO(3) exp(tr(diag(s) Q)) dQ (Haar measure of unit mass) as
//   c = C e^K,  D_i = (dc/ds_i) e^{-K},  H_ij = (d2c/ds_i ds_j) e^{-K},
// with K = max over the Kunze integrand's exponent (returned), so that no
// overflow occurs for large s.  Gauss-Kronrod adaptive quadrature of the
// one-dimensional Kunze representation c = int_{-1}^{1} 1/2 I0(a) I0(b) e^{s3 u} du,
// a = (s1-s2)(1-u)/2, b = (s1+s2)(1+u)/2, differentiated under the integral.
struct Normalizer {
    double K, C, D[3], H[9];
};
Normalizer normalizer(const double s[3]);

// Solves  dlog c / ds_i = d_i  for s (Newton on the convex function
// log c(s) - d.s from s0 with backtracking).  Returns the number of Newton
// iterations, or -1 if |residual| did not fall below tol in max_iter.
int fit_s(const double d[3], const double s0[3], double s[3], double fval[3], double tol = 1e-12,
          int max_iter = 100);

// 3x3 SVD F = U diag(sig) V^T, sig descending >= 0 (one-sided Jacobi);
// column-major 3x3 in and out.
void svd3(const double F[9], double U[9], double sig[3], double V[9]);
// Proper SVD: U, V in SO(3), F = U diag(s) V^T with s3 carrying det(F)'s sign.
void psvd(const double F[9], double U[9], double s[3], double V[9]);

// get_stat outputs (3x3 column-major; x-related blocks with leading
// dimension dim).  fit_ok = 0 fills the fit outputs with NaN.
using Stat = hpup_stat;

// Completes get_stat from the device marginals:
//   fR[iR] = sum_x f,  gR[a][iR] = sum_x x_a f,  fx[ix] = sum_R w f,
// rotation grid points by Rof(iR, R[9]), the beta index of iR (for w) and
// the omega coordinates xg[dim*ix + a]; dv = cell volume (L/N)^dim.
// fit = false stops after E[R], E[x], Var[x].
void finish_stat(int dim, int64_t nR, int64_t nx, int n2, const double* w, const double* fR, const double* gR,
                 const double* fx, const double* xg, double dv,
                 const std::function<void(int64_t, double*)>& Rof, bool fit, Stat& out);

}  // namespace mf
}  // namespace hpup
