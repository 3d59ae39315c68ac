/*
 * refcpu.h -- FP64 CPU restatement of the reference algorithm (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the parity oracle for the HIP path.  Only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Layout conventions (all MATLAB column-major, complex = interleaved re,im):
 *   B      rotation bandlimit, lmax = B-1, M = 2B-1
 *   grid f (alpha_i, beta_j, gamma_k, x) : i + 2B*(j + 2B*(k + 2B*x))
 *   F      (m+lmax, n+lmax, l, x)        : (m+lmax) + M*((n+lmax) + M*(l + B*x))
 *   d      (m+lmax, n+lmax, l, j)        : same as F with x -> j (beta index)
 *   u      (m+lmax, n+lmax, l, p)        : complex, p = 0,1,2
 *   x      Omega linear index k1 + N*(k2 + N*k3), natural FFT order of each k_p
 *   CG     cell {l1,l2} -> ((2l1+1)(2l2+1))^2 column-major, cell index l1 + l2*B
 */
#ifndef HPUP_REFCPU_H
#define HPUP_REFCPU_H
#ifdef __cplusplus
extern "C" {
#endif

void ref_grid(int B, double* alpha, double* beta, double* gamma, double* w);
void ref_rotations(int B, double* R);
void ref_wigner_d(double beta, int lmax, double* d);
void ref_wigner_d_jacobi(double beta, int l, int m, int n, double* out);
void ref_wigner_table(int B, double* d);
int  ref_clebsch_gordan(int l1, int l2, double* CG);
void ref_getu(int lmax, double* u);

void ref_fftSO3R_forward(int B, int dim, int N, const double* f, const double* d,
                         const double* w, double* F);
void ref_fftSO3R_backward(int B, int dim, int N, const double* F, const double* d,
                          double* f);

void ref_build_X(int dim, int N, double L, double* X);
void ref_build_OJO(int N, double L, const double* Jt, double* OJO);
void ref_build_MR(int B, int dim, const double* Jt, double* MR);

int  ref_derivative(int B, int dim, int N, const double* F, const double* X,
                    const double* OJO, const double* MR, const double* b,
                    const double* G, double L, const double* u,
                    const double* const* CG, double* dF, long x0, long x1);
int  ref_integrate(int B, int dim, int N, const double* F, const double* X,
                   const double* OJO, const double* MR, const double* b,
                   const double* G, double dt, double L, const double* u,
                   const double* const* CG, const char* method, double* Fnew);

/* the same derivative and integrators on the operator's exact nonzero support
 * (refcpu_structured.c); Jt replaces OJO (the gyroscopic quadratic form is
 * built from it), NULL for no gyroscopic term.  -1: X or MR lacks the
 * structure, -2: unknown method. */
int  ref_derivative_structured(int B, int dim, int N, const double* F, const double* X,
                               const double* Jt, const double* MR, const double* b,
                               const double* G, double L, const double* u,
                               const double* const* CG, double* dF);
int  ref_integrate_structured(int B, int dim, int N, const double* F, const double* X,
                              const double* Jt, const double* MR, const double* b,
                              const double* G, double dt, double L, const double* u,
                              const double* const* CG, const char* method, double* Fnew);

/* density-form reduced derivative and integrator (pendulum_reduced_den.m) */
int  ref_derivative_den(int B, int N, const double* F, const double* f, const double* X,
                        const double* mR, const double* b, const double* G, double L,
                        const double* u, const double* d, const double* w, double* dF);
int  ref_integrate_den(int B, int N, const double* F, const double* f, const double* X,
                       const double* mR, const double* b, const double* G, double dt, double L,
                       const double* u, const double* d, const double* w, const char* method,
                       double* Fnew);

/* hybrid reset, reduced (2-D omega) model */
int  ref_getLambda(int B, int N, const double* R, const double* x, double dwall,
                   double h, double r, double theta_t, double lambda_max,
                   double* lambda_full, int* indR, double* PC);
int  ref_getLambda_compact(int B, int N, const double* R, const double* x, double dwall,
                           double h, double r, double theta_t, double lambda_max,
                           double* lambda, int* indR, int* indx, int* numx, double* PC);
void ref_getOmega(int B, int N, const double* R, const double* x, const int* indR,
                  int numR, double epsilon, const double* PC, int old_dir,
                  double* Omega);
void ref_getFcL(int N, const double* x, const double* Omega, const double* lambda,
                int numR, const double* Gd, int nD, double* fcL, short* fcL_indx);
void ref_getIndRule(int N, const double* x, const double* Omega, int numR,
                    int* ind_interp, double* coeff_interp);
void ref_discrete_noise(int B, int N, const double* f, const double* lambda,
                        const int* indR, int numR, const double* fcL,
                        const short* fcL_indx, int nD, double* df);
void ref_discrete_nonoise(int B, int N, const double* f, const double* lambda,
                          const int* indR, int numR, const int* const* lambda_indx,
                          const int* lambda_numx, const int* ind_interp,
                          const double* coeff_interp, double* df);

/* hybrid reset on a 3-D omega grid (refcpu_reset3.c): per-rotation rate,
 * contact point and tangent; the pulled-back omega; df on the grid */
void ref_reset3_tables(int B, const double* R, double dwall, double h, double r, double theta_t,
                       double lambda_max, double tilt, double* lam, double* PC, double* t);
void ref_reset3_pullback(const double* Rr, const double* t, double epsilon, const double* Om,
                         double* Om_old);
void ref_reset3_df(int B, int N, double L, const double* R, const double* f, const double* lam,
                   const double* PC, const double* t, double epsilon, double tilt, int jacobian,
                   double* df);

/* statistics on the grid */
void ref_stat_moments(int B, int dim, int N, double L, const double* f,
                      const double* R, const double* x, const double* w,
                      double* ER, double* Ex, double* Varx, double* fx);

#ifdef __cplusplus
}
#endif
#endif
