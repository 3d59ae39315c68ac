/*
 * hpup.h -- C ABI of the MI355X-native spectral Fokker-Planck propagator on
 * SO(3) x R^n (libhpup.so).
 *
 * Two groups of entry points:
 *
 *  1. Gateway equivalents.  One function per MEX gateway of the reference,
 *     taking the same arrays the gateway's prhs[] carry (MATLAB column-major,
 *     complex interleaved re,im as mxGetComplexDoubles/Singles returns them),
 *     plain pointers and explicit dims, writing the array the gateway would
 *     return in plhs[0] into a caller-allocated buffer.  Host memory in, host
 *     memory out, synchronous, no device state kept between calls -- exactly
 *     the reference's calling convention, so a mexFunction shim is a few
 *     lines.  Precision is selected by the name (`...32` takes
 *     float arrays) as in the reference.
 *
 *  2. Device-resident plan API.  Tables are built once, coefficients stay in
 *     HBM in the compact layout, and any number of steps run back to back on
 *     a caller-supplied HIP stream.  This is what a time loop should call.
 *
 * All functions return HPUP_OK (0) or a negative hpup_status; the message of
 * the last failure on the calling thread is returned by hpup_last_error().
 */
#ifndef HPUP_H
#define HPUP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    HPUP_OK = 0,
    HPUP_ERR_ARG = -1,          /* bad shape, rank or value */
    HPUP_ERR_UNSUPPORTED = -2,  /* operator outside the structured support (see message) */
    HPUP_ERR_DEVICE = -3,       /* HIP runtime error */
    HPUP_ERR_NOMEM = -4,        /* device allocation failed */
    HPUP_ERR_METHOD = -5        /* unknown integrator name */
} hpup_status;

typedef enum {
    HPUP_FP64 = 0,    /* complex128 storage, FP64 arithmetic */
    HPUP_FP32 = 1,    /* complex64 storage, FP32 MFMA drift convolutions */
    HPUP_BF16X3 = 2   /* complex64 storage, 3xBF16 split-product MFMA convolutions, FP32 accumulate */
} hpup_precision;

typedef enum { HPUP_EULER = 0, HPUP_MIDPOINT = 1, HPUP_RK2 = 2, HPUP_RK4 = 3 } hpup_method;

const char* hpup_last_error(void);
int hpup_version(void);
/* Parses "euler" | "midpoint" | "RK2" | "RK4" (case-insensitive); -1 if unknown. */
int hpup_parse_method(const char* method);

/* ---------------------------------------------------------------------------
 * 1. Gateway equivalents (host buffers).
 * ------------------------------------------------------------------------- */

/* F = fftSO3R_mex('forward', f, d, w)
 * Replaces propagation/mex/sources/fftSO3R/kernel.cu:11-195 (mexFunction,
 * 'forward' branch).  f real (2B,2B,2B[,N,N[,N]]), d real (2B-1,2B-1,B,2B),
 * w real (2B); dims_f/ndims describe f (ndims 3, 5 or 6).  F complex
 * (2B-1,2B-1,B[,N,...]) interleaved; entries with max(|m|,|n|) > l are set 0. */
int hpup_fftSO3R_forward(const double* f, const int64_t* dims_f, int ndims, const double* d,
                         const double* w, double* F);

/* f = fftSO3R_mex('backward', F, d)
 * Replaces propagation/mex/sources/fftSO3R/kernel.cu:196-391 ('backward'
 * branch).  dims_F/ndims describe F (ndims 3, 5 or 6). */
int hpup_fftSO3R_backward(const double* F, const int64_t* dims_F, int ndims, const double* d,
                          double* f);

/* Fnew = pendulum_propagate_reduced(F, X, MR, bt, G, dtt, L, u, CG, method)
 * Replaces propagation/mex/sources/pendulum_propagate_reduced/kernel.cu:9-262.
 * F complex (2B-1,2B-1,B,N,N); X complex (N,N,2); MR complex (2B-1,2B-1,B,2);
 * bt real(2); G real 2x2; u complex (2B-1,2B-1,B,3); CG: B*B pointers to the
 * cells CG{l1+1,l2+1} at index l1 + l2*B (only l2 = 1 is read: MR lives on
 * degree 1, and MR content elsewhere is rejected with HPUP_ERR_UNSUPPORTED).
 * Fnew same shape as F, entries with max(|m|,|n|) > l set 0. */
int hpup_pendulum_propagate_reduced(const double* F, const int64_t* dims_F, const double* X,
                                    const double* MR, const double* bt, const double* G,
                                    double dtt, double L, const double* u,
                                    const double* const* CG, const char* method, double* Fnew);
/* pendulum_propagate_reduced32: the same in single precision (FP32 MFMA path). */
int hpup_pendulum_propagate_reduced32(const float* F, const int64_t* dims_F, const float* X,
                                      const float* MR, const float* bt, const float* G,
                                      float dtt, float L, const float* u,
                                      const float* const* CG, const char* method, float* Fnew);

/* Fnew = pendulum_propagate(F, X, OJO, MR, bt, G, dtt, L, u, CG, method)
 * Replaces propagation/mex/sources/pendulum_propagate/kernel.cu:9-519 (both
 * its small and large paths; damping and noise are always applied).
 * F complex (2B-1,2B-1,B,N,N,N); X, OJO complex (N,N,N,3); MR (...,3); bt(3);
 * G 3x3. */
int hpup_pendulum_propagate(const double* F, const int64_t* dims_F, const double* X,
                            const double* OJO, const double* MR, const double* bt,
                            const double* G, double dtt, double L, const double* u,
                            const double* const* CG, const char* method, double* Fnew);
int hpup_pendulum_propagate32(const float* F, const int64_t* dims_F, const float* X,
                              const float* OJO, const float* MR, const float* bt,
                              const float* G, float dtt, float L, const float* u,
                              const float* const* CG, const char* method, float* Fnew);

/* Fnew = gyrobias_propagate(F, X, dt, u, G1, G2)
 * Replaces propagation/mex/sources/gyrobias_propagate/kernel.cu:8-120 (called by
 * gyro_bias.m:163): one Euler step of the attitude x gyro-bias density, dF =
 * -(F (*) X_p / nx) u_p^T + sum_ij G1_ij F u_i^T u_j^T + sum_ij G2_ij c_ij(k) F
 * (integrate.cu:241-443, gyro_bias.m:235-298) on the bias box L = 2.  F complex
 * (2B-1,2B-1,B,N,N,N) with dims (6 entries); X complex (N,N,N,3); u complex
 * (2B-1,2B-1,B,3); G1, G2 real 3x3.  FP64 only, as the reference. */
int hpup_gyrobias_propagate(const double* F, const int64_t* dims, const double* X, double dt, const double* u,
                            const double* G1, const double* G2, double* Fnew);

/* Fnew = pendulum_propagate_reduced_den(F, f, X, mR, bt, G, dtt, L, u, d, w, method)
 * Replaces propagation/mex/sources/pendulum_propagate_reduced_den/kernel.cu:9-324
 * (the density form the hybrid driver calls, pendulum_reduced_hybrid.m:228):
 * the gravity moment is evaluated on the grid, -c(k_p) fftSO3R(f mR_p), with f
 * the given grid density (2B,2B,2B,N,N) for the first stage and the inverse
 * transform of the stage coefficients for RK substeps.  mR real (2,2B,2B,2B);
 * d (2B-1,2B-1,B,2B) and w (2B) as for fftSO3R_mex.  FP64 only (the
 * reference's FP32 build of this gateway does not compile). */
int hpup_pendulum_propagate_reduced_den(const double* F, const int64_t* dims_F, const double* f, const double* X,
                                        const double* mR, const double* bt, const double* G, double dtt, double L,
                                        const double* u, const double* d, const double* w, const char* method,
                                        double* Fnew);

/* ---------------------------------------------------------------------------
 * Hybrid reset gateways (reduced model, 2-D omega grid).  Index arrays are
 * 0-based as the reference's C code produces and consumes them.
 * ------------------------------------------------------------------------- */

/* [lambda, lambda_indR, PC] = getLambda_mex(R, x, d, h, r, theta_t, lambda_max, false)
 * Replaces propagation/discrete/mex/sources/pendulum_reduced_collision/src/getLambda.cu:6-92
 * (gateway mexLambda.cpp).  R real (3,3,2B,2B,2B); x real (2,N,N).  Outputs:
 * lambda (numR, N, N), indR (numR) int32, PC real (3, nR).  lambda and indR
 * must hold nR*N*N and nR entries; *numR receives the collision-region size. */
int hpup_getLambda(const double* R, int B, const double* x, int N, double d, double h, double r,
                   double theta_t, double lambda_max, double* lambda, int32_t* indR, int* numR,
                   double* PC);

/* [lambda, lambda_indR, lambda_indx, PC] = getLambda_mex(R, x, d, h, r, theta_t, lambda_max, true)
 * Replaces .../getLambda.cu:94-190 (compact branch of mexLambda.cpp): lambda
 * per collision-region rotation (numR), and for region k the ascending omega
 * indices with non-negative contact velocity in indx[k*N*N + t], t < numx[k].
 * lambda, indR, numx must hold nR entries, indx nR*N*N. */
int hpup_getLambda_compact(const double* R, int B, const double* x, int N, double d, double h, double r,
                           double theta_t, double lambda_max, double* lambda, int32_t* indR, int* numR,
                           int32_t* indx, int32_t* numx, double* PC);

/* Omega = getOmega_mex(R, x, indR, epsilon, PC, 'new'|'old')
 * Replaces .../getOmega.cu:6-137 (gateway mexOmega.cpp).  Omega real (2, numR, N, N). */
int hpup_getOmega(const double* R, int B, const double* x, int N, const int32_t* indR, int numR,
                  double epsilon, const double* PC, const char* direction, double* Omega);

/* [fcL, fcL_indx] = getFcL_mex(x, Omega, lambda, Gd, nD)
 * Replaces .../getFcL.cu:11-166 (gateway mexFcL.cpp).  fcL real, fcL_indx int16,
 * both (nD, numR, N, N). */
int hpup_getFcL(const double* x, int N, const double* Omega, const double* lambda, int numR,
                const double* Gd, int nD, double* fcL, int16_t* fcL_indx);

/* [ind_interp, coeff_interp] = getIndRule_mex(x, Omega_old)
 * Replaces .../getIndRule.cu:6-90 (gateway mexIndRule.cpp); (4, numR, N, N). */
int hpup_getIndRule(const double* x, int N, const double* Omega, int numR, int32_t* ind_interp,
                    double* coeff_interp);

/* df = pendulum_reduced_discrete_propagate(true, f, lambda, lambda_indR, fcL, fcL_indx)
 * Replaces .../propagate.cu:6-121 (gateway mexPropagate.cpp, noise branch).
 * f, df real (2B,2B,2B,N,N). */
int hpup_discrete_propagate_noise(const double* f, int B, int N, const double* lambda,
                                  const int32_t* indR, int numR, const double* fcL,
                                  const int16_t* fcL_indx, int nD, double* df);

/* df = pendulum_reduced_discrete_propagate(false, f, lambda, lambda_indR, lambda_indx,
 *                                          ind_interp, coeff_interp)
 * Replaces .../propagate.cu:123-229 (gateway mexPropagate.cpp, no-noise branch).
 * lambda (numR); lambda_indx: numR pointers to int32 lists of lengths lambda_numx. */
int hpup_discrete_propagate_nonoise(const double* f, int B, int N, const double* lambda,
                                    const int32_t* indR, int numR,
                                    const int32_t* const* lambda_indx, const int32_t* lambda_numx,
                                    const int32_t* ind_interp, const double* coeff_interp,
                                    double* df);

/* ---------------------------------------------------------------------------
 * Output statistics and the matrix-Fisher fit (host FP64).
 * ------------------------------------------------------------------------- */

/* get_stat outputs.  3x3 blocks column-major; blocks involving omega have
 * leading dimension dim (Ex dim, Varx dim x dim, ExvR dim x 3, P dim x 3,
 * Miu dim, Sigma dim x dim).  S = diag(s) of the fitted proper singular
 * values.  niter: Newton iterations of the fit, -1 if it did not reach
 * |residual| <= 1e-12 in 100 (the outputs are then those of the last
 * iterate, as the reference keeps them after its warning).  fit_ok = 0: a
 * fit output was not finite, and S, EvR, EvRvR, ExvR, P, Miu, Sigma are NaN
 * (the reference's catch branch). */
typedef struct {
    double ER[9], Ex[3], Varx[9], EvR[3], ExvR[9], EvRvR[9];
    double U[9], S[9], V[9], P[9], Miu[3], Sigma[9];
    int niter, fit_ok;
} hpup_stat;

/* c = pendulum_plot_getc(f, e, L)
 * Replaces pendulum_plot_getc (propagation/mex/sources/pendulum_plot_getc/src/mexFunction.cpp:5-33,
 * getc.cpp:7-40 get_color, 42-61 get_margin, 63-143 fftSO3, 145-175 ifftSO3_single), called by
 * pendulum_reduced_den.m:230.  f real (2B,2B,2B,N,N[,N]) with dims/ndims (5 or 6); e real Euler
 * angles (alpha, beta, gamma) with edims = (3, ns, na, 3), nedims = 4; L the omega box length.
 * The rotation marginal fR = (L/N)^(ndims-3) sum_x f is transformed (SO(3), FP64) and its inverse
 * transform is evaluated at every Euler triple; c (ns, 3) receives
 * c[s + ns k] = sum_a fR(e(:, s, a, k)). */
int hpup_pendulum_plot_getc(const double* f, const int64_t* dims, int ndims, const double* e, const int64_t* edims,
                            int nedims, double L, double* c);

/* [ER, Ex, Varx, EvR, ExvR, EvRvR, U, S, V, P, Miu, Sigma(, fx)] = get_stat(f, R, x, w)
 * Replaces the drivers' local function get_stat (propagation/pendulum_reduced.m:485-532,
 * pendulum.m:547-594, pendulum_reduced_den.m:553-600 with the fx output).
 * f real (2B,2B,2B,N,N[,N]) with dims/ndims (5 or 6); R real (3,3,2B,2B,2B);
 * x real (dim,N,..,N); w real (2B).  The grid sums run on the device in one
 * pass over f; fx (N^dim, may be NULL) receives the omega marginal. */
int hpup_get_stat(const double* f, const int64_t* dims, int ndims, const double* R, const double* x,
                  const double* w, hpup_stat* out, double* fx);

/* [U, S, V] = psvd(F): proper SVD, U, V in SO(3), F = U S V^T, S diagonal
 * with S33 carrying the sign of det F.  Replaces matrix Fisher/psvd.m:1-13. */
int hpup_psvd(const double* F, double* U, double* S, double* V);

/* [c, dc, ddc] = pdf_MF_normal(s, bool_scaled, 1, 1): normalizing constant of
 * the matrix Fisher density on SO(3) with proper singular values s, its
 * gradient (3) and Hessian (3x3); bool_scaled = 1 multiplies all three by
 * exp(-sum(s)) as the reference's scaled form does (c_bar and its
 * derivatives).  dc and ddc may be NULL.  Replaces matrix Fisher/pdf_MF_normal.m:1-211. */
int hpup_pdf_MF_normal(const double* s, int bool_scaled, double* c, double* dc, double* ddc);

/* [s, FVAL, NITER] = pdf_MF_M2S(d, s0): proper singular values whose first
 * moments E[Q_ii] equal d.  Replaces matrix Fisher/pdf_MF_M2S.m:1-102 (same
 * root and stopping rule max|FVAL| <= 1e-12, 100 iterations).  s0 may be
 * NULL (zeros).  *niter = -1 if not converged. */
int hpup_pdf_MF_M2S(const double* d, const double* s0, double* s, double* fval, int* niter);

/* ---------------------------------------------------------------------------
 * 2. Device-resident plan API.
 * ------------------------------------------------------------------------- */

typedef struct hpup_plan hpup_plan;

typedef struct {
    int B;           /* rotation bandlimit */
    int dim;         /* omega dimension: 0 (pure SO(3)), 2 (reduced) or 3 */
    int N;           /* omega grid points per axis (2*Bx) */
    double L;        /* omega box length (periodic on [-L/2, L/2)) */
    double b[3];     /* damping bt */
    double G[9];     /* noise matrix, dim x dim column-major (leading dim = dim) */
    double Jt[9];    /* scaled inertia (3-D), column-major 3x3; ignored for dim 2 */
    double dt;       /* time step */
    int method;      /* hpup_method */
    int precision;   /* hpup_precision */
} hpup_model;

/* Builds every table (Wigner-d by FP64 recursion, weights, X/OJO/MR operator
 * structure, Clebsch-Gordan terms) and allocates device work space. */
int hpup_plan_create(const hpup_model* model, hpup_plan** plan);
int hpup_plan_destroy(hpup_plan* plan);

/* Operator plans for the gradient harness.  No counterpart in the reference
 * (it has no adjoint); callers: gradient computations over the plan API.
 *   kind 0: the model's derivative operator A (as hpup_plan_create);
 *   kind 1: its adjoint A^H (dim 2 or 3, N <= 32): hpup_plan_derivative
 *           applies A^H, hpup_plan_step applies P(dt A^H), the adjoint of the
 *           forward step P(dt A) for every integrator (P has real
 *           coefficients); hpup_plan_adjoint_stage needs dim 2;
 *   kind 2 + p (p < dim): dA/db_p;  kind 2 + dim + i + dim j: dA/dG_ij
 *           (A is linear in b and G), for hpup_plan_derivative. */
int hpup_plan_create_operator(const hpup_model* model, int kind, hpup_plan** plan);

/* out = F0 + scale * A in: one Horner stage of the plan's operator (kind 0
 * or 1), the stage its integrators run (no counterpart in the reference;
 * callers: gradient sweeps recomputing forward stage values).  out may alias
 * F0 but not in. */
int hpup_plan_axpy_stage(hpup_plan* plan, const void* in, const void* F0, double scale, void* out, void* stream);

/* One reverse-mode stage of the gradient sweep, on a kind-1 (adjoint) plan
 * (no counterpart in the reference; callers: gradient computations).  For a
 * forward stage y_j = F + scale A y_{j-1} with incoming adjoint ybar (= ybar_j)
 * and forward stage value y (= y_{j-1}), all device arrays of the plan's
 * coefficient count:
 *   ybar_out = scale * A^H ybar,   fbar += ybar_out (fbar nullable: the sweep's
 *   running sum of stage adjoints), or fbar = ybar + ybar_out when fbar_init
 *   is nonzero (a step's first reverse stage, whose running sum starts at its
 *   incoming adjoint),
 *   grad[0..4] += scale * Re<ybar, (dA/dtheta) y>, theta = b_1, b_2, G_00, G_11
 *   and G_01 (the same term as G_10); grad is a device array of 5 doubles.
 * No two of ybar_out, fbar and the inputs may alias. */
int hpup_plan_adjoint_stage(hpup_plan* plan, const void* ybar, const void* y, double scale, void* ybar_out,
                            void* fbar, int fbar_init, double* grad, void* stream);

/* The stage values of one step of a forward model plan (no counterpart in the
 * reference; callers: gradient sweeps that keep them for the reverse pass).
 * With y_0 = F and the Horner stages y_j = F + c_j dt A y_{j-1} (c = 1 Euler;
 * 1/2, 1 midpoint and RK2; 1/4, 1/3, 1/2, 1 RK4), writes y_1..y_{s-1} to
 * ys[0..s-2] (s - 1 device arrays, none of them F) and, when F_next is not
 * NULL, the step's result y_s to F_next. */
int hpup_plan_step_stages(hpup_plan* plan, const void* F, void* const* ys, void* F_next, void* stream);
/* The reverse of one step on an adjoint plan (kind 1, dim 2; no counterpart
 * in the reference; callers: gradient sweeps): from the forward stage values
 * F = y_0 and ys = y_1..y_{s-1} (as hpup_plan_step_stages writes them) and
 * the incoming adjoint lam, runs the s reverse stages of
 * hpup_plan_adjoint_stage back to back: Fbar (device, coefficient count)
 * receives the adjoint of F, grad (5 device doubles) accumulates the
 * parameter terms.  Fbar must not alias F, ys or lam. */
int hpup_plan_adjoint_step(hpup_plan* adjoint_plan, const void* F, const void* const* ys, const void* lam,
                           void* Fbar, double* grad, void* stream);

/* hpup_plan_adjoint_stage without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Same arguments, kernels and results
 * as hpup_plan_adjoint_stage; the workgroups' partial sums go to a static
 * device array of the library (4096 x 5 doubles per GPU, held until the
 * process ends) instead of a plan buffer, and launches needing more than
 * 4096 workgroups are rejected with HPUP_ERR_UNSUPPORTED.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race. */
int hpup_plan_adjoint_stage_noalloc(hpup_plan* plan, const void* ybar, const void* y, double scale,
                                    void* ybar_out, void* fbar, int fbar_init, double* grad, void* stream);

/* Number of compact coefficients: nRc * nx, nRc = B(2B-1)(2B+1)/3. */
int64_t hpup_plan_num_coeffs(const hpup_plan* plan);
/* Number of grid points: (2B)^3 * nx. */
int64_t hpup_plan_num_grid(const hpup_plan* plan);

/* Device coefficient layout: compact rotation index r outermost, m-major:
 * for m = -(B-1)..B-1, l = |m|..B-1, n = -l..l, r = moff(m) + l^2 - m^2 + (n+l)
 * with moff(m) the number of (l, m', n) triples with m' < m; omega index
 * k1 + N(k2 + N k3) innermost; complex128 (FP64) or complex64 (FP32, BF16X3).
 * Grid layout: MATLAB (alpha, beta, gamma, x) real, same precision. */

/* Forward transform: grid f_dev -> coefficients F_dev (device pointers). */
int hpup_plan_forward(hpup_plan* plan, const void* f_dev, void* F_dev, void* stream);
/* Inverse transform: coefficients F_dev -> grid f_dev. */
int hpup_plan_backward(hpup_plan* plan, const void* F_dev, void* f_dev, void* stream);
/* nsteps integrator steps in place on F_dev. */
int hpup_plan_step(hpup_plan* plan, void* F_dev, int nsteps, void* stream);
/* One derivative evaluation dF = derivative(F) (both device, compact). */
int hpup_plan_derivative(hpup_plan* plan, const void* F_dev, void* dF_dev, void* stream);

/* Layout conversion between the MATLAB full layout (2B-1,2B-1,B,nx) and the
 * device compact layout; both device pointers, plan precision. */
int hpup_plan_full_to_compact(hpup_plan* plan, const void* Ffull_dev, void* F_dev, void* stream);
int hpup_plan_compact_to_full(hpup_plan* plan, const void* F_dev, void* Ffull_dev, void* stream);

/* Per-kernel timing of the Euler step: runs nreps steps on F_dev (in place)
 * recording HIP events between the kernels of each step on `stream`, and
 * writes the average milliseconds of each kernel into ms[0..n-1], n <= max.
 * Returns n (> 0) or a negative hpup_status.  hpup_plan_profile_name(i)
 * names kernel i. */
int hpup_plan_profile(hpup_plan* plan, void* F_dev, int nreps, double* ms, int max_kernels, void* stream);
const char* hpup_plan_profile_name(const hpup_plan* plan, int i);
/* Algorithmic HBM bytes one launch of kernel i must move (inputs read once,
 * outputs written once), for roofline reporting. */
double hpup_plan_profile_bytes(const hpup_plan* plan, int i);

/* Grid statistics (get_stat moments): rotation mean E[R] (3x3), omega mean
 * (dim), omega covariance (dim x dim), omega marginal fx (nx) -- all FP64
 * host outputs, from the device grid f_dev. */
int hpup_plan_stat_moments(hpup_plan* plan, const void* f_dev, double* ER, double* Ex,
                           double* Varx, double* fx, void* stream);

/* Full get_stat (moments, proper SVD, matrix-Fisher fit, MFG parameters) of
 * the device grid f_dev on the plan's own grids (rotations R(alpha, beta,
 * gamma) = Rz Ry Rz, omega x = -L/2 + L/N * index).  fx may be NULL.
 * Synchronizes `stream`. */
int hpup_plan_get_stat(hpup_plan* plan, const void* f_dev, hpup_stat* out, double* fx, void* stream);

/* ---------------------------------------------------------------------------
 * Device-resident hybrid reset (pendulum_reduced_hybrid.m:224-234 without the
 * host round trips of its gateways).
 * ------------------------------------------------------------------------- */

/* Parameters of the collision reset.  d, h, r, theta_t, lambda_max as
 * getLambda_mex takes them (getLambda.cu:192-232: wall distance, pendulum
 * length to the end disc, disc radius, half-width of the rate ramp, maximal
 * rate); epsilon the restitution coefficient as getOmega_mex takes it
 * (getOmega.cu:99-137).  tilt: the wall normal is (cos tilt, 0, sin tilt)
 * (0 is the reference's wall; dim 3 only).  jacobian: 1 multiplies the jump
 * term by 1/epsilon (the Jacobian of the map, which conserves probability;
 * dim 3 only), 0 follows the reference's law, which omits it. */
typedef struct {
    double d, h, r, theta_t, lambda_max;
    double epsilon;
    double tilt;
    int jacobian;
} hpup_reset_params;

/* Builds the reset tables of the plan's grid: the collision region (the
 * rotations with theta >= theta0 - theta_t), and per region rotation its rate,
 * contact point and tangent.  dim 2 is the reference's no-noise reduced law
 * (getLambda.cu:94-246 compact form, getOmega.cu:64-137 'old' map,
 * getIndRule.cu:50-90, propagate.cu:123-229): the chain
 * getLambda_mex(..., true) -> getOmega_mex(..., 'old') -> getIndRule_mex ->
 * pendulum_reduced_discrete_propagate(false, ...) in one device pass.  dim 3
 * extends it to a full body angular velocity (trilinear interpolation). */
int hpup_plan_set_reset(hpup_plan* plan, const hpup_reset_params* params);
/* The region's rotation indices (ascending, 0-based as lambda_indR): writes
 * up to max_rows of them into indR (nullable) and returns their number, or a
 * negative hpup_status. */
int hpup_plan_reset_region(const hpup_plan* plan, int32_t* indR, int max_rows);
/* Reset on a device grid of the plan's layout and precision.
 *   mode 0: f_out = f_in + scale df(f_in)   (f_out may be f_in)
 *   mode 1: f_out = scale df(f_in)          (zero outside the region)
 * df is the reset law's jump term (outflow and inflow on the region's rows). */
int hpup_plan_reset(hpup_plan* plan, const void* f_in, void* f_out, double scale, int mode, void* stream);
/* Adjoint of the reset (no counterpart in the reference; callers: gradient
 * computations through hybrid steps).  f_out = g + scale R^H g for the jump
 * term R (its transpose under the plain grid inner product); f_out may be g.
 * The transpose scatters with atomic adds: its last bits vary between runs. */
int hpup_plan_reset_adjoint(hpup_plan* plan, const void* g, void* f_out, double scale, void* stream);
/* Restitution derivative of the reset (no counterpart in the reference;
 * callers: gradient computations): out_dev[0] += scale <g, (dR/d epsilon) f>
 * for device grids f, g; out_dev is one device double.  The mask edges (box
 * and contact-velocity tests) are held fixed: the derivative of the jump
 * term where it is smooth in epsilon. */
int hpup_plan_reset_eps_grad(hpup_plan* plan, const void* f, const void* g, double scale, double* out_dev,
                             void* stream);
/* nsteps hybrid steps (pendulum_reduced_hybrid.m:224-234): one step of the
 * plan's integrator on F_dev, f_dev = inverse transform, f_dev += dt df(f_dev),
 * F_dev = forward transform of f_dev.  f_dev (grid) receives the last step's
 * density after the reset, the one the driver's get_stat reads.  The
 * continuous step is the plan's spectral derivative (pendulum.m's form). */
int hpup_plan_hybrid_step(hpup_plan* plan, void* F_dev, void* f_dev, int nsteps, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* HPUP_H */
