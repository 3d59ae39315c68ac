// Internal declarations shared by the HIP sources of libhpup.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <exception>
#include <string>
#include <vector>

#include "hpup.h"

namespace hpup {

// ------------------------------------------------------------------ errors --

void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define HPUP_HIP(call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::hpup::Error{HPUP_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define HPUP_REQUIRE(cond, code, msg)                  \
    do {                                               \
        if (!(cond)) throw ::hpup::Error{(code), (msg)}; \
    } while (0)

// Runs fn, turning an Error (or another exception) into its status code and
// the thread's last-error message: the body of every extern "C" entry point.
template <class F> int guard(F&& fn) {
    try {
        fn();
        return HPUP_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return HPUP_ERR_ARG;
    }
}

// ------------------------------------------------------------ index maths --

inline int64_t nrc_of(int B) { return (int64_t)B * (2 * B - 1) * (2 * B + 1) / 3; }
// Device compact rotation index, m-major: for m = -lmax..lmax, l = |m|..lmax,
// n = -l..l.  Every coefficient the drift couples to (l, m, n) -- (l, m, n+-1)
// and (l+-1, m, n+-1) -- lies within 2B+1 rows, so neighbour rows are read
// from cache.  (The reference's MEX uses l-major blocks; the layout is
// internal and converted at the boundary.)
inline int64_t sumsq(int64_t n) { return n * (n + 1) * (2 * n + 1) / 6; }
inline int64_t moff(int B, int m) {
    const int lmax = B - 1;
    const int64_t s = m <= 0 ? sumsq(lmax) - sumsq(-m) : sumsq(lmax) + sumsq(m - 1);
    return (int64_t)(m + lmax) * B * B - s;
}
inline int64_t ridx_B(int B, int l, int m, int n) { return moff(B, m) + (int64_t)l * l - (int64_t)m * m + (n + l); }
inline int64_t ipow64(int64_t a, int e) {
    int64_t r = 1;
    while (e-- > 0) r *= a;
    return r;
}

// ------------------------------------------------------------ device buf --

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        release();
        p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0;
        return *this;
    }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{HPUP_ERR_NOMEM, "hipMalloc(" + std::to_string(n) + " bytes) failed: " + hipGetErrorString(e)};
        }
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
DevBuf upload(const std::vector<T>& v) {
    DevBuf b(v.size() * sizeof(T));
    if (!v.empty()) HPUP_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return b;
}

// R = Rz(alpha) Ry(beta) Rz(gamma) on the grid alpha_i = (pi/B) i,
// beta_j = (pi/(4B)) (2j+1), gamma_k = (pi/B) k (fftSO3R.m:27-49: the angles
// rounded as the reference forms them, and each entry the same products and
// sums as the matrix product Ra Rb Rg, whose other terms are exact zeros),
// column-major.
struct EulerGrid {
    int n2;
    std::vector<double> ca, sa, cb, sb;
    explicit EulerGrid(int B) : n2(2 * B), ca(n2), sa(n2), cb(n2), sb(n2) {
        const double pi = 3.14159265358979323846;
        for (int i = 0; i < n2; i++) {
            ca[i] = std::cos(pi / B * i);
            sa[i] = std::sin(pi / B * i);
            cb[i] = std::cos(pi / (4.0 * B) * (2 * i + 1));
            sb[i] = std::sin(pi / (4.0 * B) * (2 * i + 1));
        }
    }
    void operator()(int64_t iR, double* R) const {
        const int64_t i = iR % n2, j = (iR / n2) % n2, k = iR / ((int64_t)n2 * n2);
        const double a = ca[i], s = sa[i], b = cb[j], t = sb[j], c = ca[k], u = sa[k];
        R[0] = a * b * c - s * u;
        R[1] = s * b * c + a * u;
        R[2] = -t * c;
        R[3] = -a * b * u - s * c;
        R[4] = -s * b * u + a * c;
        R[5] = t * u;
        R[6] = a * t;
        R[7] = s * t;
        R[8] = b;
    }
};

// ----------------------------------------------------------- host tables --

// Grid and weights (fftSO3R.m:27-29,63-66), Wigner-d by FP64 recursion (Wigner_d.m:10-64).
void host_grid(int B, std::vector<double>& beta, std::vector<double>& w);
void host_wigner_d(double beta, int lmax, double* d);  // (2lmax+1)^2 (lmax+1)

// Structured drift operator, built on the host in FP64 and uploaded per precision.
struct DriftSpec {
    int B = 0, dim = 0, N = 0;
    double L = 1.0;
    double b[3] = {0, 0, 0};
    double G[9] = {0};      // dim x dim column-major
    double gyro[3] = {0, 0, 0};  // ojo_p = gyro_p * x_q * x_r  ({p,q,r} = {1,2,3})
    bool has_gyro = false;
    // circulant kernels kappa_p[j] (complex), T_p F[k] = sum_a kappa_p[(k_p - a_p) mod N] F[..a_p..]
    std::vector<double> kappa;  // 2 * N * dim
    // Every kappa_p has a constant real part c0[p] (the DFT of a coordinate
    // ramp does: Re kappa = -h/2 at every index), so T_p F = c0_p (sum of F
    // along axis p) + i (Im kappa_p circulant) F: one real circulant instead of
    // a complex one.
    bool ramp = false;
    double c0[3] = {0, 0, 0};
    // Clebsch-Gordan gravity terms per compact output r: up to TG terms (src, coef[dim] complex)
    int TG = 0;
    std::vector<int32_t> cg_src;   // nRc * TG, -1 padded
    std::vector<double> cg_coef;   // nRc * TG * dim * 2
    // Lie-derivative mixing along n (u_1, u_2): up to TU terms per output
    int TU = 0;
    std::vector<int32_t> u_src;    // 2 * nRc * TU (p = 0,1)
    std::vector<double> u_coef;    // 2 * nRc * TU * 2
    // diagonal u_3 term: e3[r] = -u_3[n,n]^l (complex), 3-D only
    std::vector<double> e3;        // nRc * 2
    // the adjoint operator A^H (see adjoint_spec)
    bool adj = false;
};

// The adjoint A^H of a reduced-model (dim 2) operator, for the gradient
// harness: mixing tables transposed with conjugate coefficients (the
// Clebsch-Gordan ones negated as well, since they enter as -i c), circulant
// kernels conj(kappa(-d)), damping moved after the mixing (adj flag).
DriftSpec adjoint_spec(const DriftSpec& s);
// The derivative of the operator with respect to one model parameter, itself
// an operator of the same form: term = 2 + p gives dA/db_p, term = 2 + dim +
// i + dim j gives dA/dG_ij (the operator is linear in b and G).
DriftSpec parameter_spec(const DriftSpec& s, int term);

// Builds the structured operator from the gateway arrays (X, OJO, MR, u, CG
// as the reference's drivers pass them).  Throws HPUP_ERR_UNSUPPORTED if an
// array falls outside the structure the kernels implement.
DriftSpec build_drift_spec_from_arrays(int B, int dim, int N, double L, const double* X,
                                       const double* OJO, const double* MR, const double* bt,
                                       const double* G, const double* u,
                                       const double* const* CG, double tol = 1e-9);
// Builds the same operator from physical parameters (the drivers' recipe).
DriftSpec build_drift_spec_from_model(const hpup_model& m);

// Clebsch-Gordan cell by the reference's tridiagonal null-vector recipe (clebsch_gordan.m).
void host_clebsch_gordan(int l1, int l2, std::vector<double>& CG);

// ------------------------------------------------------------------ plan --

struct Plan;

}  // namespace hpup

struct hpup_plan;
