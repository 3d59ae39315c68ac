// Device-resident hybrid collision reset on a plan's grid (no host copies):
// the jump term of the reset law applied to the grid density f(R, Omega),
//   df(R, Omega) = J lambda(R) f(R, g(R, Omega)) [vC(g) >= 0, g in the box]
//                  - lambda(R) f(R, Omega) [vC(Omega) >= 0],
// with the pull-back g(R, Omega) = R^T (omega - ((1 + eps)/eps)(omega . t) t),
// omega = R Omega, and f(R, g) interpolated on the omega grid.
//
// dim 2 is the reference's reduced law (no-noise branch): rate and contact
// point from getLambda.cu:192-246, the 'old' map and its box and contact
// tests from getOmega.cu:64-137, the bilinear rule of getIndRule.cu:50-90,
// and the inflow and outflow of propagate.cu:123-229, with J = 1 (the
// reference omits the Jacobian).  Omega = (x1, x2) enters as R (x1, x2, 0)
// and the pull-back keeps its first two body components.
//
// dim 3 restates the same law with a full body angular velocity and a wall
// normal e_w = (cos tau, 0, sin tau) (tau = 0 is the reference's wall):
// theta = asin(r3 . e_w), PC = (h - r tan theta) r3 + r sec theta e_w,
// t = r3 x e_w / |r3 x e_w|, vC(omega) = (omega x PC) . e_w = sigma (omega . t)
// with sigma = (PC x e_w) . t, so that vC(omega-) = -vC(omega)/eps and one
// sign decides both contact tests, independently of eps (outflow where it is
// positive, inflow where it is negative, neither on the switching plane
// where the map is the identity); trilinear interpolation of the grid values
// extended by zero (the inflow fades to zero over the cell beyond the grid's
// edge instead of stopping at the edge, so it is continuous in eps), and
// J = 1/eps (|det g|, which conserves probability) or 1.
//
// Layouts: the plan's grid (nR, nx) with the rotation index fastest; the
// region's rows k (rotations with theta >= theta0 - theta_t, ascending) and a
// compact df (numR, nx) with k fastest.  The geometry is FP64 in every
// precision, so the masks and interpolation weights do not depend on the
// plan's precision; dim 2 keeps the oracle's operation order (contraction
// off), dim 3 locates cells and weights by multiplying with 1/h (the
// interpolant is continuous across cell faces, so a face assigned to either
// neighbouring cell gives the same value to rounding).
#include <algorithm>
#include <cmath>
#include <vector>

#include "drift.h"
#include "hpup_internal.h"
#include "reset_plan.h"

namespace hpup {

namespace {

// per-region geometry, structure of arrays: [c * numR + k]
enum { G_R = 0, G_T = 9, G_PC = 12, G_LAM = 15, G_N = 16 };

// sigma = (PC x e_w) . t = (h - r tan theta) |r3 x e_w|: the contact velocity
// (omega x PC) . e_w equals sigma (omega . t)
__device__ __forceinline__ double contact_sigma(const double* P, const double* ew, const double* t) {
#pragma clang fp contract(off)
    const double c0 = P[1] * ew[2] - P[2] * ew[1], c1 = P[2] * ew[0] - P[0] * ew[2], c2 = P[0] * ew[1] - P[1] * ew[0];
    return c0 * t[0] + c1 * t[1] + c2 * t[2];
}

template <class T, int DIM>
__global__ __launch_bounds__(256) void k_reset_df(const T* __restrict__ f, const double* __restrict__ geo,
                                                  const int32_t* __restrict__ indR, int numR, int64_t nR,
                                                  ResetConsts c, T* __restrict__ dfc) {
#pragma clang fp contract(off)
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int64_t ix = blockIdx.y;
    if (k >= numR) return;
    double Rr[9], t[3], P[3];
#pragma unroll
    for (int a = 0; a < 9; a++) Rr[a] = geo[(int64_t)(G_R + a) * numR + k];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        t[a] = geo[(int64_t)(G_T + a) * numR + k];
        P[a] = geo[(int64_t)(G_PC + a) * numR + k];
    }
    const double lam = geo[(int64_t)G_LAM * numR + k];
    const int64_t iR = indR[k];
    const int N = c.N;
    const int64_t i0 = ix % N, i1 = (ix / N) % N;
    double out;
    if constexpr (DIM == 2) {
        // getLambda.cu:234-246 (outflow where vC >= 0), getOmega.cu:64-137 ('old')
        const double x1 = c.xg[i0], x2 = c.xg[i1];
        const double o0 = Rr[0] * x1 + Rr[3] * x2;
        double o1 = Rr[1] * x1 + Rr[4] * x2;
        double o2 = Rr[2] * x1 + Rr[5] * x2;
        const double vC = o1 * P[2] - o2 * P[1];
        out = vC < 0.0 ? 0.0 : -(double)f[iR + nR * ix] * lam;
        const double fac = c.kfac;
        const double ot = fac * (o1 * t[1] + o2 * t[2]);
        o1 -= ot * t[1];
        o2 -= ot * t[2];
        const double a0 = Rr[0] * o0 + Rr[1] * o1 + Rr[2] * o2;
        const double a1 = Rr[3] * o0 + Rr[4] * o1 + Rr[5] * o2;
        const double vCo = o1 * P[2] - o2 * P[1];
        if (!(vCo < 0.0 || a0 < c.xlo || a0 > c.xhi || a1 < c.xlo || a1 > c.xhi)) {
            // getIndRule.cu:50-90: cell by truncation, weights from the grid values
            const int j1 = (int)((a0 / c.Lg + 0.5) * N);
            const int j2 = (int)((a1 / c.Lg + 0.5) * N);
            if (!(j1 < 0 || j1 > N - 2 || j2 < 0 || j2 > N - 2)) {
                const double g10 = c.xg[j1], g11 = c.xg[j1 + 1], g20 = c.xg[j2], g21 = c.xg[j2 + 1];
                const double a = a0 - g10, b = g11 - a0, cc = a1 - g20, e = g21 - a1;
                const double w[4] = {b * e / c.dx2, a * e / c.dx2, b * cc / c.dx2, a * cc / c.dx2};
                const int64_t q0 = j1 + (int64_t)j2 * N;
                const int64_t q[4] = {q0, q0 + 1, q0 + N, q0 + N + 1};
                double fv[4];
#pragma unroll
                for (int u = 0; u < 4; u++) fv[u] = (double)f[iR + nR * q[u]];
                double s = 0.0;
#pragma unroll
                for (int u = 0; u < 4; u++) s += fv[u] * w[u];
                out = s * lam + out;
            }
        }
    } else {
        const int64_t i2 = ix / ((int64_t)N * N);
        const double Om[3] = {c.x0 + c.hx * (double)i0, c.x0 + c.hx * (double)i1, c.x0 + c.hx * (double)i2};
        double w[3];
#pragma unroll
        for (int a = 0; a < 3; a++) w[a] = Rr[a] * Om[0] + Rr[a + 3] * Om[1] + Rr[a + 6] * Om[2];
        // vC(omega) = sigma (omega . t) and vC(omega-) = -vC(omega)/eps: one
        // sign decides both contact tests (outflow > 0, inflow < 0; on the
        // switching plane, where the map is the identity, neither)
        const double sigma = contact_sigma(P, c.ew, t);
        const double vq = sigma * (w[0] * t[0] + w[1] * t[1] + w[2] * t[2]);
        out = vq > 0.0 ? -lam * (double)f[iR + nR * ix] : 0.0;
        const double fac = c.kfac * (w[0] * t[0] + w[1] * t[1] + w[2] * t[2]);
#pragma unroll
        for (int a = 0; a < 3; a++) w[a] -= fac * t[a];
        double Oo[3];
#pragma unroll
        for (int a = 0; a < 3; a++) Oo[a] = Rr[3 * a] * w[0] + Rr[3 * a + 1] * w[1] + Rr[3 * a + 2] * w[2];
        // trilinear interpolation of the grid values extended by zero: cells
        // c in [-1, N-1], corners outside the grid weigh nothing
        int64_t cl[3];
        double s[3];
        bool support = true;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double cf = floor((Oo[a] - c.x0) * c.ihx);
            if (!(cf >= -1.0 && cf <= (double)(N - 1))) support = false;
            cl[a] = support ? (int64_t)cf : 0;
            s[a] = (Oo[a] - (c.x0 + c.hx * (double)cl[a])) * c.ihx;
        }
        if (support && vq < 0.0) {
            double fv[8], wt[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int b0 = q & 1, b1 = (q >> 1) & 1, b2 = (q >> 2) & 1;
                const int64_t j0 = cl[0] + b0, j1 = cl[1] + b1, j2 = cl[2] + b2;
                const bool in = j0 >= 0 && j0 < N && j1 >= 0 && j1 < N && j2 >= 0 && j2 < N;
                wt[q] = (b0 ? s[0] : 1 - s[0]) * (b1 ? s[1] : 1 - s[1]) * (b2 ? s[2] : 1 - s[2]);
                fv[q] = in ? (double)f[iR + nR * (j0 + N * (j1 + (int64_t)N * j2))] : 0.0;
            }
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 8; q++) v += wt[q] * fv[q];
            out += c.J * lam * v;
        }
    }
    dfc[k + (int64_t)numR * ix] = (T)out;
}

// out[iR + nR ix] = base[...] + scale dfc[k + numR ix] on the region's rows
// (base = out for an in-place update; base = nullptr: out = scale dfc)
template <class T>
__global__ __launch_bounds__(256) void k_reset_apply(const T* __restrict__ dfc, const int32_t* __restrict__ indR,
                                                     int numR, int64_t nR, const T* base, T* out, double scale) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int64_t ix = blockIdx.y;
    if (k >= numR) return;
    const int64_t i = indR[k] + nR * ix;
    const double d = scale * (double)dfc[k + (int64_t)numR * ix];
    out[i] = base ? (T)((double)base[i] + d) : (T)d;
}

// A region row's geometry: R (column-major), t, PC, lambda and (dim 3) the
// contact factor sigma = (PC x e_w) . t.
struct RowGeo {
    double R[9], t[3], P[3], lam, sigma;
};

__device__ __forceinline__ RowGeo row_geo(const double* g, int stride, int k, const ResetConsts& c) {
    RowGeo r;
#pragma unroll
    for (int a = 0; a < 9; a++) r.R[a] = g[(int64_t)(G_R + a) * stride + k];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.t[a] = g[(int64_t)(G_T + a) * stride + k];
        r.P[a] = g[(int64_t)(G_PC + a) * stride + k];
    }
    r.lam = g[(int64_t)G_LAM * stride + k];
    r.sigma = contact_sigma(r.P, c.ew, r.t);
    return r;
}

// omega point x -> grid indices (i0, i1, i2), and their advance by a fixed
// step decomposed the same way (no division in the loops)
struct OmegaIdx {
    int i0, i1, i2;
    __device__ __forceinline__ void set(int x, int N) {
        i0 = x % N;
        const int q = x / N;
        i1 = q % N;
        i2 = q / N;
    }
    __device__ __forceinline__ void advance(const OmegaIdx& d, int N) {
        i0 += d.i0;
        const int c0 = i0 >= N ? 1 : 0;
        i0 -= c0 * N;
        i1 += d.i1 + c0;
        const int c1 = i1 >= N ? 1 : 0;
        i1 -= c1 * N;
        i2 += d.i2 + c1;
    }
};

// The jump term's pieces at one (region row, omega point (i0, i1, i2)), in
// FP64: the outflow mask, the pulled-back point with its interpolation
// corners (omega point indices) and weights (inflow mask false: no corners),
// and d(pull-back)/d(eps).
template <int DIM> struct Jump {
    static constexpr int NQ = DIM == 2 ? 4 : 8;
    bool out_on, in_on;
    int q[NQ];
    double w[NQ];
    double dw[NQ];  // d w_q / d eps through the pulled-back point
};

template <int DIM>
__device__ __forceinline__ void jump_at(const RowGeo& G, const ResetConsts& c, int i0, int i1, int i2,
                                        double deps_scale, Jump<DIM>& J) {
#pragma clang fp contract(off)
    const int N = c.N;
    const double* Rr = G.R;
    const double* t = G.t;
    const double* P = G.P;
    J.in_on = false;
    if constexpr (DIM == 2) {
        const double x1 = c.xg[i0], x2 = c.xg[i1];
        const double o0 = Rr[0] * x1 + Rr[3] * x2;
        double o1 = Rr[1] * x1 + Rr[4] * x2;
        double o2 = Rr[2] * x1 + Rr[5] * x2;
        J.out_on = !(o1 * P[2] - o2 * P[1] < 0.0);
        const double dot = o1 * t[1] + o2 * t[2];
        const double ot = c.kfac * dot;
        o1 -= ot * t[1];
        o2 -= ot * t[2];
        const double a0 = Rr[0] * o0 + Rr[1] * o1 + Rr[2] * o2;
        const double a1 = Rr[3] * o0 + Rr[4] * o1 + Rr[5] * o2;
        const double vCo = o1 * P[2] - o2 * P[1];
        if (vCo < 0.0 || a0 < c.xlo || a0 > c.xhi || a1 < c.xlo || a1 > c.xhi) return;
        const int j1 = (int)((a0 / c.Lg + 0.5) * N);
        const int j2 = (int)((a1 / c.Lg + 0.5) * N);
        if (j1 < 0 || j1 > N - 2 || j2 < 0 || j2 > N - 2) return;
        J.in_on = true;
        const double g10 = c.xg[j1], g11 = c.xg[j1 + 1], g20 = c.xg[j2], g21 = c.xg[j2 + 1];
        const double a = a0 - g10, b = g11 - a0, cc = a1 - g20, e = g21 - a1;
        J.w[0] = b * e / c.dx2;
        J.w[1] = a * e / c.dx2;
        J.w[2] = b * cc / c.dx2;
        J.w[3] = a * cc / c.dx2;
        const int q0 = j1 + j2 * N;
        J.q[0] = q0;
        J.q[1] = q0 + 1;
        J.q[2] = q0 + N;
        J.q[3] = q0 + N + 1;
        // d(a0, a1)/d eps = deps_scale (o . t) (R^T t)_{0,1}, deps_scale = -dk/deps
        const double da0 = deps_scale * dot * (Rr[1] * t[1] + Rr[2] * t[2]);
        const double da1 = deps_scale * dot * (Rr[4] * t[1] + Rr[5] * t[2]);
        J.dw[0] = (-e * da0 - b * da1) / c.dx2;
        J.dw[1] = (e * da0 - a * da1) / c.dx2;
        J.dw[2] = (-cc * da0 + b * da1) / c.dx2;
        J.dw[3] = (cc * da0 + a * da1) / c.dx2;
    } else {
        const double Om[3] = {c.x0 + c.hx * (double)i0, c.x0 + c.hx * (double)i1, c.x0 + c.hx * (double)i2};
        double w[3];
        for (int a = 0; a < 3; a++) w[a] = Rr[a] * Om[0] + Rr[a + 3] * Om[1] + Rr[a + 6] * Om[2];
        const double dot = w[0] * t[0] + w[1] * t[1] + w[2] * t[2];
        const double vq = G.sigma * dot;
        J.out_on = vq > 0.0;
        if (!(vq < 0.0)) return;
        const double fac = c.kfac * dot;
        for (int a = 0; a < 3; a++) w[a] -= fac * t[a];
        double Oo[3];
        for (int a = 0; a < 3; a++) Oo[a] = Rr[3 * a] * w[0] + Rr[3 * a + 1] * w[1] + Rr[3 * a + 2] * w[2];
        int cl[3];
        double s[3], ds[3];
        for (int a = 0; a < 3; a++) {
            const double cf = floor((Oo[a] - c.x0) * c.ihx);
            if (!(cf >= -1.0 && cf <= (double)(N - 1))) return;  // beyond the zero-extended support
            cl[a] = (int)cf;
            s[a] = (Oo[a] - (c.x0 + c.hx * (double)cl[a])) * c.ihx;
            // d Oo_a / d eps = deps_scale (w . t) (R^T t)_a
            ds[a] = deps_scale * dot * (Rr[3 * a] * t[0] + Rr[3 * a + 1] * t[1] + Rr[3 * a + 2] * t[2]) * c.ihx;
        }
        J.in_on = true;
        for (int q = 0; q < 8; q++) {
            const int b0 = q & 1, b1 = (q >> 1) & 1, b2 = (q >> 2) & 1;
            const int j0 = cl[0] + b0, j1 = cl[1] + b1, j2 = cl[2] + b2;
            const bool in = j0 >= 0 && j0 < N && j1 >= 0 && j1 < N && j2 >= 0 && j2 < N;
            const double u0 = b0 ? s[0] : 1 - s[0], u1 = b1 ? s[1] : 1 - s[1], u2 = b2 ? s[2] : 1 - s[2];
            // corners outside the grid carry the zero extension: no weight, no derivative
            J.w[q] = in ? u0 * u1 * u2 : 0.0;
            J.dw[q] = in ? (b0 ? ds[0] : -ds[0]) * u1 * u2 + u0 * (b1 ? ds[1] : -ds[1]) * u2 +
                               u0 * u1 * (b2 ? ds[2] : -ds[2])
                         : 0.0;
            J.q[q] = in ? j0 + N * (j1 + N * j2) : 0;
        }
    }
}

// Transpose of the jump term, scattered into acc (numR, nx), zeroed first:
// acc[k][x] += -lambda [out] g(x); acc[k][corner q of x] += J lambda w_q g(x).
// The scatter targets collide across omega points, so it adds atomically
// (the summation order, and so the last bits, vary between runs).
template <class T, int DIM>
__global__ __launch_bounds__(256) void k_reset_adj(const T* __restrict__ g, const double* __restrict__ geo,
                                                   const int32_t* __restrict__ indR, int numR, int64_t nR,
                                                   ResetConsts c, T* acc) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int ix = (int)blockIdx.y;
    if (k >= numR) return;
    const RowGeo G = row_geo(geo, numR, k, c);
    OmegaIdx o;
    o.set(ix, c.N);
    Jump<DIM> J;
    jump_at<DIM>(G, c, o.i0, o.i1, o.i2, 0.0, J);
    const int64_t iR = indR[k];
    const double gv = (double)g[iR + nR * ix];
    if (J.out_on) atomicAdd(acc + k + (int64_t)numR * ix, (T)(-G.lam * gv));
    if (J.in_on) {
        const double a = c.J * G.lam * gv;
#pragma unroll
        for (int q = 0; q < Jump<DIM>::NQ; q++) atomicAdd(acc + k + (int64_t)numR * J.q[q], (T)(a * J.w[q]));
    }
}

// d/d eps of the jump term, against g: per block the sum of
// g(x) [dJ/d eps lambda I(f) + J lambda dI(f)/d eps] over its (k, x) items,
// written to part[block] (FP64, fixed order: lanes, then waves)
template <class T, int DIM>
__global__ __launch_bounds__(256) void k_reset_deps(const T* __restrict__ f, const T* __restrict__ g,
                                                    const double* __restrict__ geo, const int32_t* __restrict__ indR,
                                                    int numR, int64_t nR, ResetConsts c, double dJ, double dk,
                                                    double* __restrict__ part) {
    __shared__ double red[4];
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int ix = (int)blockIdx.y;
    double v = 0.0;
    if (k < numR) {
        const RowGeo G = row_geo(geo, numR, k, c);
        const double lam = G.lam;
        OmegaIdx o;
        o.set(ix, c.N);
        Jump<DIM> J;
        jump_at<DIM>(G, c, o.i0, o.i1, o.i2, -dk, J);
        if (J.in_on) {
            const int64_t iR = indR[k];
            double I = 0.0, dI = 0.0;
#pragma unroll
            for (int q = 0; q < Jump<DIM>::NQ; q++) {
                const double fq = (double)f[iR + nR * J.q[q]];
                I += J.w[q] * fq;
                dI += J.dw[q] * fq;
            }
            v = (double)g[iR + nR * ix] * lam * (dJ * I + c.J * dI);
        }
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x + (int64_t)gridDim.x * blockIdx.y] = red[0] + red[1] + red[2] + red[3];
}

// out[0] += scale * sum of n partials, one block: strided sums per thread,
// then a fixed pairwise tree (deterministic)
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ part, int64_t n, double scale,
                                                      double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] += scale * red[0];
}

// ------------------------------------------------ row-group kernels -----
//
// One workgroup per group of up to RG consecutive region rows holds the rows'
// grid values for all nx omega points in LDS, so the interpolation gathers
// (corners anywhere in the omega grid) read LDS instead of scattered global
// addresses; the staging loads run with the row index fastest, the grid's
// own fastest index, so neighbouring lanes read neighbouring rotations.
// Groups are numbered XCD-major: the groups that run at once on one XCD are
// neighbours and share the cache lines of their rows in that XCD's L2.

__device__ __forceinline__ int xcd_major_r(int b, int nb) {
    const int per = nb >> 3;
    if (b >= per << 3) return b;
    return (b & 7) * per + (b >> 3);
}

// The lanes' split of a group: thread tid works on row r = tid % RG (RG a
// power of two dividing the block) and the omega points x = tid / RG + k S,
// S = blockDim.x / RG; neighbouring lanes hold neighbouring rows (the grid's
// fastest index) at one point.
struct RowLanes {
    int r, x0, S;
    OmegaIdx p, d;  // the first point's indices and the step's
    __device__ __forceinline__ RowLanes(int RG, int N) {
        r = (int)threadIdx.x & (RG - 1);
        x0 = (int)threadIdx.x / RG;
        S = (int)blockDim.x / RG;
        p.set(x0, N);
        d.set(S, N);
    }
};

// Rows of the group into LDS: a thread's loads are issued in batches of 8
// before their LDS stores, so the staging costs one memory round trip per
// batch instead of one per element.
template <class T>
__device__ __forceinline__ void stage_rows(const T* __restrict__ src, const int32_t* __restrict__ indR, int k0, int nr,
                                           int RG, int64_t nR, int nx, T* rows) {
    constexpr int BT = 8;
    const int r = (int)threadIdx.x & (RG - 1), S = (int)blockDim.x / RG;
    if (r >= nr) return;
    const T* s = src + indR[k0 + r];
    T* row = rows + (int64_t)r * nx;
    for (int x0 = (int)threadIdx.x / RG; x0 < nx; x0 += BT * S) {
        T v[BT];
#pragma unroll
        for (int u = 0; u < BT; u++) {
            const int x = x0 + u * S;
            v[u] = s[nR * (x < nx ? x : x0)];
        }
#pragma unroll
        for (int u = 0; u < BT; u++) {
            const int x = x0 + u * S;
            if (x < nx) row[x] = v[u];
        }
    }
}

__device__ __forceinline__ void stage_geo(const double* __restrict__ geo, int numR, int k0, int nr, double* sg) {
    for (int e = threadIdx.x; e < G_N * nr; e += blockDim.x) sg[e] = geo[(int64_t)(e / nr) * numR + k0 + e % nr];
}

// out rows = (accumulate ? staged + scale df : scale df) on the group's rows
template <class T, int DIM>
__global__ __launch_bounds__(1024) void k_reset_rows(const T* f_in, T* f_out, const double* __restrict__ geo,
                                                     const int32_t* __restrict__ indR, int numR, int64_t nR, int nx,
                                                     int RG, ResetConsts c, double scale, int accumulate) {
    extern __shared__ double smem_rr[];
    double* sg = smem_rr;                       // G_N x RG
    T* rows = (T*)(smem_rr + G_N * RG);          // RG x nx
    const int k0 = xcd_major_r(blockIdx.x, gridDim.x) * RG;
    const int nr = min(RG, numR - k0);
    stage_geo(geo, numR, k0, nr, sg);
    stage_rows<T>(f_in, indR, k0, nr, RG, nR, nx, rows);
    __syncthreads();
    RowLanes L(RG, c.N);
    if (L.r >= nr) return;
    const RowGeo G = row_geo(sg, nr, L.r, c);
    const T* row = rows + (int64_t)L.r * nx;
    T* out = f_out + indR[k0 + L.r];
    for (int x = L.x0; x < nx; x += L.S, L.p.advance(L.d, c.N)) {
        Jump<DIM> J;
        jump_at<DIM>(G, c, L.p.i0, L.p.i1, L.p.i2, 0.0, J);
        double d = J.out_on ? -(double)row[x] * G.lam : 0.0;
        if (J.in_on) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < Jump<DIM>::NQ; q++) v += (double)row[J.q[q]] * J.w[q];
            d = DIM == 2 ? v * G.lam + d : d + c.J * G.lam * v;
        }
        const double base = accumulate ? (double)row[x] : 0.0;
        out[nR * x] = (T)(base + scale * d);
    }
}

// out rows = g + scale R^H g on the group's rows: the transpose scatters into
// an LDS accumulator with atomic adds (only this group's rows are targets)
template <class T, int DIM>
__global__ __launch_bounds__(1024) void k_reset_adj_rows(const T* g_in, T* g_out, const double* __restrict__ geo,
                                                         const int32_t* __restrict__ indR, int numR, int64_t nR,
                                                         int nx, int RG, ResetConsts c, double scale) {
    extern __shared__ double smem_ra[];
    double* sg = smem_ra;
    T* rows = (T*)(smem_ra + G_N * RG);
    T* acc = rows + (int64_t)RG * nx;
    const int k0 = xcd_major_r(blockIdx.x, gridDim.x) * RG;
    const int nr = min(RG, numR - k0);
    stage_geo(geo, numR, k0, nr, sg);
    stage_rows<T>(g_in, indR, k0, nr, RG, nR, nx, rows);
    for (int e = threadIdx.x; e < nr * nx; e += blockDim.x) acc[e] = T(0);
    __syncthreads();
    RowLanes L(RG, c.N);
    const bool act = L.r < nr;
    if (act) {
        const RowGeo G = row_geo(sg, nr, L.r, c);
        const T* row = rows + (int64_t)L.r * nx;
        T* ar = acc + (int64_t)L.r * nx;
        for (int x = L.x0; x < nx; x += L.S, L.p.advance(L.d, c.N)) {
            Jump<DIM> J;
            jump_at<DIM>(G, c, L.p.i0, L.p.i1, L.p.i2, 0.0, J);
            const double gv = (double)row[x];
            if (J.out_on) atomicAdd(ar + x, (T)(-G.lam * gv));
            if (J.in_on) {
                const double a = c.J * G.lam * gv;
#pragma unroll
                for (int q = 0; q < Jump<DIM>::NQ; q++)
                    if (J.w[q] != 0.0) atomicAdd(ar + J.q[q], (T)(a * J.w[q]));
            }
        }
    }
    __syncthreads();
    if (!act) return;
    const T* row = rows + (int64_t)L.r * nx;
    const T* ar = acc + (int64_t)L.r * nx;
    T* out = g_out + indR[k0 + L.r];
    for (int x = L.x0; x < nx; x += L.S) out[nR * x] = (T)((double)row[x] + scale * (double)ar[x]);
}

// <g, (dR/d eps) f> over the group's rows, f staged, g read in place; the
// block's sum (lanes, then waves, fixed order) goes to part[group]
template <class T, int DIM>
__global__ __launch_bounds__(1024) void k_reset_deps_rows(const T* __restrict__ f, const T* __restrict__ g,
                                                          const double* __restrict__ geo,
                                                          const int32_t* __restrict__ indR, int numR, int64_t nR,
                                                          int nx, int RG, ResetConsts c, double dJ, double dk,
                                                          double* __restrict__ part) {
    extern __shared__ double smem_rd[];
    __shared__ double red[16];
    double* sg = smem_rd;
    T* rows = (T*)(smem_rd + G_N * RG);
    const int grp = xcd_major_r(blockIdx.x, gridDim.x);
    const int k0 = grp * RG;
    const int nr = min(RG, numR - k0);
    stage_geo(geo, numR, k0, nr, sg);
    stage_rows<T>(f, indR, k0, nr, RG, nR, nx, rows);
    __syncthreads();
    RowLanes L(RG, c.N);
    double v = 0.0;
    if (L.r < nr) {
        const RowGeo G = row_geo(sg, nr, L.r, c);
        const T* row = rows + (int64_t)L.r * nx;
        const T* gr = g + indR[k0 + L.r];
        constexpr int BT = 4;  // g values of a batch of points loaded together
        for (int xb = L.x0; xb < nx; xb += BT * L.S) {
            T gv[BT];
#pragma unroll
            for (int u = 0; u < BT; u++) {
                const int x = xb + u * L.S;
                gv[u] = gr[nR * (x < nx ? x : xb)];
            }
#pragma unroll
            for (int u = 0; u < BT; u++) {
                if (xb + u * L.S >= nx) break;
                Jump<DIM> J;
                jump_at<DIM>(G, c, L.p.i0, L.p.i1, L.p.i2, -dk, J);
                if (J.in_on) {
                    double I = 0.0, dI = 0.0;
#pragma unroll
                    for (int q = 0; q < Jump<DIM>::NQ; q++) {
                        const double fq = (double)row[J.q[q]];
                        I += J.w[q] * fq;
                        dI += J.dw[q] * fq;
                    }
                    v += (double)gv[u] * G.lam * (dJ * I + c.J * dI);
                }
                L.p.advance(L.d, c.N);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
        part[grp] = s;
    }
}

template <class T> __global__ void k_fill_zero(T* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

dim3 grid_k(int numR, int64_t nx) { return dim3((unsigned)((numR + 255) / 256), (unsigned)nx); }

}  // namespace

void ResetGeom::build(int B, int dim_, int N_, double L, const hpup_reset_params& p) {
    HPUP_REQUIRE(dim_ == 2 || dim_ == 3, HPUP_ERR_ARG, "the reset needs an omega grid (dim 2 or 3)");
    HPUP_REQUIRE(p.epsilon > 0.0 && p.epsilon <= 1.0, HPUP_ERR_ARG, "restitution coefficient must be in (0, 1]");
    HPUP_REQUIRE(p.theta_t > 0.0 && p.h > 0.0 && p.r >= 0.0, HPUP_ERR_ARG, "bad contact geometry");
    HPUP_REQUIRE(dim_ == 3 || p.tilt == 0.0, HPUP_ERR_ARG, "the reduced model's wall has no tilt");
    HPUP_REQUIRE(N_ >= 2 && N_ <= kResetMaxN, HPUP_ERR_ARG, "omega grid out of range for the reset");
    dim = dim_;
    N = N_;
    params = p;
    const int n2 = 2 * B;
    nR = (int64_t)n2 * n2 * n2;
    nx = ipow64(N, dim);
    HPUP_REQUIRE(nx <= 65535, HPUP_ERR_ARG, "omega grid too large for the reset launch");
    // the omega grid x = -L/2 + (L/N) i, as the drivers build it
    ResetConsts& c = consts;
    c.N = N;
    for (int i = 0; i < kResetMaxN; i++) c.xg[i] = i < N ? -L / 2 + L / N * (double)i : 0.0;
    c.x0 = -L / 2;
    c.hx = L / N;
    c.ihx = 1.0 / c.hx;
    if (dim == 2) {
        // getOmega's box (the gateway's xlo/xhi) and getIndRule's L and dx^2
        const double dx = c.xg[1] - c.xg[0];
        c.xlo = c.xg[0] + 1e-10;
        c.xhi = c.xg[N - 1] - 1e-10;
        c.Lg = c.xg[N - 1] - c.xg[0] + dx;
        c.dx2 = dx * dx;
        c.kfac = (1.0 + p.epsilon) / p.epsilon;
        c.J = 1.0;
    } else {
        c.kfac = (1.0 + p.epsilon) / p.epsilon;
        c.J = p.jacobian ? 1.0 / p.epsilon : 1.0;
    }
    c.ew[0] = std::cos(p.tilt);
    c.ew[1] = 0.0;
    c.ew[2] = std::sin(p.tilt);
    // per-rotation rate, contact point and tangent (getLambda.cu:192-232)
    const double hr = std::sqrt(p.h * p.h + p.r * p.r);
    const double theta0 = std::asin(p.d / hr) - std::asin(p.r / hr);
    std::vector<int32_t> ind;
    std::vector<double> cols[G_N];
    const EulerGrid eg(B);
    for (int64_t iR = 0; iR < nR; iR++) {
        double R[9];
        eg(iR, R);
        const double r3[3] = {R[6], R[7], R[8]};
        const double th = dim == 2 ? std::asin(R[6])
                                   : std::asin(r3[0] * c.ew[0] + r3[1] * c.ew[1] + r3[2] * c.ew[2]);
        if (th < theta0 - p.theta_t) continue;  // outside the collision region
        const double lam = th > theta0 + p.theta_t
                               ? p.lambda_max
                               : p.lambda_max / 2 * std::sin(M_PI / (2 * p.theta_t) * (th - theta0)) + p.lambda_max / 2;
        const double lr3 = p.h - p.r * std::tan(th);
        double PC[3], t[3];
        if (dim == 2) {
            PC[0] = lr3 * R[6] + p.r / std::cos(th);
            PC[1] = lr3 * R[7];
            PC[2] = lr3 * R[8];
            const double nt = std::sqrt(R[7] * R[7] + R[8] * R[8]);
            t[0] = 0.0;
            t[1] = R[8] / nt;
            t[2] = -R[7] / nt;
        } else {
            for (int a = 0; a < 3; a++) PC[a] = lr3 * r3[a] + p.r / std::cos(th) * c.ew[a];
            const double cr[3] = {r3[1] * c.ew[2] - r3[2] * c.ew[1], r3[2] * c.ew[0] - r3[0] * c.ew[2],
                                  r3[0] * c.ew[1] - r3[1] * c.ew[0]};
            const double nc = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            for (int a = 0; a < 3; a++) t[a] = nc > 0 ? cr[a] / nc : 0.0;
        }
        ind.push_back((int32_t)iR);
        for (int a = 0; a < 9; a++) cols[G_R + a].push_back(R[a]);
        for (int a = 0; a < 3; a++) {
            cols[G_T + a].push_back(t[a]);
            cols[G_PC + a].push_back(PC[a]);
        }
        cols[G_LAM].push_back(lam);
    }
    numR = (int)ind.size();
    std::vector<double> g((size_t)G_N * std::max(numR, 1), 0.0);
    for (int cidx = 0; cidx < G_N; cidx++)
        for (int k = 0; k < numR; k++) g[(size_t)cidx * numR + k] = cols[cidx][k];
    indR = upload(ind.empty() ? std::vector<int32_t>{0} : ind);
    geo = upload(g);
    region = ind;
}

// rows per group of the row-group kernels: as many as fit `per` bytes of LDS
// per omega point into 128 KiB, at most 8; 0 if not even one row fits
static int rows_per_group(int64_t nx, size_t per) {
    const int64_t budget = 128 * 1024 - G_N * 8 * 8;
    const int64_t rg = std::min<int64_t>(budget / (nx * (int64_t)per), 8);
    return rg >= 8 ? 8 : rg >= 4 ? 4 : rg >= 2 ? 2 : (int)rg;  // a power of two (the kernels' lane split), or 0
}

bool ResetGeom::rows_fit(size_t per) const { return rows_per_group(nx, per) >= 1; }

template <class T> void ResetGeom::zero(void* out, hipStream_t st) const {
    const int64_t n = nR * nx;
    hipLaunchKernelGGL(k_fill_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (T*)out, n);
    HPUP_HIP(hipGetLastError());
}
template void ResetGeom::zero<double>(void*, hipStream_t) const;
template void ResetGeom::zero<float>(void*, hipStream_t) const;

template <class T>
bool ResetGeom::rows_pass(const void* fin, void* fout, double scale, bool accumulate, hipStream_t st) const {
    const int RG = rows_per_group(nx, sizeof(T));
    if (RG < 1) return false;
    if (numR == 0) return true;
    const int ng = (numR + RG - 1) / RG;
    const size_t sh = sizeof(double) * G_N * RG + sizeof(T) * (size_t)RG * nx;
    if (dim == 2) {
        allow_lds((const void*)k_reset_rows<T, 2>, sh);
        hipLaunchKernelGGL((k_reset_rows<T, 2>), dim3(ng), dim3(1024), sh, st, (const T*)fin, (T*)fout,
                           geo.as<double>(), indR.as<int32_t>(), numR, nR, (int)nx, RG, consts, scale, accumulate ? 1 : 0);
    } else {
        allow_lds((const void*)k_reset_rows<T, 3>, sh);
        hipLaunchKernelGGL((k_reset_rows<T, 3>), dim3(ng), dim3(1024), sh, st, (const T*)fin, (T*)fout,
                           geo.as<double>(), indR.as<int32_t>(), numR, nR, (int)nx, RG, consts, scale, accumulate ? 1 : 0);
    }
    HPUP_HIP(hipGetLastError());
    return true;
}

template <class T>
bool ResetGeom::adjoint_rows(const void* g, void* out, double scale, hipStream_t st) const {
    const int RG = rows_per_group(nx, 2 * sizeof(T));
    if (RG < 1) return false;
    if (numR == 0) return true;
    const int ng = (numR + RG - 1) / RG;
    const size_t sh = sizeof(double) * G_N * RG + 2 * sizeof(T) * (size_t)RG * nx;
    if (dim == 2) {
        allow_lds((const void*)k_reset_adj_rows<T, 2>, sh);
        hipLaunchKernelGGL((k_reset_adj_rows<T, 2>), dim3(ng), dim3(1024), sh, st, (const T*)g, (T*)out,
                           geo.as<double>(), indR.as<int32_t>(), numR, nR, (int)nx, RG, consts, scale);
    } else {
        allow_lds((const void*)k_reset_adj_rows<T, 3>, sh);
        hipLaunchKernelGGL((k_reset_adj_rows<T, 3>), dim3(ng), dim3(1024), sh, st, (const T*)g, (T*)out,
                           geo.as<double>(), indR.as<int32_t>(), numR, nR, (int)nx, RG, consts, scale);
    }
    HPUP_HIP(hipGetLastError());
    return true;
}

template <class T> void ResetGeom::df(const void* f, void* dfc, hipStream_t st) const {
    if (numR == 0) return;
    if (dim == 2)
        hipLaunchKernelGGL((k_reset_df<T, 2>), grid_k(numR, nx), dim3(256), 0, st, (const T*)f, geo.as<double>(),
                           indR.as<int32_t>(), numR, nR, consts, (T*)dfc);
    else
        hipLaunchKernelGGL((k_reset_df<T, 3>), grid_k(numR, nx), dim3(256), 0, st, (const T*)f, geo.as<double>(),
                           indR.as<int32_t>(), numR, nR, consts, (T*)dfc);
    HPUP_HIP(hipGetLastError());
}

template <class T>
void ResetGeom::apply(const void* dfc, const void* base, void* out, double scale, hipStream_t st) const {
    if (!base) {  // out = scale df: zero outside the region (a fill kernel, no memset node)
        const int64_t n = nR * nx;
        hipLaunchKernelGGL(k_fill_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (T*)out, n);
    }
    if (numR > 0)
        hipLaunchKernelGGL(k_reset_apply<T>, grid_k(numR, nx), dim3(256), 0, st, (const T*)dfc, indR.as<int32_t>(),
                           numR, nR, (const T*)base, (T*)out, scale);
    HPUP_HIP(hipGetLastError());
}

template <class T> void ResetGeom::adjoint(const void* g, void* acc, hipStream_t st) const {
    const int64_t n = (int64_t)numR * nx;
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (T*)acc, n);
    if (dim == 2)
        hipLaunchKernelGGL((k_reset_adj<T, 2>), grid_k(numR, nx), dim3(256), 0, st, (const T*)g, geo.as<double>(),
                           indR.as<int32_t>(), numR, nR, consts, (T*)acc);
    else
        hipLaunchKernelGGL((k_reset_adj<T, 3>), grid_k(numR, nx), dim3(256), 0, st, (const T*)g, geo.as<double>(),
                           indR.as<int32_t>(), numR, nR, consts, (T*)acc);
    HPUP_HIP(hipGetLastError());
}

template <class T>
void ResetGeom::eps_grad(const void* f, const void* g, double* part, double scale, double* out, hipStream_t st) const {
    if (numR == 0) return;
    const double eps = params.epsilon;
    const double dk = -1.0 / (eps * eps);                                        // d((1 + eps)/eps)/d eps
    const double dJ = (dim == 3 && params.jacobian) ? -1.0 / (eps * eps) : 0.0;  // d(1/eps)/d eps
    const int RG = rows_per_group(nx, sizeof(T));
    if (RG >= 1) {  // LDS row groups
        const int ng = (numR + RG - 1) / RG;
        const size_t sh = sizeof(double) * G_N * RG + sizeof(T) * (size_t)RG * nx;
        if (dim == 2) {
            allow_lds((const void*)k_reset_deps_rows<T, 2>, sh);
            hipLaunchKernelGGL((k_reset_deps_rows<T, 2>), dim3(ng), dim3(1024), sh, st, (const T*)f, (const T*)g,
                               geo.as<double>(), indR.as<int32_t>(), numR, nR, (int)nx, RG, consts, dJ, dk, part);
        } else {
            allow_lds((const void*)k_reset_deps_rows<T, 3>, sh);
            hipLaunchKernelGGL((k_reset_deps_rows<T, 3>), dim3(ng), dim3(1024), sh, st, (const T*)f, (const T*)g,
                               geo.as<double>(), indR.as<int32_t>(), numR, nR, (int)nx, RG, consts, dJ, dk, part);
        }
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, (const double*)part, (int64_t)ng, scale, out);
        HPUP_HIP(hipGetLastError());
        return;
    }
    const dim3 grid = grid_k(numR, nx);
    if (dim == 2)
        hipLaunchKernelGGL((k_reset_deps<T, 2>), grid, dim3(256), 0, st, (const T*)f, (const T*)g, geo.as<double>(),
                           indR.as<int32_t>(), numR, nR, consts, dJ, dk, part);
    else
        hipLaunchKernelGGL((k_reset_deps<T, 3>), grid, dim3(256), 0, st, (const T*)f, (const T*)g, geo.as<double>(),
                           indR.as<int32_t>(), numR, nR, consts, dJ, dk, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, (const double*)part,
                       (int64_t)grid.x * grid.y, scale, out);
    HPUP_HIP(hipGetLastError());
}

int64_t ResetGeom::eps_grad_partials() const {
    const dim3 grid = grid_k(std::max(numR, 1), nx);
    return std::max<int64_t>((int64_t)grid.x * grid.y, std::max(numR, 1));  // per-point or row-group partials
}

template bool ResetGeom::rows_pass<double>(const void*, void*, double, bool, hipStream_t) const;
template bool ResetGeom::rows_pass<float>(const void*, void*, double, bool, hipStream_t) const;
template bool ResetGeom::adjoint_rows<double>(const void*, void*, double, hipStream_t) const;
template bool ResetGeom::adjoint_rows<float>(const void*, void*, double, hipStream_t) const;
template void ResetGeom::adjoint<double>(const void*, void*, hipStream_t) const;
template void ResetGeom::adjoint<float>(const void*, void*, hipStream_t) const;
template void ResetGeom::eps_grad<double>(const void*, const void*, double*, double, double*, hipStream_t) const;
template void ResetGeom::eps_grad<float>(const void*, const void*, double*, double, double*, hipStream_t) const;
template void ResetGeom::df<double>(const void*, void*, hipStream_t) const;
template void ResetGeom::df<float>(const void*, void*, hipStream_t) const;
template void ResetGeom::apply<double>(const void*, const void*, void*, double, hipStream_t) const;
template void ResetGeom::apply<float>(const void*, const void*, void*, double, hipStream_t) const;

}  // namespace hpup
