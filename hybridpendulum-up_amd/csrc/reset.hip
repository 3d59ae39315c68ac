// Hybrid collision reset on the grid (reduced model, 2-D omega), the
// gateways of propagation/discrete/mex/sources/pendulum_reduced_collision.
// Grid f layout (nR, nx) with the rotation index fastest; per-region arrays
// are (numR, nx) with the region index fastest, as in the reference.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <strings.h>
#include <vector>

#include "hpup_internal.h"

namespace hpup {

// getLambda.cu:192-216: theta = asin(R13), category, contact point PC.
__global__ void k_theta_pc(const double* __restrict__ R, int64_t nR, double theta0, double tht, double h, double r,
                           double* __restrict__ theta, signed char* __restrict__ cat, double* __restrict__ PC) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nR) return;
    const double th = asin(R[9 * i + 6]);
    theta[i] = th;
    cat[i] = th < theta0 - tht ? 0 : (th > theta0 + tht ? 2 : 1);
    const double lr3 = h - r * tan(th);
    PC[3 * i] = lr3 * R[9 * i + 6] + r / cos(th);
    PC[3 * i + 1] = lr3 * R[9 * i + 7];
    PC[3 * i + 2] = lr3 * R[9 * i + 8];
}

// getLambda.cu:218-266: lambda per region point, zero where (omega x PC).e1 < 0.
__global__ void k_lambda(const double* __restrict__ R, const double* __restrict__ x, const double* __restrict__ PC,
                         const int32_t* __restrict__ indR, int numR, int64_t nx, const double* __restrict__ theta,
                         const signed char* __restrict__ cat, double theta0, double tht, double lmax,
                         double* __restrict__ lam) {
    const int64_t ix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (ix >= nx) return;
    const int64_t iR = indR[k];
    const double l0 = cat[iR] == 2 ? lmax : lmax / 2 * sin(M_PI / (2 * tht) * (theta[iR] - theta0)) + lmax / 2;
    const double* Rr = R + 9 * iR;
    const double o1 = Rr[1] * x[2 * ix] + Rr[4] * x[2 * ix + 1];
    const double o2 = Rr[2] * x[2 * ix] + Rr[5] * x[2 * ix + 1];
    const double vC = o1 * PC[3 * iR + 2] - o2 * PC[3 * iR + 1];
    lam[k + (int64_t)numR * ix] = vC < 0.0 ? 0.0 : l0;
}

// getLambda.cu:218-246 (compact form): lambda per region, and the flag of
// omega points with non-negative contact velocity.
__global__ void k_lambda_compact(const double* __restrict__ R, const double* __restrict__ x,
                                 const double* __restrict__ PC, const int32_t* __restrict__ indR, int numR, int64_t nx,
                                 const double* __restrict__ theta, const signed char* __restrict__ cat, double theta0,
                                 double tht, double lmax, double* __restrict__ lam, signed char* __restrict__ flag) {
    const int64_t ix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (ix >= nx) return;
    const int64_t iR = indR[k];
    if (ix == 0)
        lam[k] = cat[iR] == 2 ? lmax : lmax / 2 * sin(M_PI / (2 * tht) * (theta[iR] - theta0)) + lmax / 2;
    const double* Rr = R + 9 * iR;
    const double o1 = Rr[1] * x[2 * ix] + Rr[4] * x[2 * ix + 1];
    const double o2 = Rr[2] * x[2 * ix] + Rr[5] * x[2 * ix + 1];
    const double vC = o1 * PC[3 * iR + 2] - o2 * PC[3 * iR + 1];
    flag[(int64_t)k * nx + ix] = vC < 0.0 ? 0 : 1;
}

// getOmega.cu:64-137
__global__ void k_omega(const double* __restrict__ R, const double* __restrict__ x, const int32_t* __restrict__ indR,
                        int numR, int64_t nx, double eps, const double* __restrict__ PC, int old_dir, double xlo,
                        double xhi, double* __restrict__ Om) {
    const int64_t ix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (ix >= nx) return;
    const double* Rr = R + 9 * (int64_t)indR[k];
    const double nt = sqrt(Rr[7] * Rr[7] + Rr[8] * Rr[8]);
    const double t0 = Rr[8] / nt, t1 = -Rr[7] / nt;
    double o0 = Rr[0] * x[2 * ix] + Rr[3] * x[2 * ix + 1];
    double o1 = Rr[1] * x[2 * ix] + Rr[4] * x[2 * ix + 1];
    double o2 = Rr[2] * x[2 * ix] + Rr[5] * x[2 * ix + 1];
    const double fac = old_dir ? (1.0 + eps) / eps : (1.0 + eps);
    const double ot = fac * (o1 * t0 + o2 * t1);
    o1 -= ot * t0;
    o2 -= ot * t1;
    const double a0 = Rr[0] * o0 + Rr[1] * o1 + Rr[2] * o2;
    const double a1 = Rr[3] * o0 + Rr[4] * o1 + Rr[5] * o2;
    double* out = Om + 2 * (k + (int64_t)numR * ix);
    if (old_dir) {
        const double* P = PC + 3 * (int64_t)indR[k];
        const double vC = o1 * P[2] - o2 * P[1];
        if (vC < 0.0 || a0 < xlo || a0 > xhi || a1 < xlo || a1 > xhi) {
            out[0] = NAN;
            out[1] = NAN;
            return;
        }
    }
    out[0] = a0;
    out[1] = a1;
}

constexpr int kMaxND = 16;

// getFcL.cu:70-92: for each (region point, destination) the Gaussian density
// from every source; the nD largest kept in stable descending order.
__global__ void k_fcl_topk(const double* __restrict__ x, const double* __restrict__ Om, int numR, int64_t nx, double ig0,
                           double ig12, double ig3, double cn, int nD, double* __restrict__ fcL,
                           int16_t* __restrict__ idx) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (d >= nx) return;
    double v[kMaxND];
    int id[kMaxND];
    int cnt = 0;
    const double xd0 = x[2 * d], xd1 = x[2 * d + 1];
    for (int64_t s = 0; s < nx; s++) {
        const double* o = Om + 2 * (k + (int64_t)numR * s);
        const double d0 = xd0 - o[0], d1 = xd1 - o[1];
        const double q = ig0 * d0 * d0 + ig12 * d0 * d1 + ig3 * d1 * d1;
        const double val = exp(-0.5 * q) * cn;
        if (cnt == nD && !(val > v[nD - 1])) continue;
        int pos = cnt < nD ? cnt : nD - 1;
        while (pos > 0 && val > v[pos - 1]) {
            v[pos] = v[pos - 1];
            id[pos] = id[pos - 1];
            pos--;
        }
        v[pos] = val;
        id[pos] = (int)s;
        if (cnt < nD) cnt++;
    }
    for (int t = 0; t < nD; t++) {
        const int64_t o = t + (int64_t)nD * (k + (int64_t)numR * d);
        fcL[o] = v[t];
        idx[o] = (int16_t)id[t];
    }
}

// getFcL.cu:94-101: each source's normalizer, the sum of its kept densities
// over the destinations in ascending order (and list position t within a
// destination), the order of the reference's host loop, so the result does
// not depend on scheduling.  One thread per source; a block stages the lists
// of kFclChunk destinations of its region point in LDS and every thread scans
// them for its own source.
constexpr int kFclChunk = 64;
__global__ void k_fcl_norm(const double* __restrict__ fcL, const int16_t* __restrict__ idx, int numR, int64_t nx,
                           int nD, double* __restrict__ norm) {
    __shared__ double sv[kFclChunk * kMaxND];
    __shared__ int16_t si[kFclChunk * kMaxND];
    const int k = blockIdx.y;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    for (int64_t d0 = 0; d0 < nx; d0 += kFclChunk) {
        const int nd = (int)(nx - d0 < kFclChunk ? nx - d0 : kFclChunk);
        __syncthreads();
        for (int e = threadIdx.x; e < nd * nD; e += blockDim.x) {
            const int64_t o = (e % nD) + (int64_t)nD * (k + (int64_t)numR * (d0 + e / nD));
            sv[e] = fcL[o];
            si[e] = idx[o];
        }
        __syncthreads();
        for (int e = 0; e < nd * nD; e++)
            if (si[e] == s) acc += sv[e];
    }
    if (s < nx) norm[(int64_t)k * nx + s] = acc;
}

// getFcL.cu:94-101,160-166: per-source normalization, then lambda(source) dx^2.
__global__ void k_fcl_final(int numR, int64_t nx, int nD, const double* __restrict__ norm, const double* __restrict__ lam,
                            double dx2, double* __restrict__ fcL, const int16_t* __restrict__ idx) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (d >= nx) return;
    for (int t = 0; t < nD; t++) {
        const int64_t o = t + (int64_t)nD * (k + (int64_t)numR * d);
        const int s = idx[o];
        fcL[o] = fcL[o] / (norm[(int64_t)k * nx + s] * dx2) * (lam[k + (int64_t)numR * s] * dx2);
    }
}

// getIndRule.cu:50-90
__global__ void k_indrule(const double* __restrict__ x, const double* __restrict__ Om, int numR, int64_t nx, int N,
                          double L, double dx2, int32_t* __restrict__ ind, double* __restrict__ coeff) {
    const int64_t ix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (ix >= nx) return;
    const int64_t o = k + (int64_t)numR * ix;
    const double* om = Om + 2 * o;
    int32_t* ii = ind + 4 * o;
    double* cc = coeff + 4 * o;
    if (isnan(om[0])) {
        for (int i = 0; i < 4; i++) { ii[i] = 0; cc[i] = NAN; }
        return;
    }
    const int x1 = (int)((om[0] / L + 0.5) * N);
    const int x2 = (int)((om[1] / L + 0.5) * N);
    if (x1 < 0 || x1 > N - 2 || x2 < 0 || x2 > N - 2) {  // outside the interpolation box
        for (int i = 0; i < 4; i++) { ii[i] = 0; cc[i] = NAN; }
        return;
    }
    ii[0] = x1 + x2 * N;
    ii[1] = ii[0] + 1;
    ii[2] = ii[0] + N;
    ii[3] = ii[2] + 1;
    const double g10 = x[2 * x1], g11 = x[2 * (x1 + 1)];
    const double g20 = x[2 * (int64_t)x2 * N + 1], g21 = x[2 * (int64_t)(x2 + 1) * N + 1];
    const double a = om[0] - g10, b = g11 - om[0], c = om[1] - g20, e = g21 - om[1];
    cc[0] = b * e / dx2;
    cc[1] = a * e / dx2;
    cc[2] = b * c / dx2;
    cc[3] = a * c / dx2;
}

// propagate.cu:181-220 (noise law): df = sum_k f(src_k) fcL_k - lambda f.
// The destination omega point runs across the lanes (grid_rx) and a block
// holds one region point k: its sources all lie in the grid row iR = indR[k]
// (nx lines, reused from L2 by every destination).  With k across the lanes
// instead (contiguous fcL, fcL_indx and df accesses) each lane gathers from
// its own rotation row, and the kernel (one source at a time) measured
// 1.86 ms against 1.27 ms (B = 15, 32^2 grid, 6300 region points, nD = 10).
__global__ void k_disc_noise(const double* __restrict__ f, const double* __restrict__ lam, const int32_t* __restrict__ indR,
                             int numR, int64_t nR, int64_t nx, const double* __restrict__ fcL,
                             const int16_t* __restrict__ idx, int nD, double* __restrict__ df) {
    const int64_t ix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (ix >= nx) return;
    const int64_t iR = indR[k];
    const int64_t o0 = (int64_t)nD * (k + (int64_t)numR * ix);
    // the first kMaxND sources: all indices and weights loaded, then all
    // gathers issued, then summed in source order (one round trip each
    // instead of one dependent pair per source); sources beyond kMaxND follow
    // in the same order
    int sid[kMaxND];
    double wt[kMaxND], fv[kMaxND];
#pragma unroll
    for (int t = 0; t < kMaxND; t++)
        if (t < nD) {
            sid[t] = idx[o0 + t];
            wt[t] = fcL[o0 + t];
        }
#pragma unroll
    for (int t = 0; t < kMaxND; t++)
        if (t < nD) fv[t] = f[iR + nR * sid[t]];
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < kMaxND; t++)
        if (t < nD) s += fv[t] * wt[t];
    for (int t = kMaxND; t < nD; t++) s += f[iR + nR * idx[o0 + t]] * fcL[o0 + t];
    const int64_t i = iR + nR * ix;
    df[i] = s - f[i] * lam[k + (int64_t)numR * ix];
}

// propagate.cu:189-211 (no-noise law, inflow by bilinear pull-back).  The
// region index k runs across the lanes (grid_kx): ind_interp and coeff_interp
// (4 x numR x nx) and df's rows iR = indR[k] are then contiguous along a wave,
// and the four corner sources lie next to the destination (0.19 ms against
// 0.23 ms with the omega point across the lanes, B = 15, 32^2 grid)
__global__ void k_disc_nonoise_in(const double* __restrict__ f, const double* __restrict__ lam,
                                  const int32_t* __restrict__ indR, int numR, int64_t nR, int64_t nx,
                                  const int32_t* __restrict__ ind, const double* __restrict__ coeff,
                                  double* __restrict__ df) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int64_t ix = blockIdx.y;
    if (k >= numR) return;
    const int64_t iR = indR[k];
    const int64_t o = 4 * (k + (int64_t)numR * ix);
    const int64_t i = iR + nR * ix;
    if (isnan(coeff[o])) {
        df[i] = 0.0;
        return;
    }
    double s = 0.0;
    for (int q = 0; q < 4; q++) s += f[iR + nR * ind[o + q]] * coeff[o + q];
    df[i] = s * lam[k];
}

// propagate.cu:222-229 (no-noise law, outflow on the listed omega points)
__global__ void k_disc_nonoise_out(const double* __restrict__ f, const double* __restrict__ lam,
                                   const int32_t* __restrict__ indR, int64_t nR, const int32_t* __restrict__ ent_k,
                                   const int32_t* __restrict__ ent_x, int64_t nent, double* __restrict__ df) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    const int k = ent_k[e];
    const int64_t i = indR[k] + nR * ent_x[e];
    df[i] -= f[i] * lam[k];
}

}  // namespace hpup

using namespace hpup;

namespace {

template <class T> DevBuf up(const T* p, size_t n) {
    DevBuf b(sizeof(T) * std::max<size_t>(n, 1));
    if (n) HPUP_HIP(hipMemcpy(b.p, p, sizeof(T) * n, hipMemcpyHostToDevice));
    return b;
}

template <class T> void down(T* p, const DevBuf& b, size_t n) {
    if (n) HPUP_HIP(hipMemcpy(p, b.p, sizeof(T) * n, hipMemcpyDeviceToHost));
}

void check_range(const int32_t* a, size_t n, int64_t hi, const char* what) {
    for (size_t i = 0; i < n; i++)
        HPUP_REQUIRE(a[i] >= 0 && a[i] < hi, HPUP_ERR_ARG, std::string(what) + " out of range");
}

dim3 grid_rx(int64_t nx, int numR) { return dim3((unsigned)((nx + 127) / 128), (unsigned)std::max(numR, 1)); }
// region index across the lanes, one omega point per grid row (nx < 65536)
dim3 grid_kx(int numR, int64_t nx) {
    HPUP_REQUIRE(nx < 65536, HPUP_ERR_UNSUPPORTED, "omega grid above 65535 points");
    return dim3((unsigned)((numR + 127) / 128), (unsigned)nx);
}

}  // namespace

extern "C" {

int hpup_getLambda(const double* R, int B, const double* x, int N, double d, double h, double r, double theta_t,
                   double lambda_max, double* lambda, int32_t* indR, int* numR, double* PC) {
    return guard([&] {
        HPUP_REQUIRE(R && x && lambda && indR && numR && PC && B > 0 && N > 1, HPUP_ERR_ARG, "bad argument");
        const int64_t nR = 8LL * B * B * B, nx = (int64_t)N * N;
        const double theta0 = std::asin(d / std::sqrt(h * h + r * r)) - std::asin(r / std::sqrt(h * h + r * r));
        DevBuf Rd = up(R, 9 * nR), xd = up(x, 2 * nx), th(sizeof(double) * nR), cat(nR), PCd(sizeof(double) * 3 * nR);
        hipLaunchKernelGGL(k_theta_pc, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, 0, Rd.as<double>(), nR,
                           theta0, theta_t, h, r, th.as<double>(), cat.as<signed char>(), PCd.as<double>());
        HPUP_HIP(hipGetLastError());
        std::vector<signed char> hc(nR);
        down(hc.data(), cat, nR);
        int cnt = 0;
        for (int64_t i = 0; i < nR; i++)
            if (hc[i] != 0) indR[cnt++] = (int32_t)i;
        *numR = cnt;
        down(PC, PCd, 3 * nR);
        if (cnt == 0) return;
        DevBuf ir = up(indR, cnt), lam(sizeof(double) * cnt * nx);
        hipLaunchKernelGGL(k_lambda, grid_rx(nx, cnt), dim3(128), 0, 0, Rd.as<double>(), xd.as<double>(),
                           PCd.as<double>(), ir.as<int32_t>(), cnt, nx, th.as<double>(), cat.as<signed char>(), theta0,
                           theta_t, lambda_max, lam.as<double>());
        HPUP_HIP(hipGetLastError());
        down(lambda, lam, (size_t)cnt * nx);
    });
}

int hpup_getLambda_compact(const double* R, int B, const double* x, int N, double d, double h, double r,
                           double theta_t, double lambda_max, double* lambda, int32_t* indR, int* numR, int32_t* indx,
                           int32_t* numx, double* PC) {
    return guard([&] {
        HPUP_REQUIRE(R && x && lambda && indR && numR && indx && numx && PC && B > 0 && N > 1, HPUP_ERR_ARG,
                     "bad argument");
        const int64_t nR = 8LL * B * B * B, nx = (int64_t)N * N;
        const double theta0 = std::asin(d / std::sqrt(h * h + r * r)) - std::asin(r / std::sqrt(h * h + r * r));
        DevBuf Rd = up(R, 9 * nR), xd = up(x, 2 * nx), th(sizeof(double) * nR), cat(nR), PCd(sizeof(double) * 3 * nR);
        hipLaunchKernelGGL(k_theta_pc, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, 0, Rd.as<double>(), nR,
                           theta0, theta_t, h, r, th.as<double>(), cat.as<signed char>(), PCd.as<double>());
        HPUP_HIP(hipGetLastError());
        std::vector<signed char> hc(nR);
        down(hc.data(), cat, nR);
        int cnt = 0;
        for (int64_t i = 0; i < nR; i++)
            if (hc[i] != 0) indR[cnt++] = (int32_t)i;
        *numR = cnt;
        down(PC, PCd, 3 * nR);
        if (cnt == 0) return;
        DevBuf ir = up(indR, cnt), lam(sizeof(double) * cnt), fl((size_t)cnt * nx);
        hipLaunchKernelGGL(k_lambda_compact, grid_rx(nx, cnt), dim3(128), 0, 0, Rd.as<double>(), xd.as<double>(),
                           PCd.as<double>(), ir.as<int32_t>(), cnt, nx, th.as<double>(), cat.as<signed char>(), theta0,
                           theta_t, lambda_max, lam.as<double>(), fl.as<signed char>());
        HPUP_HIP(hipGetLastError());
        down(lambda, lam, cnt);
        std::vector<signed char> hf((size_t)cnt * nx);
        down(hf.data(), fl, hf.size());
        for (int k = 0; k < cnt; k++) {
            int c = 0;
            for (int64_t ix = 0; ix < nx; ix++)
                if (hf[(size_t)k * nx + ix]) indx[(int64_t)k * nx + c++] = (int32_t)ix;
            numx[k] = c;
        }
    });
}

int hpup_getOmega(const double* R, int B, const double* x, int N, const int32_t* indR, int numR, double epsilon,
                  const double* PC, const char* direction, double* Omega) {
    return guard([&] {
        HPUP_REQUIRE(R && x && indR && PC && direction && Omega && numR >= 0, HPUP_ERR_ARG, "bad argument");
        int old_dir;
        if (!strcasecmp(direction, "new")) old_dir = 0;
        else if (!strcasecmp(direction, "old")) old_dir = 1;
        else throw Error{HPUP_ERR_ARG, "direction must be 'new' or 'old'"};
        if (numR == 0) return;
        const int64_t nR = 8LL * B * B * B, nx = (int64_t)N * N;
        check_range(indR, numR, nR, "indR");
        DevBuf Rd = up(R, 9 * nR), xd = up(x, 2 * nx), ir = up(indR, numR), PCd = up(PC, 3 * nR),
               Om(sizeof(double) * 2 * numR * nx);
        const double xlo = x[0] + 1e-10, xhi = x[2 * (N - 1)] - 1e-10;
        hipLaunchKernelGGL(k_omega, grid_rx(nx, numR), dim3(128), 0, 0, Rd.as<double>(), xd.as<double>(),
                           ir.as<int32_t>(), numR, nx, epsilon, PCd.as<double>(), old_dir, xlo, xhi, Om.as<double>());
        HPUP_HIP(hipGetLastError());
        down(Omega, Om, (size_t)2 * numR * nx);
    });
}

int hpup_getFcL(const double* x, int N, const double* Omega, const double* lambda, int numR, const double* Gd, int nD,
                double* fcL, int16_t* fcL_indx) {
    return guard([&] {
        HPUP_REQUIRE(x && Omega && lambda && Gd && fcL && fcL_indx, HPUP_ERR_ARG, "bad argument");
        HPUP_REQUIRE(nD >= 1 && nD <= kMaxND, HPUP_ERR_UNSUPPORTED, "nD must be in [1, 16]");
        const int64_t nx = (int64_t)N * N;
        HPUP_REQUIRE(nx <= 32767 && nD <= nx, HPUP_ERR_ARG, "omega grid too large for int16 indices");
        if (numR == 0) return;
        const double det = Gd[0] * Gd[3] - Gd[2] * Gd[1];
        const double cn = 1.0 / (2 * M_PI * std::sqrt(det));
        const double ig0 = Gd[3] / det, ig12 = -Gd[2] / det - Gd[1] / det, ig3 = Gd[0] / det;
        const double dx2 = (x[2] - x[0]) * (x[2] - x[0]);
        DevBuf xd = up(x, 2 * nx), Om = up(Omega, 2 * numR * nx), lam = up(lambda, (size_t)numR * nx),
               fc(sizeof(double) * nD * numR * nx), id(sizeof(int16_t) * nD * numR * nx),
               nrm(sizeof(double) * numR * nx);
        hipLaunchKernelGGL(k_fcl_topk, grid_rx(nx, numR), dim3(128), 0, 0, xd.as<double>(), Om.as<double>(), numR, nx,
                           ig0, ig12, ig3, cn, nD, fc.as<double>(), id.as<int16_t>());
        hipLaunchKernelGGL(k_fcl_norm, grid_rx(nx, numR), dim3(128), 0, 0, fc.as<double>(), id.as<int16_t>(), numR, nx,
                           nD, nrm.as<double>());
        hipLaunchKernelGGL(k_fcl_final, grid_rx(nx, numR), dim3(128), 0, 0, numR, nx, nD, nrm.as<double>(),
                           lam.as<double>(), dx2, fc.as<double>(), id.as<int16_t>());
        HPUP_HIP(hipGetLastError());
        down(fcL, fc, (size_t)nD * numR * nx);
        down(fcL_indx, id, (size_t)nD * numR * nx);
    });
}

int hpup_getIndRule(const double* x, int N, const double* Omega, int numR, int32_t* ind_interp, double* coeff_interp) {
    return guard([&] {
        HPUP_REQUIRE(x && Omega && ind_interp && coeff_interp, HPUP_ERR_ARG, "bad argument");
        if (numR == 0) return;
        const int64_t nx = (int64_t)N * N;
        const double dx = x[2] - x[0], L = x[2 * (N - 1)] - x[0] + dx;
        DevBuf xd = up(x, 2 * nx), Om = up(Omega, 2 * numR * nx), ind(sizeof(int32_t) * 4 * numR * nx),
               co(sizeof(double) * 4 * numR * nx);
        hipLaunchKernelGGL(k_indrule, grid_rx(nx, numR), dim3(128), 0, 0, xd.as<double>(), Om.as<double>(), numR, nx,
                           N, L, dx * dx, ind.as<int32_t>(), co.as<double>());
        HPUP_HIP(hipGetLastError());
        down(ind_interp, ind, (size_t)4 * numR * nx);
        down(coeff_interp, co, (size_t)4 * numR * nx);
    });
}

int hpup_discrete_propagate_noise(const double* f, int B, int N, const double* lambda, const int32_t* indR, int numR,
                                  const double* fcL, const int16_t* fcL_indx, int nD, double* df) {
    return guard([&] {
        HPUP_REQUIRE(f && lambda && indR && fcL && fcL_indx && df && nD >= 1, HPUP_ERR_ARG, "bad argument");
        const int64_t nR = 8LL * B * B * B, nx = (int64_t)N * N;
        check_range(indR, numR, nR, "lambda_indR");
        for (size_t i = 0; i < (size_t)nD * numR * nx; i++)
            HPUP_REQUIRE(fcL_indx[i] >= 0 && fcL_indx[i] < nx, HPUP_ERR_ARG, "fcL_indx out of range");
        DevBuf fd = up(f, nR * nx), dfd(sizeof(double) * nR * nx);
        HPUP_HIP(hipMemset(dfd.p, 0, dfd.bytes));
        if (numR > 0) {
            DevBuf lam = up(lambda, (size_t)numR * nx), ir = up(indR, numR), fc = up(fcL, (size_t)nD * numR * nx),
                   id = up(fcL_indx, (size_t)nD * numR * nx);
            hipLaunchKernelGGL(k_disc_noise, grid_rx(nx, numR), dim3(128), 0, 0, fd.as<double>(), lam.as<double>(),
                               ir.as<int32_t>(), numR, nR, nx, fc.as<double>(), id.as<int16_t>(), nD, dfd.as<double>());
            HPUP_HIP(hipGetLastError());
        }
        down(df, dfd, (size_t)(nR * nx));
    });
}

int hpup_discrete_propagate_nonoise(const double* f, int B, int N, const double* lambda, const int32_t* indR, int numR,
                                    const int32_t* const* lambda_indx, const int32_t* lambda_numx,
                                    const int32_t* ind_interp, const double* coeff_interp, double* df) {
    return guard([&] {
        HPUP_REQUIRE(f && lambda && indR && df && (numR == 0 || (lambda_indx && lambda_numx && ind_interp && coeff_interp)),
                     HPUP_ERR_ARG, "bad argument");
        const int64_t nR = 8LL * B * B * B, nx = (int64_t)N * N;
        check_range(indR, numR, nR, "lambda_indR");
        for (size_t i = 0; i < (size_t)numR * nx; i++)
            if (!std::isnan(coeff_interp[4 * i])) check_range(ind_interp + 4 * i, 4, nx, "ind_interp");
        DevBuf fd = up(f, nR * nx), dfd(sizeof(double) * nR * nx);
        HPUP_HIP(hipMemset(dfd.p, 0, dfd.bytes));
        if (numR > 0) {
            DevBuf lam = up(lambda, numR), ir = up(indR, numR), ind = up(ind_interp, (size_t)4 * numR * nx),
                   co = up(coeff_interp, (size_t)4 * numR * nx);
            hipLaunchKernelGGL(k_disc_nonoise_in, grid_kx(numR, nx), dim3(128), 0, 0, fd.as<double>(), lam.as<double>(),
                               ir.as<int32_t>(), numR, nR, nx, ind.as<int32_t>(), co.as<double>(), dfd.as<double>());
            std::vector<int32_t> ek, ex;
            for (int k = 0; k < numR; k++)
                for (int t = 0; t < lambda_numx[k]; t++) {
                    HPUP_REQUIRE(lambda_indx[k][t] >= 0 && lambda_indx[k][t] < nx, HPUP_ERR_ARG, "lambda_indx out of range");
                    ek.push_back(k);
                    ex.push_back(lambda_indx[k][t]);
                }
            if (!ek.empty()) {
                DevBuf ekd = up(ek.data(), ek.size()), exd = up(ex.data(), ex.size());
                hipLaunchKernelGGL(k_disc_nonoise_out, dim3((unsigned)((ek.size() + 255) / 256)), dim3(256), 0, 0,
                                   fd.as<double>(), lam.as<double>(), ir.as<int32_t>(), nR, ekd.as<int32_t>(),
                                   exd.as<int32_t>(), (int64_t)ek.size(), dfd.as<double>());
            }
            HPUP_HIP(hipGetLastError());
        }
        down(df, dfd, (size_t)(nR * nx));
    });
}

}  // extern "C"
