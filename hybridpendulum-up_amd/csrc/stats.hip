// Grid statistics of get_stat (pendulum_reduced.m:485-532, pendulum.m:547-594)
// in ONE pass over the grid density f[iR + nR*ix] (alpha fastest, omega
// slowest):
//   fR[iR]    = sum_x f              (rotation marginal, unscaled)
//   gR[a][iR] = sum_x x_a f          (first omega moment per rotation)
//   fx[ix]    = sum_R w_j(R) f       (omega marginal)
// Every later get_stat output is a small contraction of these three: E[R],
// E[vR], E[vR vR^T] use fR; E[x vR^T] uses gR (vR is linear in R); E[x] and
// Var[x] use fx.  The pass is HBM-bound: one read of f, the partial sums are
// under 2% of it.  Accumulation is FP64 for either storage precision, and
// deterministic (fixed-order partial reductions, no atomics).
#include "hpup_internal.h"
#include "kernels_common.h"
#include "stats.h"

namespace hpup {

constexpr int kStatThreads = 256;
constexpr int kStatRT = 4;                        // rotation points per thread
constexpr int kStatTile = kStatThreads * kStatRT; // rotation points per block
constexpr int kStatXU = 4;                        // omega points per load batch

// grid: (rotation tiles, omega chunks).  Writes per-chunk partials of fR, gR
// and per-wave partials of fx:
//   fRp[c][iR], gRp[c][a][iR], fxp[tile*4 + wave][ix]
static_assert(kStatRT == 4, "ld4 loads four rotation points");
// four consecutive values as FP64 (zero when !ok; p then points at f[0])
template <class T> __device__ __forceinline__ void ld4(const T* p, bool ok, double* v) {
    if constexpr (sizeof(T) == 4) {
        const float4 q = *reinterpret_cast<const float4*>(p);
        v[0] = ok ? (double)q.x : 0.0;
        v[1] = ok ? (double)q.y : 0.0;
        v[2] = ok ? (double)q.z : 0.0;
        v[3] = ok ? (double)q.w : 0.0;
    } else {
        const double2 q0 = reinterpret_cast<const double2*>(p)[0], q1 = reinterpret_cast<const double2*>(p)[1];
        v[0] = ok ? q0.x : 0.0;
        v[1] = ok ? q0.y : 0.0;
        v[2] = ok ? q1.x : 0.0;
        v[3] = ok ? q1.y : 0.0;
    }
}

template <class T, int DIM>
__global__ __launch_bounds__(kStatThreads) void k_stat_pass(const T* __restrict__ f, int64_t nR, int64_t nx,
                                                            int n2, int64_t xper, const double* __restrict__ w,
                                                            const double* __restrict__ xg, double* __restrict__ fRp,
                                                            double* __restrict__ gRp, double* __restrict__ fxp) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t tile = blockIdx.x, c = blockIdx.y;
    const int64_t x0 = c * xper, x1 = x0 + xper < nx ? x0 + xper : nx;
    // this thread's kStatRT consecutive rotation points (nR = (2B)^3 is a
    // multiple of kStatRT, so a group is wholly in range or out): one
    // 4-element vector load per omega point, a wave's loads contiguous
    const int64_t R0 = tile * kStatTile + (int64_t)t * kStatRT;
    const bool inR = R0 < nR;
    int64_t R[kStatRT];
    double wr[kStatRT], af[kStatRT], ag[kStatRT][DIM > 0 ? DIM : 1];
#pragma unroll
    for (int r = 0; r < kStatRT; r++) {
        R[r] = R0 + r;
        wr[r] = inR ? w[(R[r] / n2) % n2] : 0.0;
        af[r] = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; a++) ag[r][a] = 0.0;
    }
    double* fxw = fxp + (tile * (kStatThreads / 64) + wave) * nx;
    for (int64_t xb = x0; xb < x1; xb += kStatXU) {
        double v[kStatXU][kStatRT];
#pragma unroll
        for (int u = 0; u < kStatXU; u++) {
            const bool ok = inR && xb + u < x1;
            ld4<T>(f + (ok ? R0 + nR * (xb + u) : 0), ok, v[u]);
        }
#pragma unroll
        for (int u = 0; u < kStatXU; u++) {
            const int64_t ix = xb + u;
            if (ix >= x1) break;
            double ws = 0.0;
#pragma unroll
            for (int r = 0; r < kStatRT; r++) {
                af[r] += v[u][r];
                ws += wr[r] * v[u][r];
#pragma unroll
                for (int a = 0; a < DIM; a++) ag[r][a] += xg[DIM * ix + a] * v[u][r];
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) ws += __shfl_xor(ws, o, 64);
            if (lane == 0) fxw[ix] = ws;
        }
    }
    if (!inR) return;
#pragma unroll
    for (int r = 0; r < kStatRT; r++) {
        fRp[c * nR + R[r]] = af[r];
#pragma unroll
        for (int a = 0; a < DIM; a++) gRp[(c * DIM + a) * nR + R[r]] = ag[r][a];
    }
}

// fixed-order sums of the partials
__global__ __launch_bounds__(256) void k_stat_finish(int64_t nR, int64_t nx, int dim, int64_t nchunk, int64_t nfx,
                                                     const double* __restrict__ fRp, const double* __restrict__ gRp,
                                                     const double* __restrict__ fxp, double* __restrict__ fR,
                                                     double* __restrict__ gR, double* __restrict__ fx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nR) {
        double s = 0.0;
        for (int64_t c = 0; c < nchunk; c++) s += fRp[c * nR + i];
        fR[i] = s;
        for (int a = 0; a < dim; a++) {
            double g = 0.0;
            for (int64_t c = 0; c < nchunk; c++) g += gRp[(c * dim + a) * nR + i];
            gR[a * nR + i] = g;
        }
    }
    if (i < nx) {
        double s = 0.0;
        for (int64_t p = 0; p < nfx; p++) s += fxp[p * nx + i];
        fx[i] = s;
    }
}

StatShape stat_shape(int B, int dim, int64_t nx) {
    StatShape s;
    const int n2 = 2 * B;
    s.nR = (int64_t)n2 * n2 * n2;
    s.nx = nx;
    s.dim = dim;
    s.ntile = (s.nR + kStatTile - 1) / kStatTile;
    // >= 2048 workgroups to fill 256 CUs, but >= 64 omega points per chunk so
    // the partial writes stay small next to the reads
    int64_t xc = (2048 + s.ntile - 1) / s.ntile;
    const int64_t cap = nx / 64 > 1 ? nx / 64 : 1;
    if (xc > cap) xc = cap;
    s.nchunk = xc;
    s.xper = (nx + xc - 1) / xc;
    s.nchunk = (nx + s.xper - 1) / s.xper;
    s.nfx = s.ntile * (kStatThreads / 64);
    s.work_doubles = s.nchunk * (1 + dim) * s.nR + s.nfx * nx;
    return s;
}

template <class T>
void launch_stat_pass(const StatShape& s, const void* f, int B, const double* w, const double* xg, double* work,
                      double* fR, double* gR, double* fx, hipStream_t st) {
    double* fRp = work;
    double* gRp = fRp + s.nchunk * s.nR;
    double* fxp = gRp + s.nchunk * s.dim * s.nR;
    const dim3 grid((unsigned)s.ntile, (unsigned)s.nchunk);
    const int n2 = 2 * B;
    switch (s.dim) {
    case 0: hipLaunchKernelGGL((k_stat_pass<T, 0>), grid, dim3(kStatThreads), 0, st, (const T*)f, s.nR, s.nx, n2, s.xper, w, xg, fRp, gRp, fxp); break;
    case 1: hipLaunchKernelGGL((k_stat_pass<T, 1>), grid, dim3(kStatThreads), 0, st, (const T*)f, s.nR, s.nx, n2, s.xper, w, xg, fRp, gRp, fxp); break;
    case 2: hipLaunchKernelGGL((k_stat_pass<T, 2>), grid, dim3(kStatThreads), 0, st, (const T*)f, s.nR, s.nx, n2, s.xper, w, xg, fRp, gRp, fxp); break;
    default: hipLaunchKernelGGL((k_stat_pass<T, 3>), grid, dim3(kStatThreads), 0, st, (const T*)f, s.nR, s.nx, n2, s.xper, w, xg, fRp, gRp, fxp); break;
    }
    const int64_t n = s.nR > s.nx ? s.nR : s.nx;
    hipLaunchKernelGGL(k_stat_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s.nR, s.nx, s.dim,
                       s.nchunk, s.nfx, fRp, gRp, fxp, fR, gR, fx);
}

template void launch_stat_pass<double>(const StatShape&, const void*, int, const double*, const double*, double*,
                                       double*, double*, double*, hipStream_t);
template void launch_stat_pass<float>(const StatShape&, const void*, int, const double*, const double*, double*,
                                      double*, double*, double*, hipStream_t);

}  // namespace hpup
