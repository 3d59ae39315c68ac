// Grid statistics of get_stat (pendulum_reduced.m:485-496, pendulum.m:547-558):
// the omega marginal fx[x] = sum_R w_j f[R,x] and the rotation marginal
// fR[R] = sum_x f[R,x], both accumulated in FP64.  E[R], E[x], Var[x] are
// small contractions of these finished on the host.
#include "hpup_internal.h"
#include "kernels_common.h"
#include "stats.h"

namespace hpup {

// one block per omega point: fx[x] = sum_{i,j,k} w_j f[i + n2 (j + n2 (k + n2 x))]
template <class T>
__global__ __launch_bounds__(256) void k_marg_x(const T* __restrict__ f, const double* __restrict__ w, int n2,
                                                double* __restrict__ fx) {
    __shared__ double red[256];
    const int64_t x = blockIdx.x;
    const int64_t nR = (int64_t)n2 * n2 * n2;
    const T* fp = f + nR * x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < nR; i += blockDim.x) s += (double)fp[i] * w[(i / n2) % n2];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) fx[x] = red[0];
}

// thread per rotation grid point: fR[R] = sum_x f[R + nR x]
template <class T>
__global__ __launch_bounds__(256) void k_marg_R(const T* __restrict__ f, int64_t nR, int64_t nx, double* __restrict__ fR) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= nR) return;
    double s = 0.0;
    for (int64_t x = 0; x < nx; x++) s += (double)f[R + nR * x];
    fR[R] = s;
}

template <class T>
void launch_marginals(const void* f, int B, int64_t nx, const double* w_dev, double* fx, double* fR, hipStream_t st) {
    const int n2 = 2 * B;
    const int64_t nR = (int64_t)n2 * n2 * n2;
    hipLaunchKernelGGL(k_marg_x<T>, dim3((unsigned)nx), dim3(256), 0, st, (const T*)f, w_dev, n2, fx);
    hipLaunchKernelGGL(k_marg_R<T>, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, st, (const T*)f, nR, nx, fR);
}

template void launch_marginals<double>(const void*, int, int64_t, const double*, double*, double*, hipStream_t);
template void launch_marginals<float>(const void*, int, int64_t, const double*, double*, double*, hipStream_t);

}  // namespace hpup
