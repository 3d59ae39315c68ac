// SO(3) x R^n forward / inverse transforms (fftSO3R.m:90-106, ifftSO3R.m:16-37).
//
// Forward:  grid f (alpha,beta,gamma,x)
//   1. k_fwd_ag   per (x-chunk, beta_j): real 2-D DFT over (alpha, gamma) with
//                 the flip(fftshift) sign convention (frequency -m stored at m),
//                 Hermitian half computed, the other half by conjugation
//                 -> G[(m,n)][j][x]
//   2. k_beta_fwd per (m,n): F^l_mn(x) = sum_j w_j d^l_mn(beta_j) G[(m,n)][j][x]
//   3. omega DFT of every compact row (slab kernel on axes 1,2; axis kernel on 3)
// Inverse: the same stages reversed (1/nx on omega, unnormalized over alpha,
// gamma, real part), Nyquist bins of alpha and gamma implicitly zero.
#include "drift.h"
#include "kernels_common.h"

namespace hpup {

// ---------------------------------------------------------- forward 2-D --

template <class T>
__global__ __launch_bounds__(256) void k_fwd_ag(XformDev<T> X, const T* __restrict__ f, const T* __restrict__ wR,
                                                int wstride, cT<T>* __restrict__ G, int XC) {
    extern __shared__ char smem_raw[];
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    cT<T>* sH = reinterpret_cast<cT<T>*>(smem_raw);      // [xl][g][m>=0]  XC*n2*B
    cT<T>* tw = sH + (size_t)XC * n2 * B;                 // n2
    T* sf = reinterpret_cast<T*>(tw + n2);                // [xl][g][a]     XC*n2*n2
    const int j = blockIdx.y;
    const int64_t x0 = (int64_t)blockIdx.x * XC;
    const int tid = threadIdx.x;
    for (int t = tid; t < n2; t += blockDim.x) tw[t] = X.twB[t];
    for (int i = tid; i < XC * n2 * n2; i += blockDim.x) {
        const int xl = i / (n2 * n2), g = (i / n2) % n2, a = i % n2;
        const int64_t x = x0 + xl;
        const int64_t iR = a + (int64_t)n2 * (j + (int64_t)n2 * g);
        T v = x < X.nx ? f[iR + (int64_t)n2 * n2 * n2 * x] : T(0);
        if (wR) v *= wR[iR * wstride];  // transform of the product f(R, x) w(R)
        sf[i] = v;
    }
    __syncthreads();
    // alpha DFT, m >= 0:  H = sum_a f e^{+i m alpha_a}
    for (int i = tid; i < XC * n2 * B; i += blockDim.x) {
        const int m = i % B, g = (i / B) % n2, xl = i / (B * n2);
        const T* row = sf + ((size_t)xl * n2 + g) * n2;
        cT<T> acc = czero<T>();
        int t = 0;
        for (int a = 0; a < n2; a++) {
            const cT<T> w = tw[t];
            acc.x = fma(row[a], w.x, acc.x);
            acc.y = fma(row[a], w.y, acc.y);
            t += m;
            if (t >= n2) t -= n2;
        }
        sH[i] = acc;
    }
    __syncthreads();
    // gamma DFT: G[m][n] = sum_g H[g][m] e^{+i n gamma_g}; G[-m][-n] = conj G[m][n]
    const int64_t nx = X.nx;
    for (int i = tid; i < XC * B * M; i += blockDim.x) {
        const int xl = i % XC, mn = i / XC;
        const int m = mn % B, n = mn / B - lmax;
        const int64_t x = x0 + xl;
        if (x >= nx) continue;
        cT<T> acc = czero<T>();
        const int step = ((n % n2) + n2) % n2;
        int t = 0;
        for (int g = 0; g < n2; g++) {
            cfma<T>(acc, sH[((size_t)xl * n2 + g) * B + m], tw[t]);
            t += step;
            if (t >= n2) t -= n2;
        }
        G[((int64_t)(m + lmax) + M * (n + lmax)) * n2 * nx + (int64_t)j * nx + x] = acc;
        if (m > 0) G[((int64_t)(-m + lmax) + M * (-n + lmax)) * n2 * nx + (int64_t)j * nx + x] = cconj<T>(acc);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_beta_fwd(XformDev<T> X, const cT<T>* __restrict__ G, cT<T>* __restrict__ F) {
    extern __shared__ char smem_raw[];
    T* sW = reinterpret_cast<T*>(smem_raw);
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    const int mn = blockIdx.y;
    const int m = mn % M - lmax, n = mn / M - lmax;
    const int lmin = max(abs(m), abs(n));
    for (int i = threadIdx.x; i < B * n2; i += blockDim.x) sW[i] = X.Wf[(size_t)mn * B * n2 + i];
    __syncthreads();
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= X.nx) return;
    const cT<T>* Gp = G + (int64_t)mn * n2 * X.nx + x;
    for (int l0 = lmin; l0 <= lmax; l0 += 8) {
        cT<T> acc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = czero<T>();
        for (int j = 0; j < n2; j++) {
            const cT<T> g = Gp[(int64_t)j * X.nx];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                if (l0 + t > lmax) break;
                const T w = sW[(l0 + t) * n2 + j];
                acc[t].x = fma(w, g.x, acc[t].x);
                acc[t].y = fma(w, g.y, acc[t].y);
            }
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (l0 + t > lmax) break;
            F[d_ridx(B, l0 + t, m, n) * X.nx + x] = acc[t];
        }
    }
}

// ---------------------------------------------------------- inverse -------

template <class T>
__global__ __launch_bounds__(256) void k_beta_inv(XformDev<T> X, const cT<T>* __restrict__ F, cT<T>* __restrict__ G) {
    extern __shared__ char smem_raw[];
    T* sW = reinterpret_cast<T*>(smem_raw);
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    const int mn = blockIdx.y;
    const int m = mn % M - lmax, n = mn / M - lmax;
    const int lmin = max(abs(m), abs(n));
    for (int i = threadIdx.x; i < B * n2; i += blockDim.x) sW[i] = X.Wi[(size_t)mn * B * n2 + i];
    __syncthreads();
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= X.nx) return;
    cT<T>* Gp = G + (int64_t)mn * n2 * X.nx + x;
    for (int j0 = 0; j0 < n2; j0 += 8) {
        cT<T> acc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = czero<T>();
        for (int l = lmin; l <= lmax; l++) {
            const cT<T> fv = F[d_ridx(B, l, m, n) * X.nx + x];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                if (j0 + t >= n2) break;
                const T w = sW[(j0 + t) * B + l];
                acc[t].x = fma(w, fv.x, acc[t].x);
                acc[t].y = fma(w, fv.y, acc[t].y);
            }
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (j0 + t >= n2) break;
            Gp[(int64_t)(j0 + t) * X.nx] = acc[t];
        }
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_inv_ag(XformDev<T> X, const cT<T>* __restrict__ G, T* __restrict__ f, int XC) {
    extern __shared__ char smem_raw[];
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    cT<T>* sH = reinterpret_cast<cT<T>*>(smem_raw);      // [m][g][xl]
    cT<T>* tw = sH + (size_t)M * n2 * XC;                 // n2
    const int j = blockIdx.y;
    const int64_t x0 = (int64_t)blockIdx.x * XC;
    const int64_t nx = X.nx;
    const int tid = threadIdx.x;
    for (int t = tid; t < n2; t += blockDim.x) tw[t] = X.twB[t];
    __syncthreads();
    // over n -> gamma: H[m][g] = sum_n G[m][n] e^{-i n gamma_g}  (G rows read through the cache)
    for (int i = tid; i < M * n2 * XC; i += blockDim.x) {
        const int xl = i % XC, g = (i / XC) % n2, mi = i / (XC * n2);
        const int64_t x = x0 + xl;
        cT<T> acc = czero<T>();
        if (x < nx) {
            for (int ni = 0; ni < M; ni++) {
                const int n = ni - lmax;
                const int t = (int)((((int64_t)n * g) % n2 + n2) % n2);
                const cT<T> gv = G[((int64_t)mi + (int64_t)M * ni) * n2 * nx + (int64_t)j * nx + x];
                cfma<T>(acc, gv, cconj<T>(tw[t]));
            }
        }
        sH[i] = acc;
    }
    __syncthreads();
    // over m -> alpha, real part
    for (int i = tid; i < XC * n2 * n2; i += blockDim.x) {
        const int a = i % n2, g = (i / n2) % n2, xl = i / (n2 * n2);
        const int64_t x = x0 + xl;
        if (x >= nx) continue;
        T acc = T(0);
        for (int mi = 0; mi < M; mi++) {
            const int m = mi - lmax;
            const int t = (int)((((int64_t)m * a) % n2 + n2) % n2);
            const cT<T> h = sH[((size_t)mi * n2 + g) * XC + xl];
            const cT<T> w = tw[t];  // e^{-i m alpha} = conj(w): Re(h conj(w)) = h.x w.x + h.y w.y
            acc = fma(h.x, w.x, fma(h.y, w.y, acc));
        }
        f[a + (int64_t)n2 * (j + (int64_t)n2 * (g + (int64_t)n2 * x))] = acc;
    }
}

// ---------------------------------------------------------- omega DFT -----

// In-place 2-D DFT over (k1, k2) of every N x N slab, twiddle-indexed.
template <class T>
__global__ __launch_bounds__(1024) void k_slab_dft(cT<T>* __restrict__ F, const cT<T>* __restrict__ twN, int N, T scale) {
    extern __shared__ char smem_raw[];
    const int NN = N * N;
    cT<T>* sA = reinterpret_cast<cT<T>*>(smem_raw);
    cT<T>* sB = sA + NN;
    cT<T>* tw = sB + NN;
    const int64_t base = (int64_t)blockIdx.x * NN;
    for (int e = threadIdx.x; e < NN; e += blockDim.x) sA[e] = F[base + e];
    for (int e = threadIdx.x; e < N; e += blockDim.x) tw[e] = twN[e];
    __syncthreads();
    for (int e = threadIdx.x; e < NN; e += blockDim.x) {
        const int k1 = e % N, k2 = e / N;
        cT<T> acc = czero<T>();
        int t = 0;
        for (int a = 0; a < N; a++) {
            cfma<T>(acc, tw[t], sA[k2 * N + a]);
            t += k1;
            if (t >= N) t -= N;
        }
        sB[e] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NN; e += blockDim.x) {
        const int k1 = e % N, k2 = e / N;
        cT<T> acc = czero<T>();
        int t = 0;
        for (int a = 0; a < N; a++) {
            cfma<T>(acc, tw[t], sB[a * N + k1]);
            t += k2;
            if (t >= N) t -= N;
        }
        F[base + e] = cscale<T>(acc, scale);
    }
}

// ---------------------------------------------------------- layouts -------

template <class T>
__global__ void k_full_to_compact(int B, int64_t nx, const cT<T>* __restrict__ full, cT<T>* __restrict__ comp) {
    __shared__ cT<T> tile[32][33];
    const int lmax = B - 1, M = 2 * B - 1;
    const int64_t nrc = (int64_t)B * M * (2 * B + 1) / 3;
    const int64_t r0 = (int64_t)blockIdx.x * 32, x0 = (int64_t)blockIdx.y * 32;
    const int64_t nfr = (int64_t)M * M * B;
    for (int yy = threadIdx.y; yy < 32; yy += blockDim.y) {
        const int64_t r = r0 + threadIdx.x, x = x0 + yy;
        cT<T> v = czero<T>();
        if (r < nrc && x < nx) {
            int l, m, n;
            d_rlmn(B, r, l, m, n);
            v = full[(m + lmax) + (int64_t)M * ((n + lmax) + (int64_t)M * l) + nfr * x];
        }
        tile[yy][threadIdx.x] = v;
    }
    __syncthreads();
    for (int yy = threadIdx.y; yy < 32; yy += blockDim.y) {
        const int64_t r = r0 + yy, x = x0 + threadIdx.x;
        if (r < nrc && x < nx) comp[r * nx + x] = tile[threadIdx.x][yy];
    }
}

template <class T>
__global__ void k_compact_to_full(int B, int64_t nx, const cT<T>* __restrict__ comp, cT<T>* __restrict__ full) {
    __shared__ cT<T> tile[32][33];
    const int lmax = B - 1, M = 2 * B - 1;
    const int64_t nfr = (int64_t)M * M * B;
    const int64_t f0 = (int64_t)blockIdx.x * 32, x0 = (int64_t)blockIdx.y * 32;
    for (int yy = threadIdx.y; yy < 32; yy += blockDim.y) {
        const int64_t fr = f0 + yy, x = x0 + threadIdx.x;
        cT<T> v = czero<T>();
        if (fr < nfr && x < nx) {
            const int m = (int)(fr % M) - lmax, n = (int)((fr / M) % M) - lmax, l = (int)(fr / ((int64_t)M * M));
            if (abs(m) <= l && abs(n) <= l) v = comp[d_ridx(B, l, m, n) * nx + x];
        }
        tile[yy][threadIdx.x] = v;
    }
    __syncthreads();
    for (int yy = threadIdx.y; yy < 32; yy += blockDim.y) {
        const int64_t fr = f0 + threadIdx.x, x = x0 + yy;
        if (fr < nfr && x < nx) full[fr + nfr * x] = tile[threadIdx.x][yy];
    }
}

// ---------------------------------------------------------- launchers -----

static int pick_xc(size_t per_x, size_t budget, int64_t nx) {
    int xc = (int)std::max<size_t>(1, budget / per_x);
    if (xc > 16) xc = 16;
    if (xc > nx) xc = (int)nx;
    return std::max(xc, 1);
}

template <class T>
void launch_fwd_ag(const XformDev<T>& X, const void* f, const void* wR, int wstride, void* G, hipStream_t st) {
    const int n2 = 2 * X.B;
    const size_t per_x = sizeof(T) * n2 * n2 + sizeof(cT<T>) * n2 * X.B;
    const int XC = pick_xc(per_x, kLdsTile - sizeof(cT<T>) * n2, X.nx);
    const size_t sh = per_x * XC + sizeof(cT<T>) * n2;
    allow_lds((const void*)k_fwd_ag<T>, sh);
    dim3 grid((unsigned)((X.nx + XC - 1) / XC), (unsigned)n2);
    hipLaunchKernelGGL(k_fwd_ag<T>, grid, dim3(256), sh, st, X, (const T*)f, (const T*)wR, wstride, (cT<T>*)G, XC);
}

template <class T>
void launch_beta_fwd(const XformDev<T>& X, const void* G, void* F, hipStream_t st) {
    const int M = 2 * X.B - 1;
    dim3 grid((unsigned)((X.nx + 255) / 256), (unsigned)(M * M));
    hipLaunchKernelGGL(k_beta_fwd<T>, grid, dim3(256), sizeof(T) * X.B * 2 * X.B, st, X, (const cT<T>*)G,
                       (cT<T>*)F);
}

template <class T>
void launch_beta_inv(const XformDev<T>& X, const void* F, void* G, hipStream_t st) {
    const int M = 2 * X.B - 1;
    dim3 grid((unsigned)((X.nx + 255) / 256), (unsigned)(M * M));
    hipLaunchKernelGGL(k_beta_inv<T>, grid, dim3(256), sizeof(T) * X.B * 2 * X.B, st, X, (const cT<T>*)F,
                       (cT<T>*)G);
}

template <class T>
void launch_inv_ag(const XformDev<T>& X, const void* G, void* f, hipStream_t st) {
    const int n2 = 2 * X.B, M = 2 * X.B - 1;
    const size_t per_x = sizeof(cT<T>) * ((size_t)M * n2);
    const int XC = pick_xc(per_x, kLdsTile - sizeof(cT<T>) * n2, X.nx);
    const size_t sh = per_x * XC + sizeof(cT<T>) * n2;
    allow_lds((const void*)k_inv_ag<T>, sh);
    dim3 grid((unsigned)((X.nx + XC - 1) / XC), (unsigned)n2);
    hipLaunchKernelGGL(k_inv_ag<T>, grid, dim3(256), sh, st, X, (const cT<T>*)G, (T*)f, XC);
}

template <class T>
void launch_omega_dft(const XformDev<T>& X, void* F, void* tmp, bool inverse, hipStream_t st) {
    (void)tmp;
    if (X.dim == 0) return;
    const int N = X.N, NN = N * N;
    HPUP_REQUIRE(NN <= 4096, HPUP_ERR_UNSUPPORTED, "omega grid larger than 64 per axis");
    const cT<T>* Mat = inverse ? X.idftN : X.dftN;
    // twiddles: row 1 of the matrix (unscaled)
    const T sc2 = inverse ? T(1) / T(NN) : T(1);
    const int64_t nslab = X.nrc * (X.dim == 3 ? N : 1);
    const int nt = NN <= 256 ? 256 : 1024;
    const size_t sh = sizeof(cT<T>) * (2 * (size_t)NN + N);
    allow_lds((const void*)k_slab_dft<T>, sh);
    hipLaunchKernelGGL(k_slab_dft<T>, dim3((unsigned)nslab), dim3(nt), sh, st, (cT<T>*)F, inverse ? X.twNi : X.twNf,
                       N, sc2);
    if (X.dim == 3)
        launch_axis_outer<T>(F, F, Mat, N, (int64_t)NN, X.nrc, T(1), st);
}

template <class T>
void launch_full_to_compact(int B, int64_t nx, const void* full, void* compact, hipStream_t st) {
    const int64_t nrc = nrc_of(B);
    dim3 grid((unsigned)((nrc + 31) / 32), (unsigned)((nx + 31) / 32));
    hipLaunchKernelGGL(k_full_to_compact<T>, grid, dim3(32, 8), 0, st, B, nx, (const cT<T>*)full,
                       (cT<T>*)compact);
}

template <class T>
void launch_compact_to_full(int B, int64_t nx, const void* compact, void* full, hipStream_t st) {
    const int M = 2 * B - 1;
    const int64_t nfr = (int64_t)M * M * B;
    dim3 grid((unsigned)((nfr + 31) / 32), (unsigned)((nx + 31) / 32));
    hipLaunchKernelGGL(k_compact_to_full<T>, grid, dim3(32, 8), 0, st, B, nx, (const cT<T>*)compact,
                       (cT<T>*)full);
}

#define HPUP_INST(T)                                                                                  \
    template void launch_fwd_ag<T>(const XformDev<T>&, const void*, const void*, int, void*, hipStream_t); \
    template void launch_beta_fwd<T>(const XformDev<T>&, const void*, void*, hipStream_t);           \
    template void launch_beta_inv<T>(const XformDev<T>&, const void*, void*, hipStream_t);           \
    template void launch_inv_ag<T>(const XformDev<T>&, const void*, void*, hipStream_t);             \
    template void launch_omega_dft<T>(const XformDev<T>&, void*, void*, bool, hipStream_t);          \
    template void launch_full_to_compact<T>(int, int64_t, const void*, void*, hipStream_t);          \
    template void launch_compact_to_full<T>(int, int64_t, const void*, void*, hipStream_t);
HPUP_INST(double)
HPUP_INST(float)
#undef HPUP_INST

}  // namespace hpup
