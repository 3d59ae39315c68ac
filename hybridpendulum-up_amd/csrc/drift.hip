// Spectral update (derivative of the Fokker-Planck operator) on the compact
// coefficient layout [r][k3][k2][k1] (omega index innermost).
//
// The reference's derivative (pendulum.m:388-527, pendulum_reduced.m:362-465)
// is evaluated in its exact nonzero structure:
//   dF = T1[V1] + T2[V2] + b1 c1 T1F + b2 c2 T2F + (e3 + b3 c3) T3F + W
//   V1 = -U1 F - a2 c2 T3F - a3 c3 T2F,  V2 = -U2 F - a1 c1 T3F
//   W  = -sum_p c_p CG_p[F] + D F
// where T_p is the 1-D circulant along omega axis p (F (*) X_p / nx), c_p the
// omega-derivative factor i*kc(k_p), U_p the n -> n+-1 Lie-derivative mixing,
// e3 the diagonal u_3 factor, a_p the gyroscopic coefficients, CG_p the banded
// Clebsch-Gordan product and D the diffusion factor.  Three passes:
//   A  rotation mixing, pointwise in omega:       W, -U1 F, -U2 F
//   B  circulant along axis 3:                      Z3 = T3 F
//   C  fused axis-1/axis-2 slab kernel + integrator epilogue.
#include "drift.h"
#include "kernels_common.h"

namespace hpup {

// ------------------------------------------------------------- pass A ---

template <class T>
__global__ __launch_bounds__(256) void k_rotmix(DriftDev<T> D, const cT<T>* __restrict__ F,
                                                cT<T>* __restrict__ W, cT<T>* __restrict__ V1,
                                                cT<T>* __restrict__ V2) {
    const int64_t r = blockIdx.x;
    const int64_t x = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
    if (x >= D.nx) return;
    const int N = D.N;
    int k[3];
    k[0] = (int)(x % N);
    k[1] = (int)((x / N) % N);
    k[2] = D.dim == 3 ? (int)(x / ((int64_t)N * N)) : 0;
    T kc[3], kd[3];
    for (int p = 0; p < D.dim; p++) { kc[p] = D.kc[k[p]]; kd[p] = D.kd[k[p]]; }

    cT<T> acc[3] = {czero<T>(), czero<T>(), czero<T>()};
    for (int t = 0; t < D.TG; t++) {
        const int s = D.cg_src[r * D.TG + t];
        if (s < 0) break;
        const cT<T> f = F[(int64_t)s * D.nx + x];
        const cT<T>* c = D.cg_coef + (r * D.TG + t) * D.dim;
        for (int p = 0; p < D.dim; p++) cfma<T>(acc[p], c[p], f);
    }
    const cT<T> Fr = F[r * D.nx + x];
    T dd = T(0);
    for (int i = 0; i < D.dim; i++)
        for (int j = 0; j < D.dim; j++) {
            const T g = D.G[i + D.dim * j];
            dd += (i == j) ? g * kd[i] : -g * kc[i] * kc[j];
        }
    cT<T> w = cscale<T>(Fr, dd);
    for (int p = 0; p < D.dim; p++) w = csub<T>(w, cimul<T>(acc[p], kc[p]));
    for (int p = 0; p < D.dim; p++)  // gravity evaluated on the grid (density form)
        if (D.grav[p]) w = csub<T>(w, cimul<T>(D.grav[p][r * D.nx + x], kc[p]));
    W[r * D.nx + x] = w;

    for (int q = 0; q < 2; q++) {
        cT<T> a = czero<T>();
        const int64_t row = (q * D.nrc + r) * D.TU;
        for (int t = 0; t < D.TU; t++) {
            const int s = D.u_src[row + t];
            if (s < 0) break;
            cfma<T>(a, D.u_coef[row + t], F[(int64_t)s * D.nx + x]);
        }
        (q == 0 ? V1 : V2)[r * D.nx + x] = a;
    }
}

// ------------------------------------------------------------- pass B ---

// out[L][k][j] = scale * sum_a M[k][a] in[L][a][j],  j in [0, inner).
template <class T>
__global__ __launch_bounds__(256) void k_axis_outer(const cT<T>* in, cT<T>* out,
                                                    const cT<T>* __restrict__ Mat, int N, int64_t inner,
                                                    T scale) {
    extern __shared__ char smem_raw[];
    cT<T>* sM = reinterpret_cast<cT<T>*>(smem_raw);
    cT<T>* sT = sM + N * N;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * 64 + tx;
    const int64_t L = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * 64 + tx;
    for (int i = tid; i < N * N; i += 256) sM[i] = Mat[i];
    for (int a = ty; a < N; a += 4) sT[a * 64 + tx] = (j < inner) ? in[(L * N + a) * inner + j] : czero<T>();
    __syncthreads();
    if (j >= inner) return;
    for (int k = ty; k < N; k += 4) {
        cT<T> acc = czero<T>();
        for (int a = 0; a < N; a++) cfma<T>(acc, sM[k * N + a], sT[a * 64 + tx]);
        out[(L * N + k) * inner + j] = cscale<T>(acc, scale);
    }
}

// ------------------------------------------------------------- pass C ---

template <class T, int EPT>
__global__ __launch_bounds__(256) void k_drift_slab(DriftDev<T> D, const cT<T>* Fin,
                                                     const cT<T>* __restrict__ Z3, const cT<T>* __restrict__ W,
                                                     const cT<T>* __restrict__ V1g, const cT<T>* __restrict__ V2g,
                                                     const cT<T>* F0, cT<T>* Fout, cT<T>* ACC, StageCoef<T> sc) {
    extern __shared__ char smem_raw[];
    const int N = D.N, NN = N * N;
    cT<T>* sA = reinterpret_cast<cT<T>*>(smem_raw);
    cT<T>* sB = sA + NN;
    cT<T>* k1s = sB + NN;
    cT<T>* k2s = k1s + N;
    const int nz = D.dim == 3 ? N : 1;
    const int64_t s = blockIdx.x;
    const int64_t r = s / nz;
    const int k3 = (int)(s % nz);
    const int64_t base = s * NN;
    const int tid = threadIdx.x, nt = blockDim.x;

    for (int e = tid; e < NN; e += nt) sA[e] = Fin[base + e];
    for (int e = tid; e < N; e += nt) { k1s[e] = D.kap[e]; k2s[e] = D.kap[N + e]; }
    __syncthreads();

    const bool d3 = D.dim == 3;
    const T kc3 = d3 ? D.kc[k3] : T(0);
    cT<T> P1[EPT], P2[EPT], Zr[EPT];
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int e = tid + i * nt;
        P1[i] = czero<T>(); P2[i] = czero<T>(); Zr[i] = czero<T>();
        if (e >= NN) continue;
        const int k1 = e % N, k2 = e / N;
        int i1 = k1, i2 = k2;
        for (int a = 0; a < N; a++) {
            cfma<T>(P1[i], k1s[i1], sA[k2 * N + a]);
            cfma<T>(P2[i], k2s[i2], sA[a * N + k1]);
            i1 = i1 == 0 ? N - 1 : i1 - 1;
            i2 = i2 == 0 ? N - 1 : i2 - 1;
        }
    }
    cT<T> V1[EPT], V2[EPT];
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int e = tid + i * nt;
        if (e >= NN) continue;
        const int k1 = e % N, k2 = e / N;
        V1[i] = V1g[base + e];
        V2[i] = V2g[base + e];
        if (d3) {
            Zr[i] = Z3[base + e];
            if (D.has_gyro) {
                const T c1 = D.kc[k1], c2 = D.kc[k2];
                V1[i] = csub<T>(V1[i], cimul<T>(Zr[i], c2 * D.gyro[1]));
                V1[i] = csub<T>(V1[i], cimul<T>(P2[i], kc3 * D.gyro[2]));
                V2[i] = csub<T>(V2[i], cimul<T>(Zr[i], c1 * D.gyro[0]));
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int e = tid + i * nt;
        if (e >= NN) continue;
        sA[e] = V1[i];
        sB[e] = V2[i];
    }
    __syncthreads();
    const cT<T> e3 = d3 ? D.e3[r] : czero<T>();
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int e = tid + i * nt;
        if (e >= NN) continue;
        const int k1 = e % N, k2 = e / N;
        cT<T> q = W[base + e];
        int i1 = k1, i2 = k2;
        for (int a = 0; a < N; a++) {
            cfma<T>(q, k1s[i1], sA[k2 * N + a]);
            cfma<T>(q, k2s[i2], sB[a * N + k1]);
            i1 = i1 == 0 ? N - 1 : i1 - 1;
            i2 = i2 == 0 ? N - 1 : i2 - 1;
        }
        q = cadd<T>(q, cimul<T>(P1[i], D.b[0] * D.kc[k1]));
        q = cadd<T>(q, cimul<T>(P2[i], D.b[1] * D.kc[k2]));
        if (d3) {
            q = cadd<T>(q, cmul<T>(Zr[i], e3));
            q = cadd<T>(q, cimul<T>(Zr[i], D.b[2] * kc3));
        }
        if (sc.deriv_only) {
            Fout[base + e] = q;
            continue;
        }
        const cT<T> f0 = sc.read_f0 ? F0[base + e] : Fin[base + e];
        const cT<T> ac = sc.read_acc ? ACC[base + e] : czero<T>();
        cT<T> upd = cadd<T>(cscale<T>(ac, sc.ca), cscale<T>(q, sc.cd));
        Fout[base + e] = cadd<T>(f0, cscale<T>(upd, sc.dt));
        if (sc.write_acc) ACC[base + e] = cadd<T>(cscale<T>(ac, sc.ea), cscale<T>(q, sc.ed));
    }
}

// ---------------------------------------------------------- launchers ---

template <class T>
void launch_rotmix(const DriftDev<T>& D, const void* F, void* W, void* V1, void* V2, hipStream_t st) {
    dim3 grid((unsigned)D.nrc, (unsigned)((D.nx + 255) / 256));
    hipLaunchKernelGGL(k_rotmix<T>, grid, dim3(256), 0, st, D, (const cT<T>*)F, (cT<T>*)W, (cT<T>*)V1,
                       (cT<T>*)V2);
}

template <class T>
void launch_axis_outer(const void* in, void* out, const void* Mat, int N, int64_t inner, int64_t nlines,
                       T scale, hipStream_t st) {
    dim3 grid((unsigned)((inner + 63) / 64), (unsigned)nlines);
    size_t sh = sizeof(cT<T>) * ((size_t)N * N + 64 * (size_t)N);
    allow_lds((const void*)k_axis_outer<T>, sh);
    hipLaunchKernelGGL(k_axis_outer<T>, grid, dim3(64, 4), sh, st, (const cT<T>*)in, (cT<T>*)out,
                       (const cT<T>*)Mat, N, inner, scale);
}

template <class T>
void launch_drift_slab(const DriftDev<T>& D, const void* Fin, const void* Z3, const void* W, const void* V1,
                       const void* V2, const void* F0, void* Fout, void* ACC, const StageCoef<T>& sc,
                       hipStream_t st) {
    const int N = D.N, NN = N * N;
    const int64_t nslab = D.nrc * (D.dim == 3 ? N : 1);
    const size_t sh = sizeof(cT<T>) * (2 * (size_t)NN + 2 * (size_t)N);
    HPUP_REQUIRE(NN <= 16 * 256, HPUP_ERR_UNSUPPORTED, "omega grid larger than 64 per axis");
    const int ept = (NN + 255) / 256;
    auto go = [&](auto kern) {
        allow_lds((const void*)kern, sh);
        hipLaunchKernelGGL(kern, dim3((unsigned)nslab), dim3(256), sh, st, D, (const cT<T>*)Fin, (const cT<T>*)Z3,
                           (const cT<T>*)W, (const cT<T>*)V1, (const cT<T>*)V2, (const cT<T>*)F0, (cT<T>*)Fout,
                           (cT<T>*)ACC, sc);
    };
    if (ept <= 1) go(k_drift_slab<T, 1>);
    else if (ept <= 2) go(k_drift_slab<T, 2>);
    else if (ept <= 4) go(k_drift_slab<T, 4>);
    else if (ept <= 8) go(k_drift_slab<T, 8>);
    else go(k_drift_slab<T, 16>);
}

// out[r][x] += scale sum_t coef[r][t] in[src[r][t]][x]  (rows of nx points; src -1 padded)
template <class T>
__global__ __launch_bounds__(256) void k_row_mix_add(const cT<T>* __restrict__ in, cT<T>* __restrict__ out,
                                                     const int32_t* __restrict__ src, const cT<T>* __restrict__ coef,
                                                     int TS, int64_t nx, T scale) {
    const int64_t r = blockIdx.y;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    cT<T> acc = czero<T>();
    for (int t = 0; t < TS; t++) {
        const int32_t sr = src[r * TS + t];
        if (sr < 0) break;
        cfma<T>(acc, coef[r * TS + t], in[(int64_t)sr * nx + x]);
    }
    out[r * nx + x] = cadd<T>(out[r * nx + x], cscale<T>(acc, scale));
}

template <class T>
void launch_row_mix_add(const void* in, void* out, const int32_t* src, const void* coef, int TS, int64_t nrows,
                        int64_t nx, T scale, hipStream_t st) {
    HPUP_REQUIRE(nrows <= 65535, HPUP_ERR_UNSUPPORTED, "too many rows for one launch");
    dim3 grid((unsigned)((nx + 255) / 256), (unsigned)nrows);
    hipLaunchKernelGGL(k_row_mix_add<T>, grid, dim3(256), 0, st, (const cT<T>*)in, (cT<T>*)out, src,
                       (const cT<T>*)coef, TS, nx, scale);
}

template void launch_row_mix_add<double>(const void*, void*, const int32_t*, const void*, int, int64_t, int64_t, double,
                                         hipStream_t);

template void launch_rotmix<double>(const DriftDev<double>&, const void*, void*, void*, void*, hipStream_t);
template void launch_rotmix<float>(const DriftDev<float>&, const void*, void*, void*, void*, hipStream_t);
template void launch_axis_outer<double>(const void*, void*, const void*, int, int64_t, int64_t, double, hipStream_t);
template void launch_axis_outer<float>(const void*, void*, const void*, int, int64_t, int64_t, float, hipStream_t);
template void launch_drift_slab<double>(const DriftDev<double>&, const void*, const void*, const void*, const void*,
                                        const void*, const void*, void*, void*, const StageCoef<double>&, hipStream_t);
template void launch_drift_slab<float>(const DriftDev<float>&, const void*, const void*, const void*, const void*,
                                       const void*, const void*, void*, void*, const StageCoef<float>&, hipStream_t);

}  // namespace hpup
