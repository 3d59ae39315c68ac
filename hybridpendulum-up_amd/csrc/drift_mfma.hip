// MFMA drift kernels for omega grids up to 32 points per axis.
//
// k_drift_fused evaluates the whole derivative of one (r, k3) slab and applies
// the integrator epilogue (see drift.hip for the operator split):
//   P1 = T1 F, P2 = T2 F                                    (MFMA)
//   W  = -sum_p c_p CG_p[F] + D F,  V1 = -U1 F, V2 = -U2 F   (neighbour rows,
//        within 2B+1 rows of r in the m-major layout, read through L2)
//   V1 -= a2 c2 Z3 + a3 c3 P2,  V2 -= a1 c1 Z3              (gyroscopic)
//   Q1 = T1 V1, Q2 = T2 V2                                  (MFMA)
//   dF = Q1 + Q2 + b1 c1 P1 + b2 c2 P2 + (e3 + b3 c3) Z3 + W
// The complex 30x30 circulants act as real 64x64 GEMMs on the (re | im)
// stacked slab: T2 as C2 X (circulant as the A operand), T1 as X C1^T (slab as
// the A operand), so both outputs land in the same lane layout.  Each of the
// four waves owns one 16x16 quadrant of the (zero-padded) 32x32 slab; lane l,
// accumulator v hold element (k2, k1) = (16 k2t + orow(l >> 4, v), 16 k1t + (l & 15)).
//
// The GEMMs are a policy: GemmF (FP32 or FP64 MFMA 16x16x4, K = 4 per
// instruction) or GemmBf16x3 (v_mfma_f32_16x16x32_bf16 on hi/lo bf16 splits of
// both operands, hi*hi + hi*lo + lo*hi accumulated in FP32).
//
// k_axis3 computes Z3 = T3 F (circulant along the outermost omega axis) as the
// GEMM C3 [F_re; F_im] over 64-column chunks of the k1,k2 plane.
#include "drift.h"
#include "kernels_common.h"

namespace hpup {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <class T> struct Mfma;
// D = A(16xK) B(Kx16) + C.  The 16x16x4 instructions take A[l&15][l>>4] and
// B[l>>4][l&15] from lane l; they differ in where output row i lands:
// accumulator v of lane l holds D[orow(l>>4, v)][l&15] (measured on gfx950).
template <> struct Mfma<float> {
    using acc = f32x4;
    static __device__ __forceinline__ acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
    static __device__ __forceinline__ int orow(int lq, int v) { return 4 * lq + v; }
    static __device__ __forceinline__ acc fma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<double> {
    using acc = f64x4;
    static __device__ __forceinline__ acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ int orow(int lq, int v) { return lq + 4 * v; }
    static __device__ __forceinline__ acc fma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

constexpr int kPad = 33;  // LDS row stride of the 32x32 planes

// Real-ified circulant entry for GEMM row/column `o` (output index along the
// axis), reduction index k in [0, 64): (re | im) stacked input, `im_out`
// selects the imaginary output half:  [[Cr, -Ci], [Ci, Cr]].
template <class T>
__device__ __forceinline__ T circ_entry(const cT<T>* kap, int N, int o, int k, bool im_out) {
    const int a = k & 31;
    cT<T> c = czero<T>();
    if (o < N && a < N) c = kap[o >= a ? o - a : o - a + N];
    const bool re_in = k < 32;
    return im_out ? (re_in ? c.y : c.x) : (re_in ? c.x : -c.y);
}

// ------------------------------------------------------------ policies ---

// FP32 / FP64 MFMA, 16 reduction steps of K = 4.
template <class T> struct GemmF {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc;
    T a2re[16], a2im[16], b1re[16], b1im[16];
    int k2t, k1t, lr, lq;
    static __device__ __forceinline__ int orow(int q, int v) { return MF::orow(q, v); }
    __device__ __forceinline__ void init(const cT<T>* kap1, const cT<T>* kap2, int N, int k2t_, int k1t_, int lr_,
                                         int lq_) {
        k2t = k2t_; k1t = k1t_; lr = lr_; lq = lq_;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const int k = 4 * s + lq;
            a2re[s] = circ_entry<T>(kap2, N, 16 * k2t + lr, k, false);
            a2im[s] = circ_entry<T>(kap2, N, 16 * k2t + lr, k, true);
            b1re[s] = circ_entry<T>(kap1, N, 16 * k1t + lr, k, false);
            b1im[s] = circ_entry<T>(kap1, N, 16 * k1t + lr, k, true);
        }
    }
    // Y = X C1^T along k1 (slab rows as the A operand)
    __device__ __forceinline__ void t1(const T* pr, const T* pi, cT<T>* out) const {
        acc_t are = MF::zero(), aim = MF::zero();
        const int row = (16 * k2t + lr) * kPad;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const T av = (s < 8 ? pr : pi)[row + ((4 * s + lq) & 31)];
            are = MF::fma(av, b1re[s], are);
            aim = MF::fma(av, b1im[s], aim);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<T>(are[v], aim[v]);
    }
    // Y = C2 X along k2 (slab columns as the B operand)
    __device__ __forceinline__ void t2(const T* pr, const T* pi, cT<T>* out) const {
        acc_t are = MF::zero(), aim = MF::zero();
        const int col = 16 * k1t + lr;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const T bv = (s < 8 ? pr : pi)[((4 * s + lq) & 31) * kPad + col];
            are = MF::fma(a2re[s], bv, are);
            aim = MF::fma(a2im[s], bv, aim);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<T>(are[v], aim[v]);
    }
};

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// x = hi + lo + O(2^-18 |x|), hi and lo bf16 (round to nearest even)
__device__ __forceinline__ void split8(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const __bf16 h = (__bf16)v[t];
        hi[t] = h;
        lo[t] = (__bf16)(v[t] - (float)h);
    }
}

// 3xBF16: two reduction steps of K = 32 (re half, im half); the 16x16x32 bf16
// instruction takes A[l&15][8(l>>4) + t] and B[8(l>>4) + t][l&15], t < 8.
struct GemmBf16x3 {
    bf16x8 a2re_h[2], a2re_l[2], a2im_h[2], a2im_l[2], b1re_h[2], b1re_l[2], b1im_h[2], b1im_l[2];
    int k2t, k1t, lr, lq;
    static __device__ __forceinline__ int orow(int q, int v) { return 4 * q + v; }
    __device__ __forceinline__ void init(const float2* kap1, const float2* kap2, int N, int k2t_, int k1t_, int lr_,
                                         int lq_) {
        k2t = k2t_; k1t = k1t_; lr = lr_; lq = lq_;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            float v[4][8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int k = 32 * s + 8 * lq + t;
                v[0][t] = circ_entry<float>(kap2, N, 16 * k2t + lr, k, false);
                v[1][t] = circ_entry<float>(kap2, N, 16 * k2t + lr, k, true);
                v[2][t] = circ_entry<float>(kap1, N, 16 * k1t + lr, k, false);
                v[3][t] = circ_entry<float>(kap1, N, 16 * k1t + lr, k, true);
            }
            split8(v[0], a2re_h[s], a2re_l[s]);
            split8(v[1], a2im_h[s], a2im_l[s]);
            split8(v[2], b1re_h[s], b1re_l[s]);
            split8(v[3], b1im_h[s], b1im_l[s]);
        }
    }
    __device__ __forceinline__ void t1(const float* pr, const float* pi, float2* out) const {
        f32x4 are = {0.f, 0.f, 0.f, 0.f}, aim = {0.f, 0.f, 0.f, 0.f};
        const int row = (16 * k2t + lr) * kPad + 8 * lq;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const float* p = (s == 0 ? pr : pi) + row;
            float v[8];
#pragma unroll
            for (int t = 0; t < 8; t++) v[t] = p[t];
            bf16x8 h, l;
            split8(v, h, l);
            are = mfma_bf16(l, b1re_h[s], are);
            are = mfma_bf16(h, b1re_l[s], are);
            are = mfma_bf16(h, b1re_h[s], are);
            aim = mfma_bf16(l, b1im_h[s], aim);
            aim = mfma_bf16(h, b1im_l[s], aim);
            aim = mfma_bf16(h, b1im_h[s], aim);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<float>(are[v], aim[v]);
    }
    __device__ __forceinline__ void t2(const float* pr, const float* pi, float2* out) const {
        f32x4 are = {0.f, 0.f, 0.f, 0.f}, aim = {0.f, 0.f, 0.f, 0.f};
        const int col = 16 * k1t + lr;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const float* p = (s == 0 ? pr : pi) + 8 * lq * kPad + col;
            float v[8];
#pragma unroll
            for (int t = 0; t < 8; t++) v[t] = p[t * kPad];
            bf16x8 h, l;
            split8(v, h, l);
            are = mfma_bf16(a2re_l[s], h, are);
            are = mfma_bf16(a2re_h[s], l, are);
            are = mfma_bf16(a2re_h[s], h, are);
            aim = mfma_bf16(a2im_l[s], h, aim);
            aim = mfma_bf16(a2im_h[s], l, aim);
            aim = mfma_bf16(a2im_h[s], h, aim);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<float>(are[v], aim[v]);
    }
};

// workgroup -> logical index so that each XCD (workgroups are dealt round
// robin over 8 XCDs) processes one contiguous range of logical indices
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb) {
    const int64_t q = nb / 8, rem = nb % 8, x = b % 8, i = b / 8;
    return x * q + (x < rem ? x : rem) + i;
}

// ------------------------------------------------------- fused kernel ---

template <class T, class G>
__global__ __launch_bounds__(256) void k_drift_fused(DriftDev<T> D, MixDev<T> MX, const cT<T>* Fin,
                                                     const cT<T>* __restrict__ Z3, const cT<T>* F0, cT<T>* Fout,
                                                     cT<T>* ACC, StageCoef<T> sc) {
    __shared__ T pa_re[32 * kPad], pa_im[32 * kPad], pb_re[32 * kPad], pb_im[32 * kPad];
    __shared__ cT<T> kap[2][32];

    const int N = D.N, NN = N * N, dim = D.dim;
    const int nz = dim == 3 ? N : 1;
    const int64_t sl = xcd_remap(blockIdx.x, gridDim.x);  // = k3 * nrc + r
    const int64_t r = sl % D.nrc;
    const int k3 = (int)(sl / D.nrc);
    const int64_t base = (r * nz + k3) * NN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int k2t = w >> 1, k1t = w & 1, lr = lane & 15, lq = lane >> 4;

    for (int idx = tid; idx < 1024; idx += 256) {
        const int a2 = idx >> 5, a1 = idx & 31;
        const cT<T> v = (a2 < N && a1 < N) ? Fin[base + a2 * N + a1] : czero<T>();
        pa_re[a2 * kPad + a1] = v.x;
        pa_im[a2 * kPad + a1] = v.y;
    }
    if (tid < 64) {
        const int p = tid >> 5, j = tid & 31;
        kap[p][j] = j < N ? D.kap[p * N + j] : czero<T>();
    }
    __syncthreads();

    G gm;
    gm.init(kap[0], kap[1], N, k2t, k1t, lr, lq);

    cT<T> P1[4], P2[4];
    gm.t1(pa_re, pa_im, P1);
    gm.t2(pa_re, pa_im, P2);

    // rotation mixing at this lane's four elements
    const int k1 = 16 * k1t + lr;
    const bool k1ok = k1 < N;
    const T kc1 = k1ok ? D.kc[k1] : T(0);
    const T kd1 = k1ok ? D.kd[k1] : T(0);
    const T kc3 = dim == 3 ? D.kc[k3] : T(0);
    const T kd3 = dim == 3 ? D.kd[k3] : T(0);
    cT<T> Wv[4], V1[4], V2[4], Zv[4];
    cT<T> ag[4][3];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        V1[v] = czero<T>(); V2[v] = czero<T>(); Zv[v] = czero<T>();
        ag[v][0] = czero<T>(); ag[v][1] = czero<T>(); ag[v][2] = czero<T>();
    }
    const int NC = dim + 2;
    for (int t = 0; t < MX.TS; t++) {
        const int src = MX.src[r * MX.TS + t];
        if (src < 0) break;
        const cT<T>* cf = MX.coef + (r * MX.TS + t) * NC;
        const cT<T> cu1 = cf[dim], cu2 = cf[dim + 1];
        const cT<T> cg0 = cf[0], cg1 = cf[1], cg2 = dim == 3 ? cf[2] : czero<T>();
        const cT<T>* Fs = Fin + ((int64_t)src * nz + k3) * NN;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k2 = 16 * k2t + G::orow(lq, v);
            if (!(k1ok && k2 < N)) continue;
            const cT<T> fv = Fs[k2 * N + k1];
            cfma<T>(ag[v][0], cg0, fv);
            cfma<T>(ag[v][1], cg1, fv);
            if (dim == 3) cfma<T>(ag[v][2], cg2, fv);
            cfma<T>(V1[v], cu1, fv);
            cfma<T>(V2[v], cu2, fv);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int k2 = 16 * k2t + G::orow(lq, v);
        Wv[v] = czero<T>();
        if (!(k1ok && k2 < N)) continue;
        const T kc2 = D.kc[k2], kd2 = D.kd[k2];
        const T kc[3] = {kc1, kc2, kc3}, kd[3] = {kd1, kd2, kd3};
        T dd = T(0);
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) {
                const T g = D.G[i + dim * j];
                dd += (i == j) ? g * kd[i] : -g * kc[i] * kc[j];
            }
        const cT<T> fh = cmk<T>(pa_re[k2 * kPad + k1], pa_im[k2 * kPad + k1]);
        cT<T> wv = cscale<T>(fh, dd);
        wv = csub<T>(wv, cimul<T>(ag[v][0], kc1));
        wv = csub<T>(wv, cimul<T>(ag[v][1], kc2));
        for (int p = 0; p < dim; p++)  // gravity evaluated on the grid (density form)
            if (D.grav[p]) wv = csub<T>(wv, cimul<T>(D.grav[p][base + k2 * N + k1], kc[p]));
        if (dim == 3) {
            wv = csub<T>(wv, cimul<T>(ag[v][2], kc3));
            Zv[v] = Z3[base + k2 * N + k1];
            if (D.has_gyro) {
                V1[v] = csub<T>(V1[v], cimul<T>(Zv[v], kc2 * D.gyro[1]));
                V1[v] = csub<T>(V1[v], cimul<T>(P2[v], kc3 * D.gyro[2]));
                V2[v] = csub<T>(V2[v], cimul<T>(Zv[v], kc1 * D.gyro[0]));
            }
        }
        Wv[v] = wv;
    }
    __syncthreads();  // all reads of the F planes are done
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int k2 = 16 * k2t + G::orow(lq, v);
        pa_re[k2 * kPad + k1] = V1[v].x;
        pa_im[k2 * kPad + k1] = V1[v].y;
        pb_re[k2 * kPad + k1] = V2[v].x;
        pb_im[k2 * kPad + k1] = V2[v].y;
    }
    __syncthreads();
    cT<T> Q1[4], Q2[4];
    gm.t1(pa_re, pa_im, Q1);
    gm.t2(pb_re, pb_im, Q2);

    const cT<T> e3 = dim == 3 ? D.e3[r] : czero<T>();
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int k2 = 16 * k2t + G::orow(lq, v);
        if (!(k1ok && k2 < N)) continue;
        cT<T> q = cadd<T>(Q1[v], Q2[v]);
        q = cadd<T>(q, Wv[v]);
        q = cadd<T>(q, cimul<T>(P1[v], D.b[0] * kc1));
        q = cadd<T>(q, cimul<T>(P2[v], D.b[1] * D.kc[k2]));
        if (dim == 3) {
            q = cadd<T>(q, cmul<T>(Zv[v], e3));
            q = cadd<T>(q, cimul<T>(Zv[v], D.b[2] * kc3));
        }
        const int64_t o = base + k2 * N + k1;
        if (sc.deriv_only) {
            Fout[o] = q;
            continue;
        }
        const cT<T> f0 = sc.read_f0 ? F0[o] : Fin[o];
        const cT<T> ac = sc.read_acc ? ACC[o] : czero<T>();
        const cT<T> upd = cadd<T>(cscale<T>(ac, sc.ca), cscale<T>(q, sc.cd));
        Fout[o] = cadd<T>(f0, cscale<T>(upd, sc.dt));
        if (sc.write_acc) ACC[o] = cadd<T>(cscale<T>(ac, sc.ea), cscale<T>(q, sc.ed));
    }
}

// ------------------------------------------------------------ axis 3 ----

template <class T>
__global__ __launch_bounds__(256) void k_axis3_mfma(DriftDev<T> D, const cT<T>* __restrict__ in,
                                                    cT<T>* __restrict__ out) {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc;
    constexpr int CP = 65;
    __shared__ T sre[32 * CP], sim[32 * CP];
    __shared__ cT<T> kap[32];
    const int N = D.N;
    const int64_t NN = (int64_t)N * N;
    const int64_t r = blockIdx.y;
    const int64_t j0 = (int64_t)blockIdx.x * 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lq = lane >> 4;
    for (int idx = tid; idx < 32 * 64; idx += 256) {
        const int a = idx >> 6, jj = idx & 63;
        const int64_t j = j0 + jj;
        const cT<T> v = (a < N && j < NN) ? in[(r * N + a) * NN + j] : czero<T>();
        sre[a * CP + jj] = v.x;
        sim[a * CP + jj] = v.y;
    }
    if (tid < 32) kap[tid] = tid < N ? D.kap[2 * N + tid] : czero<T>();
    __syncthreads();
    acc_t acc[4];
#pragma unroll
    for (int it = 0; it < 4; it++) acc[it] = MF::zero();
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const int k = 4 * s + lq;
        const T bv = (s < 8 ? sre : sim)[(k & 31) * CP + 16 * w + lr];
#pragma unroll
        for (int it = 0; it < 4; it++)
            acc[it] = MF::fma(circ_entry<T>(kap, N, 16 * (it & 1) + lr, k, (it >> 1) != 0), bv, acc[it]);
    }
    const int64_t j = j0 + 16 * w + lr;
    if (j >= NN) return;
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k3 = 16 * h + MF::orow(lq, v);
            if (k3 < N) out[(r * N + k3) * NN + j] = cmk<T>(acc[h][v], acc[h + 2][v]);
        }
}

__global__ __launch_bounds__(256) void k_axis3_bf16x3(DriftDev<float> D, const float2* __restrict__ in,
                                                      float2* __restrict__ out) {
    constexpr int CP = 65;
    __shared__ float sre[32 * CP], sim[32 * CP];
    __shared__ float2 kap[32];
    const int N = D.N;
    const int64_t NN = (int64_t)N * N;
    const int64_t r = blockIdx.y;
    const int64_t j0 = (int64_t)blockIdx.x * 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lq = lane >> 4;
    for (int idx = tid; idx < 32 * 64; idx += 256) {
        const int a = idx >> 6, jj = idx & 63;
        const int64_t j = j0 + jj;
        const float2 v = (a < N && j < NN) ? in[(r * N + a) * NN + j] : czero<float>();
        sre[a * CP + jj] = v.x;
        sim[a * CP + jj] = v.y;
    }
    if (tid < 32) kap[tid] = tid < N ? D.kap[2 * N + tid] : czero<float>();
    __syncthreads();
    f32x4 acc[4];
#pragma unroll
    for (int it = 0; it < 4; it++) acc[it] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 2; s++) {
        float bv[8];
        const float* p = (s == 0 ? sre : sim) + 8 * lq * CP + 16 * w + lr;
#pragma unroll
        for (int t = 0; t < 8; t++) bv[t] = p[t * CP];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
#pragma unroll
        for (int it = 0; it < 4; it++) {
            float av[8];
#pragma unroll
            for (int t = 0; t < 8; t++) av[t] = circ_entry<float>(kap, N, 16 * (it & 1) + lr, 32 * s + 8 * lq + t,
                                                                  (it >> 1) != 0);
            bf16x8 ah, al;
            split8(av, ah, al);
            acc[it] = mfma_bf16(al, bh, acc[it]);
            acc[it] = mfma_bf16(ah, bl, acc[it]);
            acc[it] = mfma_bf16(ah, bh, acc[it]);
        }
    }
    const int64_t j = j0 + 16 * w + lr;
    if (j >= NN) return;
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k3 = 16 * h + 4 * lq + v;
            if (k3 < N) out[(r * N + k3) * NN + j] = cmk<float>(acc[h][v], acc[h + 2][v]);
        }
}

// ---------------------------------------------------------- launchers ---

template <class T>
void launch_drift_fused(const DriftDev<T>& D, const MixDev<T>& MX, const void* Fin, const void* Z3, const void* F0,
                        void* Fout, void* ACC, const StageCoef<T>& sc, bool bf16x3, hipStream_t st) {
    HPUP_REQUIRE(D.N <= 32, HPUP_ERR_UNSUPPORTED, "fused MFMA drift needs N <= 32");
    const int64_t nslab = D.nrc * (D.dim == 3 ? D.N : 1);
    if constexpr (sizeof(T) == 4) {
        if (bf16x3) {
            hipLaunchKernelGGL((k_drift_fused<float, GemmBf16x3>), dim3((unsigned)nslab), dim3(256), 0, st, D, MX,
                               (const cT<T>*)Fin, (const cT<T>*)Z3, (const cT<T>*)F0, (cT<T>*)Fout, (cT<T>*)ACC, sc);
            return;
        }
    }
    HPUP_REQUIRE(!bf16x3, HPUP_ERR_UNSUPPORTED, "3xBF16 needs FP32 storage");
    hipLaunchKernelGGL((k_drift_fused<T, GemmF<T>>), dim3((unsigned)nslab), dim3(256), 0, st, D, MX,
                       (const cT<T>*)Fin, (const cT<T>*)Z3, (const cT<T>*)F0, (cT<T>*)Fout, (cT<T>*)ACC, sc);
}

template <class T>
void launch_axis3_mfma(const DriftDev<T>& D, const void* in, void* out, bool bf16x3, hipStream_t st) {
    HPUP_REQUIRE(D.N <= 32 && D.dim == 3, HPUP_ERR_UNSUPPORTED, "axis-3 MFMA kernel needs dim 3, N <= 32");
    const int64_t NN = (int64_t)D.N * D.N;
    dim3 grid((unsigned)((NN + 63) / 64), (unsigned)D.nrc);
    if constexpr (sizeof(T) == 4) {
        if (bf16x3) {
            hipLaunchKernelGGL(k_axis3_bf16x3, grid, dim3(256), 0, st, D, (const float2*)in, (float2*)out);
            return;
        }
    }
    HPUP_REQUIRE(!bf16x3, HPUP_ERR_UNSUPPORTED, "3xBF16 needs FP32 storage");
    hipLaunchKernelGGL(k_axis3_mfma<T>, grid, dim3(256), 0, st, D, (const cT<T>*)in, (cT<T>*)out);
}

template void launch_drift_fused<float>(const DriftDev<float>&, const MixDev<float>&, const void*, const void*,
                                        const void*, void*, void*, const StageCoef<float>&, bool, hipStream_t);
template void launch_drift_fused<double>(const DriftDev<double>&, const MixDev<double>&, const void*, const void*,
                                         const void*, void*, void*, const StageCoef<double>&, bool, hipStream_t);
template void launch_axis3_mfma<float>(const DriftDev<float>&, const void*, void*, bool, hipStream_t);
template void launch_axis3_mfma<double>(const DriftDev<double>&, const void*, void*, bool, hipStream_t);

}  // namespace hpup
