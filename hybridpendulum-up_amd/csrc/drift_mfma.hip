// MFMA drift kernels for omega grids up to 32 points per axis.
//
// k_drift_fused evaluates the whole derivative of each (r, k3) slab it visits
// and applies the integrator epilogue (see drift.hip for the operator split):
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
// instruction) or GemmBf16x3 (v_mfma_f32_16x16x32_bf16 on three-piece bf16
// splits of both operands, the six products of order <= 2 accumulated in FP32).
//
// k_axis3 computes Z3 = T3 F (circulant along the outermost omega axis) as the
// GEMM C3 [F_re; F_im] over 64-column chunks of the k1,k2 plane.
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

#include "drift.h"
#include "kernels_common.h"
#include "mfma.h"

namespace hpup {

// Partial sums of the gradient stage for callers that cannot allocate:
// kGradMaxWG x 5 doubles (160 KiB), held by every device the library runs on
// until the process ends.  Host code never uses this name as an address: the
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[kGradMaxWG * 5];

double* grad_scratch(int device) {
    static std::mutex mu;
    static double* ptr[kMaxDevices] = {};
    HPUP_REQUIRE(device >= 0 && device < kMaxDevices, HPUP_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[device]) {
        int cur = -1;
        HPUP_HIP(hipGetDevice(&cur));
        HPUP_REQUIRE(cur == device, HPUP_ERR_ARG, "grad_scratch: the device must be current");
        void* a = nullptr;
        HPUP_HIP(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
        HPUP_REQUIRE(a != nullptr, HPUP_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        ptr[device] = static_cast<double*>(a);
    }
    return ptr[device];
}

// LDS row stride of the 32x32 planes.  With 34 (= 2 mod 32) the row-wise
// operand reads of T1 (lane (lr, lq) reads row lr, column 4s + lq) and the
// column-wise reads of T2 (lane reads row 8 lq + t, column lr) each hit 32
// distinct banks per half-wave.
constexpr int kPad = 34;

// ------------------------------------------------------------ policies ---

// Circulant fragments from LDS.  Entry (o, a) of a circulant with kernel
// kappa is kappa[(o - a) mod N]; the table holds, for d = o - a + TILE in
// [0, 2 TILE), the pair (re, im) of kappa[(o - a) mod N] and the pair (-im, re)
// used for the imaginary input half.  Rows o >= N or columns a >= N read
// finite values that meet zero-padded data or are discarded, so the table
// needs no padding logic.  CircTab<T, TILE>::fill builds it for one axis
// (threads 0 .. 2 TILE - 1 of the workgroup).
template <class T, int TILE = 32> struct CircTab {
    cT<T> re_in[2 * TILE], im_in[2 * TILE];
    cT<T> rim[2 * TILE];  // (-Im kappa, Im kappa): the real circulant of the ramp path
    __device__ __forceinline__ void fill(const cT<T>* kap, int N, int tid) {
        if (tid < 2 * TILE) {
            const int d = tid - TILE;
            const cT<T> c = kap[((d % N) + N) % N];
            re_in[tid] = c;
            im_in[tid] = cmk<T>(-c.y, c.x);
            rim[tid] = cmk<T>(-c.y, c.y);
        }
    }
};

// Sum over the four lanes l, l ^ 16, l ^ 32, l ^ 48 (the lq groups).
template <class T> __device__ __forceinline__ T sum_lq(T x) {
    x += __shfl_xor(x, 16, 64);
    x += __shfl_xor(x, 32, 64);
    return x;
}

// FP32 / FP64 MFMA; fragments read from LDS, planes of TILE rows at stride
// TP = TILE + 2.  General circulants: TILE / 2 reduction steps of K = 4 over
// the (re | im) stacked slab.  RAMP (kappa = c0 + i (real circulant)): TILE / 4
// steps of two MFMAs (-Im C . X_im, Im C . X_re) plus c0 times the axis sums
// of X, which the operand loads accumulate on the way (the four lq lanes of a
// row or column hold a quarter each).
template <class T, bool RAMP = false, int TILE = 32> struct GemmF {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc;
    static constexpr int TP = TILE + 2;  // plane row stride
    static constexpr int KS = TILE / 4;  // K = 4 steps per re / im half
    const CircTab<T, TILE>* c1;
    const CircTab<T, TILE>* c2;
    int k2t, k1t, lr, lq, b1, b2;
    // 32-point tiles: K = 4 steps per re/im half that reach the data, ks =
    // ceil(n / 4) for n data points per axis (the rest of the tile is zero),
    // and T2's lane stride in its reduction order
    int ks;
    static __device__ __forceinline__ int orow(int q, int v) { return MF::orow(q, v); }
    // T2's reduction index of step u (< KS) in each re/im half (64-point tiles)
    static __device__ __forceinline__ int k2of(int u) { return (u & 7) + 32 * (u >> 3); }
    __device__ __forceinline__ void init(const CircTab<T, TILE>* tab1, const CircTab<T, TILE>* tab2, int k2t_,
                                         int k1t_, int lr_, int lq_, int n = TILE) {
        c1 = tab1; c2 = tab2;
        k2t = k2t_; k1t = k1t_; lr = lr_; lq = lq_;
        ks = TILE == 32 ? (n + 3) / 4 : KS;
        // T1 reduces over k = 4 u + lq; T2 over k = ks lq + u (32-point tiles)
        // or 8 lq + k2of(u) (64-point tiles), in each re/im half.  The order
        // of a GEMM's reduction is free; with TP = 2 mod 32 these orders keep
        // the data reads conflict-free at ks = 8 (a few 2-way conflicts for
        // smaller ks).  d = o - k + TILE.
        b1 = 16 * k1t + lr + TILE - lq;
        b2 = 16 * k2t + lr + TILE - (TILE == 32 ? ks : 8) * lq;
    }
    // T2's reduction index (relative to its lane's start) of step u
    __device__ __forceinline__ int k2u(int u) const { return TILE == 32 ? u : k2of(u); }
    // Y = X C1^T along k1 (slab rows as the A operand); c0: Re kappa_1 (RAMP)
    __device__ __forceinline__ void t1(const T* pr, const T* pi, cT<T>* out, T c0 = T(0)) const {
        acc_t are = MF::zero(), aim = MF::zero();
        const int row = (16 * k2t + lr) * TP;
        if constexpr (RAMP) {
            T sr = T(0), si = T(0);
#pragma unroll 1
            for (int s = 0; s < ks; s++) {
                const int col = 4 * s + lq;
                const T xr = pr[row + col], xi = pi[row + col];
                const cT<T> f = c1->rim[b1 - 4 * s];
                are = MF::fma(xi, f.x, are);
                aim = MF::fma(xr, f.y, aim);
                sr += xr;
                si += xi;
            }
            // row 16 k2t + lr summed over its TILE columns; output row
            // orow(lq, v) is held by the lanes with lr = orow(lq, v)
            sr = sum_lq<T>(sr);
            si = sum_lq<T>(si);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int src = orow(lq, v);
                out[v] = cmk<T>(are[v] + c0 * __shfl(sr, src, 64), aim[v] + c0 * __shfl(si, src, 64));
            }
            return;
        }
#pragma unroll 4
        for (int u = 0; u < ks; u++) {  // real half
            const T av = pr[row + 4 * u + lq];
            const cT<T> f = c1->re_in[b1 - 4 * u];
            are = MF::fma(av, f.x, are);
            aim = MF::fma(av, f.y, aim);
        }
#pragma unroll 4
        for (int u = 0; u < ks; u++) {  // imaginary half
            const T av = pi[row + 4 * u + lq];
            const cT<T> f = c1->im_in[b1 - 4 * u];
            are = MF::fma(av, f.x, are);
            aim = MF::fma(av, f.y, aim);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<T>(are[v], aim[v]);
    }
    // Y = C2 X along k2 (slab columns as the B operand); c0: Re kappa_2 (RAMP)
    __device__ __forceinline__ void t2(const T* pr, const T* pi, cT<T>* out, T c0 = T(0)) const {
        acc_t are = MF::zero(), aim = MF::zero();
        const int col = 16 * k1t + lr;
        if constexpr (RAMP) {
            T sr = T(0), si = T(0);
            const int r0 = (TILE == 32 ? ks : 8) * lq;
#pragma unroll 1
            for (int s = 0; s < ks; s++) {
                const int rw = (r0 + k2u(s)) * TP + col;
                const T xr = pr[rw], xi = pi[rw];
                const cT<T> f = c2->rim[b2 - k2u(s)];
                are = MF::fma(f.x, xi, are);
                aim = MF::fma(f.y, xr, aim);
                sr += xr;
                si += xi;
            }
            // column 16 k1t + lr summed over its TILE rows: this lane's own column
            sr = sum_lq<T>(sr);
            si = sum_lq<T>(si);
#pragma unroll
            for (int v = 0; v < 4; v++) out[v] = cmk<T>(are[v] + c0 * sr, aim[v] + c0 * si);
            return;
        }
        const int r0 = (TILE == 32 ? ks : 8) * lq;
#pragma unroll 4
        for (int u = 0; u < ks; u++) {  // real half
            const T bv = pr[(r0 + k2u(u)) * TP + col];
            const cT<T> f = c2->re_in[b2 - k2u(u)];
            are = MF::fma(f.x, bv, are);
            aim = MF::fma(f.y, bv, aim);
        }
#pragma unroll 4
        for (int u = 0; u < ks; u++) {  // imaginary half
            const T bv = pi[(r0 + k2u(u)) * TP + col];
            const cT<T> f = c2->im_in[b2 - k2u(u)];
            are = MF::fma(f.x, bv, are);
            aim = MF::fma(f.y, bv, aim);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<T>(are[v], aim[v]);
    }
    // The same products with three real GEMMs instead of four (Gauss):
    // P1 = Re C Re X, P2 = Im C Im X, P3 = (Re C + Im C)(Re X + Im X), and
    // C X = P1 - P2 + i (P3 - P1 - P2).  The tables' rim[d].x holds Re C + Im C.
    __device__ __forceinline__ void t1_gauss(const T* pr, const T* pi, cT<T>* out) const {
        acc_t p1 = MF::zero(), p2 = MF::zero(), p3 = MF::zero();
        const int row = (16 * k2t + lr) * TP;
#pragma unroll 4
        for (int u = 0; u < ks; u++) {
            const T xr = pr[row + 4 * u + lq], xi = pi[row + 4 * u + lq];
            const cT<T> f = c1->re_in[b1 - 4 * u];
            const T sc = c1->rim[b1 - 4 * u].x;
            p1 = MF::fma(xr, f.x, p1);
            p2 = MF::fma(xi, f.y, p2);
            p3 = MF::fma(xr + xi, sc, p3);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<T>(p1[v] - p2[v], p3[v] - p1[v] - p2[v]);
    }
    __device__ __forceinline__ void t2_gauss(const T* pr, const T* pi, cT<T>* out) const {
        acc_t p1 = MF::zero(), p2 = MF::zero(), p3 = MF::zero();
        const int col = 16 * k1t + lr;
        const int r0 = (TILE == 32 ? ks : 8) * lq;
#pragma unroll 4
        for (int u = 0; u < ks; u++) {
            const int rw = (r0 + k2u(u)) * TP + col;
            const T xr = pr[rw], xi = pi[rw];
            const cT<T> f = c2->re_in[b2 - k2u(u)];
            const T sc = c2->rim[b2 - k2u(u)].x;
            p1 = MF::fma(f.x, xr, p1);
            p2 = MF::fma(f.y, xi, p2);
            p3 = MF::fma(sc, xr + xi, p3);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<T>(p1[v] - p2[v], p3[v] - p1[v] - p2[v]);
    }
};

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// x = hi + mid + lo + O(2^-26 |x|), each piece bf16 (round to nearest even):
// every piece holds the leading 8 bits of what the previous ones left
__device__ __forceinline__ void split8(const float* v, bf16x8& hi, bf16x8& mid, bf16x8& lo) {
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const __bf16 h = (__bf16)v[t];
        const float r1 = v[t] - (float)h;
        const __bf16 m = (__bf16)r1;
        hi[t] = h;
        mid[t] = m;
        lo[t] = (__bf16)(r1 - (float)m);
    }
}

// sum of the six split products a_i b_j with i + j <= 2 (hi = 0, mid = 1,
// lo = 2), smallest first: every term of order 2^-16 or larger relative to
// a b, so the product is good to about 2^-24, the FP32 rounding itself
__device__ __forceinline__ f32x4 mfma_split6(bf16x8 ah, bf16x8 am, bf16x8 al, bf16x8 bh, bf16x8 bm, bf16x8 bl,
                                             f32x4 c) {
    c = mfma_bf16(al, bh, c);
    c = mfma_bf16(am, bm, c);
    c = mfma_bf16(ah, bl, c);
    c = mfma_bf16(am, bh, c);
    c = mfma_bf16(ah, bm, c);
    return mfma_bf16(ah, bh, c);
}

// 3xBF16: each FP32 operand split into three bf16 pieces, the six products
// of total order <= 2 (mfma_split6) accumulated in FP32; two reduction steps
// of K = 32 (re half, im half); the 16x16x32 bf16 instruction takes
// A[l&15][8(l>>4) + t] and B[8(l>>4) + t][l&15], t < 8.
struct GemmBf16x3 {
    bf16x8 a2re_h[2], a2re_m[2], a2re_l[2], a2im_h[2], a2im_m[2], a2im_l[2];
    bf16x8 b1re_h[2], b1re_m[2], b1re_l[2], b1im_h[2], b1im_m[2], b1im_l[2];
    int k2t, k1t, lr, lq;
    static __device__ __forceinline__ int orow(int q, int v) { return 4 * q + v; }
    __device__ __forceinline__ void init(const CircTab<float>* tab1, const CircTab<float>* tab2, int k2t_, int k1t_,
                                         int lr_, int lq_, int = 32) {
        k2t = k2t_; k1t = k1t_; lr = lr_; lq = lq_;
        // entry (o, k) of the real-ified circulant, as GemmF reads it
        auto ent = [](const CircTab<float>* t, int o, int k, bool im_out) {
            const float2 c = (k < 32 ? t->re_in : t->im_in)[o - (k & 31) + 32];
            return im_out ? c.y : c.x;
        };
#pragma unroll
        for (int s = 0; s < 2; s++) {
            float v[4][8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int k = 32 * s + 8 * lq + t;
                v[0][t] = ent(tab2, 16 * k2t + lr, k, false);
                v[1][t] = ent(tab2, 16 * k2t + lr, k, true);
                v[2][t] = ent(tab1, 16 * k1t + lr, k, false);
                v[3][t] = ent(tab1, 16 * k1t + lr, k, true);
            }
            split8(v[0], a2re_h[s], a2re_m[s], a2re_l[s]);
            split8(v[1], a2im_h[s], a2im_m[s], a2im_l[s]);
            split8(v[2], b1re_h[s], b1re_m[s], b1re_l[s]);
            split8(v[3], b1im_h[s], b1im_m[s], b1im_l[s]);
        }
    }
    __device__ __forceinline__ void t1(const float* pr, const float* pi, float2* out, float = 0.f) const {
        f32x4 are = {0.f, 0.f, 0.f, 0.f}, aim = {0.f, 0.f, 0.f, 0.f};
        const int row = (16 * k2t + lr) * kPad + 8 * lq;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const float* p = (s == 0 ? pr : pi) + row;
            float v[8];
#pragma unroll
            for (int t = 0; t < 8; t++) v[t] = p[t];
            bf16x8 h, m, l;
            split8(v, h, m, l);
            are = mfma_split6(h, m, l, b1re_h[s], b1re_m[s], b1re_l[s], are);
            aim = mfma_split6(h, m, l, b1im_h[s], b1im_m[s], b1im_l[s], aim);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) out[v] = cmk<float>(are[v], aim[v]);
    }
    __device__ __forceinline__ void t2(const float* pr, const float* pi, float2* out, float = 0.f) const {
        f32x4 are = {0.f, 0.f, 0.f, 0.f}, aim = {0.f, 0.f, 0.f, 0.f};
        const int col = 16 * k1t + lr;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const float* p = (s == 0 ? pr : pi) + 8 * lq * kPad + col;
            float v[8];
#pragma unroll
            for (int t = 0; t < 8; t++) v[t] = p[t * kPad];
            bf16x8 h, m, l;
            split8(v, h, m, l);
            are = mfma_split6(a2re_h[s], a2re_m[s], a2re_l[s], h, m, l, are);
            aim = mfma_split6(a2im_h[s], a2im_m[s], a2im_l[s], h, m, l, aim);
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

// Data the kernels touch once (T3 F, F0, ACC, the output; F and T3 F in the
// axis-3 kernel).  FP64 streams them non-temporally, which keeps the re-read
// neighbour rows in L2; for FP32 the plain forms measured faster.
template <class T> __device__ __forceinline__ cT<T> ld_once(const cT<T>* p) {
    if constexpr (sizeof(T) == 8) return ld_stream<T>(p);
    else return *p;
}
template <class T> __device__ __forceinline__ void st_once(cT<T>* p, cT<T> v) {
    if constexpr (sizeof(T) == 8) st_stream<T>(p, v);
    else *p = v;
}

// ------------------------------------------------------- fused kernel ---
//
// Persistent: each workgroup builds its circulant tables once and walks a
// strided sequence of slabs inside its XCD's share (see the slab order below).  Three LDS plane pairs (slab, V1, V2) need
// two barriers per slab: writes to the slab planes follow the previous slab's
// second barrier, writes to the V planes follow this slab's first barrier.

template <class T> struct MixBatch;  // source rows whose loads are issued together
template <> struct MixBatch<float> { static constexpr int n = 3; };
template <> struct MixBatch<double> { static constexpr int n = 2; };

// p + ob bytes: a 32-bit unsigned offset from a workgroup-uniform base, so
// the access takes the scalar-base + 32-bit vector-offset address form
// instead of a 64-bit address per lane
template <class P> __device__ __forceinline__ P* at_b(P* p, unsigned ob) {
    using C = typename std::conditional<std::is_const<P>::value, const char, char>::type;
    return reinterpret_cast<P*>(reinterpret_cast<C*>(p) + ob);
}

// Tables read at workgroup-uniform indices (mixing sources and coefficients,
// per-row and per-k3 constants) are read through the constant address space:
// nothing writes them during a launch, so they become scalar loads, which do
// not wait behind the vector loads of the slab data.
template <class T> using kc_ptr = __attribute__((address_space(4))) const T*;
template <class T> __device__ __forceinline__ kc_ptr<T> as_const(const T* p) { return (kc_ptr<T>)p; }
template <class T> __device__ __forceinline__ cT<T> kc_cplx(kc_ptr<T> p, int64_t i) {
    return cmk<T>(p[2 * i], p[2 * i + 1]);
}

// One mixing source with structured coefficients: W += (a + i b_v) F with
// a = Re(-i CG_1) kc1 and b_v = Im(-i CG_2) kc2 (CG_3 = 0), V1 += i Im(u_1) F,
// V2 += Re(u_2) F -- nine FP operations per element instead of sixteen.
template <class T, int DIM>
__device__ __forceinline__ void mix_structured(kc_ptr<T> coefT, int64_t cb, int fl, T kc1, const T* kc2v,
                                               const cT<T>* f, cT<T>* Wv, cT<T>* V1, cT<T>* V2) {
    if (fl & 1) {
        const T a = coefT[2 * cb] * kc1, g1 = coefT[2 * (cb + 1) + 1];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const T b = g1 * kc2v[v];
            Wv[v].x = fma(a, f[v].x, fma(-b, f[v].y, Wv[v].x));
            Wv[v].y = fma(a, f[v].y, fma(b, f[v].x, Wv[v].y));
        }
    }
    if (fl & 2) {
        const T ui = coefT[2 * (cb + DIM) + 1];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            V1[v].x = fma(-ui, f[v].y, V1[v].x);
            V1[v].y = fma(ui, f[v].x, V1[v].y);
        }
    }
    if (fl & 4) {
        const T ur = coefT[2 * (cb + DIM + 1)];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            V2[v].x = fma(ur, f[v].x, V2[v].x);
            V2[v].y = fma(ur, f[v].y, V2[v].y);
        }
    }
}

// Waves per SIMD the fused kernel is compiled for (the register budget):
// 4 for FP32 (at most 128 VGPRs), 2 for FP64 and 3xBF16.
template <class T, class G> struct FusedWaves { static constexpr int n = 2; };
template <bool R> struct FusedWaves<float, GemmF<float, R>> { static constexpr int n = 4; };

// GRAV: the gravity term comes from the grid (density form, D.grav) instead
// of the Clebsch-Gordan table; a separate instantiation so that the common
// case carries none of its addressing.
// ADJ (reduced model): the adjoint operator A^H.  The host hands over
// transposed mixing tables and the conjugate-reversed circulants; the damping
// term i b_p c_p T_p F becomes T_p^H(-i b_p c_p F), which joins the V planes,
// so the first GEMM pair is not needed.
// GRAD (with ADJ): one reverse-mode stage of the gradient sweep.  Besides
// A^H ybar it reduces, against the forward stage value y (D.gy), the
// parameter terms Re<ybar, (dA/dtheta) y>: the damping ones as
// Re<T_p^H(-i K_p ybar), y>, formed in the GEMM pair the adjoint leaves free
// (planes A and D hold -i K_1 ybar and -i K_2 ybar), the diffusion ones
// elementwise; each workgroup writes its five sums to D.gpart.
// SMIX: structured mixing coefficients (DriftDev::smix), each multiplied by
// its one nonzero component (the adjoint's transposed tables keep the
// classes: conjugation leaves a real coefficient real, an imaginary one
// imaginary).
template <class T, class G, int DIM, bool GRAV, bool ADJ = false, bool GRAD = false, bool SMIX = false>
__global__ __launch_bounds__(256, (GRAD ? (sizeof(T) == 4 ? 3 : 2) : FusedWaves<T, G>::n)) void k_drift_fused(DriftDev<T> D, MixDev<T> MX, const cT<T>* Fin,
                                                     const cT<T>* __restrict__ Z3, const cT<T>* F0, cT<T>* Fout,
                                                     cT<T>* ACC, StageCoef<T> sc, int64_t nslab) {
    __shared__ T pa_re[32 * kPad], pa_im[32 * kPad], pb_re[32 * kPad], pb_im[32 * kPad], pc_re[32 * kPad],
        pc_im[32 * kPad];
    __shared__ CircTab<T> tabs[2];
    __shared__ T pd_re[GRAD ? 32 * kPad : 1], pd_im[GRAD ? 32 * kPad : 1];
    __shared__ double gred[GRAD ? 4 * 5 : 1];
    double ga[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // GRAD: b_1, b_2, G_00, G_11, G_01 (= G_10) terms
    constexpr int NC = DIM + 2;
    constexpr int MB = MixBatch<T>::n;

    const int N = D.N, NN = N * N;
    const int nz = DIM == 3 ? N : 1;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int k2t = w >> 1, k1t = w & 1, lr = lane & 15, lq = lane >> 4;
    tabs[0].fill(D.kap, N, tid);
    tabs[1].fill(D.kap + N, N, tid);
    __syncthreads();
    G gm;
    gm.init(&tabs[0], &tabs[1], k2t, k1t, lr, lq, N);

    // this lane's four elements (k2, k1) of every slab: k1 = 16 k1t + lr and
    // k2 = 16 k2t + orow(lq, v); their omega factors stay in registers, their
    // indices are recomputed per slab
    const int k1 = 16 * k1t + lr;
    const bool k1ok = k1 < N;
    const T kc1 = k1ok ? D.kc[k1] : T(0), kd1 = k1ok ? D.kd[k1] : T(0);
    T kc2v[4], kd2v[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int k2 = 16 * k2t + G::orow(lq, v);
        const bool ok = k1ok && k2 < N;
        kc2v[v] = ok ? D.kc[k2] : T(0);
        kd2v[v] = ok ? D.kd[k2] : T(0);
    }
    // D(k) = sum_i G_ii kd(k_i) - sum_{i != j} G_ij kc(k_i) kc(k_j)
    const T g00 = D.G[0], g11 = D.G[DIM + 1], g01 = D.G[1] + D.G[DIM];
    const T g22 = DIM == 3 ? D.G[8] : T(0), g02 = DIM == 3 ? D.G[2] + D.G[6] : T(0),
            g12 = DIM == 3 ? D.G[5] + D.G[7] : T(0);

    // Slab order.  Workgroups are dealt round robin over the 8 XCDs, each with
    // its own L2: XCD x walks its own contiguous eighth of the slabs (sl =
    // k3 * nrc + r), its workgroups side by side, so the neighbour rows (within
    // 2B + 1 of r) that the mixing reads were loaded into that XCD's L2 moments
    // earlier.  The launch uses a multiple of 8 workgroups (or fewer than 8).
    // slab counts fit 32 bits (the launcher checks): 32-bit bookkeeping keeps
    // the loop state in few scalar registers
    const int nrc = (int)D.nrc;
    const int nxcd = gridDim.x >= 8 ? 8 : 1;
    const int per = (int)gridDim.x / nxcd;  // workgroups per XCD
    const int xcd = (int)blockIdx.x % nxcd;
    const int hi = (int)(nslab * (xcd + 1) / nxcd);
    int sl = (int)(nslab * xcd / nxcd) + (int)blockIdx.x / nxcd;
    const int dr = per % nrc;
    const int dk = per / nrc;
    int r = sl % nrc;
    int k3 = sl / nrc;
    while (sl < hi) {
        // zo is 0, produced by an opaque move each iteration: indices derived
        // from it are recomputed per slab instead of being hoisted out of the
        // loop (one register each for every address), which keeps the kernel
        // within the register budget of its occupancy
        int zo;
        asm volatile("v_mov_b32 %0, 0" : "=v"(zo));
        const int tidz = tid + zo;
        // the lane's elements: out-of-range ones (ok false, in the 32 x 32
        // padding) use a clamped in-range offset and are discarded
        const int lanez = tidz & 63, wz = tidz >> 6;
        const int k1w = 16 * (wz & 1) + (lanez & 15), lqz = lanez >> 4;
        int k2w[4];
        unsigned offb[4];  // byte offsets of the elements in a slab
        bool ok[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            k2w[v] = 16 * (wz >> 1) + G::orow(lqz, v);
            ok[v] = k1w < N && k2w[v] < N;
            offb[v] = (unsigned)((k2w[v] < N ? k2w[v] : N - 1) * N + (k1w < N ? k1w : N - 1)) *
                      (unsigned)sizeof(cT<T>);
        }
        const int64_t base = ((int64_t)r * nz + k3) * NN;
        const cT<T>* Fb = Fin + base;  // this slab
        const T kc3 = DIM == 3 ? as_const<T>(D.kc)[k3] : T(0), kd3 = DIM == 3 ? as_const<T>(D.kd)[k3] : T(0);

        // rotation mixing: sources in batches of MB rows, structurally zero
        // terms skipped by the table's flags (the branches are uniform over
        // the workgroup).  The Clebsch-Gordan terms enter W as
        // (sum_p kc_p (-i CG_p)) F_src with the kc of this lane's element (the
        // table holds -i CG_p), so one accumulator holds all of them.
        const int TS = MX.TS;
        const kc_ptr<int32_t> srcT = as_const<int32_t>(MX.src) + (int64_t)r * TS;
        const kc_ptr<T> coefT = as_const<T>((const T*)MX.coef) + 2 * (int64_t)r * TS * NC;
        auto load_batch = [&](int t0, int* src, cT<T> (*fv)[4]) {
#pragma unroll
            for (int u = 0; u < MB; u++) src[u] = t0 + u < TS ? srcT[t0 + u] : -1;
#pragma unroll
            for (int u = 0; u < MB; u++) {
                const cT<T>* Fs = Fin + ((int64_t)(src[u] < 0 ? 0 : (src[u] & 0xffffff)) * nz + k3) * NN;
#pragma unroll
                for (int v = 0; v < 4; v++) fv[u][v] = *at_b(Fs, offb[v]);
            }
        };
        cT<T> V1[4], V2[4], Wv[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            V1[v] = czero<T>();
            V2[v] = czero<T>();
            Wv[v] = czero<T>();
        }
        auto mix_batch = [&](int t0, const int* src, cT<T> (*fv)[4]) {
#pragma unroll
            for (int u = 0; u < MB; u++) {
                if (src[u] < 0) continue;
                const int fl = src[u] >> 24;
                const int64_t cb = (int64_t)(t0 + u) * NC;
                if constexpr (SMIX) {
                    mix_structured<T, DIM>(coefT, cb, fl, kc1, kc2v, fv[u], Wv, V1, V2);
                    continue;
                }
                if (fl & 1) {
                    const cT<T> g0 = kc_cplx<T>(coefT, cb), g1 = kc_cplx<T>(coefT, cb + 1),
                                g2 = DIM == 3 ? kc_cplx<T>(coefT, cb + 2) : czero<T>();
                    // kc1 (-i CG_1) + kc3 (-i CG_3) is the same for the lane's four elements
                    const cT<T> e = cadd<T>(cscale<T>(g0, kc1), cscale<T>(g2, kc3));
#pragma unroll
                    for (int v = 0; v < 4; v++) cfma<T>(Wv[v], cadd<T>(e, cscale<T>(g1, kc2v[v])), fv[u][v]);
                }
                if (fl & 2) {
                    const cT<T> cu1 = kc_cplx<T>(coefT, cb + DIM);
#pragma unroll
                    for (int v = 0; v < 4; v++) cfma<T>(V1[v], cu1, fv[u][v]);
                }
                if (fl & 4) {
                    const cT<T> cu2 = kc_cplx<T>(coefT, cb + DIM + 1);
#pragma unroll
                    for (int v = 0; v < 4; v++) cfma<T>(V2[v], cu2, fv[u][v]);
                }
            }
        };

        // this lane's elements of the slab and of T3 F, issued together.  The
        // lanes' elements tile the 32 x 32 plane exactly once, so they also
        // fill plane A (zero in the padding)
        cT<T> fo[4], Zv[4], Sv[4];
        const bool ldz = DIM == 3 && (!ADJ || D.has_gyro);  // the adjoint reads Z3 for its gyroscopic term only
#pragma unroll
        for (int v = 0; v < 4; v++) {
            fo[v] = *at_b(Fb, offb[v]);
            Zv[v] = ldz ? ld_once<T>(at_b(Z3 + base, offb[v])) : czero<T>();
            Sv[v] = (ADJ && DIM == 3) ? ld_once<T>(at_b(D.s3 + base, offb[v])) : czero<T>();
        }
        cT<T> yv[4];
        if constexpr (GRAD) {
#pragma unroll
            for (int v = 0; v < 4; v++) yv[v] = ld_once<T>(at_b((const cT<T>*)D.gy + base, offb[v]));
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const cT<T> a = cimul<T>(fo[v], -kc1), d = cimul<T>(fo[v], -kc2v[v]);
                pa_re[k2w[v] * kPad + k1w] = ok[v] ? a.x : T(0);
                pa_im[k2w[v] * kPad + k1w] = ok[v] ? a.y : T(0);
                pd_re[k2w[v] * kPad + k1w] = ok[v] ? d.x : T(0);
                pd_im[k2w[v] * kPad + k1w] = ok[v] ? d.y : T(0);
            }
        } else if (!ADJ || (DIM == 3 && D.has_gyro)) {  // the 3-D adjoint's gyroscopic term needs T2^H F
#pragma unroll
            for (int v = 0; v < 4; v++) {
                pa_re[k2w[v] * kPad + k1w] = ok[v] ? fo[v].x : T(0);
                pa_im[k2w[v] * kPad + k1w] = ok[v] ? fo[v].y : T(0);
            }
        }
        // (issuing the first batch together with the slab loads measured 0.3-0.8%
        // slower here, unlike in the compact kernel)
        for (int t0 = 0; t0 < TS; t0 += MB) {
            int src[MB];
            cT<T> fv[MB][4];
            load_batch(t0, src, fv);
            mix_batch(t0, src, fv);
        }
        // W += D(k) F - sum_p c_p (grid gravity term)
        const T dd0 = g00 * kd1 + g22 * kd3 - g02 * kc1 * kc3;
        const T dd1 = -g01 * kc1 - g12 * kc3;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const T dd = dd0 + g11 * kd2v[v] + dd1 * kc2v[v];
            cT<T> wv = cadd<T>(Wv[v], cscale<T>(fo[v], dd));
            const T kcp[3] = {kc1, kc2v[v], kc3};
            if constexpr (GRAV)
#pragma unroll
                for (int p = 0; p < DIM; p++)  // gravity evaluated on the grid (density form)
                    wv = csub<T>(wv, cimul<T>(*at_b(D.grav[p] + base, offb[v]), kcp[p]));
            Wv[v] = wv;
        }
        __syncthreads();  // plane A complete; the V planes of the previous slab are consumed

        cT<T> P1[4], P2[4];
        if constexpr (ADJ) {
            if constexpr (GRAD) {
                cT<T> H1[4], H2[4];  // T_p^H (-i K_p ybar)
                gm.t1(pa_re, pa_im, H1, D.c0[0]);
                gm.t2(pd_re, pd_im, H2, D.c0[1]);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    if (!ok[v]) continue;
                    const double yr = yv[v].x, yi = yv[v].y;
                    ga[0] += (double)H1[v].x * yr + (double)H1[v].y * yi;
                    ga[1] += (double)H2[v].x * yr + (double)H2[v].y * yi;
                    const double q = (double)fo[v].x * yr + (double)fo[v].y * yi;  // Re(conj(ybar) y)
                    ga[2] += (double)kd1 * q;
                    ga[3] += (double)kd2v[v] * q;
                    ga[4] -= (double)kc1 * (double)kc2v[v] * q;
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                P1[v] = P2[v] = czero<T>();
                V1[v] = csub<T>(V1[v], cimul<T>(fo[v], D.b[0] * kc1));
                V2[v] = csub<T>(V2[v], cimul<T>(fo[v], D.b[1] * kc2v[v]));
            }
            if (DIM == 3 && D.has_gyro) gm.t2(pa_re, pa_im, P2, D.c0[1]);
        } else {
            gm.t1(pa_re, pa_im, P1, D.c0[0]);
            gm.t2(pa_re, pa_im, P2, D.c0[1]);
        }
        if (DIM == 3 && D.has_gyro) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                V1[v] = csub<T>(V1[v], cimul<T>(Zv[v], kc2v[v] * D.gyro[1]));
                V1[v] = csub<T>(V1[v], cimul<T>(P2[v], kc3 * D.gyro[2]));
                V2[v] = csub<T>(V2[v], cimul<T>(Zv[v], kc1 * D.gyro[0]));
            }
        }
        // every term but T1 V1 + T2 V2, summed now so that P, Z and W retire:
        // S = W + b1 c1 P1 + b2 c2 P2 + (e3 + b3 c3) Z3
        cT<T> S[4];
        if constexpr (ADJ) {
            // the damping went into the V planes; 3-D: the axis-3 term comes
            // whole from the scaled axis-3 pass, T3^H [(e3^* - i b3 kc3) F]
#pragma unroll
            for (int v = 0; v < 4; v++)
                S[v] = DIM == 3 ? cadd<T>(Wv[v], Sv[v]) : Wv[v];
        } else {
            const cT<T> e3 = DIM == 3 ? kc_cplx<T>(as_const<T>((const T*)D.e3), r) : czero<T>();
#pragma unroll
            for (int v = 0; v < 4; v++) {
                cT<T> q = cadd<T>(Wv[v], cimul<T>(P1[v], D.b[0] * kc1));
                q = cadd<T>(q, cimul<T>(P2[v], D.b[1] * kc2v[v]));
                if (DIM == 3) {
                    q = cadd<T>(q, cmul<T>(Zv[v], e3));
                    q = cadd<T>(q, cimul<T>(Zv[v], D.b[2] * kc3));
                }
                S[v] = q;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {  // V planes, zero in the padding
            const int k2 = k2w[v];
            pb_re[k2 * kPad + k1w] = ok[v] ? V1[v].x : T(0);
            pb_im[k2 * kPad + k1w] = ok[v] ? V1[v].y : T(0);
            pc_re[k2 * kPad + k1w] = ok[v] ? V2[v].x : T(0);
            pc_im[k2 * kPad + k1w] = ok[v] ? V2[v].y : T(0);
        }
        __syncthreads();  // V planes complete; plane A is consumed
        cT<T> Q1[4], Q2[4];
        gm.t1(pb_re, pb_im, Q1, D.c0[0]);
        gm.t2(pc_re, pc_im, Q2, D.c0[1]);

        // epilogue: the stage's streams F0 (Horner stages 2-4 and the stage
        // sums) and ACC (stage sums of density-form plans only) are loaded for
        // all four elements at once (padded elements read their clamped
        // in-range address), so their latency is paid once; stores are masked
        // per element
        cT<T> q[4], f0v[4], acv[4];
#pragma unroll
        for (int v = 0; v < 4; v++) q[v] = cadd<T>(cadd<T>(Q1[v], Q2[v]), S[v]);
        if (sc.deriv_only) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (ok[v]) st_once<T>(at_b(Fout + base, offb[v]), cscale<T>(q[v], sc.dt));
            if constexpr (GRAD) {
                if (D.gacc) {  // the sweep's running sum: gacc += dt q (or input + dt q)
                    cT<T>* A = (cT<T>*)D.gacc + base;
                    cT<T> av[4];
                    if (D.gacc_init) {
#pragma unroll
                        for (int v = 0; v < 4; v++) av[v] = fo[v];  // the stage input, already loaded
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; v++) av[v] = ld_once<T>(at_b(A, offb[v]));
                    }
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        av[v] = cadd<T>(av[v], cscale<T>(q[v], sc.dt));
                        asm volatile("" : "+v"(av[v].x), "+v"(av[v].y));
                    }
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        if (ok[v]) st_once<T>(at_b(A, offb[v]), av[v]);
                }
            }
        } else {
            // a stage without F0 takes the slab itself; the loads of the four
            // elements are issued together, under workgroup-uniform conditions
            // only, and every loaded value is consumed.  These streams and the
            // outputs are touched once (non-temporal for FP64).
            if (sc.read_f0) {
#pragma unroll
                for (int v = 0; v < 4; v++) f0v[v] = ld_once<T>(at_b(F0 + base, offb[v]));
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++) f0v[v] = fo[v];  // the slab itself, already loaded
            }
            if (sc.read_acc) {
#pragma unroll
                for (int v = 0; v < 4; v++) acv[v] = ld_once<T>(at_b(ACC + base, offb[v]));
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++) acv[v] = czero<T>();
            }
            // every loaded value is consumed before the first store: a load
            // still in flight at the loop's back edge (its use sunk into a
            // masked store) makes the next slab wait for vmcnt(0), i.e. for
            // all of this slab's stores as well
            cT<T> fn[4], an[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const cT<T> upd = cadd<T>(cscale<T>(acv[v], sc.ca), cscale<T>(q[v], sc.cd));
                fn[v] = cadd<T>(f0v[v], cscale<T>(upd, sc.dt));
                an[v] = cadd<T>(cscale<T>(acv[v], sc.ea), cscale<T>(q[v], sc.ed));
                asm volatile("" : "+v"(fn[v].x), "+v"(fn[v].y), "+v"(an[v].x), "+v"(an[v].y));
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (ok[v]) {
                    st_once<T>(at_b(Fout + base, offb[v]), fn[v]);
                    if (sc.write_acc) st_once<T>(at_b(ACC + base, offb[v]), an[v]);
                }
            }
        }
        sl += per;
        r += dr;
        k3 += dk;
        if (r >= nrc) {
            r -= nrc;
            k3++;
        }
    }
    if constexpr (GRAD) {  // the workgroup's five sums, in a fixed order
#pragma unroll
        for (int i = 0; i < 5; i++) {
            double x = ga[i];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            if (lane == 0) gred[w * 5 + i] = x;
        }
        __syncthreads();
        if (tid < 5) D.gpart[(int64_t)blockIdx.x * 5 + tid] = gred[tid] + gred[5 + tid] + gred[10 + tid] + gred[15 + tid];
    }
}

// --------------------------------------------- fused kernel, 64 points ---
//
// Omega grids of 33 to 64 points per axis, reduced model (dim 2), forward
// operator.  The slab of rotation row r is its whole (zero-padded) 64 x 64
// omega plane; a workgroup of 16 waves covers it, wave w owning the 16 x 16
// output tile (k2t, k1t) = (w >> 2, w & 3), lane l and accumulator v element
// (16 k2t + orow(l >> 4, v), 16 k1t + (l & 15)), four per lane as in
// k_drift_fused, whose terms and epilogue this kernel shares.  The circulants
// are real 128 x 128 GEMMs on the (re | im) stacked plane.
//
// LDS: FP32 keeps three plane pairs (slab, V1, V2: two barriers per slab, as
// k_drift_fused).  FP64 has room for two pairs only (4 x 64 x 66 x 8 B =
// 135 KB of the 160 KB) and alternates their roles X, Y between slabs, with
// three barriers per slab:
//   F -> X; sync; P1, P2 from X; V2 -> Y; sync; Q2 from Y; V1 -> X; sync; Q1 from X.
// The next slab writes its F into Y (read last by Q2, before the third sync)
// and its V2 into X (read last by Q1, before the next slab's first sync).
constexpr int kFused64Threads = 1024;
template <class T> struct Fused64 {
    static constexpr int pairs = sizeof(T) == 4 ? 3 : 2;
    static constexpr int plane = 64 * GemmF<T, true, 64>::TP;  // elements of one real plane
    static constexpr size_t lds = 2 * sizeof(CircTab<T, 64>) + (size_t)2 * pairs * plane * sizeof(T);
};

template <class T, bool RAMP, bool GRAV, bool SMIX>
__global__ __launch_bounds__(kFused64Threads, 1) void k_drift_fused64(DriftDev<T> D, MixDev<T> MX,
                                                                      const cT<T>* Fin, const cT<T>* F0,
                                                                      cT<T>* Fout, cT<T>* ACC, StageCoef<T> sc,
                                                                      int nslab) {
    using G = GemmF<T, RAMP, 64>;
    constexpr int TP = G::TP, PL = Fused64<T>::plane, NPAIR = Fused64<T>::pairs;
    constexpr int NC = 4;  // per source: -i CG_1, -i CG_2, -u_1, -u_2
    // source rows whose loads are issued together: one for FP64, whose
    // register budget at 16 waves per CU (128 VGPRs) holds no more
    constexpr int MB = sizeof(T) == 8 ? 1 : MixBatch<T>::n;
    extern __shared__ __align__(16) unsigned char smem64[];
    CircTab<T, 64>* tabs = reinterpret_cast<CircTab<T, 64>*>(smem64);
    T* planes = reinterpret_cast<T*>(smem64 + 2 * sizeof(CircTab<T, 64>));  // pair i: re at 2 i PL, im at (2 i + 1) PL

    const int N = D.N, NN = N * N;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int k2t = w >> 2, k1t = w & 3, lr = lane & 15, lq = lane >> 4;
    tabs[0].fill(D.kap, N, tid);
    tabs[1].fill(D.kap + N, N, tid);
    __syncthreads();
    G gm;
    gm.init(&tabs[0], &tabs[1], k2t, k1t, lr, lq);

    // this lane's omega factors: kc of its elements and the diffusion
    // D(k) = G_00 kd(k1) + G_11 kd(k2) - (G_01 + G_10) kc(k1) kc(k2)
    const int k1 = 16 * k1t + lr;
    const bool k1ok = k1 < N;
    const T kc1 = k1ok ? D.kc[k1] : T(0), kd1 = k1ok ? D.kd[k1] : T(0);
    const T g00 = D.G[0], g11 = D.G[3], g01 = D.G[1] + D.G[2];
    T kc2v[4], ddv[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int k2 = 16 * k2t + G::orow(lq, v);
        const bool ok = k1ok && k2 < N;
        kc2v[v] = ok ? D.kc[k2] : T(0);
        const T kd2 = ok ? D.kd[k2] : T(0);
        ddv[v] = g00 * kd1 + g11 * kd2 - g01 * kc1 * kc2v[v];
    }

    // XCD x walks its own contiguous share of the rows (see k_drift_fused)
    const int nxcd = gridDim.x >= 8 ? 8 : 1;
    const int per = (int)gridDim.x / nxcd;
    const int xcd = (int)blockIdx.x % nxcd;
    const int hi = (int)((int64_t)nslab * (xcd + 1) / nxcd);
    int par = 0;
    for (int r = (int)((int64_t)nslab * xcd / nxcd) + (int)blockIdx.x / nxcd; r < hi; r += per) {
        int zo;  // 0, opaque: the indices below are recomputed per slab (see k_drift_fused)
        asm volatile("v_mov_b32 %0, 0" : "=v"(zo));
        const int tidz = tid + zo;
        const int lanez = tidz & 63, wz = tidz >> 6;
        const int k1w = 16 * (wz & 3) + (lanez & 15), lqz = lanez >> 4;
        int k2w[4];
        unsigned offb[4];
        bool ok[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            k2w[v] = 16 * (wz >> 2) + G::orow(lqz, v);
            ok[v] = k1w < N && k2w[v] < N;
            offb[v] = (unsigned)((k2w[v] < N ? k2w[v] : N - 1) * N + (k1w < N ? k1w : N - 1)) *
                      (unsigned)sizeof(cT<T>);
        }
        const int64_t base = (int64_t)r * NN;
        T* xre = planes + 2 * (NPAIR == 3 ? 0 : par) * PL;
        T* xim = xre + PL;
        T* yre = planes + 2 * (NPAIR == 3 ? 1 : 1 - par) * PL;
        T* yim = yre + PL;
        T* zre = planes + 2 * (NPAIR == 3 ? 2 : 0) * PL;  // three pairs only
        T* zim = zre + PL;

        cT<T> fo[4];
#pragma unroll
        for (int v = 0; v < 4; v++) fo[v] = *at_b(Fin + base, offb[v]);
#pragma unroll
        for (int v = 0; v < 4; v++) {  // the lanes' elements tile the 64 x 64 plane once
            xre[k2w[v] * TP + k1w] = ok[v] ? fo[v].x : T(0);
            xim[k2w[v] * TP + k1w] = ok[v] ? fo[v].y : T(0);
        }

        // W starts as D(k) F - sum_p c_p (grid gravity term); the mixing adds
        // the rotation terms
        cT<T> V1[4], V2[4], Wv[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            V1[v] = czero<T>();
            V2[v] = czero<T>();
            Wv[v] = cscale<T>(fo[v], ddv[v]);
            if constexpr (GRAV) {
                Wv[v] = csub<T>(Wv[v], cimul<T>(*at_b(D.grav[0] + base, offb[v]), kc1));
                Wv[v] = csub<T>(Wv[v], cimul<T>(*at_b(D.grav[1] + base, offb[v]), kc2v[v]));
            }
        }
        const int TS = MX.TS;
        const kc_ptr<int32_t> srcT = as_const<int32_t>(MX.src) + (int64_t)r * TS;
        const kc_ptr<T> coefT = as_const<T>((const T*)MX.coef) + 2 * (int64_t)r * TS * NC;
        for (int t0 = 0; t0 < TS; t0 += MB) {
            int src[MB];
            cT<T> fv[MB][4];
#pragma unroll
            for (int u = 0; u < MB; u++) src[u] = t0 + u < TS ? srcT[t0 + u] : -1;
#pragma unroll
            for (int u = 0; u < MB; u++) {
                const cT<T>* Fs = Fin + (int64_t)(src[u] < 0 ? 0 : (src[u] & 0xffffff)) * NN;
#pragma unroll
                for (int v = 0; v < 4; v++) fv[u][v] = *at_b(Fs, offb[v]);
            }
#pragma unroll
            for (int u = 0; u < MB; u++) {
                if (src[u] < 0) continue;
                const int fl = src[u] >> 24;
                const int64_t cb = (int64_t)(t0 + u) * NC;
                if constexpr (SMIX) {
                    mix_structured<T, 2>(coefT, cb, fl, kc1, kc2v, fv[u], Wv, V1, V2);
                    continue;
                }
                if (fl & 1) {
                    const cT<T> g0 = kc_cplx<T>(coefT, cb), g1 = kc_cplx<T>(coefT, cb + 1);
                    const cT<T> e = cscale<T>(g0, kc1);
#pragma unroll
                    for (int v = 0; v < 4; v++) cfma<T>(Wv[v], cadd<T>(e, cscale<T>(g1, kc2v[v])), fv[u][v]);
                }
                if (fl & 2) {
                    const cT<T> cu1 = kc_cplx<T>(coefT, cb + 2);
#pragma unroll
                    for (int v = 0; v < 4; v++) cfma<T>(V1[v], cu1, fv[u][v]);
                }
                if (fl & 4) {
                    const cT<T> cu2 = kc_cplx<T>(coefT, cb + 3);
#pragma unroll
                    for (int v = 0; v < 4; v++) cfma<T>(V2[v], cu2, fv[u][v]);
                }
            }
        }
        __syncthreads();  // plane X complete; the planes the previous slab's GEMMs read are consumed

        // V planes (zero in the padding) written before the P GEMMs, so that
        // the V registers retire early: FP32 both (planes Y, Z), FP64 V2 (Y)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            yre[k2w[v] * TP + k1w] = ok[v] ? (NPAIR == 3 ? V1[v].x : V2[v].x) : T(0);
            yim[k2w[v] * TP + k1w] = ok[v] ? (NPAIR == 3 ? V1[v].y : V2[v].y) : T(0);
            if constexpr (NPAIR == 3) {
                zre[k2w[v] * TP + k1w] = ok[v] ? V2[v].x : T(0);
                zim[k2w[v] * TP + k1w] = ok[v] ? V2[v].y : T(0);
            }
        }
        // S = W + b1 c1 P1 + b2 c2 P2
        cT<T> P[4], S[4];
        gm.t1(xre, xim, P, D.c0[0]);
#pragma unroll
        for (int v = 0; v < 4; v++) S[v] = cadd<T>(Wv[v], cimul<T>(P[v], D.b[0] * kc1));
        gm.t2(xre, xim, P, D.c0[1]);
#pragma unroll
        for (int v = 0; v < 4; v++) S[v] = cadd<T>(S[v], cimul<T>(P[v], D.b[1] * kc2v[v]));
        cT<T> Q1[4], Q2[4];
        if constexpr (NPAIR == 3) {
            __syncthreads();  // V planes complete; plane X consumed
            gm.t1(yre, yim, Q1, D.c0[0]);
            gm.t2(zre, zim, Q2, D.c0[1]);
        } else {
            __syncthreads();  // V2 in Y complete; plane X consumed
            gm.t2(yre, yim, Q2, D.c0[1]);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                xre[k2w[v] * TP + k1w] = ok[v] ? V1[v].x : T(0);
                xim[k2w[v] * TP + k1w] = ok[v] ? V1[v].y : T(0);
            }
            __syncthreads();  // V1 in X complete
            gm.t1(xre, xim, Q1, D.c0[0]);
            par ^= 1;
        }

        // epilogue (as k_drift_fused): all loads issued together and consumed
        // before the first store; stores masked per element
        cT<T> q[4], f0v[4], acv[4];
#pragma unroll
        for (int v = 0; v < 4; v++) q[v] = cadd<T>(cadd<T>(Q1[v], Q2[v]), S[v]);
        if (sc.deriv_only) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (ok[v]) st_once<T>(at_b(Fout + base, offb[v]), cscale<T>(q[v], sc.dt));
        } else {
            if (sc.read_f0) {
#pragma unroll
                for (int v = 0; v < 4; v++) f0v[v] = ld_once<T>(at_b(F0 + base, offb[v]));
            } else if constexpr (sizeof(T) == 8) {
                // the slab itself, read again (an L2 hit): FP64 does not keep
                // it in registers across the slab
#pragma unroll
                for (int v = 0; v < 4; v++) f0v[v] = ld_once<T>(at_b(Fin + base, offb[v]));
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++) f0v[v] = fo[v];
            }
            if (sc.read_acc) {
#pragma unroll
                for (int v = 0; v < 4; v++) acv[v] = ld_once<T>(at_b(ACC + base, offb[v]));
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++) acv[v] = czero<T>();
            }
            cT<T> fn[4], an[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const cT<T> upd = cadd<T>(cscale<T>(acv[v], sc.ca), cscale<T>(q[v], sc.cd));
                fn[v] = cadd<T>(f0v[v], cscale<T>(upd, sc.dt));
                an[v] = cadd<T>(cscale<T>(acv[v], sc.ea), cscale<T>(q[v], sc.ed));
                asm volatile("" : "+v"(fn[v].x), "+v"(fn[v].y), "+v"(an[v].x), "+v"(an[v].y));
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (ok[v]) {
                    st_once<T>(at_b(Fout + base, offb[v]), fn[v]);
                    if (sc.write_acc) st_once<T>(at_b(ACC + base, offb[v]), an[v]);
                }
            }
        }
    }
}

// ------------------------------------------ fused kernel, compact layout ---
//
// The same operator and epilogue as k_drift_fused (forward, real circulants,
// structured mixing) with the elementwise phases on a compact layout: lane
// tid of the workgroup holds the slab elements e = tid + 256 i, i < EPL
// (k2 = e / N, k1 = e % N), so no lane carries padding of the 32 x 32 MFMA
// tile, and the slab, T3 F, mixing-source and epilogue accesses of a wave are
// contiguous.  Only the circulant GEMMs use the tile layout.  The two layouts
// meet in LDS, three barriers per slab:
//   compact: load F, T3 F; F -> A; mixing; V1 (without the gyroscopic P2
//            term) -> B, V2 -> C; S' = W + (e3 + i b3 c3) T3 F in registers
//   sync 1
//   tile:    P1 = T1 A, P2 = T2 A; B -= i kc3 gyro3 P2 (own elements);
//            D = i (b1 kc1 P1 + b2 kc2 P2)
//   sync 2
//   tile:    Q = T1 B + T2 C -> A (valid elements only: the padding of A,
//            B and C stays zero from the start)
//   sync 3
//   compact: dF = A + D + S'; epilogue.
// A compact lane writes and reads only its own elements, so the next slab's
// F -> A needs no further barrier; the tile lanes read A, B, C and write D
// only between barriers.
template <class T> struct CompactWaves { static constexpr int n = sizeof(T) == 4 ? 4 : 2; };

template <class T, int DIM, int EPL>
__global__ __launch_bounds__(256, CompactWaves<T>::n) void k_drift_fused_c(DriftDev<T> D, MixDev<T> MX,
                                                                          const cT<T>* Fin, const cT<T>* __restrict__ Z3,
                                                                          const cT<T>* F0, cT<T>* Fout, cT<T>* ACC,
                                                                          StageCoef<T> sc, int64_t nslab) {
    using G = GemmF<T, true, 32>;
    constexpr int PL = 32 * kPad;
    __shared__ T pa_re[PL], pa_im[PL], pb_re[PL], pb_im[PL], pc_re[PL], pc_im[PL], pd_re[PL], pd_im[PL];
    __shared__ CircTab<T> tabs[2];
    constexpr int NC = DIM + 2;
    constexpr int MB = MixBatch<T>::n * (EPL <= 2 ? 2 : 1);  // source rows loaded together

    const int N = D.N, NN = N * N;
    const int nz = DIM == 3 ? N : 1;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int k2t = w >> 1, k1t = w & 1, lr = lane & 15, lq = lane >> 4;
    tabs[0].fill(D.kap, N, tid);
    tabs[1].fill(D.kap + N, N, tid);
    for (int i = tid; i < PL; i += 256) {  // the padding stays zero for the whole launch
        pa_re[i] = pa_im[i] = pb_re[i] = pb_im[i] = pc_re[i] = pc_im[i] = T(0);
    }
    __syncthreads();
    G gm;
    gm.init(&tabs[0], &tabs[1], k2t, k1t, lr, lq, N);

    // tile layout: lane's four elements (16 k2t + orow(lq, v), 16 k1t + lr)
    const int k1m = 16 * k1t + lr;
    const T kc1m = k1m < N ? D.kc[k1m] : T(0);
    T kc2m[4];
    bool okm[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int k2 = 16 * k2t + G::orow(lq, v);
        okm[v] = k1m < N && k2 < N;
        kc2m[v] = okm[v] ? D.kc[k2] : T(0);
    }
    // compact layout: elements e = tid + 256 i
    T kc1c[EPL], kc2c[EPL], ddc[EPL];
    int pos[EPL];  // LDS index k2 kPad + k1 (0 for e >= NN)
    unsigned offc[EPL];
    bool okc[EPL];
    const T g00 = D.G[0], g11 = D.G[DIM + 1], g01 = D.G[1] + D.G[DIM];
    const T g22 = DIM == 3 ? D.G[8] : T(0), g02 = DIM == 3 ? D.G[2] + D.G[6] : T(0),
            g12 = DIM == 3 ? D.G[5] + D.G[7] : T(0);
#pragma unroll
    for (int i = 0; i < EPL; i++) {
        const int e = tid + 256 * i;
        okc[i] = e < NN;
        const int ec = okc[i] ? e : NN - 1;
        const int k2 = ec / N, k1 = ec % N;
        pos[i] = k2 * kPad + k1;
        offc[i] = (unsigned)ec * (unsigned)sizeof(cT<T>);
        kc1c[i] = D.kc[k1];
        kc2c[i] = D.kc[k2];
        ddc[i] = g00 * D.kd[k1] + g11 * D.kd[k2] - g01 * kc1c[i] * kc2c[i];
    }

    const int nrc = (int)D.nrc;
    const int nxcd = gridDim.x >= 8 ? 8 : 1;
    const int per = (int)gridDim.x / nxcd;
    const int xcd = (int)blockIdx.x % nxcd;
    const int hi = (int)(nslab * (xcd + 1) / nxcd);
    int sl = (int)(nslab * xcd / nxcd) + (int)blockIdx.x / nxcd;
    const int dr = per % nrc;
    const int dk = per / nrc;
    int r = sl % nrc;
    int k3 = sl / nrc;
    while (sl < hi) {
        const int64_t base = ((int64_t)r * nz + k3) * NN;
        const cT<T>* Fb = Fin + base;
        const T kc3 = DIM == 3 ? as_const<T>(D.kc)[k3] : T(0), kd3 = DIM == 3 ? as_const<T>(D.kd)[k3] : T(0);

        // ---- compact: slab, T3 F and the first batch of mixing sources
        // (issued together), plane A, mixing
        const int TS = MX.TS;
        const kc_ptr<int32_t> srcT = as_const<int32_t>(MX.src) + (int64_t)r * TS;
        const kc_ptr<T> coefT = as_const<T>((const T*)MX.coef) + 2 * (int64_t)r * TS * NC;
        auto load_batch = [&](int t0, int* src, cT<T> (*fv)[EPL]) {
#pragma unroll
            for (int u = 0; u < MB; u++) src[u] = t0 + u < TS ? srcT[t0 + u] : -1;
#pragma unroll
            for (int u = 0; u < MB; u++) {
                const cT<T>* Fs = Fin + ((int64_t)(src[u] < 0 ? 0 : (src[u] & 0xffffff)) * nz + k3) * NN;
#pragma unroll
                for (int i = 0; i < EPL; i++) fv[u][i] = *at_b(Fs, offc[i]);
            }
        };
        cT<T> fo[EPL], Zv[EPL];
#pragma unroll
        for (int i = 0; i < EPL; i++) {
            fo[i] = *at_b(Fb, offc[i]);
            Zv[i] = DIM == 3 ? ld_once<T>(at_b(Z3 + base, offc[i])) : czero<T>();
        }
        int src0[MB];
        cT<T> fv0[MB][EPL];
        load_batch(0, src0, fv0);
#pragma unroll
        for (int i = 0; i < EPL; i++)
            if (okc[i]) {
                pa_re[pos[i]] = fo[i].x;
                pa_im[pos[i]] = fo[i].y;
            }
        cT<T> V1[EPL], V2[EPL], Wv[EPL];
#pragma unroll
        for (int i = 0; i < EPL; i++) {
            T dd = ddc[i];
            if (DIM == 3) dd += g22 * kd3 - (g02 * kc1c[i] + g12 * kc2c[i]) * kc3;
            Wv[i] = cscale<T>(fo[i], dd);
            V1[i] = czero<T>();
            V2[i] = czero<T>();
        }
        auto mix_batch = [&](int t0, const int* src, cT<T> (*fv)[EPL]) {
#pragma unroll
            for (int u = 0; u < MB; u++) {
                if (src[u] < 0) continue;
                const int fl = src[u] >> 24;
                const int64_t cb = (int64_t)(t0 + u) * NC;
                if (fl & 1) {  // W += (a + i b) F, a = Re(-i CG_1) kc1, b = Im(-i CG_2) kc2
                    const T g0 = coefT[2 * cb], g1 = coefT[2 * (cb + 1) + 1];
#pragma unroll
                    for (int i = 0; i < EPL; i++) {
                        const T a = g0 * kc1c[i], b = g1 * kc2c[i];
                        Wv[i].x = fma(a, fv[u][i].x, fma(-b, fv[u][i].y, Wv[i].x));
                        Wv[i].y = fma(a, fv[u][i].y, fma(b, fv[u][i].x, Wv[i].y));
                    }
                }
                if (fl & 2) {
                    const T ui = coefT[2 * (cb + DIM) + 1];
#pragma unroll
                    for (int i = 0; i < EPL; i++) {
                        V1[i].x = fma(-ui, fv[u][i].y, V1[i].x);
                        V1[i].y = fma(ui, fv[u][i].x, V1[i].y);
                    }
                }
                if (fl & 4) {
                    const T ur = coefT[2 * (cb + DIM + 1)];
#pragma unroll
                    for (int i = 0; i < EPL; i++) {
                        V2[i].x = fma(ur, fv[u][i].x, V2[i].x);
                        V2[i].y = fma(ur, fv[u][i].y, V2[i].y);
                    }
                }
            }
        };
        mix_batch(0, src0, fv0);
        for (int t0 = MB; t0 < TS; t0 += MB) {
            int src[MB];
            cT<T> fv[MB][EPL];
            load_batch(t0, src, fv);
            mix_batch(t0, src, fv);
        }
        // gyroscopic terms from T3 F; S' = W + (e3 + i b3 kc3) T3 F
        cT<T> Sp[EPL];
        const cT<T> e3 = DIM == 3 ? kc_cplx<T>(as_const<T>((const T*)D.e3), r) : czero<T>();
#pragma unroll
        for (int i = 0; i < EPL; i++) {
            if (DIM == 3 && D.has_gyro) {
                V1[i] = csub<T>(V1[i], cimul<T>(Zv[i], kc2c[i] * D.gyro[1]));
                V2[i] = csub<T>(V2[i], cimul<T>(Zv[i], kc1c[i] * D.gyro[0]));
            }
            Sp[i] = Wv[i];
            if (DIM == 3) {
                Sp[i] = cadd<T>(Sp[i], cmul<T>(Zv[i], e3));
                Sp[i] = cadd<T>(Sp[i], cimul<T>(Zv[i], D.b[2] * kc3));
            }
        }
#pragma unroll
        for (int i = 0; i < EPL; i++)
            if (okc[i]) {
                pb_re[pos[i]] = V1[i].x;
                pb_im[pos[i]] = V1[i].y;
                pc_re[pos[i]] = V2[i].x;
                pc_im[pos[i]] = V2[i].y;
            }
        __syncthreads();  // 1: A, B, C complete

        // ---- tile: P GEMMs, the gyroscopic P2 term into B, the damping terms into D
        {
            cT<T> P1[4], P2[4];
            gm.t1(pa_re, pa_im, P1, D.c0[0]);
            gm.t2(pa_re, pa_im, P2, D.c0[1]);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (!okm[v]) continue;
                const int p = (16 * k2t + G::orow(lq, v)) * kPad + k1m;
                if (DIM == 3 && D.has_gyro) {
                    const cT<T> g = cimul<T>(P2[v], kc3 * D.gyro[2]);
                    pb_re[p] -= g.x;
                    pb_im[p] -= g.y;
                }
                const cT<T> s = cadd<T>(cimul<T>(P1[v], D.b[0] * kc1m), cimul<T>(P2[v], D.b[1] * kc2m[v]));
                pd_re[p] = s.x;
                pd_im[p] = s.y;
            }
        }
        __syncthreads();  // 2: B and D complete; A consumed
        {
            cT<T> Q1[4], Q2[4];
            gm.t1(pb_re, pb_im, Q1, D.c0[0]);
            gm.t2(pc_re, pc_im, Q2, D.c0[1]);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (!okm[v]) continue;
                const int p = (16 * k2t + G::orow(lq, v)) * kPad + k1m;
                pa_re[p] = Q1[v].x + Q2[v].x;
                pa_im[p] = Q1[v].y + Q2[v].y;
            }
        }
        __syncthreads();  // 3: Q in A complete

        // ---- compact: dF = Q + D + S', epilogue (as k_drift_fused)
        cT<T> q[EPL], f0v[EPL], acv[EPL];
#pragma unroll
        for (int i = 0; i < EPL; i++)
            q[i] = cadd<T>(cadd<T>(cmk<T>(pa_re[pos[i]], pa_im[pos[i]]), cmk<T>(pd_re[pos[i]], pd_im[pos[i]])), Sp[i]);
        if (sc.deriv_only) {
#pragma unroll
            for (int i = 0; i < EPL; i++)
                if (okc[i]) st_once<T>(at_b(Fout + base, offc[i]), cscale<T>(q[i], sc.dt));
        } else {
            if (sc.read_f0) {
#pragma unroll
                for (int i = 0; i < EPL; i++) f0v[i] = ld_once<T>(at_b(F0 + base, offc[i]));
            } else {
#pragma unroll
                for (int i = 0; i < EPL; i++) f0v[i] = fo[i];
            }
            if (sc.read_acc) {
#pragma unroll
                for (int i = 0; i < EPL; i++) acv[i] = ld_once<T>(at_b(ACC + base, offc[i]));
            } else {
#pragma unroll
                for (int i = 0; i < EPL; i++) acv[i] = czero<T>();
            }
            cT<T> fn[EPL], an[EPL];
#pragma unroll
            for (int i = 0; i < EPL; i++) {
                const cT<T> upd = cadd<T>(cscale<T>(acv[i], sc.ca), cscale<T>(q[i], sc.cd));
                fn[i] = cadd<T>(f0v[i], cscale<T>(upd, sc.dt));
                an[i] = cadd<T>(cscale<T>(acv[i], sc.ea), cscale<T>(q[i], sc.ed));
                asm volatile("" : "+v"(fn[i].x), "+v"(fn[i].y), "+v"(an[i].x), "+v"(an[i].y));
            }
#pragma unroll
            for (int i = 0; i < EPL; i++) {
                if (okc[i]) {
                    st_once<T>(at_b(Fout + base, offc[i]), fn[i]);
                    if (sc.write_acc) st_once<T>(at_b(ACC + base, offc[i]), an[i]);
                }
            }
        }
        sl += per;
        r += dr;
        k3 += dk;
        if (r >= nrc) {
            r -= nrc;
            k3++;
        }
    }
}

// out[i] += scale * sum over the workgroups' partials, one workgroup of
// kGradRedThreads: thread t sums partials t, t + kGradRedThreads, ... in that
// order, then a fixed pairwise tree combines the threads, so the result does
// not depend on timing (the partials are independent loads, all in flight)
constexpr int kGradRedThreads = 256;
__global__ __launch_bounds__(kGradRedThreads) void k_grad_reduce(const double* __restrict__ part, int nwg, double scale,
                                                                 double* __restrict__ out) {
    __shared__ double red[5][kGradRedThreads];
    const int t = threadIdx.x;
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int g = t; g < nwg; g += kGradRedThreads)
#pragma unroll
        for (int i = 0; i < 5; i++) s[i] += part[(int64_t)g * 5 + i];
#pragma unroll
    for (int i = 0; i < 5; i++) red[i][t] = s[i];
    __syncthreads();
    for (int h = kGradRedThreads / 2; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int i = 0; i < 5; i++) red[i][t] += red[i][t + h];
        __syncthreads();
    }
    if (t < 5) out[t] += scale * red[t][0];
}

// ------------------------------------------------------------ axis 3 ----

// T3 F: one workgroup per rotation row r and chunk of 64 (k1, k2) columns.
// The chunk's loads are unconditional (out-of-range elements read a clamped
// address and are zeroed), all eight issued before the first wait, with
// 32-bit offsets from the row's base.  F and T3 F are touched once here:
// FP64 streams them non-temporally (measured faster; FP32 measured slower).
template <class T, bool SCALED = false>
__global__ __launch_bounds__(256) void k_axis3_mfma(DriftDev<T> D, const cT<T>* __restrict__ in,
                                                    cT<T>* __restrict__ out) {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc;
    constexpr int CP = 65;
    __shared__ T sre[32 * CP], sim[32 * CP];
    __shared__ CircTab<T> tab;
    const int N = D.N, NN = N * N;
    const int r = blockIdx.y, c = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lq = lane >> 4;
    const cT<T>* rin = in + (int64_t)r * N * NN;
    cT<T>* rout = out + (int64_t)r * N * NN;
    cT<T> x[8];  // row a = w + 4 i, column 64 c + lane
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int a = w + 4 * i, j = 64 * c + lane;
        const bool ok = a < N && j < NN;
        const cT<T> v = ld_once<T>(at_b(rin, ok ? (unsigned)(a * NN + j) * (unsigned)sizeof(cT<T>) : 0u));
        x[i] = ok ? v : czero<T>();
    }
    tab.fill(D.kap + 2 * N, N, tid);
    if constexpr (SCALED) {
        // row a (= k3) of the input times e3[r] - i b3 kc(a)
        const cT<T> e = kc_cplx<T>(as_const<T>((const T*)D.e3), r);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int a = w + 4 * i;
            const T kca = a < N ? D.kc[a] : T(0);
            x[i] = cmul<T>(x[i], cmk<T>(e.x, e.y - D.b[2] * kca));
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int a = w + 4 * i;
        sre[a * CP + lane] = x[i].x;
        sim[a * CP + lane] = x[i].y;
    }
    __syncthreads();
    // acc[h + 2 * im_out]: output rows k3 = 16 h + ..., real / imaginary part
    acc_t acc[4];
#pragma unroll
    for (int it = 0; it < 4; it++) acc[it] = MF::zero();
    if (D.ramp) {
        // kappa_3 = c0 + i (real circulant): 8 steps, plus c0 times the k3 sum
        // of this lane's column (its four lq lanes hold a quarter each)
        T sr = T(0), si = T(0);
#pragma unroll 4
        for (int s = 0; s < 8; s++) {
            const int k = 4 * s + lq;
            const T xr = sre[k * CP + 16 * w + lr], xi = sim[k * CP + 16 * w + lr];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const cT<T> f = tab.rim[16 * h + lr + 32 - k];
                acc[h] = MF::fma(f.x, xi, acc[h]);
                acc[h + 2] = MF::fma(f.y, xr, acc[h + 2]);
            }
            sr += xr;
            si += xi;
        }
        sr = sum_lq<T>(sr) * D.c0[2];
        si = sum_lq<T>(si) * D.c0[2];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                acc[h][v] += sr;
                acc[h + 2][v] += si;
            }
    } else {
#pragma unroll 4
        for (int s = 0; s < 16; s++) {
            const int k = 4 * s + lq;
            const T bv = (s < 8 ? sre : sim)[(k & 31) * CP + 16 * w + lr];
            const cT<T>* t = s < 8 ? tab.re_in : tab.im_in;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const cT<T> f = t[16 * h + lr + 32 - (k & 31)];
                acc[h] = MF::fma(f.x, bv, acc[h]);
                acc[h + 2] = MF::fma(f.y, bv, acc[h + 2]);
            }
        }
    }
    const int j = 64 * c + 16 * w + lr;
    if (j >= NN) return;
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k3 = 16 * h + MF::orow(lq, v);
            if (k3 < N) st_once<T>(at_b(rout, (unsigned)(k3 * NN + j) * (unsigned)sizeof(cT<T>)), cmk<T>(acc[h][v], acc[h + 2][v]));
        }
}

__global__ __launch_bounds__(256) void k_axis3_bf16x3(DriftDev<float> D, const float2* __restrict__ in,
                                                      float2* __restrict__ out) {
    constexpr int CP = 65;
    __shared__ float sre[32 * CP], sim[32 * CP];
    __shared__ CircTab<float> tab;
    const int N = D.N, NN = N * N;
    const int r = blockIdx.y, c = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lq = lane >> 4;
    const float2* rin = in + (int64_t)r * N * NN;
    float2* rout = out + (int64_t)r * N * NN;
    float2 x[8];  // unconditional clamped loads, as in k_axis3_mfma
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int a = w + 4 * i, j = 64 * c + lane;
        const bool ok = a < N && j < NN;
        const float2 v = *at_b(rin, ok ? (unsigned)(a * NN + j) * (unsigned)sizeof(float2) : 0u);
        x[i] = ok ? v : czero<float>();
    }
    tab.fill(D.kap + 2 * N, N, tid);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int a = w + 4 * i;
        sre[a * CP + lane] = x[i].x;
        sim[a * CP + lane] = x[i].y;
    }
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
        bf16x8 bh, bm, bl;
        split8(bv, bh, bm, bl);
#pragma unroll
        for (int it = 0; it < 4; it++) {
            float av[8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int k = 32 * s + 8 * lq + t;
                const float2 c = (k < 32 ? tab.re_in : tab.im_in)[16 * (it & 1) + lr + 32 - (k & 31)];
                av[t] = (it >> 1) ? c.y : c.x;
            }
            bf16x8 ah, am, al;
            split8(av, ah, am, al);
            acc[it] = mfma_split6(ah, am, al, bh, bm, bl, acc[it]);
        }
    }
    const int j = 64 * c + 16 * w + lr;
    if (j >= NN) return;
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k3 = 16 * h + 4 * lq + v;
            if (k3 < N) *at_b(rout, (unsigned)(k3 * NN + j) * (unsigned)sizeof(float2)) = cmk<float>(acc[h][v], acc[h + 2][v]);
        }
}

// ---------------------------------------------------------- launchers ---

// Workgroups of a persistent fused kernel: enough to fill every CU at the
// kernel's occupancy, never more than there are slabs.  The resident count is
// looked up once per (kernel, device).
template <class K> static unsigned persistent_grid(K kernel, int64_t nslab, int threads = 256, size_t lds = 0) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> resident;
    int dev = 0;
    HPUP_HIP(hipGetDevice(&dev));
    int cached = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = resident.find({(const void*)kernel, dev});
        if (it != resident.end()) cached = it->second;
    }
    if (!cached) {
        int ncu = 0, per_cu = 0;
        HPUP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HPUP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
        cached = std::max(1, ncu * std::max(per_cu, 1));
        std::lock_guard<std::mutex> lk(mu);
        resident[{(const void*)kernel, dev}] = cached;
    }
    int64_t g = std::min<int64_t>(nslab, cached);
    if (g >= 8) g -= g % 8;  // whole workgroups per XCD
    return (unsigned)g;
}

// the compact-layout kernel serves omega grids up to kCompactMaxN points per
// axis (at most two elements per lane: the padding of the MFMA tile is then
// at least half of it)
#ifndef HPUP_COMPACT_MAXN
#define HPUP_COMPACT_MAXN 22
#endif
constexpr int kCompactMaxN = HPUP_COMPACT_MAXN;

template <class T, int DIM, int EPL>
static void launch_compact_t(const DriftDev<T>& D, const MixDev<T>& MX, const void* Fin, const void* Z3,
                             const void* F0, void* Fout, void* ACC, const StageCoef<T>& sc, int64_t nslab,
                             hipStream_t st) {
    auto kern = k_drift_fused_c<T, DIM, EPL>;
    const unsigned grid = persistent_grid(kern, nslab);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, D, MX, (const cT<T>*)Fin, (const cT<T>*)Z3,
                       (const cT<T>*)F0, (cT<T>*)Fout, (cT<T>*)ACC, sc, nslab);
}
template <class T, int DIM>
static void launch_compact(const DriftDev<T>& D, const MixDev<T>& MX, const void* Fin, const void* Z3, const void* F0,
                           void* Fout, void* ACC, const StageCoef<T>& sc, int64_t nslab, hipStream_t st) {
    const int NN = D.N * D.N;
    if (NN <= 256) launch_compact_t<T, DIM, 1>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
    else if (NN <= 512) launch_compact_t<T, DIM, 2>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
    else if (NN <= 768) launch_compact_t<T, DIM, 3>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
    else launch_compact_t<T, DIM, 4>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
}

template <class T, bool RAMP, bool GRAV, bool SMIX>
static void launch_fused64_t(const DriftDev<T>& D, const MixDev<T>& MX, const void* Fin, const void* F0, void* Fout,
                             void* ACC, const StageCoef<T>& sc, int64_t nslab, hipStream_t st) {
    auto kern = k_drift_fused64<T, RAMP, GRAV, SMIX>;
    const size_t sh = Fused64<T>::lds;
    allow_lds((const void*)kern, sh);
    const unsigned grid = persistent_grid(kern, nslab, kFused64Threads, sh);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kFused64Threads), sh, st, D, MX, (const cT<T>*)Fin, (const cT<T>*)F0,
                       (cT<T>*)Fout, (cT<T>*)ACC, sc, (int)nslab);
}

template <class T, class G, int DIM, bool GRAV, bool ADJ = false, bool GRAD = false, bool SMIX = false>
static unsigned launch_fused_t(const DriftDev<T>& D, const MixDev<T>& MX, const void* Fin, const void* Z3,
                               const void* F0, void* Fout, void* ACC, const StageCoef<T>& sc, int64_t nslab,
                               hipStream_t st) {
    auto kern = k_drift_fused<T, G, DIM, GRAV, ADJ, GRAD, SMIX>;
    const unsigned grid = persistent_grid(kern, nslab);
    // the gradient variant writes five partials per workgroup
    HPUP_REQUIRE(!GRAD || grid <= (unsigned)kGradMaxWG, HPUP_ERR_UNSUPPORTED, "gradient stage: partial buffer too small");
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, D, MX, (const cT<T>*)Fin, (const cT<T>*)Z3,
                       (const cT<T>*)F0, (cT<T>*)Fout, (cT<T>*)ACC, sc, nslab);
    return grid;
}

// forward operator: the structured-mixing instantiation where the tables allow it
template <class T, class G, int DIM, bool GRAV>
static void launch_fwd_t(const DriftDev<T>& D, const MixDev<T>& MX, const void* Fin, const void* Z3, const void* F0,
                         void* Fout, void* ACC, const StageCoef<T>& sc, int64_t nslab, hipStream_t st) {
    if (D.smix)
        launch_fused_t<T, G, DIM, GRAV, false, false, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
    else
        launch_fused_t<T, G, DIM, GRAV>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
}

template <class T>
void launch_drift_fused(const DriftDev<T>& D, const MixDev<T>& MX, const void* Fin, const void* Z3, const void* F0,
                        void* Fout, void* ACC, const StageCoef<T>& sc, bool bf16x3, hipStream_t st) {
    HPUP_REQUIRE(D.dim == 2 || D.dim == 3, HPUP_ERR_UNSUPPORTED, "fused MFMA drift needs dim 2 or 3");
    HPUP_REQUIRE(D.N <= 32 || (D.dim == 2 && D.N <= 64), HPUP_ERR_UNSUPPORTED,
                 "fused MFMA drift needs N <= 32 (N <= 64 for dim 2)");
    const int64_t nslab = D.nrc * (D.dim == 3 ? D.N : 1);
    HPUP_REQUIRE(nslab < ((int64_t)1 << 30), HPUP_ERR_UNSUPPORTED, "fused MFMA drift: too many slabs");
    const bool grav = D.grav[0] || D.grav[1] || D.grav[2];
    if (D.N > 32) {
        HPUP_REQUIRE(!bf16x3, HPUP_ERR_UNSUPPORTED, "3xBF16 drift needs N <= 32");
        HPUP_REQUIRE(!D.adj, HPUP_ERR_UNSUPPORTED, "the adjoint operator needs N <= 32");
        if (D.ramp && D.smix) {
            if (grav) launch_fused64_t<T, true, true, true>(D, MX, Fin, F0, Fout, ACC, sc, nslab, st);
            else launch_fused64_t<T, true, false, true>(D, MX, Fin, F0, Fout, ACC, sc, nslab, st);
        } else if (D.ramp) {
            if (grav) launch_fused64_t<T, true, true, false>(D, MX, Fin, F0, Fout, ACC, sc, nslab, st);
            else launch_fused64_t<T, true, false, false>(D, MX, Fin, F0, Fout, ACC, sc, nslab, st);
        } else {
            if (grav) launch_fused64_t<T, false, true, false>(D, MX, Fin, F0, Fout, ACC, sc, nslab, st);
            else launch_fused64_t<T, false, false, false>(D, MX, Fin, F0, Fout, ACC, sc, nslab, st);
        }
        return;
    }
    if constexpr (sizeof(T) == 4) {
        if (bf16x3) {
            HPUP_REQUIRE(!grav, HPUP_ERR_UNSUPPORTED, "3xBF16 drift has no density-form gravity");
            HPUP_REQUIRE(!D.adj, HPUP_ERR_UNSUPPORTED, "the adjoint operator has no 3xBF16 form");
            if (D.dim == 3)
                launch_fwd_t<float, GemmBf16x3, 3, false>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
            else
                launch_fwd_t<float, GemmBf16x3, 2, false>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
            return;
        }
    }
    HPUP_REQUIRE(!bf16x3, HPUP_ERR_UNSUPPORTED, "3xBF16 needs FP32 storage");
    HPUP_REQUIRE(!grav || D.dim == 2, HPUP_ERR_UNSUPPORTED, "density-form gravity is a 2-D path");
    if (D.adj) {
        // model plans have coordinate-ramp circulants; only that form is built
        HPUP_REQUIRE(!grav && D.ramp, HPUP_ERR_UNSUPPORTED,
                     "the adjoint operator is built for model plans on their own omega grid");
        if (D.dim == 3) {
            HPUP_REQUIRE(!D.gpart, HPUP_ERR_UNSUPPORTED, "parameter gradients: reduced model (dim 2) only");
            HPUP_REQUIRE(D.s3, HPUP_ERR_ARG, "3-D adjoint stage needs its axis-3 term");
            if (D.smix) launch_fused_t<T, GemmF<T, true>, 3, false, true, false, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
            else launch_fused_t<T, GemmF<T, true>, 3, false, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
            return;
        }
        if (D.gpart) {
            HPUP_REQUIRE(D.gy && sc.deriv_only, HPUP_ERR_ARG, "gradient stage needs the forward stage value");
            const unsigned nwg =
                D.smix ? launch_fused_t<T, GemmF<T, true>, 2, false, true, true, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc,
                                                                                      nslab, st)
                       : launch_fused_t<T, GemmF<T, true>, 2, false, true, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc,
                                                                                nslab, st);
            hipLaunchKernelGGL(k_grad_reduce, dim3(1), dim3(kGradRedThreads), 0, st, (const double*)D.gpart, (int)nwg,
                               (double)D.gscale, D.gout);
            return;
        }
        if (D.smix) launch_fused_t<T, GemmF<T, true>, 2, false, true, false, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
        else launch_fused_t<T, GemmF<T, true>, 2, false, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
        return;
    }
    if (D.ramp && D.smix && !grav && D.N <= kCompactMaxN) {
        if (D.dim == 3) launch_compact<T, 3>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
        else launch_compact<T, 2>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
        return;
    }
    if (D.ramp) {
        if (D.dim == 3)
            launch_fwd_t<T, GemmF<T, true>, 3, false>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
        else if (grav)
            launch_fwd_t<T, GemmF<T, true>, 2, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
        else
            launch_fwd_t<T, GemmF<T, true>, 2, false>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
        return;
    }
    if (D.dim == 3)
        launch_fused_t<T, GemmF<T>, 3, false>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
    else if (grav)
        launch_fused_t<T, GemmF<T>, 2, true>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
    else
        launch_fused_t<T, GemmF<T>, 2, false>(D, MX, Fin, Z3, F0, Fout, ACC, sc, nslab, st);
}

template <class T> const char* drift_fused_name(const DriftDev<T>& D, bool grav, bool bf16x3) {
    if (D.N > 32) return "k_drift_fused64";
    if (!bf16x3 && !D.adj && D.ramp && D.smix && !grav && D.N <= kCompactMaxN) return "k_drift_fused_c";
    return "k_drift_fused";
}

template <class T>
void launch_axis3_mfma(const DriftDev<T>& D, const void* in, void* out, bool bf16x3, hipStream_t st, bool scaled) {
    HPUP_REQUIRE(D.N <= 32 && D.dim == 3, HPUP_ERR_UNSUPPORTED, "axis-3 MFMA kernel needs dim 3, N <= 32");
    const int64_t NN = (int64_t)D.N * D.N;
    HPUP_REQUIRE(D.nrc <= 65535, HPUP_ERR_UNSUPPORTED, "axis-3 kernel: more than 65535 rotation rows");
    const int nch = (int)((NN + 63) / 64);
    if (scaled) {
        HPUP_REQUIRE(!bf16x3, HPUP_ERR_UNSUPPORTED, "the adjoint operator has no 3xBF16 form");
        dim3 grid((unsigned)nch, (unsigned)D.nrc);
        hipLaunchKernelGGL((k_axis3_mfma<T, true>), grid, dim3(256), 0, st, D, (const cT<T>*)in, (cT<T>*)out);
        return;
    }
    if constexpr (sizeof(T) == 4) {
        if (bf16x3) {
            dim3 grid((unsigned)nch, (unsigned)D.nrc);
            hipLaunchKernelGGL(k_axis3_bf16x3, grid, dim3(256), 0, st, D, (const float2*)in, (float2*)out);
            return;
        }
    }
    HPUP_REQUIRE(!bf16x3, HPUP_ERR_UNSUPPORTED, "3xBF16 needs FP32 storage");
    dim3 grid((unsigned)nch, (unsigned)D.nrc);
    hipLaunchKernelGGL((k_axis3_mfma<T, false>), grid, dim3(256), 0, st, D, (const cT<T>*)in, (cT<T>*)out);
}

// ------------------------------------------------- omega DFTs (chirp) --
//
// The length-N DFT as a chirp-Toeplitz product (Bluestein's identity
// -2 k a = (k - a)^2 - k^2 - a^2):
//   X_k = p_k sum_a c_{k-a} (p_a x_a),  p_k = e^{-+ i pi k^2 / N}, c_d = e^{+- i pi d^2 / N}.
// The Toeplitz c_{k-a} is exactly what the circulant tables hold when they are
// filled with c_d for d = -32 .. 31 directly, so the transform runs on the
// same MFMA GEMMs as the drift circulants, with the chirps applied on the way
// into and out of LDS.  bs holds c_d at [d + 32] (64 entries) and then p_k
// (zero for k >= N), 96 entries; the inverse tables are the conjugates.

template <class T>
__device__ __forceinline__ void fill_toeplitz(CircTab<T>& t, const cT<T>* bs, int tid) {
    if (tid < 64) {
        const cT<T> c = bs[tid];
        t.re_in[tid] = c;
        t.im_in[tid] = cmk<T>(-c.y, c.x);
        t.rim[tid] = cmk<T>(c.x + c.y, T(0));  // the Gauss products' Re c + Im c
    }
}

// Omega DFTs of the coefficients of a real grid.  Before the omega DFT the
// coefficients satisfy F^l_{-m,-n}(x) = s conj F^l_mn(x), s = (-1)^(m-n)
// (the forward beta stage writes that mirror), so after it
// F^l_{-m,-n}(k) = s conj F^l_mn(-k): the forward DFTs need only the rows
// with n > 0, or n = 0 and m >= 0 (the half list hr: pairs (row, mirror row
// | bit 30 when s = -1)), and the last pass writes the mirror rows.  The
// inverse transform's real output depends on the coefficients only through
// their Hermitian part (F^l_mn(k) + s conj F^l_{-m,-n}(-k)) / 2, which the
// first inverse pass forms on the half rows, so the rest of the inverse runs
// on those rows only.
enum OmegaRows {
    kRowsAll = 0,     // every row
    kRowsHalf = 1,    // the half rows
    kRowsMirror = 2,  // the half rows, output written to the mirror rows too
    kRowsSym = 3      // the half rows, input taken as the Hermitian part
};

struct HalfRow {
    int r, rm;  // row and mirror row
    bool neg;   // s = -1
};
__device__ __forceinline__ HalfRow half_row(const int32_t* hr, int64_t i) {
    const int32_t b = hr[2 * i + 1];
    return {hr[2 * i], b & 0x3fffffff, ((b >> 30) & 1) != 0};
}
// s conj z
template <class T> __device__ __forceinline__ cT<T> sconj(cT<T> z, bool neg) {
    return neg ? cmk<T>(-z.x, z.y) : cmk<T>(z.x, -z.y);
}

// 2-D DFT over (k1, k2) of N x N slabs of in, times `scale`, into out (in
// place when they are equal: a slab is read whole before it is written, and
// the mirror rows a kRowsMirror pass writes are read by no other slab).  Slab
// sl is row sl / nz at k3 = sl % nz (rows from the half list unless kRowsAll).
template <class T, int MODE>
__global__ __launch_bounds__(256) void k_slab_dft_mfma(const cT<T>* in, cT<T>* out, const cT<T>* __restrict__ bs,
                                                       int N, T scale, int nslab, int nz, const int32_t* __restrict__ hr) {
    __shared__ T pa_re[32 * kPad], pa_im[32 * kPad], pb_re[32 * kPad], pb_im[32 * kPad];
    __shared__ CircTab<T> tab;
    __shared__ cT<T> pk[32];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int k2t = w >> 1, k1t = w & 1, lr = lane & 15, lq = lane >> 4;
    fill_toeplitz<T>(tab, bs, tid);
    if (tid < 32) pk[tid] = bs[64 + tid];
    __syncthreads();
    // reductions trimmed to the N data points (plane A is zero beyond N along
    // k1; plane B's rows k2 >= N are zero through p_k2 = 0)
    GemmF<T, false> gm;
    gm.init(&tab, &tab, k2t, k1t, lr, lq, N);
    const int NN = N * N;
    const int k1 = 16 * k1t + lr;
    const cT<T> p1 = pk[k1];
    // slab sl: its base, and for the mirror modes the mirror slab's (at -k3)
    auto slab = [&](int sl, int64_t& b, int64_t& bm, bool& neg, bool& self) {
        if (MODE == kRowsAll) {
            b = (int64_t)sl * NN;
            bm = b;
            neg = false;
            self = true;
            return;
        }
        const int i = sl / nz, k3 = sl - i * nz;
        const HalfRow h = half_row(hr, i);
        b = ((int64_t)h.r * nz + k3) * NN;
        bm = ((int64_t)h.rm * nz + (k3 == 0 ? 0 : nz - k3)) * NN;
        neg = h.neg;
        self = h.rm == h.r;
    };
    // the slab's four elements per thread (clamped in-range loads; with the
    // mirror slab's element at (-k2, -k1) for kRowsSym); the next slab's are
    // loaded while this one is transformed
    cT<T> x[4], y[4];
    auto load = [&](int sl) {
        int64_t b, bm;
        bool ng, sf;
        slab(sl, b, bm, ng, sf);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int idx = tid + 256 * i, a2 = idx >> 5, a1 = idx & 31;
            const bool ok = a2 < N && a1 < N;
            x[i] = in[b + (ok ? a2 * N + a1 : 0)];
            if (MODE == kRowsSym) {
                const int m2 = a2 == 0 ? 0 : N - a2, m1 = a1 == 0 ? 0 : N - a1;
                y[i] = in[bm + (ok ? m2 * N + m1 : 0)];
            }
        }
    };
    if ((int)blockIdx.x < nslab) load(blockIdx.x);
    for (int sl = blockIdx.x; sl < nslab; sl += gridDim.x) {
        int64_t b, bm;
        bool neg, self;
        slab(sl, b, bm, neg, self);
#pragma unroll
        for (int i = 0; i < 4; i++) {  // pre-chirp along k1, zero outside the slab
            const int idx = tid + 256 * i, a2 = idx >> 5, a1 = idx & 31;
            cT<T> u = x[i];
            if (MODE == kRowsSym) {
                const cT<T> c = sconj<T>(y[i], neg);
                u = cmk<T>(T(0.5) * (u.x + c.x), T(0.5) * (u.y + c.y));
            }
            const cT<T> v = a2 < N && a1 < N ? cmul<T>(u, pk[a1]) : czero<T>();
            pa_re[a2 * kPad + a1] = v.x;
            pa_im[a2 * kPad + a1] = v.y;
        }
        __syncthreads();  // plane A complete; plane B of the previous slab consumed
        // another slab: no overlap with this slab's output, also in place
        if (sl + (int)gridDim.x < nslab) load(sl + gridDim.x);
        cT<T> R[4];
        gm.t1_gauss(pa_re, pa_im, R);  // three real GEMMs (Gauss), 2-3% faster than four
#pragma unroll
        for (int v = 0; v < 4; v++) {  // along k1 done; pre-chirp along k2 (p = 0 in the padding)
            const int k2 = 16 * k2t + GemmF<T, false>::orow(lq, v);
            const cT<T> yv = cmul<T>(cmul<T>(R[v], p1), pk[k2]);
            pb_re[k2 * kPad + k1] = yv.x;
            pb_im[k2 * kPad + k1] = yv.y;
        }
        __syncthreads();  // plane B complete; plane A consumed
        gm.t2_gauss(pb_re, pb_im, R);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k2 = 16 * k2t + GemmF<T, false>::orow(lq, v);
            if (k1 < N && k2 < N) {
                const cT<T> z = cscale<T>(cmul<T>(R[v], pk[k2]), scale);
                out[b + k2 * N + k1] = z;
                if (MODE == kRowsMirror && !self)
                    out[bm + (k2 == 0 ? 0 : N - k2) * N + (k1 == 0 ? 0 : N - k1)] = sconj<T>(z, neg);
            }
        }
    }
}

// DFT along the outermost axis: lines of N points at stride `inner`, one per
// (line block L, column j); 64-column chunks as the B operand of the Toeplitz
// GEMM (the k_axis3_mfma layout), times `scale`.  Out of place or in place.
// Line block blockIdx.y is a row (from the half list unless kRowsAll); the
// mirror of column j = k2 N + k1 is (-k2) N + (-k1) (3-D rows, inner = N^2).
template <class T, int MODE>
__global__ __launch_bounds__(256) void k_axis_dft_mfma(const cT<T>* in, cT<T>* out, const cT<T>* __restrict__ bs,
                                                       int N, int64_t inner, T scale, const int32_t* __restrict__ hr) {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc;
    constexpr int CP = 65;
    __shared__ T sre[32 * CP], sim[32 * CP];
    __shared__ CircTab<T> tab;
    __shared__ cT<T> pk[32];
    int64_t L = blockIdx.y, Lm = L;
    bool neg = false;
    if (MODE != kRowsAll) {
        const HalfRow h = half_row(hr, blockIdx.y);
        L = h.r;
        Lm = h.rm;
        neg = h.neg;
    }
    const int64_t j0 = (int64_t)blockIdx.x * 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lr = lane & 15, lq = lane >> 4;
    // mirror column of j (clamped in range)
    auto mirror_col = [&](int64_t j) {
        const int jj = (int)(j < inner ? j : inner - 1), c2 = jj / N, c1 = jj - c2 * N;
        return (int64_t)(c2 == 0 ? 0 : N - c2) * N + (c1 == 0 ? 0 : N - c1);
    };
    // the chunk's eight elements per thread (row a = w + 4 i, column lane):
    // unconditional clamped loads, all issued before the first use
    cT<T> x[8], y[8];
    {
        const int64_t j = j0 + lane, jm = MODE == kRowsSym ? mirror_col(j) : 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int a = w + 4 * i;
            const bool ok = a < N && j < inner;
            x[i] = in[ok ? (L * N + a) * inner + j : 0];
            if (MODE == kRowsSym) y[i] = in[ok ? (Lm * N + (a == 0 ? 0 : N - a)) * inner + jm : 0];
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int a = w + 4 * i;
            if (MODE == kRowsSym) {
                const cT<T> c = sconj<T>(y[i], neg);
                x[i] = cmk<T>(T(0.5) * (x[i].x + c.x), T(0.5) * (x[i].y + c.y));
            }
            if (!(a < N && j < inner)) x[i] = czero<T>();
        }
    }
    fill_toeplitz<T>(tab, bs, tid);
    if (tid < 32) pk[tid] = bs[64 + tid];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int a = w + 4 * i;
        const cT<T> v = a < N ? cmul<T>(x[i], pk[a]) : czero<T>();
        sre[a * CP + lane] = v.x;
        sim[a * CP + lane] = v.y;
    }
    __syncthreads();
    acc_t acc[4];
#pragma unroll
    for (int it = 0; it < 4; it++) acc[it] = MF::zero();
    // K = 4 steps over the rows that hold data (rows a >= N are zero), for the
    // real and then the imaginary half
    const int ks = (N + 3) / 4;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        const T* sp = half == 0 ? sre : sim;
        const cT<T>* t = half == 0 ? tab.re_in : tab.im_in;
#pragma unroll 4
        for (int s = 0; s < ks; s++) {
            const int k = 4 * s + lq;
            const T bv = sp[k * CP + 16 * w + lr];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const cT<T> f = t[16 * h + lr + 32 - k];
                acc[h] = MF::fma(f.x, bv, acc[h]);
                acc[h + 2] = MF::fma(f.y, bv, acc[h + 2]);
            }
        }
    }
    const int64_t j = j0 + 16 * w + lr;
    if (j >= inner) return;
    const int64_t jm = MODE == kRowsMirror ? mirror_col(j) : 0;
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k = 16 * h + MF::orow(lq, v);
            if (k < N) {
                const cT<T> z = cscale<T>(cmul<T>(cmk<T>(acc[h][v], acc[h + 2][v]), pk[k]), scale);
                out[(L * N + k) * inner + j] = z;
                if (MODE == kRowsMirror && Lm != L) out[(Lm * N + (k == 0 ? 0 : N - k)) * inner + jm] = sconj<T>(z, neg);
            }
        }
}

template <class T>
void launch_slab_dft_mfma(const void* in, void* out, int64_t nslab, int N, const void* bs, T scale,
                          hipStream_t st) {
    HPUP_REQUIRE(N <= 32, HPUP_ERR_UNSUPPORTED, "chirp MFMA DFT needs N <= 32");
    HPUP_REQUIRE(nslab < ((int64_t)1 << 31), HPUP_ERR_UNSUPPORTED, "chirp MFMA DFT: too many slabs");
    const unsigned grid = (unsigned)std::min<int64_t>(nslab, 4096);
    hipLaunchKernelGGL((k_slab_dft_mfma<T, kRowsAll>), dim3(grid), dim3(256), 0, st, (const cT<T>*)in, (cT<T>*)out,
                       (const cT<T>*)bs, N, scale, (int)nslab, 1, nullptr);
}

template <class T>
void launch_axis_dft_mfma(const void* in, void* out, int N, int64_t inner, int64_t nlines, const void* bs, T scale,
                          hipStream_t st) {
    HPUP_REQUIRE(N <= 32, HPUP_ERR_UNSUPPORTED, "chirp MFMA DFT needs N <= 32");
    dim3 grid((unsigned)((inner + 63) / 64), (unsigned)nlines);
    hipLaunchKernelGGL((k_axis_dft_mfma<T, kRowsAll>), grid, dim3(256), 0, st, (const cT<T>*)in, (cT<T>*)out,
                       (const cT<T>*)bs, N, inner, scale, nullptr);
}

template <class T, int MODE>
static void slab_half(const XformDev<T>& X, const void* in, void* out, const cT<T>* bs, T scale, hipStream_t st) {
    const int N = X.N, nz = X.dim == 3 ? N : 1;
    const int64_t nslab = X.nhalf * nz;
    HPUP_REQUIRE(nslab < ((int64_t)1 << 31), HPUP_ERR_UNSUPPORTED, "chirp MFMA DFT: too many slabs");
    const unsigned grid = (unsigned)std::min<int64_t>(nslab, 4096);
    hipLaunchKernelGGL((k_slab_dft_mfma<T, MODE>), dim3(grid), dim3(256), 0, st, (const cT<T>*)in, (cT<T>*)out, bs,
                       N, scale, (int)nslab, nz, X.hrows);
}
template <class T, int MODE>
static void axis_half(const XformDev<T>& X, const void* in, void* out, const cT<T>* bs, T scale, hipStream_t st) {
    const int N = X.N;
    const int64_t inner = (int64_t)N * N;
    HPUP_REQUIRE(X.nhalf <= 65535, HPUP_ERR_UNSUPPORTED, "axis DFT: more than 65535 rows");
    dim3 grid((unsigned)((inner + 63) / 64), (unsigned)X.nhalf);
    hipLaunchKernelGGL((k_axis_dft_mfma<T, MODE>), grid, dim3(256), 0, st, (const cT<T>*)in, (cT<T>*)out, bs, N,
                       inner, scale, X.hrows);
}

template <class T>
void launch_omega_half(const XformDev<T>& X, const void* in, void* out, bool inverse, hipStream_t st) {
    HPUP_REQUIRE((X.dim == 2 || X.dim == 3) && X.N <= 32 && X.bsF && X.hrows, HPUP_ERR_UNSUPPORTED,
                 "half-row omega DFT: dim 2 or 3, N <= 32");
    const int N = X.N;
    const T s2 = inverse ? T(1) / T(N * N) : T(1), s3 = inverse ? T(1) / T(N) : T(1);
    const cT<T>* bs = inverse ? X.bsI : X.bsF;
    if (!inverse) {
        HPUP_REQUIRE(in == out, HPUP_ERR_ARG, "forward half-row omega DFT runs in place");
        if (X.dim == 3) {
            slab_half<T, kRowsHalf>(X, in, out, bs, s2, st);
            axis_half<T, kRowsMirror>(X, out, out, bs, s3, st);
        } else {
            slab_half<T, kRowsMirror>(X, in, out, bs, s2, st);
        }
        return;
    }
    HPUP_REQUIRE(in != out, HPUP_ERR_ARG, "inverse half-row omega DFT runs out of place");
    if (X.dim == 3) {
        axis_half<T, kRowsSym>(X, in, out, bs, s3, st);
        slab_half<T, kRowsHalf>(X, out, out, bs, s2, st);
    } else {
        slab_half<T, kRowsSym>(X, in, out, bs, s2, st);
    }
}

template void launch_omega_half<float>(const XformDev<float>&, const void*, void*, bool, hipStream_t);
template void launch_omega_half<double>(const XformDev<double>&, const void*, void*, bool, hipStream_t);
template void launch_slab_dft_mfma<float>(const void*, void*, int64_t, int, const void*, float, hipStream_t);
template void launch_slab_dft_mfma<double>(const void*, void*, int64_t, int, const void*, double, hipStream_t);
template void launch_axis_dft_mfma<float>(const void*, void*, int, int64_t, int64_t, const void*, float, hipStream_t);
template void launch_axis_dft_mfma<double>(const void*, void*, int, int64_t, int64_t, const void*, double,
                                           hipStream_t);

template void launch_drift_fused<float>(const DriftDev<float>&, const MixDev<float>&, const void*, const void*,
                                        const void*, void*, void*, const StageCoef<float>&, bool, hipStream_t);
template void launch_drift_fused<double>(const DriftDev<double>&, const MixDev<double>&, const void*, const void*,
                                         const void*, void*, void*, const StageCoef<double>&, bool, hipStream_t);
template const char* drift_fused_name<float>(const DriftDev<float>&, bool, bool);
template const char* drift_fused_name<double>(const DriftDev<double>&, bool, bool);
template void launch_axis3_mfma<float>(const DriftDev<float>&, const void*, void*, bool, hipStream_t, bool);
template void launch_axis3_mfma<double>(const DriftDev<double>&, const void*, void*, bool, hipStream_t, bool);

}  // namespace hpup
