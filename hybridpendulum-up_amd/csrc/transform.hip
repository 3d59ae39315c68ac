// SO(3) x R^n forward / inverse transforms (fftSO3R.m:90-106, ifftSO3R.m:16-37).
//
// Forward:  grid f (alpha,beta,gamma,x)
//   1. k_fwd_ag   per (x-chunk, beta_j): real 2-D DFT over (alpha, gamma) with
//                 the flip(fftshift) sign convention (frequency -m stored at m),
//                 Hermitian half computed, the other half by conjugation
//                 -> G[(m,n)][j][x] (stored Omega-blocked, see g_off)
//   2. k_beta_fwd per (m,n): F^l_mn(x) = sum_j w_j d^l_mn(beta_j) G[(m,n)][j][x]
//   3. omega DFT of every compact row (slab kernel on axes 1,2; axis kernel on 3)
// Inverse: the same stages reversed (1/nx on omega, unnormalized over alpha,
// gamma, real part), Nyquist bins of alpha and gamma implicitly zero; the
// intermediate (H, the Hermitian part) in the order of its (alpha, gamma)
// readers (see h_base).
#include <algorithm>

#include "drift.h"
#include "kernels_common.h"
#include "mfma.h"

namespace hpup {

// Offset of G[(m,n)][j][x] (mn = m + lmax + M (n + lmax)) in the intermediate
// buffer: Omega points in blocks of kGX, each block holding [(m,n)][j][x % kGX].
// A (beta_j, Omega chunk) tile of the angular stages then spans one block
// (tens of KB) instead of one page per (m, n) (the plain [(m,n)][j][x] order
// put its 841 rows 6.5 MB apart at B = 15, 30^3, and the address translations
// dominated), while the beta stages still read and write runs of kGX
// consecutive Omega points.
__device__ __forceinline__ int64_t g_off(int B, int mn, int j, int64_t x) {
    const int64_t M = 2 * B - 1, n2 = 2 * B;
    return ((x / kGX) * (M * M * n2) + (int64_t)mn * n2 + j) * kGX + (x % kGX);
}

// The forward (alpha, gamma) stages fill the half of G that the beta stage
// reads: the rows mn >= (M^2 - 1)/2, i.e. n > 0, or n = 0 and m >= 0.  The
// grid is real, so G[-m][-n] = conj G[m][n]: a value computed at (m >= 0, n)
// goes to (m, n) for n >= 0 and, conjugated, to (-m, -n) for n < 0 and
// m > 0 (at m = 0 the n < 0 half is the conjugate of the n > 0 half).
// Gx: G at (mn = 0, j, x); ni = n + lmax.
template <class T>
__device__ __forceinline__ void put_g_half(cT<T>* Gx, int B, int m, int ni, cT<T> z) {
    const int lmax = B - 1, M = 2 * B - 1;
    const int64_t rs = (int64_t)2 * B * kGX;
    if (ni >= lmax) Gx[((int64_t)(m + lmax) + M * ni) * rs] = z;
    else if (m > 0) Gx[((int64_t)(lmax - m) + M * (M - 1 - ni)) * rs] = cconj<T>(z);
}

// (-1)^d conj z: the forward coefficient at (-m, -n) from the one at (m, n),
// d = m - n (before the omega DFT, at the same omega point)
template <class T> __device__ __forceinline__ cT<T> mirror_coef(cT<T> z, int d) {
    return (d & 1) ? cmk<T>(-z.x, z.y) : cmk<T>(z.x, -z.y);
}

// The inverse transform's real output depends on G only through its
// Hermitian part, H[m][n] = (G[m][n] + conj G[-m][-n]) / 2 at each (beta_j,
// omega point): Re sum_mn G[m][n] e^{-i(m a + n g)} = sum_mn H[m][n] e^{-i(m a + n g)}.
// The inverse beta stage therefore writes H on the same half as the forward
// (rows mn >= (M^2 - 1)/2), and the (alpha, gamma) stages read the other half
// as conjugates.
//
// The inverse intermediate H (the Hermitian part on the stored half of (m, n),
// written by the inverse beta stage) in the order of its readers: for each
// pair of omega points and each beta_j the half's rows contiguous,
// [x / 2][j][mh][x % 2], mh = mn - (M^2 - 1)/2 < h_rows(B) (padded to a
// multiple of 8, so that a 128-byte line holds 8 rows of a pair in FP32).  An
// (alpha, gamma) block of the inverse then reads whole lines (in the forward
// order, [(m, n)][j][x % 64], it used 16 bytes of each 128-byte line), and a
// beta workgroup writes whole lines by taking 8 consecutive rows.
__host__ __device__ __forceinline__ int h_rows(int B) {
    const int M = 2 * B - 1;
    return ((M * M + 1) / 2 + 7) / 8 * 8;
}
// H at (mh = 0, j, x); row mh is at [2 mh]
__device__ __forceinline__ int64_t h_base(int B, int j, int64_t x) {
    return (((x >> 1) * (2 * B) + j) * (int64_t)h_rows(B)) * 2 + (x & 1);
}
// Value at (mi, ni) (mn = mi + M ni) from Hx = H + h_base(B, j, x): the
// unstored half is read as the conjugate of the row (-m, -n), with one load
// at the selected row and a sign select
template <class T> __device__ __forceinline__ cT<T> get_h(const cT<T>* Hx, int B, int mi, int ni) {
    const int M = 2 * B - 1, mn = mi + M * ni, h0 = (M * M - 1) / 2;
    const bool stored = mn >= h0;
    const cT<T> v = Hx[2 * (stored ? mn - h0 : h0 - mn)];
    return cmk<T>(v.x, stored ? v.y : -v.y);
}

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
        put_g_half<T>(G + g_off(B, 0, j, x), B, m, n + lmax, acc);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_beta_fwd(XformDev<T> X, const cT<T>* __restrict__ G, cT<T>* __restrict__ F) {
    extern __shared__ char smem_raw[];
    T* sW = reinterpret_cast<T*>(smem_raw);
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    const int mn = (M * M - 1) / 2 + blockIdx.y;  // the half of G the forward stages fill
    const int m = mn % M - lmax, n = mn / M - lmax;
    const int lmin = max(abs(m), abs(n));
    for (int i = threadIdx.x; i < B * n2; i += blockDim.x) sW[i] = X.Wf[((size_t)mn * B + i / n2) * X.wfs + i % n2];
    __syncthreads();
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= X.nx) return;
    const cT<T>* Gp = G + g_off(B, mn, 0, x);
    for (int l0 = lmin; l0 <= lmax; l0 += 8) {
        cT<T> acc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = czero<T>();
        for (int j = 0; j < n2; j++) {
            const cT<T> g = Gp[(int64_t)j * kGX];
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
            // F^l_{-m,-n} = (-1)^(m-n) conj F^l_mn (a real grid; d^l_{-m,-n} = (-1)^(m-n) d^l_mn)
            if (mn != (M * M - 1) / 2) F[d_ridx(B, l0 + t, -m, -n) * X.nx + x] = mirror_coef<T>(acc[t], m - n);
        }
    }
}

// Reduction order of the beta GEMMs: the K = 4 steps in groups of KG = min(KS,
// 4); lane group lq takes index (s / KG) 4 KG + KG lq + s % KG at step s, so
// that its KG consecutive steps read consecutive weights of a row (the tables'
// rows are zero-padded to 4 KS entries, whole groups)
template <int KS> struct KGroup { static constexpr int n = KS < 4 ? KS : 4; };
template <int KS> __device__ __forceinline__ int kstep_index(int s, int lq) {
    constexpr int KG = KGroup<KS>::n;
    return (s / KG) * 4 * KG + KG * lq + s % KG;
}

// The beta stages as MFMA GEMMs per (m, n) block over the omega points.
// Forward: F^l_mn(x) = sum_j Wf[mn][l][j] G[mn][j][x] -- A = Wf (rows l, RT
// tiles of 16; K = j in KS steps of 4), B = G (K = j, columns x), one real
// GEMM on the real and one on the imaginary part of G.  A workgroup of four
// waves takes one (m, n) and one kGX-block of omega points (each wave a
// 16-point column tile), whose G chunk [j][x % kGX] is contiguous; rows l <
// max(|m|, |n|) and the padding carry zero weights and are not stored.
// MIRROR: also write the rows (l, -m, -n) (otherwise the half-row omega DFT
// writes them after its own transform)
template <class T, int RT, int KS, bool MIRROR>
__global__ __launch_bounds__(256) void k_beta_fwd_mfma(XformDev<T> X, const cT<T>* __restrict__ G,
                                                       cT<T>* __restrict__ F) {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc;
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    const int mn = (M * M - 1) / 2 + blockIdx.y;  // the half of G the forward stages fill
    const int m = mn % M - lmax, n = mn / M - lmax;
    const int lmin = max(abs(m), abs(n));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    const int64_t nx = X.nx;
    const int64_t x = (int64_t)blockIdx.x * kGX + 16 * w + lr;  // this lane's column
    const bool xok = x < nx;
    const cT<T>* Gp = G + g_off(B, mn, 0, xok ? x : nx - 1);
    // reduction order: step s takes j = kstep_index(s, lq), so that a lane's
    // weights of KG consecutive steps are consecutive in a row (one vector load)
    cT<T> g[KS];
#pragma unroll
    for (int s = 0; s < KS; s++) {
        const int j = kstep_index<KS>(s, lq);
        const cT<T> v = Gp[(int64_t)(j < n2 ? j : 0) * kGX];
        g[s] = (j < n2 && xok) ? v : czero<T>();
    }
    const T* W = X.Wf + (size_t)mn * B * X.wfs;
    acc_t re[RT], im[RT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
        re[rt] = MF::zero();
        im[rt] = MF::zero();
        const int l = 16 * rt + lr;
        const bool lok = l >= lmin && l < B;
        const T* Wl = W + (size_t)(l < B ? l : lmax) * X.wfs;
#pragma unroll
        for (int c = 0; c < KS / KGroup<KS>::n; c++) {
            constexpr int KG = KGroup<KS>::n;
            typedef T VecT __attribute__((ext_vector_type(KG)));
            const VecT wv = *reinterpret_cast<const VecT*>(Wl + 4 * KG * c + KG * lq);
#pragma unroll
            for (int t = 0; t < KG; t++) {
                const int s = KG * c + t;
                const T a = lok ? wv[t] : T(0);  // zero beyond n2 in the table
                re[rt] = MF::fma(a, g[s].x, re[rt]);
                im[rt] = MF::fma(a, g[s].y, im[rt]);
            }
        }
    }
    if (!xok) return;
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int l = 16 * rt + MF::orow(lq, v);
            if (l >= lmin && l < B) {
                const cT<T> z = cmk<T>(re[rt][v], im[rt][v]);
                F[d_ridx(B, l, m, n) * nx + x] = z;
                if (MIRROR && mn != (M * M - 1) / 2) F[d_ridx(B, l, -m, -n) * nx + x] = mirror_coef<T>(z, m - n);
            }
        }
}

// Inverse: H[mn][j][x] = sum_l Wi[mn][j][l] Fs^l_mn(x) -- A = Wi (rows j, RT
// tiles; K = l in KS steps), B = the coefficient rows of (m, n) (Fs: their
// Hermitian part).  A workgroup of eight waves takes eight consecutive rows
// mh of the stored half and 16 omega points, wave w the row 8 blockIdx.y + w,
// so that together they write whole lines of H (h_base).
// SYM: form the Hermitian part from the rows (l, m, n) and (l, -m, -n)
// (otherwise the half-row omega DFT has formed it on the rows (l, m, n))
template <class T, int RT, int KS, bool SYM>
__global__ __launch_bounds__(512) void k_beta_inv_mfma(XformDev<T> X, const cT<T>* __restrict__ F,
                                                       cT<T>* __restrict__ G) {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc;
    // one 16-row tile of j at a time, [j % 16][x % 16][w ^ ((x % 16) / 2)], stored
    // as whole lines (the swizzle spreads the 16 omega points of a wave's write
    // over distinct banks; a line's reads stay within one 128-byte row group)
    __shared__ cT<T> tile[16 * 16 * 8];
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1, h0 = (M * M - 1) / 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    // the padding rows (mh > h0) compute on the clamped row and store nothing
    const int mh = 8 * blockIdx.y + w, mhc = mh <= h0 ? mh : h0;
    const int mn = h0 + mhc;
    const int m = mn % M - lmax, n = mn / M - lmax;
    const int lmin = max(abs(m), abs(n));
    const int64_t nx = X.nx;
    const int64_t x = (int64_t)blockIdx.x * 16 + lr;
    const bool xok = x < nx;
    const int64_t xc = xok ? x : nx - 1;
    // (F^l_mn + (-1)^(m-n) conj F^l_{-m,-n}) / 2: with Wi[-m,-n] = (-1)^(m-n) Wi[m,n]
    // its contraction is H[m][n]
    // FP32 with 8 K steps (bandlimits 17-32): the forward stage's grouped
    // reduction order (step s takes l = kstep_index(s, lq), one vector load per
    // KG weights of a row; 1% faster at B = 30); otherwise l = 4 s + lq (the
    // grouped order measured 2% slower at B = 15 in FP32, 10% in FP64)
    constexpr bool GROUPED = sizeof(T) == 4 && KS >= 8;
    cT<T> f[KS];
#pragma unroll
    for (int s = 0; s < KS; s++) {
        const int l = GROUPED ? kstep_index<KS>(s, lq) : 4 * s + lq;
        const bool lok = l >= lmin && l < B;
        const cT<T> v = F[d_ridx(B, lok ? l : lmax, m, n) * nx + xc];
        if (SYM) {
            const cT<T> u = mirror_coef<T>(F[d_ridx(B, lok ? l : lmax, -m, -n) * nx + xc], m - n);
            f[s] = (lok && xok) ? cmk<T>(T(0.5) * (v.x + u.x), T(0.5) * (v.y + u.y)) : czero<T>();
        } else {
            f[s] = (lok && xok) ? v : czero<T>();
        }
    }
    const T* W = X.Wi + (size_t)mn * n2 * X.wis;
    acc_t re[RT], im[RT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
        re[rt] = MF::zero();
        im[rt] = MF::zero();
        const int j = 16 * rt + lr;
        if constexpr (GROUPED) {
            // rows j >= n2 (tile padding) read row n2 - 1 and use zeros; the
            // rows hold wis >= 4 KS entries, zero beyond B
            constexpr int KG = KGroup<KS>::n;
            typedef T VecT __attribute__((ext_vector_type(KG)));
            const T* Wj = W + (size_t)(j < n2 ? j : n2 - 1) * X.wis;
#pragma unroll
            for (int c = 0; c < KS / KG; c++) {
                const VecT wv = *reinterpret_cast<const VecT*>(Wj + 4 * KG * c + KG * lq);
#pragma unroll
                for (int t = 0; t < KG; t++) {
                    const int s = KG * c + t, l = kstep_index<KS>(s, lq);
                    const T a = (j < n2 && l >= lmin && l < B) ? wv[t] : T(0);
                    re[rt] = MF::fma(a, f[s].x, re[rt]);
                    im[rt] = MF::fma(a, f[s].y, im[rt]);
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const int l = 4 * s + lq;
                const T a = (j < n2 && l >= lmin && l < B) ? W[(size_t)j * X.wis + l] : T(0);
                re[rt] = MF::fma(a, f[s].x, re[rt]);
                im[rt] = MF::fma(a, f[s].y, im[rt]);
            }
        }
    }
    // H lines: (j, pair of omega points) -> the 8 rows of the workgroup, both
    // points of the pair per row (FP32: one 16-byte piece per row, FP64: one
    // per row and point); consecutive threads take consecutive pieces of a line
    constexpr int PR = sizeof(T) == 4 ? 1 : 2;  // pieces per (row, pair)
    const int64_t x0 = (int64_t)blockIdx.x * 16;
    const int mrows = h0 + 1 - 8 * (int)blockIdx.y;  // rows of this workgroup that exist
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
        if (16 * rt >= n2) break;
#pragma unroll
        for (int v = 0; v < 4; v++)
            tile[(MF::orow(lq, v) * 16 + lr) * 8 + (w ^ (lr >> 1))] = cmk<T>(re[rt][v], im[rt][v]);
        __syncthreads();
#pragma unroll
        for (int p = threadIdx.x; p < 16 * 8 * 8 * PR; p += 512) {
            const int r8 = (p / PR) % 8, xp = (p / (PR * 8)) % 8, jl = p / (PR * 64);
            const int j = 16 * rt + jl;
            const int64_t xa = x0 + 2 * xp;
            if (j < n2 && r8 < mrows && xa < nx) {
                cT<T>* dst = G + h_base(B, j, xa) + 2 * (8 * (int)blockIdx.y + r8);
                if (PR == 1) {
                    const cT<T> a = tile[(jl * 16 + 2 * xp) * 8 + (r8 ^ xp)],
                                b = tile[(jl * 16 + 2 * xp + 1) * 8 + (r8 ^ xp)];
                    if (xa + 1 < nx) {
                        typedef T V4 __attribute__((ext_vector_type(4)));
                        V4 q;
                        q.x = a.x; q.y = a.y; q.z = b.x; q.w = b.y;
                        *reinterpret_cast<V4*>(dst) = q;
                    } else {
                        dst[0] = a;
                    }
                } else {
                    const int xo = p % 2;
                    if (xa + xo < nx) dst[xo] = tile[(jl * 16 + 2 * xp + xo) * 8 + (r8 ^ xp)];
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------- inverse -------

template <class T>
__global__ __launch_bounds__(256) void k_beta_inv(XformDev<T> X, const cT<T>* __restrict__ F, cT<T>* __restrict__ G) {
    extern __shared__ char smem_raw[];
    T* sW = reinterpret_cast<T*>(smem_raw);
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    const int mn = (M * M - 1) / 2 + blockIdx.y;  // the stored half of H (see get_h)
    const int m = mn % M - lmax, n = mn / M - lmax;
    const int lmin = max(abs(m), abs(n));
    for (int i = threadIdx.x; i < B * n2; i += blockDim.x) sW[i] = X.Wi[((size_t)mn * n2 + i / B) * X.wis + i % B];
    __syncthreads();
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= X.nx) return;
    const int mh = mn - (M * M - 1) / 2;
    for (int j0 = 0; j0 < n2; j0 += 8) {
        cT<T> acc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = czero<T>();
        for (int l = lmin; l <= lmax; l++) {
            const cT<T> v = F[d_ridx(B, l, m, n) * X.nx + x];
            const cT<T> u = mirror_coef<T>(F[d_ridx(B, l, -m, -n) * X.nx + x], m - n);
            const cT<T> fv = cmk<T>(T(0.5) * (v.x + u.x), T(0.5) * (v.y + u.y));
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
            G[h_base(B, j0 + t, x) + 2 * mh] = acc[t];
        }
    }
}

// Inverse over (alpha, gamma) for one beta_j and XC omega points:
//   stage 1: H[m][g] = sum_n G[m][n] e^{-i n gamma_g}: a task per (m, x,
//            chunk of GC gammas) streams its G row once, the GC outputs in
//            registers, twiddle indices advanced incrementally;
//   stage 2: f[a][g] = Re sum_m H[m][g] e^{-i m alpha_a} from LDS.
template <class T, int GC>
__global__ __launch_bounds__(256) void k_inv_ag(XformDev<T> X, const cT<T>* __restrict__ G, T* __restrict__ f, int XC) {
    extern __shared__ char smem_raw[];
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    cT<T>* sH = reinterpret_cast<cT<T>*>(smem_raw);      // [m][g][xl]
    cT<T>* tw = sH + (size_t)M * n2 * XC;                 // n2: e^{i pi t / B}
    const int j = blockIdx.y;
    const int64_t x0 = (int64_t)blockIdx.x * XC;
    const int64_t nx = X.nx;
    const int tid = threadIdx.x;
    for (int t = tid; t < n2; t += blockDim.x) tw[t] = X.twB[t];
    __syncthreads();
    const int ngc = (n2 + GC - 1) / GC;
    for (int task = tid; task < M * ngc * XC; task += blockDim.x) {
        const int xl = task % XC, gc = (task / XC) % ngc, mi = task / (XC * ngc);
        const int64_t x = x0 + xl;
        const int g0 = gc * GC;
        cT<T> acc[GC];
#pragma unroll
        for (int k = 0; k < GC; k++) acc[k] = czero<T>();
        if (x < nx) {
            const cT<T>* Hx = G + h_base(B, j, x);
            for (int ni = 0; ni < M; ni++) {
                const cT<T> gv = get_h<T>(Hx, B, mi, ni);
                const int n = ni - lmax;
                const int dn = ((n % n2) + n2) % n2;  // (n g) mod n2 advances by dn per gamma
                int t = (int)(((int64_t)dn * g0) % n2);
#pragma unroll
                for (int k = 0; k < GC; k++) {
                    cfma<T>(acc[k], gv, cconj<T>(tw[t]));
                    t += dn;
                    if (t >= n2) t -= n2;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < GC; k++)
            if (g0 + k < n2) sH[((size_t)mi * n2 + g0 + k) * XC + xl] = acc[k];
    }
    __syncthreads();
    // over m -> alpha, real part
    for (int i = tid; i < XC * n2 * n2; i += blockDim.x) {
        const int a = i % n2, g = (i / n2) % n2, xl = i / (n2 * n2);
        const int64_t x = x0 + xl;
        if (x >= nx) continue;
        int t = ((-lmax * a) % n2 + n2) % n2;  // (m a) mod n2 for m = -lmax, then + a per m
        T acc = T(0);
        for (int mi = 0; mi < M; mi++) {
            const cT<T> h = sH[((size_t)mi * n2 + g) * XC + xl];
            const cT<T> w = tw[t];  // e^{-i m alpha} = conj(w): Re(h conj(w)) = h.x w.x + h.y w.y
            acc = fma(h.x, w.x, fma(h.y, w.y, acc));
            t += a;
            if (t >= n2) t -= n2;
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

// ------------------------------------- (alpha, gamma) stages for B <= 16 --
//
// Work items hold one row of the stage's input in registers and sweep a
// frequency x angle twiddle table E[r][t] = e^{i pi (r - lmax) t / B}, zero
// padded to 32 x 32, so every loop runs a fixed 32 (or 8) steps without
// guards.  The table row and column of each step depend only on loop
// counters and on the wave's part of the block, so the table is read through
// the scalar cache (uniform addresses) instead of LDS.  A stage whose work
// items fill only part of the block splits its output range into chunks of
// 8 over wave-aligned parts.  Blocks are numbered so that consecutive omega
// chunks of one beta_j land on the same XCD (the hardware deals blocks to the
// 8 XCDs round-robin): a cache line shared by neighbouring chunks is then
// fetched into one L2 only.

constexpr int kAgMax = 32;  // 2B limit of the register-row kernels

// table in the constant address space: uniform indices become scalar loads
template <class T> using cTab = __attribute__((address_space(4))) const T*;
template <class T> __device__ __forceinline__ cT<T> etw(cTab<T> e, int i) { return cmk<T>(e[2 * i], e[2 * i + 1]); }

// block b of nb -> work index, a bijection with XCD k = b % 8 owning a
// contiguous range of work indices (the last nb % 8 blocks keep their index)
__device__ __forceinline__ int xcd_major(int b, int nb) {
    const int per = nb >> 3;
    if (b >= per << 3) return b;
    return (b & 7) * per + (b >> 3);
}

// Work of block blockIdx.x of an (alpha, gamma) kernel: the slice beta_j and
// the omega points x0 .. x0 + XC - 1.  The blocks take groups of XT x-tiles
// (the points of one 128-byte line of G) for every j in turn, so that the
// blocks an XCD runs side by side share the G lines of neighbouring x-tiles
// and also the f lines that straddle the slices j and j + 1 of a (gamma, x) row
// (with j slowest, every such line was fetched twice: 1.47x the f bytes).
template <class T, int XC>
__device__ __forceinline__ void ag_block(int nxc, int n2, int& j, int64_t& x0) {
    constexpr int XT = XC * (int)sizeof(cT<T>) >= 128 ? 1 : 128 / (XC * (int)sizeof(cT<T>));
    const int w = xcd_major(blockIdx.x, gridDim.x);
    const int full = nxc / XT, head = full * XT * n2;
    int xt;
    if (w < head) {
        const int r = w % (XT * n2);
        j = r / XT;
        xt = (w / (XT * n2)) * XT + r % XT;
    } else {  // the last nxc % XT x-tiles
        const int tr = nxc - full * XT, r = w - head;
        j = r / tr;
        xt = full * XT + r % tr;
    }
    x0 = (int64_t)xt * XC;
}

// tasks of a stage -> parts of the block: <= 64 -> 4, <= 128 -> 2, else 1
__device__ __forceinline__ int stage_parts(int T) { return T <= 64 ? 4 : (T <= 128 ? 2 : 1); }

template <class T, int XC>
__global__ __launch_bounds__(256) void k_fwd_ag_r(XformDev<T> X, const T* __restrict__ f, const T* __restrict__ wR,
                                                  int wstride, cT<T>* __restrict__ G, int nxc) {
    extern __shared__ char smem_raw[];
    const int B = X.B, n2 = 2 * B, lmax = B - 1, M = 2 * B - 1;
    const int rp = n2 + 1;                                   // padded row of sf
    T* sf = reinterpret_cast<T*>(smem_raw);                  // [xl][g][a]  XC*n2*rp
    cT<T>* sH = reinterpret_cast<cT<T>*>(smem_raw);          // [m>=0][g][xl], aliases sf
    const cTab<T> E = (cTab<T>)(const T*)X.twMB;
    int j;
    int64_t x0;
    ag_block<T, XC>(nxc, n2, j, x0);
    const int64_t nx = X.nx;
    const int tid = threadIdx.x;
    const int64_t n2c = (int64_t)n2 * n2 * n2;
    for (int i = tid; i < XC * n2 * n2; i += 256) {
        const int xl = i / (n2 * n2), g = (i / n2) % n2, a = i % n2;
        const int64_t x = x0 + xl;
        const int64_t iR = a + (int64_t)n2 * (j + (int64_t)n2 * g);
        T v = x < nx ? f[iR + n2c * x] : T(0);
        if (wR) v *= wR[iR * wstride];  // transform of the product f(R, x) w(R)
        sf[(xl * n2 + g) * rp + a] = v;
    }
    __syncthreads();
    // alpha DFT, m >= 0: H[m][g] = sum_a f[a] e^{+i m alpha_a}; item (g, xl)
    {
        const int TA = n2 * XC, SA = stage_parts(TA), PA = 256 / SA;
        const int part = __builtin_amdgcn_readfirstlane(tid / PA), task = tid % PA;
        const bool act = task < TA;
        const int g = task / XC, xl = task % XC;
        T row[kAgMax];
#pragma unroll
        for (int a = 0; a < kAgMax; a++) row[a] = (act && a < n2) ? sf[(xl * n2 + g) * rp + a] : T(0);
        __syncthreads();  // sH overwrites sf
        for (int m = part; m < B; m += SA) {
            const cTab<T> e = E + 2 * (m + lmax) * kAgMax;
            cT<T> acc = czero<T>();
#pragma unroll
            for (int a = 0; a < kAgMax; a++) {
                const cT<T> tw = etw<T>(e, a);
                acc.x = fma(row[a], tw.x, acc.x);
                acc.y = fma(row[a], tw.y, acc.y);
            }
            if (act) sH[(m * n2 + g) * XC + xl] = acc;
        }
    }
    __syncthreads();
    // gamma DFT: G[m][n] = sum_g H[m][g] e^{+i n gamma_g}, G[-m][-n] = conj; item (m, xl)
    {
        const int TB = B * XC, SB = stage_parts(TB), PB = 256 / SB;
        const int part = __builtin_amdgcn_readfirstlane(tid / PB), task = tid % PB;
        const bool act = task < TB;
        const int m = task / XC, xl = task % XC;
        const int64_t x = x0 + xl;
        cT<T> h[kAgMax];
#pragma unroll
        for (int g = 0; g < kAgMax; g++) h[g] = (act && g < n2) ? sH[(m * n2 + g) * XC + xl] : czero<T>();
        const bool st = act && x < nx;
        cT<T>* Gx = G + g_off(B, 0, j, min<int64_t>(x, nx - 1));
        for (int ni = part; ni < M; ni += SB) {
            const cTab<T> e = E + 2 * ni * kAgMax;
            cT<T> acc = czero<T>();
#pragma unroll
            for (int g = 0; g < kAgMax; g++) cfma<T>(acc, h[g], etw<T>(e, g));
            if (st) put_g_half<T>(Gx, B, m, ni, acc);
        }
    }
}

template <class T, int XC>
__global__ __launch_bounds__(256) void k_inv_ag_r(XformDev<T> X, const cT<T>* __restrict__ G, T* __restrict__ f,
                                                  int nxc) {
    extern __shared__ char smem_raw[];
    const int B = X.B, n2 = 2 * B, M = 2 * B - 1;
    cT<T>* sH = reinterpret_cast<cT<T>*>(smem_raw);  // [m][g][xl]
    const cTab<T> ET = (cTab<T>)(const T*)X.twMB + 2 * kAgMax * kAgMax;  // [t][r]
    int j;
    int64_t x0;
    ag_block<T, XC>(nxc, n2, j, x0);
    const int64_t nx = X.nx;
    const int tid = threadIdx.x;
    // H[m][g] = sum_n G[m][n] e^{-i n gamma_g}; item (m, xl) holds its G row
    {
        const int T1 = M * XC, S1 = stage_parts(T1), P1 = 256 / S1;
        const int part = __builtin_amdgcn_readfirstlane(tid / P1), task = tid % P1;
        const bool act = task < T1;
        const int mi = task / XC, xl = task % XC;
        const int64_t x = x0 + xl;
        const bool ld = act && x < nx;
        // clamped, branch-free loads; invalid entries are zeroed after the load
        const cT<T>* Hx = G + h_base(B, j, min<int64_t>(x, nx - 1));
        cT<T> gr[kAgMax];
#pragma unroll
        for (int ni = 0; ni < kAgMax; ni++) {
            const cT<T> v = get_h<T>(Hx, B, min(mi, M - 1), min(ni, M - 1));
            gr[ni] = (ld && ni < M) ? v : czero<T>();
        }
        for (int g = part; g < n2; g += S1) {
            const cTab<T> e = ET + 2 * g * kAgMax;
            cT<T> acc = czero<T>();
#pragma unroll
            for (int ni = 0; ni < kAgMax; ni++) cfma<T>(acc, gr[ni], cconj<T>(etw<T>(e, ni)));
            if (act) sH[(mi * n2 + g) * XC + xl] = acc;
        }
    }
    __syncthreads();
    // f[a][g] = Re sum_m H[m][g] e^{-i m alpha_a}; item (g, xl) holds its H column
    {
        const int T2 = n2 * XC, S2 = stage_parts(T2), P2 = 256 / S2;
        const int part = __builtin_amdgcn_readfirstlane(tid / P2), task = tid % P2;
        const bool act = task < T2;
        const int g = task / XC, xl = task % XC;
        const int64_t x = x0 + xl;
        const bool st = act && x < nx;
        cT<T> hr[kAgMax];
#pragma unroll
        for (int mi = 0; mi < kAgMax; mi++) hr[mi] = (act && mi < M) ? sH[(mi * n2 + g) * XC + xl] : czero<T>();
        T* fo = f + (int64_t)n2 * (j + (int64_t)n2 * (g + (int64_t)n2 * min<int64_t>(x, nx - 1)));
        for (int a = part; a < n2; a += S2) {
            const cTab<T> e = ET + 2 * a * kAgMax;
            T acc = T(0);
#pragma unroll
            for (int mi = 0; mi < kAgMax; mi++) {
                const cT<T> tw = etw<T>(e, mi);  // Re(h conj(tw))
                acc = fma(hr[mi].x, tw.x, fma(hr[mi].y, tw.y, acc));
            }
            if (st) fo[a] = acc;
        }
    }
}

// The f tile of a forward (alpha, gamma) block, [xl][g][a] at row stride rp:
// f(alpha_a, beta_j, gamma_g, x0 + xl) (times the grid weight wR when given),
// zero beyond nx.  A row of 2B alpha values is contiguous in f, so the tile
// is read as 16-byte (8-byte for 2B = 30 in FP32) pieces at 32-bit offsets from
// the block's base, all issued before the first use: one memory round trip
// per tile.  The weights are applied in a second pass; a weight branch inside
// the load loop had made every load wait for its data (one round trip per
// element).
template <class T, int XC, int NT, int n2>
__device__ __forceinline__ void stage_f_tile(const T* __restrict__ f, const T* __restrict__ wR, int wstride, int j,
                                             int64_t x0, int64_t nx, T* sf, int rp) {
    constexpr int VE = n2 % (16 / (int)sizeof(T)) == 0 ? 16 / (int)sizeof(T) : 8 / (int)sizeof(T);
    static_assert(VE >= 1 && n2 % VE == 0, "rows of whole pieces");
    typedef T VecT __attribute__((ext_vector_type(VE)));
    constexpr int RV = n2 / VE, TOT = XC * n2 * RV, NL = (TOT + NT - 1) / NT;
    const char* fb = reinterpret_cast<const char*>(f + (int64_t)n2 * j + (int64_t)n2 * n2 * n2 * x0);
    const int xlmax = (int)(nx - 1 - x0 < XC - 1 ? nx - 1 - x0 : XC - 1);  // last omega point of the tile
    const int tid = threadIdx.x;
    VecT v[NL];
#pragma unroll
    for (int u = 0; u < NL; u++) {
        const int i = tid + u * NT;
        const int ic = i < TOT ? i : TOT - 1;
        const int xl = ic / (n2 * RV), g = (ic / RV) % n2, pc = ic % RV;
        const int xc = xl < xlmax ? xl : xlmax;
        const unsigned off = (unsigned)((pc * VE + n2 * n2 * g + n2 * n2 * n2 * xc) * (int)sizeof(T));
        v[u] = *reinterpret_cast<const VecT*>(fb + off);
    }
    if (wR) {  // transform of the product f(R, x) w(R)
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int i = tid + u * NT;
            const int ic = i < TOT ? i : TOT - 1;
            const int g = (ic / RV) % n2, pc = ic % RV;
#pragma unroll
            for (int e = 0; e < VE; e++)
                v[u][e] *= wR[(pc * VE + e + (int64_t)n2 * (j + (int64_t)n2 * g)) * wstride];
        }
    }
#pragma unroll
    for (int u = 0; u < NL; u++) {
        const int i = tid + u * NT;
        if (i < TOT) {
            const int xl = i / (n2 * RV), g = (i / RV) % n2, pc = i % RV;
            T* dst = sf + (xl * n2 + g) * rp + pc * VE;
#pragma unroll
            for (int e = 0; e < VE; e++) dst[e] = xl <= xlmax ? v[u][e] : T(0);
        }
    }
}

// The output tile of an inverse (alpha, gamma) block from LDS ([xl][g][a], row
// stride n2) to f(alpha_a, beta_j, gamma_g, x0 + xl): a row of 2B reals is
// contiguous in f, so it is written in 16-byte pieces (8-byte for 2B = 30 in
// FP32), one store per piece instead of one per value.
template <class T, int XC, int NT, int n2>
__device__ __forceinline__ void store_f_rows(const T* so, T* __restrict__ f, int j, int64_t x0, int64_t nx) {
    constexpr int VE = n2 % (16 / (int)sizeof(T)) == 0 ? 16 / (int)sizeof(T) : 8 / (int)sizeof(T);
    static_assert(VE >= 1 && n2 % VE == 0, "rows of whole pieces");
    typedef T VecT __attribute__((ext_vector_type(VE)));
    constexpr int RV = n2 / VE;
    char* fb = reinterpret_cast<char*>(f + (int64_t)n2 * j + (int64_t)n2 * n2 * n2 * x0);
    for (int e = threadIdx.x; e < XC * n2 * RV; e += NT) {
        const int xl = e / (n2 * RV), g = (e / RV) % n2, pc = e % RV;
        if (x0 + xl < nx) {
            const unsigned off = (unsigned)((pc * VE + n2 * n2 * g + n2 * n2 * n2 * xl) * (int)sizeof(T));
            *reinterpret_cast<VecT*>(fb + off) = *reinterpret_cast<const VecT*>(so + (xl * n2 + g) * n2 + pc * VE);
        }
    }
}

// ------------------------------------------- (alpha, gamma) stages, 2B = 30 --
//
// At 2B = 30 the two angular DFTs of each (beta_j, omega) tile are length-30
// DFTs of a whole row or column, done as Good-Thomas prime-factor FFTs
// (30 = 2 * 3 * 5, coprime factors, no twiddle factors): x[n] is read at
// n = 15 n1 + 10 n2 + 6 n3 (mod 30) as a 2 x 3 x 5 array, transformed along
// each axis, and X[k] written back through the same map (15, 10 and 6 are
// also the CRT idempotents of 2, 3 and 5).  DIR = -1: e^{-2 pi i n k / 30},
// DIR = +1: e^{+2 pi i n k / 30}.  About 20x fewer operations than the
// direct DFTs, so the stages become bound by their G and f traffic.

template <class T, int DIR> __device__ __forceinline__ cT<T> irot(cT<T> a) {  // DIR i a
    return DIR > 0 ? cmk<T>(-a.y, a.x) : cmk<T>(a.y, -a.x);
}

template <class T, int DIR> __device__ __forceinline__ void dft3(cT<T>& x0, cT<T>& x1, cT<T>& x2) {
    const T h = T(0.86602540378443864676);  // sin(2 pi / 3)
    const cT<T> t = cadd<T>(x1, x2), d = csub<T>(x1, x2);
    const cT<T> m = cmk<T>(x0.x - T(0.5) * t.x, x0.y - T(0.5) * t.y);
    const cT<T> r = irot<T, DIR>(cscale<T>(d, h));
    x0 = cadd<T>(x0, t);
    x1 = cadd<T>(m, r);
    x2 = csub<T>(m, r);
}

template <class T, int DIR> __device__ __forceinline__ void dft5(cT<T>* x) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);  // cos(2 pi / 5), cos(4 pi / 5)
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);   // sin(2 pi / 5), sin(4 pi / 5)
    const cT<T> a1 = cadd<T>(x[1], x[4]), b1 = csub<T>(x[1], x[4]);
    const cT<T> a2 = cadd<T>(x[2], x[3]), b2 = csub<T>(x[2], x[3]);
    const cT<T> p1 = cmk<T>(x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y);
    const cT<T> p2 = cmk<T>(x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y);
    const cT<T> q1 = irot<T, DIR>(cmk<T>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    const cT<T> q2 = irot<T, DIR>(cmk<T>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    x[0] = cadd<T>(x[0], cadd<T>(a1, a2));
    x[1] = cadd<T>(p1, q1);
    x[4] = csub<T>(p1, q1);
    x[2] = cadd<T>(p2, q2);
    x[3] = csub<T>(p2, q2);
}

template <class T, int DIR> __device__ __forceinline__ void fft30(cT<T>* x) {
    cT<T> a[2][3][5];
#pragma unroll
    for (int n1 = 0; n1 < 2; n1++)
#pragma unroll
        for (int n2 = 0; n2 < 3; n2++)
#pragma unroll
            for (int n3 = 0; n3 < 5; n3++) a[n1][n2][n3] = x[(15 * n1 + 10 * n2 + 6 * n3) % 30];
#pragma unroll
    for (int n2 = 0; n2 < 3; n2++)
#pragma unroll
        for (int n3 = 0; n3 < 5; n3++) {
            const cT<T> u = a[0][n2][n3], v = a[1][n2][n3];
            a[0][n2][n3] = cadd<T>(u, v);
            a[1][n2][n3] = csub<T>(u, v);
        }
#pragma unroll
    for (int n1 = 0; n1 < 2; n1++)
#pragma unroll
        for (int n3 = 0; n3 < 5; n3++) dft3<T, DIR>(a[n1][0][n3], a[n1][1][n3], a[n1][2][n3]);
#pragma unroll
    for (int n1 = 0; n1 < 2; n1++)
#pragma unroll
        for (int n2 = 0; n2 < 3; n2++) dft5<T, DIR>(a[n1][n2]);
#pragma unroll
    for (int k1 = 0; k1 < 2; k1++)
#pragma unroll
        for (int k2 = 0; k2 < 3; k2++)
#pragma unroll
            for (int k3 = 0; k3 < 5; k3++) x[(15 * k1 + 10 * k2 + 6 * k3) % 30] = a[k1][k2][k3];
}

template <class T, int DIR> __device__ __forceinline__ void dft4(cT<T>* x) {
    const cT<T> a = cadd<T>(x[0], x[2]), b = csub<T>(x[0], x[2]);
    const cT<T> c = cadd<T>(x[1], x[3]), d = irot<T, DIR>(csub<T>(x[1], x[3]));
    x[0] = cadd<T>(a, c);
    x[2] = csub<T>(a, c);
    x[1] = cadd<T>(b, d);
    x[3] = csub<T>(b, d);
}

// the length-2B transform of the register-row FFT kernels (2B = 30)
template <class T, int DIR, int NF> __device__ __forceinline__ void fftn(cT<T>* x) {
    static_assert(NF == 30, "register-row FFTs are built for 2B = 30");
    fft30<T, DIR>(x);
}

// Inverse (alpha, gamma) stage at 2B = 30: H[m][g] = sum_n G[m][n] e^{-i n gamma_g}
// (item (m, xl): its G row, frequency n at index n mod 30), then
// f[a][g] = Re sum_m H[m][g] e^{-i m alpha_a} (item (g, xl): its H column).
// Row stride (elements) of the FFT kernels' H tile [m][g][xl]: padded so that
// the lanes of a wave, which write or read one g of eight (FP32) or four
// (FP64) rows m at once, spread over all LDS banks (2-way for 8-byte and
// 4-way for 16-byte elements, the minimum for a wave's 512 or 1024 bytes)
template <class T> __host__ __device__ constexpr int fft_sh(int XC, int n2 = 30) {
    return n2 * XC + (sizeof(T) == 4 ? 8 : 4);
}

template <class T, int B, int XC, int NT>
__global__ __launch_bounds__(NT) void k_inv_ag_fft(XformDev<T> X, const cT<T>* __restrict__ G, T* __restrict__ f,
                                                   int nxc) {
    extern __shared__ char smem_raw[];
    constexpr int n2 = 2 * B, M = 2 * B - 1, lmax = B - 1, SH = fft_sh<T>(XC, n2);
    cT<T>* sH = reinterpret_cast<cT<T>*>(smem_raw);  // [m >= 0][g][xl], row stride SH
    int j;
    int64_t x0;
    ag_block<T, XC>(nxc, n2, j, x0);
    const int64_t nx = X.nx;
    // the Hermitian part of G (get_h): H[-m][g] = conj H[m][g], so only
    // the rows m >= 0 are transformed over n
    for (int task = threadIdx.x; task < B * XC; task += NT) {
        const int m = task / XC, xl = task % XC;
        const int64_t x = x0 + xl;
        const bool ld = x < nx;
        // clamped, branch-free loads; a column beyond nx is zeroed after the load
        const cT<T>* Hx = G + h_base(B, j, min<int64_t>(x, nx - 1));
        cT<T> v[n2];
#pragma unroll
        for (int ni = 0; ni < M; ni++) {
            const cT<T> g = get_h<T>(Hx, B, lmax + m, ni);
            v[(ni - lmax + n2) % n2] = ld ? g : czero<T>();
        }
        v[B] = czero<T>();
        fftn<T, -1, n2>(v);
#pragma unroll
        for (int g = 0; g < n2; g++) sH[m * SH + g * XC + xl] = v[g];
    }
    __syncthreads();
    // one column (g, xl) per thread; the real rows go through LDS ([xl][g][a],
    // aliasing sH) so that f is written in whole-row pieces
    static_assert(n2 * XC <= NT, "one column per thread");
    const int task = threadIdx.x;
    const bool act = task < n2 * XC;
    const int g = act ? task / XC : 0, xl = act ? task % XC : 0;
    cT<T> v[n2];
#pragma unroll
    for (int mi = 0; mi < M; mi++) {
        const int m = mi - lmax;
        const cT<T> h = sH[(m >= 0 ? m : -m) * SH + g * XC + xl];
        v[(mi - lmax + n2) % n2] = m >= 0 ? h : cconj<T>(h);
    }
    v[B] = czero<T>();
    fftn<T, -1, n2>(v);
    __syncthreads();  // so overwrites sH
    T* so = reinterpret_cast<T*>(smem_raw);
    if (act) {
#pragma unroll
        for (int a = 0; a < n2; a++) so[(xl * n2 + g) * n2 + a] = v[a].x;
    }
    __syncthreads();
    store_f_rows<T, XC, NT, n2>(so, f, j, x0, nx);
}

// Forward (alpha, gamma) stage at 2B = 30: H[m][g] = sum_a f[a][g] e^{+i m alpha_a},
// m >= 0 (item (g, xl): its f row), then G[m][n] = sum_g H[m][g] e^{+i n gamma_g}
// and G[-m][-n] = conj (item (m, xl): its H row).
template <class T, int B, int XC, int NT>
__global__ __launch_bounds__(NT) void k_fwd_ag_fft(XformDev<T> X, const T* __restrict__ f, const T* __restrict__ wR,
                                                   int wstride, cT<T>* __restrict__ G, int nxc) {
    extern __shared__ char smem_raw[];
    constexpr int n2 = 2 * B, lmax = B - 1, M = 2 * B - 1, rp = n2 + 1, SH = fft_sh<T>(XC, n2);
    T* sf = reinterpret_cast<T*>(smem_raw);          // [xl][g][a]  XC*n2*rp
    cT<T>* sH = reinterpret_cast<cT<T>*>(smem_raw);  // [m>=0][g][xl], row stride SH, aliases sf
    int j;
    int64_t x0;
    ag_block<T, XC>(nxc, n2, j, x0);
    const int64_t nx = X.nx;
    const int tid = threadIdx.x;
    stage_f_tile<T, XC, NT, n2>(f, wR, wstride, j, x0, nx, sf, rp);
    __syncthreads();
    constexpr int TA = n2 * XC;
    static_assert(TA <= NT, "one alpha row per thread");
    cT<T> v[n2];
    const bool actA = tid < TA;
    const int gA = tid / XC, xlA = tid % XC;
    if (actA) {
#pragma unroll
        for (int a = 0; a < n2; a++) v[a] = cmk<T>(sf[(xlA * n2 + gA) * rp + a], T(0));
    }
    __syncthreads();  // sH overwrites sf
    if (actA) {
        fftn<T, 1, n2>(v);
#pragma unroll
        for (int m = 0; m < B; m++) sH[m * SH + gA * XC + xlA] = v[m];
    }
    __syncthreads();
    for (int task = tid; task < B * XC; task += NT) {
        const int m = task / XC, xl = task % XC;
        const int64_t x = x0 + xl;
#pragma unroll
        for (int g = 0; g < n2; g++) v[g] = sH[m * SH + g * XC + xl];
        fftn<T, 1, n2>(v);
        if (x < nx) {
            cT<T>* Gx = G + g_off(B, 0, j, x);
#pragma unroll
            for (int ni = 0; ni < M; ni++) put_g_half<T>(Gx, B, m, ni, v[(ni - lmax + n2) % n2]);
        }
    }
}

// ------------------------------------------- (alpha, gamma) stages, 2B = 60 --
//
// A whole 60-point vector per thread needs more than 120 VGPRs (one wave per
// SIMD), so each length-60 DFT is spread over the four lanes of a quad: lane
// q holds x[4t + q], t < 15, transforms it with a 15-point Good-Thomas FFT
// (15 = 3 x 5: input n = 5 n1 + 3 n2, output k = 10 k1 + 6 k2, mod 15),
// multiplies by W^(q k), W = e^{DIR 2 pi i / 60}, and a radix-4 DFT across
// the quad in two lane exchanges finishes the transform:
//   X[k + 15 p] = sum_q W4^(q p) W^(q k) Y_q[k],  W4 = W^15 = DIR i.
// With p = 2 p1 + p0 and q = 2 q1 + q0 the exchange with lane q ^ 2 forms
// sum_q1 (-1)^(q1 p0) y_(2 q1 + q0), lane q keeping p0 = q1; lane 3 then
// multiplies by W4 (q0 = p0 = 1); the exchange with lane q ^ 1 forms the sum
// over q0 with p1 = q0.  Lane L ends with X[k + 15 rev(L)], k < 15, rev the
// 2-bit reversal.

template <class T, int DIR> __device__ __forceinline__ void fft15(cT<T>* x) {
    cT<T> a[3][5];
#pragma unroll
    for (int n1 = 0; n1 < 3; n1++)
#pragma unroll
        for (int n2 = 0; n2 < 5; n2++) a[n1][n2] = x[(5 * n1 + 3 * n2) % 15];
#pragma unroll
    for (int n2 = 0; n2 < 5; n2++) dft3<T, DIR>(a[0][n2], a[1][n2], a[2][n2]);
#pragma unroll
    for (int n1 = 0; n1 < 3; n1++) dft5<T, DIR>(a[n1]);
#pragma unroll
    for (int k1 = 0; k1 < 3; k1++)
#pragma unroll
        for (int k2 = 0; k2 < 5; k2++) x[(10 * k1 + 6 * k2) % 15] = a[k1][k2];
}

// the value of lane (l ^ 1) or (l ^ 2) of the quad (DPP quad_perm [1,0,3,2] =
// 0xB1, [2,3,0,1] = 0x4E); every lane of the quad must be active
template <int CTRL> __device__ __forceinline__ float qperm(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL> __device__ __forceinline__ double qperm(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL, class T> __device__ __forceinline__ cT<T> qperm_c(cT<T> v) {
    return cmk<T>(qperm<CTRL>(v.x), qperm<CTRL>(v.y));
}
__device__ __forceinline__ int rev2(int q) { return ((q & 1) << 1) | (q >> 1); }

// tw[t] = e^{+2 pi i t / 60}, t < 60 (LDS)
template <class T, int DIR> __device__ __forceinline__ void fft60_quad(cT<T>* y, int q, const cT<T>* tw) {
    fft15<T, DIR>(y);
#pragma unroll
    for (int k = 1; k < 15; k++) {
        cT<T> w = tw[q * k];
        if (DIR < 0) w.y = -w.y;
        y[k] = cmul<T>(y[k], w);
    }
    const bool hi = q & 2, lo = q & 1;
#pragma unroll
    for (int k = 0; k < 15; k++) {
        const cT<T> p = qperm_c<0x4E, T>(y[k]);
        y[k] = hi ? csub<T>(p, y[k]) : cadd<T>(y[k], p);
        if (q == 3) y[k] = irot<T, DIR>(y[k]);
    }
#pragma unroll
    for (int k = 0; k < 15; k++) {
        const cT<T> p = qperm_c<0xB1, T>(y[k]);
        y[k] = lo ? csub<T>(p, y[k]) : cadd<T>(y[k], p);
    }
}

// Forward (alpha, gamma) stage at 2B = 60, block (beta_j, XC omega points),
// one quad lane per thread in each stage:
//  A. rows g and g + 30 of the f tile as one complex vector z = f_g + i f_(g+30),
//     transformed over alpha (DIR +1); the packed spectrum Z_g goes to LDS;
//  B. per (xl, m < 30): H[m][g] = (Z_g[m] + conj Z_g[-m]) / 2 for g < 30 and
//     (Z_(g-30)[m] - conj Z_(g-30)[-m]) / (2i) for g >= 30 (the spectra of
//     the two real rows), transformed over gamma (DIR +1): G[m][n] for
//     n = -29..29, and G[-m][-n] = conj G[m][n].
template <class T, int XC, int NT>
__global__ __launch_bounds__(NT) void k_fwd_ag_q60(XformDev<T> X, const T* __restrict__ f, const T* __restrict__ wR,
                                                   int wstride, cT<T>* __restrict__ G, int nxc) {
    constexpr int B = 30, n2 = 60, lmax = 29, M = 59, rp = n2 + 1, ZS = n2;
    constexpr int TQ = XC * 30 * 4;
    static_assert(TQ <= NT, "one quad lane per thread");
    extern __shared__ char smem_raw[];
    cT<T>* tw = reinterpret_cast<cT<T>*>(smem_raw);
    T* sf = reinterpret_cast<T*>(smem_raw + 64 * sizeof(cT<T>));       // [xl][g][a], row stride rp
    cT<T>* sZ = reinterpret_cast<cT<T>*>(smem_raw + 64 * sizeof(cT<T>));  // [xl][g < 30][K], aliases sf
    int j;
    int64_t x0;
    ag_block<T, XC>(nxc, n2, j, x0);
    const int64_t nx = X.nx;
    const int tid = threadIdx.x;
    if (tid < n2) tw[tid] = X.twB[tid];
    stage_f_tile<T, XC, NT, n2>(f, wR, wstride, j, x0, nx, sf, rp);
    __syncthreads();
    // quads beyond TQ run on clamped in-range data and store nothing
    const int q = tid & 3, P = rev2(q);
    const bool act = tid < TQ;
    const int pr = act ? tid >> 2 : 0;
    cT<T> y[15];
    {
        const int xl = pr / 30, g = pr % 30;
        const T* r0 = sf + (xl * n2 + g) * rp + q;
        const T* r1 = r0 + 30 * rp;
#pragma unroll
        for (int t = 0; t < 15; t++) y[t] = cmk<T>(r0[4 * t], r1[4 * t]);
    }
    fft60_quad<T, 1>(y, q, tw);
    __syncthreads();  // sZ overwrites sf
    if (act) {
        cT<T>* zr = sZ + pr * ZS + 15 * P;
#pragma unroll
        for (int k = 0; k < 15; k++) zr[k] = y[k];
    }
    __syncthreads();
    // quads of neighbouring omega points take neighbouring lanes, so each G
    // store instruction writes runs of XC consecutive points
    const int xl = pr % XC, m = pr / XC, mneg = m == 0 ? 0 : n2 - m;
#pragma unroll
    for (int t = 0; t < 15; t++) {
        const int g = 4 * t + q, gr = g < 30 ? g : g - 30;
        const cT<T>* zr = sZ + (xl * 30 + gr) * ZS;
        const cT<T> a = zr[m], b = cconj<T>(zr[mneg]);
        const cT<T> s = g < 30 ? cadd<T>(a, b) : csub<T>(a, b);
        // (a + b) / 2, or (a - b) / (2i) = -i (a - b) / 2
        y[t] = g < 30 ? cmk<T>(T(0.5) * s.x, T(0.5) * s.y) : cmk<T>(T(0.5) * s.y, T(-0.5) * s.x);
    }
    fft60_quad<T, 1>(y, q, tw);
    const int64_t x = x0 + xl;
    if (act && x < nx) {
        cT<T>* Gx = G + g_off(B, 0, j, x);
#pragma unroll
        for (int k = 0; k < 15; k++) {
            const int K = k + 15 * P;  // frequency n = K or K - 60; the Nyquist bin K = 30 is dropped
            if (K == 30) continue;
            put_g_half<T>(Gx, B, m, K <= lmax ? K + lmax : K - 31, y[k]);
        }
    }
}

// Inverse (alpha, gamma) stage at 2B = 60, block (beta_j, XC omega points):
//  1. per (xl, mi < 59): v[K] = G[mi][ni] at K = ni - 29 (mod 60), v[30] = 0,
//     transformed over n (DIR -1): H[mi][g] to LDS;
//  2. per (xl, g): v[K] = H[mi][g] at K = mi - 29 (mod 60), v[30] = 0,
//     transformed over m (DIR -1); f[a][g] = Re v[a], staged in LDS and
//     written row by row (60 consecutive reals per (g, xl)).
template <class T, int XC, int NT>
__global__ __launch_bounds__(NT) void k_inv_ag_q60(XformDev<T> X, const cT<T>* __restrict__ G, T* __restrict__ f,
                                                   int nxc) {
    constexpr int B = 30, n2 = 60, M = 59, lmax = 29, HS = n2;
    constexpr int T1 = XC * B * 4, T2 = XC * n2 * 4;
    static_assert(T1 <= NT && T2 <= NT, "one quad lane per thread");
    extern __shared__ char smem_raw[];
    cT<T>* tw = reinterpret_cast<cT<T>*>(smem_raw);
    cT<T>* sH = reinterpret_cast<cT<T>*>(smem_raw + 64 * sizeof(cT<T>));  // [xl][m >= 0][g]
    T* so = reinterpret_cast<T*>(smem_raw + 64 * sizeof(cT<T>));          // [xl][g][a], aliases sH
    int j;
    int64_t x0;
    ag_block<T, XC>(nxc, n2, j, x0);
    const int64_t nx = X.nx;
    const int tid = threadIdx.x;
    if (tid < n2) tw[tid] = X.twB[tid];
    const int q = tid & 3, P = rev2(q);
    cT<T> y[15];
    // stage 1 on the rows m >= 0 of the Hermitian part of G (get_h); the
    // rows m < 0 of H are the conjugates
    const bool act1 = tid < T1;
    {
        const int pr = act1 ? tid >> 2 : 0;
        const int xl = pr % XC, m = pr / XC;
        const int64_t x = x0 + xl;
        const bool ld = x < nx;
        // clamped loads; a column beyond nx is zeroed after the load
        const cT<T>* Hx = G + h_base(B, j, min<int64_t>(x, nx - 1));
#pragma unroll
        for (int t = 0; t < 15; t++) {
            const int ni = (4 * t + q + 29) % n2;  // 59 at the Nyquist bin K = 30
            const cT<T> g = get_h<T>(Hx, B, lmax + m, ni < M ? ni : 0);
            y[t] = (ld && ni < M) ? g : czero<T>();
        }
    }
    __syncthreads();  // twiddles
    fft60_quad<T, -1>(y, q, tw);
    if (act1) {
        const int pr = tid >> 2, xl = pr % XC, m = pr / XC;
        cT<T>* hr = sH + (xl * B + m) * HS + 15 * P;
#pragma unroll
        for (int k = 0; k < 15; k++) hr[k] = y[k];
    }
    __syncthreads();
    const bool act2 = tid < T2;
    const int pr2 = act2 ? tid >> 2 : 0;
    const int xl2 = pr2 % XC, g2 = pr2 / XC;
#pragma unroll
    for (int t = 0; t < 15; t++) {
        const int mi = (4 * t + q + 29) % n2;  // 59 at the Nyquist bin
        const int m = mi - lmax;
        const cT<T> h = sH[(xl2 * B + (m >= 0 ? (mi < M ? m : 0) : -m)) * HS + g2];
        y[t] = mi < M ? (m >= 0 ? h : cconj<T>(h)) : czero<T>();
    }
    fft60_quad<T, -1>(y, q, tw);
    __syncthreads();  // so overwrites sH
    if (act2) {
        T* orow = so + (xl2 * n2 + g2) * n2 + 15 * P;
#pragma unroll
        for (int k = 0; k < 15; k++) orow[k] = y[k].x;
    }
    __syncthreads();
    store_f_rows<T, XC, NT, n2>(so, f, j, x0, nx);
}

// ---------------------------------------------------------- launchers -----

static int pick_xc(size_t per_x, size_t budget, int64_t nx) {
    int xc = (int)std::max<size_t>(1, budget / per_x);
    if (xc > 16) xc = 16;
    if (xc > nx) xc = (int)nx;
    return std::max(xc, 1);
}

// omega points per block and block size of the 2B = 30 FFT kernels (the
// inverse's XC doubles for FP32)
#ifndef HPUP_FFT_INV_XC
#define HPUP_FFT_INV_XC 4
#endif
#ifndef HPUP_FFT_FWD_XC
#define HPUP_FFT_FWD_XC 8
#endif
#ifndef HPUP_FFT_NT
#define HPUP_FFT_NT 256
#endif
constexpr int kFftInvXC = HPUP_FFT_INV_XC, kFftFwdXC = HPUP_FFT_FWD_XC, kFftNT = HPUP_FFT_NT;

// omega points per block of the register-row (alpha, gamma) kernels
constexpr int kFwdXC = 8;
template <class T> constexpr int inv_xc() { return sizeof(T) == 8 ? 4 : 8; }

template <class T, int B, int XC, int NT>
static void fwd_ag_fft(const XformDev<T>& X, const void* f, const void* wR, int wstride, void* G, hipStream_t st) {
    constexpr int n2 = 2 * B;
    static_assert(n2 * XC <= NT, "one alpha row per thread");
    const int nxc = (int)((X.nx + XC - 1) / XC);
    HPUP_REQUIRE((int64_t)nxc * n2 < (1ll << 31), HPUP_ERR_UNSUPPORTED, "omega grid too large");
    const size_t sh = std::max(sizeof(T) * XC * n2 * (n2 + 1), sizeof(cT<T>) * B * fft_sh<T>(XC, n2));
    allow_lds((const void*)k_fwd_ag_fft<T, B, XC, NT>, sh);
    hipLaunchKernelGGL((k_fwd_ag_fft<T, B, XC, NT>), dim3((unsigned)(nxc * n2)), dim3(NT), sh, st, X, (const T*)f,
                       (const T*)wR, wstride, (cT<T>*)G, nxc);
}

template <class T, int B, int XC, int NT>
static void inv_ag_fft(const XformDev<T>& X, const void* G, void* f, hipStream_t st) {
    constexpr int n2 = 2 * B, M = 2 * B - 1;
    const int nxc = (int)((X.nx + XC - 1) / XC);
    HPUP_REQUIRE((int64_t)nxc * n2 < (1ll << 31), HPUP_ERR_UNSUPPORTED, "omega grid too large");
    // rows m >= 0 of H, then the real output tile
    const size_t sh = std::max(sizeof(cT<T>) * B * fft_sh<T>(XC, n2), sizeof(T) * XC * n2 * n2);
    allow_lds((const void*)k_inv_ag_fft<T, B, XC, NT>, sh);
    hipLaunchKernelGGL((k_inv_ag_fft<T, B, XC, NT>), dim3((unsigned)(nxc * n2)), dim3(NT), sh, st, X,
                       (const cT<T>*)G, (T*)f, nxc);
}

// omega points per block of the quad-FFT kernels (their blocks hold one
// quad lane per thread: 120 XC forward, 240 XC inverse)
#ifndef HPUP_Q60_FWD_XC32
#define HPUP_Q60_FWD_XC32 4
#endif
#ifndef HPUP_Q60_FWD_XC64
#define HPUP_Q60_FWD_XC64 2
#endif
#ifndef HPUP_Q60_INV_XC32
#define HPUP_Q60_INV_XC32 2
#endif
#ifndef HPUP_Q60_INV_XC64
#define HPUP_Q60_INV_XC64 1
#endif
constexpr int kQ60FwdXC32 = HPUP_Q60_FWD_XC32, kQ60FwdXC64 = HPUP_Q60_FWD_XC64, kQ60InvXC32 = HPUP_Q60_INV_XC32,
              kQ60InvXC64 = HPUP_Q60_INV_XC64;

template <class T, int XC, int NT>
static void fwd_ag_q60(const XformDev<T>& X, const void* f, const void* wR, int wstride, void* G, hipStream_t st) {
    constexpr int n2 = 60;
    const int nxc = (int)((X.nx + XC - 1) / XC);
    HPUP_REQUIRE((int64_t)nxc * n2 < (1ll << 31), HPUP_ERR_UNSUPPORTED, "omega grid too large");
    const size_t sh = 64 * sizeof(cT<T>) + std::max(sizeof(T) * XC * n2 * (n2 + 1), sizeof(cT<T>) * XC * 30 * n2);
    allow_lds((const void*)k_fwd_ag_q60<T, XC, NT>, sh);
    hipLaunchKernelGGL((k_fwd_ag_q60<T, XC, NT>), dim3((unsigned)(nxc * n2)), dim3(NT), sh, st, X, (const T*)f,
                       (const T*)wR, wstride, (cT<T>*)G, nxc);
}

template <class T, int XC, int NT>
static void inv_ag_q60(const XformDev<T>& X, const void* G, void* f, hipStream_t st) {
    constexpr int n2 = 60, B = 30;
    const int nxc = (int)((X.nx + XC - 1) / XC);
    HPUP_REQUIRE((int64_t)nxc * n2 < (1ll << 31), HPUP_ERR_UNSUPPORTED, "omega grid too large");
    const size_t sh = 64 * sizeof(cT<T>) + std::max(sizeof(cT<T>) * XC * B * n2, sizeof(T) * XC * n2 * n2);
    allow_lds((const void*)k_inv_ag_q60<T, XC, NT>, sh);
    hipLaunchKernelGGL((k_inv_ag_q60<T, XC, NT>), dim3((unsigned)(nxc * n2)), dim3(NT), sh, st, X,
                       (const cT<T>*)G, (T*)f, nxc);
}

template <class T>
void launch_fwd_ag(const XformDev<T>& X, const void* f, const void* wR, int wstride, void* G, hipStream_t st) {
    const int n2 = 2 * X.B;
    if (n2 == 30) {
        fwd_ag_fft<T, 15, kFftFwdXC, kFftNT>(X, f, wR, wstride, G, st);
        return;
    }
    if (n2 == 60) {  // quad FFTs, kQ60* omega points per block
        if constexpr (sizeof(T) == 8)
            fwd_ag_q60<T, kQ60FwdXC64, (120 * kQ60FwdXC64 + 63) / 64 * 64>(X, f, wR, wstride, G, st);
        else
            fwd_ag_q60<T, kQ60FwdXC32, (120 * kQ60FwdXC32 + 63) / 64 * 64>(X, f, wR, wstride, G, st);
        return;
    }
    if (n2 <= kAgMax) {
        constexpr int XC = kFwdXC;
        const int nxc = (int)((X.nx + XC - 1) / XC);
        HPUP_REQUIRE((int64_t)nxc * n2 < (1ll << 31), HPUP_ERR_UNSUPPORTED, "omega grid too large");
        const size_t sh = std::max(sizeof(T) * XC * n2 * (n2 + 1), sizeof(cT<T>) * X.B * n2 * XC);
        allow_lds((const void*)k_fwd_ag_r<T, XC>, sh);
        hipLaunchKernelGGL((k_fwd_ag_r<T, XC>), dim3((unsigned)(nxc * n2)), dim3(256), sh, st, X, (const T*)f,
                           (const T*)wR, wstride, (cT<T>*)G, nxc);
        return;
    }
    const size_t per_x = sizeof(T) * n2 * n2 + sizeof(cT<T>) * n2 * X.B;
    const int XC = pick_xc(per_x, kLdsTile - sizeof(cT<T>) * n2, X.nx);
    const size_t sh = per_x * XC + sizeof(cT<T>) * n2;
    allow_lds((const void*)k_fwd_ag<T>, sh);
    dim3 grid((unsigned)((X.nx + XC - 1) / XC), (unsigned)n2);
    hipLaunchKernelGGL(k_fwd_ag<T>, grid, dim3(256), sh, st, X, (const T*)f, (const T*)wR, wstride, (cT<T>*)G, XC);
}

// MFMA beta stages for B <= 32 (row tiles and K steps fixed per bandlimit
// class); larger bandlimits keep the FMA kernels
template <class T, int RT, int KS>
static void beta_fwd_mfma(const XformDev<T>& X, const void* G, void* F, bool mirror, hipStream_t st) {
    const int M = 2 * X.B - 1;
    dim3 grid((unsigned)((X.nx + kGX - 1) / kGX), (unsigned)((M * M + 1) / 2));  // half of (m, n)
    if (mirror)
        hipLaunchKernelGGL((k_beta_fwd_mfma<T, RT, KS, true>), grid, dim3(256), 0, st, X, (const cT<T>*)G, (cT<T>*)F);
    else
        hipLaunchKernelGGL((k_beta_fwd_mfma<T, RT, KS, false>), grid, dim3(256), 0, st, X, (const cT<T>*)G, (cT<T>*)F);
}
template <class T, int RT, int KS>
static void beta_inv_mfma(const XformDev<T>& X, const void* F, void* G, bool sym, hipStream_t st) {
    // 16 omega points by 8 rows of the stored half of H per workgroup
    HPUP_REQUIRE(X.wis >= 4 * KS, HPUP_ERR_ARG, "inverse beta weight rows narrower than the K steps");
    dim3 grid((unsigned)((X.nx + 15) / 16), (unsigned)(h_rows(X.B) / 8));
    if (sym)
        hipLaunchKernelGGL((k_beta_inv_mfma<T, RT, KS, true>), grid, dim3(512), 0, st, X, (const cT<T>*)F, (cT<T>*)G);
    else
        hipLaunchKernelGGL((k_beta_inv_mfma<T, RT, KS, false>), grid, dim3(512), 0, st, X, (const cT<T>*)F, (cT<T>*)G);
}

template <class T>
void launch_beta_fwd(const XformDev<T>& X, const void* G, void* F, hipStream_t st, bool mirror) {
    static_assert(kGX == 64, "one workgroup of four 16-point tiles per G block");
    if (X.B <= 8) return beta_fwd_mfma<T, 1, 4>(X, G, F, mirror, st);     // l < 16, j < 16
    if (X.B <= 16) return beta_fwd_mfma<T, 1, 8>(X, G, F, mirror, st);    // l < 16, j < 32
    if (X.B <= 32) return beta_fwd_mfma<T, 2, 16>(X, G, F, mirror, st);   // l < 32, j < 64
    HPUP_REQUIRE(mirror, HPUP_ERR_UNSUPPORTED, "beta stage without mirror rows: bandlimit <= 32");
    const int M = 2 * X.B - 1;
    dim3 grid((unsigned)((X.nx + 255) / 256), (unsigned)((M * M + 1) / 2));  // half of (m, n), mirrored
    hipLaunchKernelGGL(k_beta_fwd<T>, grid, dim3(256), sizeof(T) * X.B * 2 * X.B, st, X, (const cT<T>*)G,
                       (cT<T>*)F);
}

template <class T>
void launch_beta_inv(const XformDev<T>& X, const void* F, void* G, hipStream_t st, bool sym) {
    if (X.B <= 8) return beta_inv_mfma<T, 1, 2>(X, F, G, sym, st);     // j < 16, l < 8
    if (X.B <= 16) return beta_inv_mfma<T, 2, 4>(X, F, G, sym, st);    // j < 32, l < 16
    if (X.B <= 32) return beta_inv_mfma<T, 4, 8>(X, F, G, sym, st);    // j < 64, l < 32
    HPUP_REQUIRE(sym, HPUP_ERR_UNSUPPORTED, "beta stage on Hermitian rows: bandlimit <= 32");
    const int M = 2 * X.B - 1;
    dim3 grid((unsigned)((X.nx + 255) / 256), (unsigned)((M * M + 1) / 2));  // the stored half of H
    hipLaunchKernelGGL(k_beta_inv<T>, grid, dim3(256), sizeof(T) * X.B * 2 * X.B, st, X, (const cT<T>*)F,
                       (cT<T>*)G);
}

template <class T>
void launch_inv_ag(const XformDev<T>& X, const void* G, void* f, hipStream_t st) {
    const int n2 = 2 * X.B, M = 2 * X.B - 1;
    if (n2 == 30) {
        inv_ag_fft<T, 15, kFftInvXC * (sizeof(T) == 8 ? 1 : 2), kFftNT>(X, G, f, st);
        return;
    }
    if (n2 == 60) {  // quad FFTs, kQ60* omega points per block
        if constexpr (sizeof(T) == 8) inv_ag_q60<T, kQ60InvXC64, (240 * kQ60InvXC64 + 63) / 64 * 64>(X, G, f, st);
        else inv_ag_q60<T, kQ60InvXC32, (240 * kQ60InvXC32 + 63) / 64 * 64>(X, G, f, st);
        return;
    }
    if (n2 <= kAgMax) {
        constexpr int XC = inv_xc<T>();
        const int nxc = (int)((X.nx + XC - 1) / XC);
        HPUP_REQUIRE((int64_t)nxc * n2 < (1ll << 31), HPUP_ERR_UNSUPPORTED, "omega grid too large");
        const size_t sh = sizeof(cT<T>) * M * n2 * XC;
        allow_lds((const void*)k_inv_ag_r<T, XC>, sh);
        hipLaunchKernelGGL((k_inv_ag_r<T, XC>), dim3((unsigned)(nxc * n2)), dim3(256), sh, st, X, (const cT<T>*)G,
                           (T*)f, nxc);
        return;
    }
    // gamma chunks of 16 for 2B > 32
    constexpr int GC = 16;
    const int ngc = (n2 + GC - 1) / GC;
    const size_t per_x = sizeof(cT<T>) * ((size_t)M * n2);
    int XC = std::max(1, 256 / (M * ngc));
    XC = std::min<int>(XC, pick_xc(per_x, kLdsTile - sizeof(cT<T>) * n2, X.nx));
    const size_t sh = per_x * XC + sizeof(cT<T>) * n2;
    dim3 grid((unsigned)((X.nx + XC - 1) / XC), (unsigned)n2);
    allow_lds((const void*)k_inv_ag<T, GC>, sh);
    hipLaunchKernelGGL((k_inv_ag<T, GC>), grid, dim3(256), sh, st, X, (const cT<T>*)G, (T*)f, XC);
}

// out-of-place omega DFT: on the chirp path the first pass reads in and
// writes out (the k3 pass of a 3-D grid, then the (k1, k2) pass in place on
// out: DFTs along different axes commute); otherwise a copy and the in-place
// transform
template <class T>
void launch_omega_dft_to(const XformDev<T>& X, const void* in, void* out, bool inverse, hipStream_t st) {
    const int N = X.N, NN = N * N;
    if ((X.dim == 2 || X.dim == 3) && N <= 32 && X.bsF && in != out) {
        const cT<T>* bs = inverse ? X.bsI : X.bsF;
        const T s2 = inverse ? T(1) / T(NN) : T(1);
        if (X.dim == 3) {
            launch_axis_dft_mfma<T>(in, out, N, (int64_t)NN, X.nrc, bs, inverse ? T(1) / T(N) : T(1), st);
            launch_slab_dft_mfma<T>(out, out, X.nrc * N, N, bs, s2, st);
        } else {
            launch_slab_dft_mfma<T>(in, out, X.nrc, N, bs, s2, st);
        }
        return;
    }
    if (in != out)
        HPUP_HIP(hipMemcpyAsync(out, in, sizeof(cT<T>) * (size_t)(X.nrc * X.nx), hipMemcpyDeviceToDevice, st));
    launch_omega_dft<T>(X, out, nullptr, inverse, st);
}

template <class T>
void launch_omega_dft(const XformDev<T>& X, void* F, void* tmp, bool inverse, hipStream_t st) {
    (void)tmp;
    if (X.dim == 0) return;
    const int N = X.N, NN = N * N;
    HPUP_REQUIRE(NN <= 4096, HPUP_ERR_UNSUPPORTED, "omega grid larger than 64 per axis");
    if (N <= 32 && X.bsF) {  // chirp-Toeplitz products on MFMA
        const cT<T>* bs = inverse ? X.bsI : X.bsF;
        const int64_t nslab = X.nrc * (X.dim == 3 ? N : 1);
        launch_slab_dft_mfma<T>(F, F, nslab, N, bs, inverse ? T(1) / T(NN) : T(1), st);
        if (X.dim == 3) launch_axis_dft_mfma<T>(F, F, N, (int64_t)NN, X.nrc, bs, inverse ? T(1) / T(N) : T(1), st);
        return;
    }
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
    template void launch_beta_fwd<T>(const XformDev<T>&, const void*, void*, hipStream_t, bool);           \
    template void launch_beta_inv<T>(const XformDev<T>&, const void*, void*, hipStream_t, bool);           \
    template void launch_inv_ag<T>(const XformDev<T>&, const void*, void*, hipStream_t);             \
    template void launch_omega_dft<T>(const XformDev<T>&, void*, void*, bool, hipStream_t);          \
    template void launch_omega_dft_to<T>(const XformDev<T>&, const void*, void*, bool, hipStream_t); \
    template void launch_full_to_compact<T>(int, int64_t, const void*, void*, hipStream_t);          \
    template void launch_compact_to_full<T>(int, int64_t, const void*, void*, hipStream_t);
HPUP_INST(double)
HPUP_INST(float)
#undef HPUP_INST

}  // namespace hpup
