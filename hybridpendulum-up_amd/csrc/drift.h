// Device views of the structured drift operator and the launchers of its kernels.
#pragma once

#include "hpup_internal.h"
#include "kernels_common.h"

namespace hpup {

// LDS bytes one workgroup may allocate on the current device (160 KB on
// gfx950); kernels size their tiles to kLdsTile and only the large-grid
// fallbacks go beyond it.
constexpr size_t kLdsTile = 64 * 1024;
size_t lds_limit();
// Raises a kernel's dynamic-LDS cap when one launch needs more than 64 KB.
void allow_lds(const void* kernel, size_t bytes);

template <class T> struct DriftDev {
    int B, dim, N;
    int64_t nx, nrc;
    int TG, TU;
    const int32_t* cg_src;   // nrc*TG
    const cT<T>* cg_coef;    // nrc*TG*dim
    const int32_t* u_src;    // 2*nrc*TU
    const cT<T>* u_coef;     // 2*nrc*TU
    const cT<T>* e3;         // nrc
    const cT<T>* kap;        // dim*N circulant kernels
    const T* kc;             // N: 2 pi k'/L, Nyquist -> 0
    const T* kd;             // N: -(2 pi k/L)^2
    T G[9];
    T b[3];
    T gyro[3];
    int has_gyro;
    // density-form gravity: transforms of f mR_p on the coefficient layout
    // (nullptr when the gravity term comes from the Clebsch-Gordan table)
    const cT<T>* grav[3];
    // every kap_p = c0[p] + i (real circulant): the real-circulant path
    int ramp;
    T c0[3];
    // structured mixing coefficients: -i CG_1 real, -i CG_2 imaginary,
    // CG_3 zero, u_1 imaginary, u_2 real (the pendulum models)
    int smix;
    // the adjoint operator (transposed tables, damping through the V planes)
    int adj;
    // 3-D adjoint: T3^H [(e3 - i b3 kc(k3)) F] from the scaled axis-3 pass
    // (e3 of the adjoint tables is the conjugate of the forward one)
    const cT<T>* s3;
    // gradient stage (adj only, gpart non-null): forward stage value y, the
    // sweep's running sum, the workgroups' partial sums (kGradMaxWG x 5), and
    // out[5] += gscale * sums
    const void* gy;
    void* gacc;      // nullable: gacc += the stage output
    int gacc_init;   // gacc = the stage input + the stage output instead (the sweep's first stage)
    double* gpart;
    double* gout;
    double gscale;
};
constexpr int kGradMaxWG = 4096;
constexpr int kMaxDevices = 64;
// Device address of the gradient stage's static partial-sum array (kGradMaxWG
// x 5 doubles) on `device`, which must be current; resolved once per device.
double* grad_scratch(int device);

// Integrator epilogue of the slab kernel:
//   Fout = F0 + dt (ca ACC + cd dF);  ACC <- ea ACC + ed dF
template <class T> struct StageCoef {
    T dt, ca, cd, ea, ed;
    int read_f0, read_acc, write_acc, deriv_only;
};

// Rotation-mixing sources per output row r for the fused kernel: up to TS
// distinct source rows; per source dim + 2 complex coefficients
// [-i CG_1 .. -i CG_dim, -u_1, -u_2].  An entry of src holds the source row
// in bits 0-23 and which of its terms are nonzero in bits 24 (Clebsch-Gordan),
// 25 (u_1) and 26 (u_2); -1 pads.
template <class T> struct MixDev {
    int TS;
    const int32_t* src;   // nrc*TS, -1 padded
    const cT<T>* coef;    // nrc*TS*(dim+2)
};

// bf16x3: circulants as 3xBF16 split products (FP32 storage only)
template <class T>
void launch_drift_fused(const DriftDev<T>& D, const MixDev<T>& MX, const void* Fin, const void* Z3, const void* F0,
                        void* Fout, void* ACC, const StageCoef<T>& sc, bool bf16x3, hipStream_t st);
// name of the kernel launch_drift_fused runs for D (grav: density-form gravity)
template <class T> const char* drift_fused_name(const DriftDev<T>& D, bool grav, bool bf16x3);
// out = T3 in; scaled: out = T3 [(e3[r] - i b3 kc(k3)) in] (the 3-D adjoint's
// axis-3 term, on an adjoint plan's tables)
template <class T>
void launch_axis3_mfma(const DriftDev<T>& D, const void* in, void* out, bool bf16x3, hipStream_t st,
                       bool scaled = false);

// Omega DFTs as chirp-Toeplitz MFMA products (N <= 32).  bs: 96 entries, the
// Toeplitz kernel c_d at [d + 32] for d in [-32, 32) and then the chirp p_k
// (zero for k >= N); see k_slab_dft_mfma.
template <class T>
void launch_slab_dft_mfma(const void* in, void* out, int64_t nslab, int N, const void* bs, T scale,
                          hipStream_t st);
template <class T>
void launch_axis_dft_mfma(const void* in, void* out, int N, int64_t inner, int64_t nlines, const void* bs, T scale,
                          hipStream_t st);

template <class T>
void launch_rotmix(const DriftDev<T>& D, const void* F, void* W, void* V1, void* V2, hipStream_t st);
// out[r][x] += scale sum_t coef[r][t] in[src[r][t]][x] over rows of nx points (src -1 padded, TS per row)
template <class T>
void launch_row_mix_add(const void* in, void* out, const int32_t* src, const void* coef, int TS, int64_t nrows,
                        int64_t nx, T scale, hipStream_t st);
template <class T>
void launch_axis_outer(const void* in, void* out, const void* Mat, int N, int64_t inner, int64_t nlines,
                       T scale, hipStream_t st);
template <class T>
void launch_drift_slab(const DriftDev<T>& D, const void* Fin, const void* Z3, const void* W, const void* V1,
                       const void* V2, const void* F0, void* Fout, void* ACC, const StageCoef<T>& sc,
                       hipStream_t st);

// ---------------------------------------------------------- transforms --

template <class T> struct XformDev {
    int B, dim, N;
    int64_t nx, nrc;
    const T* Wf;        // forward beta weights  [mn][l][j] = w_j d^l_mn(beta_j), rows of wfs
    const T* Wi;        // inverse beta weights  [mn][j][l] = (2l+1) d^l_mn(beta_j), rows of wis
    int wfs, wis;       // row strides (zero-padded to whole reduction groups)
    const cT<T>* twB;   // e^{i pi t / B}, t in [0, 2B)
    const cT<T>* twMB;  // [32][32] e^{i pi (r - (B-1)) t / B} for r < 2B-1, t < 2B, zero elsewhere, then its
                        // transpose [t][r] (B <= 16)
    const cT<T>* dftN;  // forward DFT matrix N x N (row-major [k][a]) e^{-2 pi i k a/N}
    const cT<T>* idftN; // inverse DFT matrix (1/N scaled) e^{+2 pi i k a/N}/N
    const cT<T>* twNf;  // e^{-2 pi i t/N}, t in [0, N)
    const cT<T>* twNi;  // e^{+2 pi i t/N}
    const cT<T>* bsF;   // chirp-Toeplitz tables of the forward / inverse omega DFT (N <= 32),
    const cT<T>* bsI;   // 96 entries each (launch_slab_dft_mfma)
    // the rows (l, m, n) with n > 0, or n = 0 and m >= 0: pairs (row, row of
    // (l, -m, -n) | bit 30 when (-1)^(m-n) = -1), nhalf of them
    const int32_t* hrows;
    int64_t nhalf;
};

// Elements of the angular/beta intermediate: the forward's G, (2B-1)^2 2B rows
// of the Omega points rounded up to whole blocks of kGX (its layout is
// Omega-blocked), or the inverse's H, pairs of Omega points by 2B by the
// stored half of (m, n) padded to 8 rows, whichever is larger
constexpr int kGX = 64;
inline int64_t g_elems(int B, int64_t nx) {
    const int64_t M = 2 * B - 1, n2 = 2 * B;
    const int64_t g = M * M * n2 * ((nx + kGX - 1) / kGX) * kGX;
    const int64_t h = (nx + 1) / 2 * 2 * n2 * (((M * M + 1) / 2 + 7) / 8 * 8);
    return g > h ? g : h;
}
template <class T>
// wR (nullable): transform the product f(R, x) wR[R * wstride] (grid weight)
void launch_fwd_ag(const XformDev<T>& X, const void* f, const void* wR, int wstride, void* G, hipStream_t st);
// mirror = false: only the rows with n > 0, or n = 0 and m >= 0 are written
// (for launch_omega_half, which writes the others); bandlimit <= 32
template <class T>
void launch_beta_fwd(const XformDev<T>& X, const void* G, void* F, hipStream_t st, bool mirror = true);
// sym = false: F's rows already hold the Hermitian part (launch_omega_half's
// inverse), so only the rows (l, m, n) of the stored half are read
template <class T>
void launch_beta_inv(const XformDev<T>& X, const void* F, void* G, hipStream_t st, bool sym = true);
template <class T>
void launch_inv_ag(const XformDev<T>& X, const void* G, void* f, hipStream_t st);
// in-place omega DFT of the compact rows (forward sign -1, inverse +1 with 1/nx)
template <class T>
void launch_omega_dft(const XformDev<T>& X, void* F, void* tmp, bool inverse, hipStream_t st);
// the same from in to out (out = in allowed)
template <class T>
void launch_omega_dft_to(const XformDev<T>& X, const void* in, void* out, bool inverse, hipStream_t st);
// Omega DFTs of the coefficients of a real grid on the half rows only (N <= 32).
// Forward (in place): the input holds the mirror relation F_{-m,-n} = (-1)^(m-n)
// conj F_mn on its half rows' partners implicitly (only the half rows are read)
// and every row is written.  Inverse (out of place): out's half rows get the
// inverse DFT of the Hermitian part (F_mn(k) + (-1)^(m-n) conj F_{-m,-n}(-k)) / 2;
// the other rows of out are not written.
template <class T>
void launch_omega_half(const XformDev<T>& X, const void* in, void* out, bool inverse, hipStream_t st);

// layout conversion full (M,M,B,nx) <-> compact [r][x]
template <class T>
void launch_full_to_compact(int B, int64_t nx, const void* full, void* compact, hipStream_t st);
template <class T>
void launch_compact_to_full(int B, int64_t nx, const void* compact, void* full, hipStream_t st);

}  // namespace hpup
