// C ABI of libhpup: the plan (device-resident) API and the gateway
// equivalents of the reference's MEX entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <strings.h>

#include "drift.h"
#include "getc.h"
#include "hpup_internal.h"
#include "matfisher.h"
#include "reset_plan.h"
#include "stats.h"

namespace hpup {

static thread_local std::string g_last_error;
void set_error(const std::string& m) { g_last_error = m; }

size_t lds_limit() {
    int dev = 0, v = 0;
    HPUP_HIP(hipGetDevice(&dev));
    HPUP_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return (size_t)v;
}

void allow_lds(const void* kernel, size_t bytes) {
    if (bytes <= kLdsTile) return;
    HPUP_REQUIRE(bytes <= lds_limit(), HPUP_ERR_UNSUPPORTED,
                 "kernel needs " + std::to_string(bytes) + " bytes of LDS, device allows " + std::to_string(lds_limit()));
    HPUP_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

using cd = std::complex<double>;
static const double kPi = 3.14159265358979323846;

// --------------------------------------------------------- table upload --

template <class T> std::vector<T> cast_vec(const std::vector<double>& v) { return std::vector<T>(v.begin(), v.end()); }

// Transform tables from a Wigner-d table d (M,M,B,2B) and weights w (2B).
template <class T> struct XformTables {
    DevBuf Wf, Wi, twB, twMB, dftN, idftN, twNf, twNi, bsF, bsI, hrows;
    XformDev<T> view{};
    void build(int B, int dim, int N, const double* d, const double* w) {
        const int M = 2 * B - 1, n2 = 2 * B;
        // row strides of the weight tables: whole K groups of the MFMA beta
        // kernels' reduction (zero beyond the data), the row length beyond B = 32
        const int wfs = B <= 8 ? 16 : B <= 16 ? 32 : B <= 32 ? 64 : n2;
        const int wis = B <= 8 ? 8 : B <= 16 ? 16 : B <= 32 ? 32 : B;
        std::vector<double> wf((size_t)M * M * B * wfs, 0.0), wi((size_t)M * M * n2 * wis, 0.0);
        for (int mn = 0; mn < M * M; mn++)
            for (int l = 0; l < B; l++)
                for (int j = 0; j < n2; j++) {
                    const int mi = mn % M, ni = mn / M;
                    const double dv = d[mi + (size_t)M * (ni + (size_t)M * (l + (size_t)B * j))];
                    wf[((size_t)mn * B + l) * wfs + j] = w[j] * dv;
                    wi[((size_t)mn * n2 + j) * wis + l] = (2.0 * l + 1.0) * dv;
                }
        view.wfs = wfs;
        view.wis = wis;
        Wf = upload(cast_vec<T>(wf));
        Wi = upload(cast_vec<T>(wi));
        std::vector<T> tb(2 * n2), dn, idn, tf, ti;
        for (int t = 0; t < n2; t++) { tb[2 * t] = (T)std::cos(kPi * t / B); tb[2 * t + 1] = (T)std::sin(kPi * t / B); }
        twB = upload(tb);
        if (n2 <= 32) {  // [r][t] then [t][r]: frequency r - (B-1), grid angle pi t / B, zero padded to 32 x 32
            std::vector<T> tm(2 * 2 * 32 * 32, T(0));
            for (int r = 0; r < M; r++)
                for (int t = 0; t < n2; t++) {
                    const int e = (((r - (B - 1)) * t) % n2 + n2) % n2;
                    for (int c = 0; c < 2; c++) {
                        tm[2 * (r * 32 + t) + c] = tb[2 * e + c];
                        tm[2 * (32 * 32 + t * 32 + r) + c] = tb[2 * e + c];
                    }
                }
            twMB = upload(tm);
        }
        const int Nn = std::max(N, 1);
        dn.resize(2 * (size_t)Nn * Nn); idn.resize(2 * (size_t)Nn * Nn); tf.resize(2 * Nn); ti.resize(2 * Nn);
        for (int k = 0; k < Nn; k++) {
            for (int a = 0; a < Nn; a++) {
                const double ang = 2.0 * kPi * (double)((int64_t)k * a % Nn) / Nn;
                dn[2 * ((size_t)k * Nn + a)] = (T)std::cos(ang);
                dn[2 * ((size_t)k * Nn + a) + 1] = (T)-std::sin(ang);
                idn[2 * ((size_t)k * Nn + a)] = (T)(std::cos(ang) / Nn);
                idn[2 * ((size_t)k * Nn + a) + 1] = (T)(std::sin(ang) / Nn);
            }
            const double ang = 2.0 * kPi * k / Nn;
            tf[2 * k] = (T)std::cos(ang); tf[2 * k + 1] = (T)-std::sin(ang);
            ti[2 * k] = (T)std::cos(ang); ti[2 * k + 1] = (T)std::sin(ang);
        }
        dftN = upload(dn); idftN = upload(idn); twNf = upload(tf); twNi = upload(ti);
        if (Nn <= 32) {  // chirp-Toeplitz tables: c_d = e^{+i pi d^2/N} (d in [-32, 32)), p_k = e^{-i pi k^2/N}
            std::vector<T> bf(2 * 96, T(0)), bi(2 * 96, T(0));
            auto ang = [&](int64_t q) { return kPi * (double)(((q % (2 * Nn)) + 2 * Nn) % (2 * Nn)) / Nn; };
            for (int d = -32; d < 32; d++) {
                const double a = ang((int64_t)d * d);
                bf[2 * (d + 32)] = (T)std::cos(a); bf[2 * (d + 32) + 1] = (T)std::sin(a);
                bi[2 * (d + 32)] = (T)std::cos(a); bi[2 * (d + 32) + 1] = (T)-std::sin(a);
            }
            for (int k = 0; k < Nn; k++) {
                const double a = ang((int64_t)k * k);
                bf[2 * (64 + k)] = (T)std::cos(a); bf[2 * (64 + k) + 1] = (T)-std::sin(a);
                bi[2 * (64 + k)] = (T)std::cos(a); bi[2 * (64 + k) + 1] = (T)std::sin(a);
            }
            bsF = upload(bf); bsI = upload(bi);
        }
        // the half rows of the omega DFTs: (l, m, n) with n > 0, or n = 0 and
        // m >= 0, each with the row of (l, -m, -n) (bit 30: (-1)^(m-n) = -1)
        std::vector<int32_t> hr;
        for (int l = 0; l < B; l++)
            for (int n = 0; n <= l; n++)
                for (int m = (n == 0 ? 0 : -l); m <= l; m++)
                    hr.insert(hr.end(), {(int32_t)ridx_B(B, l, m, n),
                                         (int32_t)ridx_B(B, l, -m, -n) | (((m - n) & 1) ? (1 << 30) : 0)});
        hrows = upload(hr);
        view.hrows = hrows.as<int32_t>();
        view.nhalf = (int64_t)hr.size() / 2;
        view.B = B; view.dim = dim; view.N = Nn;
        view.nx = ipow64(Nn, dim); view.nrc = nrc_of(B);
        view.Wf = Wf.as<T>(); view.Wi = Wi.as<T>(); view.twB = twB.as<cT<T>>();
        view.twMB = n2 <= 32 ? twMB.as<cT<T>>() : nullptr;
        view.dftN = dftN.as<cT<T>>(); view.idftN = idftN.as<cT<T>>();
        view.twNf = twNf.as<cT<T>>(); view.twNi = twNi.as<cT<T>>();
        view.bsF = Nn <= 32 ? bsF.as<cT<T>>() : nullptr;
        view.bsI = Nn <= 32 ? bsI.as<cT<T>>() : nullptr;
    }
};

template <class T> struct DriftTables {
    DevBuf cg_src, cg_coef, u_src, u_coef, e3, kap, kc, kd, M3, mix_src, mix_coef;
    DriftDev<T> view{};
    MixDev<T> mix{};
    bool smix = false;
    // Merges the Clebsch-Gordan and u_1/u_2 term lists of each row by source row.
    void build_mix(const DriftSpec& s) {
        const int64_t nrc = nrc_of(s.B);
        const int NC = s.dim + 2;
        std::vector<std::vector<std::pair<int32_t, std::vector<double>>>> rows(nrc);
        auto slot = [&](int64_t r, int32_t src) -> std::vector<double>& {
            for (auto& e : rows[r])
                if (e.first == src) return e.second;
            rows[r].push_back({src, std::vector<double>(2 * NC, 0.0)});
            return rows[r].back().second;
        };
        for (int64_t r = 0; r < nrc; r++) {
            for (int t = 0; t < s.TG; t++) {
                const int32_t src = s.cg_src[r * s.TG + t];
                if (src < 0) break;
                auto& c = slot(r, src);
                for (int p = 0; p < s.dim; p++) {  // stored as -i CG_p
                    c[2 * p] += s.cg_coef[((r * s.TG + t) * s.dim + p) * 2 + 1];
                    c[2 * p + 1] -= s.cg_coef[((r * s.TG + t) * s.dim + p) * 2];
                }
            }
            for (int q = 0; q < 2; q++)
                for (int t = 0; t < s.TU; t++) {
                    const int64_t i = (q * nrc + r) * s.TU + t;
                    const int32_t src = s.u_src[i];
                    if (src < 0) break;
                    auto& c = slot(r, src);
                    c[2 * (s.dim + q)] += s.u_coef[2 * i];
                    c[2 * (s.dim + q) + 1] += s.u_coef[2 * i + 1];
                }
        }
        int TS = 1;
        for (auto& rw : rows) TS = std::max<int>(TS, (int)rw.size());
        HPUP_REQUIRE(nrc < ((int64_t)1 << 24), HPUP_ERR_UNSUPPORTED, "fused mixing table: more than 2^24 rotation rows");
        std::vector<int32_t> src(nrc * TS, -1);
        std::vector<double> coef(nrc * TS * NC * 2, 0.0);
        for (int64_t r = 0; r < nrc; r++)
            for (size_t t = 0; t < rows[r].size(); t++) {
                const std::vector<double>& c = rows[r][t].second;
                auto nz = [&](int k) { return c[2 * k] != 0.0 || c[2 * k + 1] != 0.0; };
                int fl = 0;
                for (int p = 0; p < s.dim; p++) fl |= nz(p) ? 1 : 0;
                fl |= nz(s.dim) ? 2 : 0;
                fl |= nz(s.dim + 1) ? 4 : 0;
                src[r * TS + t] = rows[r][t].first | (fl << 24);
                std::copy(c.begin(), c.end(), coef.begin() + (r * TS + t) * NC * 2);
            }
        // Structured mixing: in the pendulum models -i CG_1 is real, -i CG_2
        // imaginary, CG_3 zero (a diagonal inertia), u_1 imaginary and u_2 real
        // (to rounding of the transforms that produce them); the kernels then
        // multiply by the one nonzero component
        {
            double mx[5] = {0, 0, 0, 0, 0}, off[5] = {0, 0, 0, 0, 0};
            const int part[5] = {1, 0, -1, 0, 1};  // the component that must vanish (-1: both)
            for (size_t e = 0; e < coef.size() / 2; e++) {
                const int k = (int)(e % NC), cls = k < s.dim ? k : 3 + (k - s.dim);
                const double re = coef[2 * e], im = coef[2 * e + 1];
                mx[cls] = std::max(mx[cls], std::hypot(re, im));
                const double z = part[cls] == 0 ? std::abs(re) : part[cls] == 1 ? std::abs(im) : std::hypot(re, im);
                off[cls] = std::max(off[cls], z);
            }
            double gmax = 0.0;
            for (int c = 0; c < 5; c++) gmax = std::max(gmax, mx[c]);
            smix = true;
            for (int c = 0; c < 5; c++)
                if (off[c] > 1e-13 * gmax) smix = false;
        }
        mix_src = upload(src);
        mix_coef = upload(cast_vec<T>(coef));
        mix.TS = TS;
        mix.src = mix_src.as<int32_t>();
        mix.coef = mix_coef.as<cT<T>>();
    }
    void build(const DriftSpec& s) {
        const int N = s.N;
        build_mix(s);
        cg_src = upload(s.cg_src);
        cg_coef = upload(cast_vec<T>(s.cg_coef));
        u_src = upload(s.u_src);
        u_coef = upload(cast_vec<T>(s.u_coef));
        e3 = upload(cast_vec<T>(s.e3));
        kap = upload(cast_vec<T>(s.kappa));
        std::vector<double> kcv(N), kdv(N);
        const int Bx = N / 2;
        for (int k = 0; k < N; k++) {
            const double kk = k < Bx ? k : k - N;
            const double kp = k < Bx ? k : (k == Bx ? 0 : k - N);
            kcv[k] = 2.0 * kPi * kp / s.L;
            kdv[k] = -(2.0 * kPi * kk / s.L) * (2.0 * kPi * kk / s.L);
        }
        kc = upload(cast_vec<T>(kcv));
        kd = upload(cast_vec<T>(kdv));
        if (s.dim == 3) {
            std::vector<double> m3(2 * (size_t)N * N);
            for (int k = 0; k < N; k++)
                for (int a = 0; a < N; a++) {
                    const int j = ((k - a) % N + N) % N;
                    m3[2 * ((size_t)k * N + a)] = s.kappa[2 * (2 * N + j)];
                    m3[2 * ((size_t)k * N + a) + 1] = s.kappa[2 * (2 * N + j) + 1];
                }
            M3 = upload(cast_vec<T>(m3));
        }
        DriftDev<T>& v = view;
        v.B = s.B; v.dim = s.dim; v.N = N;
        v.nx = ipow64(N, s.dim); v.nrc = nrc_of(s.B);
        v.TG = s.TG; v.TU = s.TU;
        v.cg_src = cg_src.as<int32_t>(); v.cg_coef = cg_coef.as<cT<T>>();
        v.u_src = u_src.as<int32_t>(); v.u_coef = u_coef.as<cT<T>>();
        v.e3 = e3.as<cT<T>>(); v.kap = kap.as<cT<T>>();
        v.kc = kc.as<T>(); v.kd = kd.as<T>();
        for (int i = 0; i < 9; i++) v.G[i] = (T)s.G[i];
        for (int i = 0; i < 3; i++) { v.b[i] = (T)s.b[i]; v.gyro[i] = (T)s.gyro[i]; v.c0[i] = (T)s.c0[i]; }
        v.ramp = s.ramp ? 1 : 0;
        v.smix = smix ? 1 : 0;
        v.adj = s.adj ? 1 : 0;
        v.has_gyro = s.has_gyro ? 1 : 0;
        v.grav[0] = v.grav[1] = v.grav[2] = nullptr;
        v.s3 = nullptr;
    }
};

// ---------------------------------------------------------------- plan --

// One device pass of get_stat over the grid f (precision T): fR[nR],
// gR[dim][nR], fx[nx] copied to the host.  work/out grow as needed.
template <class T>
static void grid_marginals(int B, int dim, int64_t nx, const void* f, const double* w_dev, const double* xg_dev,
                           DevBuf& work, DevBuf& out, std::vector<double>& hR, std::vector<double>& hG,
                           std::vector<double>& hx, hipStream_t st) {
    const StatShape s = stat_shape(B, dim, nx);
    HPUP_REQUIRE(dim == 0 || xg_dev, HPUP_ERR_ARG, "omega coordinates missing");
    const size_t nout = (size_t)(s.nR * (1 + dim) + nx);
    if (work.bytes < sizeof(double) * s.work_doubles) work.alloc(sizeof(double) * s.work_doubles);
    if (out.bytes < sizeof(double) * nout) out.alloc(sizeof(double) * nout);
    double* o = out.as<double>();
    launch_stat_pass<T>(s, f, B, w_dev, xg_dev, work.as<double>(), o, o + s.nR, o + s.nR * (1 + dim), st);
    HPUP_HIP(hipGetLastError());
    hR.resize(s.nR);
    hG.resize((size_t)dim * s.nR);
    hx.resize(nx);
    HPUP_HIP(hipMemcpyAsync(hR.data(), o, sizeof(double) * s.nR, hipMemcpyDeviceToHost, st));
    if (dim) HPUP_HIP(hipMemcpyAsync(hG.data(), o + s.nR, sizeof(double) * dim * s.nR, hipMemcpyDeviceToHost, st));
    HPUP_HIP(hipMemcpyAsync(hx.data(), o + s.nR * (1 + dim), sizeof(double) * nx, hipMemcpyDeviceToHost, st));
    HPUP_HIP(hipStreamSynchronize(st));
}

struct PlanBase {
    hpup_model model{};
    int64_t nx = 1, nrc = 0, ncoef = 0, ngrid = 0;
    std::vector<double> w;  // beta weights (FP64)
    DevBuf w_dev;
    virtual ~PlanBase() = default;
    virtual void forward(const void* f, void* F, hipStream_t st) = 0;
    virtual void backward(const void* F, void* f, hipStream_t st) = 0;
    virtual void step(void* F, int nsteps, hipStream_t st) = 0;
    virtual void derivative(const void* F, void* dF, hipStream_t st) = 0;
    virtual void adjoint_stage(const void* ybar, const void* y, double scale, void* out, void* fbar,
                               int fbar_init, double* gout, hipStream_t st, bool noalloc = false) = 0;
    virtual void axpy_stage(const void* in, const void* F0, double scale, void* out, hipStream_t st) = 0;
    virtual void step_stages(const void* F, void* const* ys, void* Fnext, hipStream_t st) = 0;
    virtual void adjoint_step(const void* F, const void* const* ys, const void* lam, void* Fbar, double* gout,
                              hipStream_t st) = 0;
    virtual void full_to_compact(const void* a, void* b, hipStream_t st) = 0;
    virtual void compact_to_full(const void* a, void* b, hipStream_t st) = 0;
    // device pass of get_stat: rotation marginal, first omega moment per
    // rotation, omega marginal (host copies)
    virtual void marginals(const void* f, const double* xg, std::vector<double>& fR, std::vector<double>& gR,
                           std::vector<double>& fx, hipStream_t st) = 0;
    std::vector<double> xg_host;  // omega coordinates [ix][a]
    DevBuf xg_dev;
    void get_stat(const void* f, hpup_stat& out, double* fx, bool fit, hipStream_t st);
    virtual int profile(void* F, int nreps, double* ms, int maxk, hipStream_t st) = 0;
    virtual const char* profile_name(int i) const = 0;
    virtual double profile_bytes(int i) const = 0;
    // hybrid collision reset on the grid (set_reset builds its tables)
    std::unique_ptr<ResetGeom> reset;
    void set_reset(const hpup_reset_params& prm) {
        auto g = std::make_unique<ResetGeom>();
        g->build(model.B, model.dim, model.N, model.L, prm);
        reset = std::move(g);
    }
    virtual void reset_grid(const void* fin, void* out, double scale, int mode, hipStream_t st) = 0;
    virtual void reset_adjoint(const void* g, void* out, double scale, hipStream_t st) = 0;
    virtual void reset_eps_grad(const void* f, const void* g, double scale, double* out, hipStream_t st) = 0;
    virtual void hybrid_step(void* F, void* f, int nsteps, hipStream_t st) = 0;
};

template <class T> struct PlanT : PlanBase {
    XformTables<T> xf;
    DriftTables<T> dr;
    bool has_drift = false;
    bool bf16x3 = false;  // circulants as 3xBF16 split products (FP32 storage)
    DevBuf G, Wb, V1, V2, Z3, S3, Fs, Fs2, ACC, tmp, st_work, st_out, gpart, dfc, epart, Ya, Yb;

    void init_xform(const double* d, const double* wv) {
        const hpup_model& m = model;
        xf.build(m.B, m.dim, m.N, d, wv);
    }
    void init_drift(const DriftSpec& s) {
        dr.build(s);
        has_drift = true;
    }
    void ensure(DevBuf& b, size_t bytes) {
        if (b.bytes < bytes) b.alloc(bytes);
    }
    size_t cbytes() const { return sizeof(cT<T>) * (size_t)ncoef; }

    void forward(const void* f, void* F, hipStream_t st) override { forward_weighted(f, nullptr, 0, F, st); }
    // forward transform of the grid product f(R, x) wR[R * wstride]
    void forward_weighted(const void* f, const T* wR, int wstride, void* F, hipStream_t st) {
        const int M = 2 * model.B - 1;
        ensure(G, sizeof(cT<T>) * (size_t)g_elems(model.B, nx));
        launch_fwd_ag<T>(xf.view, f, wR, wstride, G.p, st);
        if (half_omega()) {  // the omega DFTs on half of the rows write the other half
            launch_beta_fwd<T>(xf.view, G.p, F, st, false);
            launch_omega_half<T>(xf.view, F, F, false, st);
        } else {
            launch_beta_fwd<T>(xf.view, G.p, F, st);
            launch_omega_dft<T>(xf.view, F, nullptr, false, st);
        }
        HPUP_HIP(hipGetLastError());
    }
    void backward(const void* F, void* f, hipStream_t st) override {
        const int M = 2 * model.B - 1;
        ensure(G, sizeof(cT<T>) * (size_t)g_elems(model.B, nx));
        ensure(tmp, cbytes());
        if (half_omega()) {  // the Hermitian part on half of the rows
            launch_omega_half<T>(xf.view, F, tmp.p, true, st);
            launch_beta_inv<T>(xf.view, tmp.p, G.p, st, false);
        } else {
            launch_omega_dft_to<T>(xf.view, F, tmp.p, true, st);
            launch_beta_inv<T>(xf.view, tmp.p, G.p, st);
        }
        launch_inv_ag<T>(xf.view, G.p, f, st);
        HPUP_HIP(hipGetLastError());
    }
    // omega DFTs of the transforms on half of the rows (chirp MFMA kernels)
    bool half_omega() const {
        return (model.dim == 2 || model.dim == 3) && model.N <= 32 && model.B <= 32 && xf.view.bsF && xf.view.hrows;
    }
    // MFMA kernels cover omega grids up to 32 points per axis, and up to 64 for
    // the reduced model (dim 2); larger grids use the LDS-tiled FMA kernels
    // (three passes).
    bool fused() const { return model.N <= 32 || model.dim == 0 || (model.dim == 2 && model.N <= 64); }

    // Density-form gravity (pendulum_reduced_den.m:453-516): the term is
    // -sum_p c(k_p) fftSO3R(f mR_p) on the grid density f, with f the caller's
    // grid for the first stage and the inverse transform of the stage input
    // for the RK substeps.
    bool den = false;
    const void* den_f = nullptr;  // grid density consistent with the next stage input
    DevBuf mR_dev, fgrid, gravb[3];
    void den_gravity(const void* Fin, DriftDev<T>& D, hipStream_t st) {
        const void* f = den_f;
        den_f = nullptr;
        if (!f) {
            ensure(fgrid, sizeof(T) * (size_t)ngrid);
            backward(Fin, fgrid.p, st);
            f = fgrid.p;
        }
        for (int p = 0; p < model.dim; p++) {
            ensure(gravb[p], cbytes());
            forward_weighted(f, mR_dev.as<T>() + p, model.dim, gravb[p].p, st);
            D.grav[p] = gravb[p].as<cT<T>>();
        }
    }

    // One derivative evaluation with its integrator epilogue.  ev (nullable)
    // receives an event before and after each kernel (nk + 1 events).
    void stage(const void* Fin, const void* F0, void* Fout, const StageCoef<T>& sc, hipStream_t st,
               hipEvent_t* ev = nullptr) {
        DriftDev<T> D = dr.view;
        // the fused kernel reads neighbouring rotation rows of Fin while other
        // workgroups write Fout, so the two must not alias (the three-pass
        // path reads Fin only through its own slab and may run in place)
        HPUP_REQUIRE(!(fused() && Fin == Fout), HPUP_ERR_ARG, "fused drift stage cannot run in place");
        auto mark = [&](int k) {
            if (ev) HPUP_HIP(hipEventRecord(ev[k], st));
        };
        if (den) den_gravity(Fin, D, st);
        if (fused()) {
            mark(0);
            if (D.dim == 3 && D.adj) {
                // T3^H F for the gyroscopic terms; T3^H [(e3^* - i b3 kc3) F] for the axis-3 term
                ensure(S3, cbytes());
                if (D.has_gyro) launch_axis3_mfma<T>(D, Fin, Z3.p, false, st);
                launch_axis3_mfma<T>(D, Fin, S3.p, false, st, true);
                D.s3 = S3.as<cT<T>>();
            } else if (D.dim == 3) {
                launch_axis3_mfma<T>(D, Fin, Z3.p, bf16x3, st);
            }
            mark(1);
            launch_drift_fused<T>(D, dr.mix, Fin, Z3.p, F0, Fout, ACC.p, sc, bf16x3, st);
            mark(2);
            return;
        }
        HPUP_REQUIRE(!D.adj, HPUP_ERR_UNSUPPORTED, "the adjoint operator needs the fused path (N <= 32)");
        mark(0);
        launch_rotmix<T>(D, Fin, Wb.p, V1.p, V2.p, st);
        mark(1);
        if (D.dim == 3) launch_axis_outer<T>(Fin, Z3.p, dr.M3.p, D.N, (int64_t)D.N * D.N, D.nrc, T(1), st);
        mark(2);
        launch_drift_slab<T>(D, Fin, Z3.p, Wb.p, V1.p, V2.p, F0, Fout, ACC.p, sc, st);
        mark(3);
    }
    struct StageRef {
        const void* in;
        const void* f0;
        void* out;
        StageCoef<T> sc;
    };
    // The stages of one step of a Runge-Kutta method on the state F
    // (pendulum.m:346-385); no fused stage writes its own input, so RK4's
    // middle stages alternate Fs and Fs2 on the fused path.
    // The drift operator is linear and autonomous, so each integrator's step
    // is the polynomial P(dt A) of its stability function: 1 + z + z^2/2 for
    // midpoint and RK2 (pendulum.m:357-371), sum_{j<=4} z^j/j! for RK4
    // (pendulum.m:373-385).  The stages evaluate it in Horner form,
    //   y <- F + (dt / j) A y,  j = s, s-1, ..., 1,
    // the reference's polynomial with one stream (F) besides each stage's input
    // and output, where the stage sums also read and write an accumulator.
    // Density-form plans keep the stage sums: their first stage takes gravity
    // from the caller's grid density, not from the stage input, so the step is
    // not a polynomial in one operator.
    int rk_stages(cT<T>* F, StageRef* sr) {
        const T dt = (T)model.dt;
        void* Fm = Fs.p;
        if (model.method == HPUP_RK4 && fused()) {
            ensure(Fs2, cbytes());
            Fm = Fs2.p;
        }
        if (!den) {
            switch (model.method) {
            case HPUP_MIDPOINT:
            case HPUP_RK2:
                sr[0] = {F, nullptr, Fs.p, StageCoef<T>{dt / 2, 0, 1, 0, 0, 0, 0, 0, 0}};
                sr[1] = {Fs.p, F, F, StageCoef<T>{dt, 0, 1, 0, 0, 1, 0, 0, 0}};
                return 2;
            case HPUP_RK4:
                sr[0] = {F, nullptr, Fs.p, StageCoef<T>{dt / 4, 0, 1, 0, 0, 0, 0, 0, 0}};
                sr[1] = {Fs.p, F, Fm, StageCoef<T>{dt / 3, 0, 1, 0, 0, 1, 0, 0, 0}};
                sr[2] = {Fm, F, Fs.p, StageCoef<T>{dt / 2, 0, 1, 0, 0, 1, 0, 0, 0}};
                sr[3] = {Fs.p, F, F, StageCoef<T>{dt, 0, 1, 0, 0, 1, 0, 0, 0}};
                return 4;
            default:
                throw Error{HPUP_ERR_METHOD, "unknown integrator"};
            }
        }
        ensure(ACC, cbytes());
        switch (model.method) {
        case HPUP_MIDPOINT:
            sr[0] = {F, nullptr, Fs.p, StageCoef<T>{dt / 2, 0, 1, 0, 0, 0, 0, 0, 0}};
            sr[1] = {Fs.p, F, F, StageCoef<T>{dt, 0, 1, 0, 0, 1, 0, 0, 0}};
            return 2;
        case HPUP_RK2:
            sr[0] = {F, nullptr, Fs.p, StageCoef<T>{dt, 0, 1, 0, 1, 0, 0, 1, 0}};
            sr[1] = {Fs.p, F, F, StageCoef<T>{dt / 2, 1, 1, 0, 0, 1, 1, 0, 0}};
            return 2;
        case HPUP_RK4:
            sr[0] = {F, nullptr, Fs.p, StageCoef<T>{dt / 2, 0, 1, 0, 1, 0, 0, 1, 0}};
            sr[1] = {Fs.p, F, Fm, StageCoef<T>{dt / 2, 0, 1, 1, 2, 1, 1, 1, 0}};
            sr[2] = {Fm, F, Fs.p, StageCoef<T>{dt, 0, 1, 1, 2, 1, 1, 1, 0}};
            sr[3] = {Fs.p, F, F, StageCoef<T>{dt / 6, 1, 1, 0, 0, 1, 1, 0, 0}};
            return 4;
        default:
            throw Error{HPUP_ERR_METHOD, "unknown integrator"};
        }
    }
    // Runs nsteps steps of the plan's method; ev (nullable) receives nk + 1
    // events per stage.  Euler steps alternate between F and Fs on the fused
    // path, and an odd count ends with a copy back into F.
    void run_steps(cT<T>* F, int nsteps, hipStream_t st, hipEvent_t* ev, int nk) {
        const bool fz = fused();
        if (model.method == HPUP_EULER) {
            const StageCoef<T> sc{(T)model.dt, 0, 1, 0, 0, 0, 0, 0, 0};
            cT<T>* cur = F;
            cT<T>* nxt = fz ? Fs.as<cT<T>>() : F;
            for (int it = 0; it < nsteps; it++) {
                stage(cur, nullptr, nxt, sc, st, ev ? ev + (size_t)it * (nk + 1) : nullptr);
                std::swap(cur, nxt);
            }
            if (fz && cur != F) HPUP_HIP(hipMemcpyAsync(F, cur, cbytes(), hipMemcpyDeviceToDevice, st));
            return;
        }
        StageRef sr[4];
        const int ns = rk_stages(F, sr);
        for (int it = 0; it < nsteps; it++)
            for (int k = 0; k < ns; k++)
                stage(sr[k].in, sr[k].f0, sr[k].out, sr[k].sc, st,
                      ev ? ev + ((size_t)it * ns + k) * (nk + 1) : nullptr);
    }
    int stages_per_step() {
        if (model.method == HPUP_EULER) return 1;
        StageRef sr[4];
        return rk_stages(nullptr, sr);
    }
    void ensure_drift_buffers(bool rk) {
        HPUP_REQUIRE(has_drift, HPUP_ERR_ARG, "plan has no drift operator (dim must be 2 or 3)");
        HPUP_REQUIRE(model.N <= 32 || !bf16x3, HPUP_ERR_UNSUPPORTED, "3xBF16 drift needs an omega grid of at most 32");
        if (!fused()) { ensure(Wb, cbytes()); ensure(V1, cbytes()); ensure(V2, cbytes()); }
        if (model.dim == 3) ensure(Z3, cbytes());
        if (rk) ensure(Fs, cbytes());
    }
    void step(void* Fv, int nsteps, hipStream_t st) override {
        ensure_drift_buffers(true);
        run_steps((cT<T>*)Fv, nsteps, st, nullptr, 0);
        HPUP_HIP(hipGetLastError());
    }
    void derivative(const void* F, void* dF, hipStream_t st) override {
        ensure_drift_buffers(false);
        stage(F, nullptr, dF, StageCoef<T>{1, 0, 1, 0, 0, 0, 0, 0, 1}, st);
        HPUP_HIP(hipGetLastError());
    }
    // out = F0 + scale A in: one Horner stage (the fused kernel's epilogue)
    void axpy_stage(const void* in, const void* F0, double scale, void* out, hipStream_t st) override {
        HPUP_REQUIRE(in && F0 && out, HPUP_ERR_ARG, "null argument");
        ensure_drift_buffers(false);
        stage(in, F0, out, StageCoef<T>{(T)scale, 0, 1, 0, 0, 1, 0, 0, 0}, st);
        HPUP_HIP(hipGetLastError());
    }
    // The Horner stage scales h_j = c_j dt of the plan's integrator: one step
    // is y_j = F + h_j A y_{j-1}, j = 1..s, y_0 = F, F_next = y_s; rounded
    // exactly as the step's own stages form them (dt / 4, dt / 3, ... in the
    // plan precision), so stage values taken here equal the step's bit for bit
    int horner(double* c) const {
        const T dt = (T)model.dt;
        switch (model.method) {
        case HPUP_EULER: c[0] = (double)dt; return 1;
        case HPUP_MIDPOINT:
        case HPUP_RK2: c[0] = (double)(dt / 2); c[1] = (double)dt; return 2;
        case HPUP_RK4:
            c[0] = (double)(dt / 4); c[1] = (double)(dt / 3); c[2] = (double)(dt / 2); c[3] = (double)dt;
            return 4;
        default: throw Error{HPUP_ERR_METHOD, "unknown integrator"};
        }
    }
    // ys[0..s-2] = y_1..y_{s-1} of one step from F; Fnext (nullable) = y_s
    void step_stages(const void* F, void* const* ys, void* Fnext, hipStream_t st) override {
        HPUP_REQUIRE(!den && !dr.view.adj, HPUP_ERR_UNSUPPORTED, "step stages: a forward model plan");
        HPUP_REQUIRE(F && (ys || model.method == HPUP_EULER), HPUP_ERR_ARG, "null argument");
        double c[4];
        const int ns = horner(c);
        const void* prev = F;
        for (int j = 1; j < ns; j++) {
            HPUP_REQUIRE(ys[j - 1] && ys[j - 1] != F, HPUP_ERR_ARG, "stage buffers must be distinct from F");
            axpy_stage(prev, F, c[j - 1], ys[j - 1], st);
            prev = ys[j - 1];
        }
        if (Fnext) {
            HPUP_REQUIRE(Fnext != prev, HPUP_ERR_ARG, "F_next must not alias the last stage value");
            axpy_stage(prev, F, c[ns - 1], Fnext, st);
        }
    }
    // The whole reverse step on an adjoint plan from the forward stage values
    // y_0 = F, y_1..y_{s-1} = ys: ybar_s = lam, ybar_{j-1} = c_j dt A^H ybar_j,
    // Fbar = sum_j ybar_j, gout += the parameter terms of every stage
    void adjoint_step(const void* F, const void* const* ys, const void* lam, void* Fbar, double* gout,
                      hipStream_t st) override {
        HPUP_REQUIRE(F && lam && Fbar && gout && (ys || model.method == HPUP_EULER), HPUP_ERR_ARG, "null argument");
        double c[4];
        const int ns = horner(c);
        ensure(Ya, cbytes());
        ensure(Yb, cbytes());
        const void* ybar = lam;
        for (int j = ns - 1; j >= 0; j--) {
            const void* y = j == 0 ? F : ys[j - 1];
            void* out = (ybar == Ya.p) ? Yb.p : Ya.p;
            adjoint_stage(ybar, y, c[j], out, Fbar, j == ns - 1 ? 1 : 0, gout, st, false);
            ybar = out;
        }
    }
    // One reverse-mode stage of the gradient sweep on an adjoint plan:
    // out = scale A^H ybar, fbar += out (when fbar is given; fbar = ybar + out
    // with fbar_init), gout[0..4] += scale Re<ybar, (dA/dtheta) y> for
    // theta = b_1, b_2, G_00, G_11 and G_01 (= the G_10 term).
    // noalloc: the workgroups' partial sums go to the static device array
    // (grad_scratch) instead of the plan's buffer, and nothing is allocated
    void adjoint_stage(const void* ybar, const void* y, double scale, void* out, void* fbar, int fbar_init,
                       double* gout, hipStream_t st, bool noalloc) override {
        HPUP_REQUIRE(has_drift && fused() && dr.view.adj, HPUP_ERR_UNSUPPORTED,
                     "adjoint_stage needs an adjoint operator plan (operator kind 1, N <= 32)");
        HPUP_REQUIRE(ybar && y && out && gout, HPUP_ERR_ARG, "null argument");
        HPUP_REQUIRE(ybar != out && y != out, HPUP_ERR_ARG, "adjoint stage cannot run in place");
        HPUP_REQUIRE(fbar != ybar && fbar != y && fbar != out, HPUP_ERR_ARG, "fbar must not alias the stage arrays");
        HPUP_REQUIRE(model.dim == 2, HPUP_ERR_UNSUPPORTED, "parameter gradients: reduced model (dim 2) only");
        double* part;
        if (noalloc) {
            int dev = 0;
            HPUP_HIP(hipGetDevice(&dev));
            part = grad_scratch(dev);
        } else {
            ensure(gpart, sizeof(double) * kGradMaxWG * 5);
            part = gpart.as<double>();
        }
        DriftDev<T> D = dr.view;
        D.gy = y;
        D.gacc = fbar;
        D.gacc_init = fbar_init ? 1 : 0;
        D.gpart = part;
        D.gout = gout;
        D.gscale = scale;
        launch_drift_fused<T>(D, dr.mix, ybar, Z3.p, nullptr, out, ACC.p,
                              StageCoef<T>{(T)scale, 0, 1, 0, 0, 0, 0, 0, 1}, false, st);
        HPUP_HIP(hipGetLastError());
    }
    // Per-kernel timing of nreps steps of the plan's method: an event before
    // and after every kernel launch, all steps back to back after a few
    // untimed ones (the device clocks up under sustained load), one
    // synchronization at the end; ms[k] is the average per launch.
    int profile(void* Fv, int nreps, double* ms, int maxk, hipStream_t st) override {
        ensure_drift_buffers(true);
        HPUP_REQUIRE(nreps >= 1 && maxk >= 3, HPUP_ERR_ARG, "profile: need nreps >= 1 and room for 3 kernels");
        HPUP_REQUIRE(!den, HPUP_ERR_UNSUPPORTED, "profile: density-form plans are not profiled");
        const int nk = fused() ? 2 : 3;
        const int ns = stages_per_step();
        cT<T>* F = (cT<T>*)Fv;
        std::vector<hipEvent_t> ev((size_t)nreps * ns * (nk + 1), nullptr);
        try {
            for (auto& e : ev) HPUP_HIP(hipEventCreate(&e));
            run_steps(F, 3, st, nullptr, nk);
            run_steps(F, nreps, st, ev.data(), nk);
            HPUP_HIP(hipEventSynchronize(ev.back()));
            double acc[3] = {0, 0, 0};
            for (size_t l = 0; l < (size_t)nreps * ns; l++)
                for (int k = 0; k < nk; k++) {
                    float t = 0.f;
                    HPUP_HIP(hipEventElapsedTime(&t, ev[l * (nk + 1) + k], ev[l * (nk + 1) + k + 1]));
                    acc[k] += t;
                }
            for (int k = 0; k < nk; k++) ms[k] = acc[k] / ((double)nreps * ns);
        } catch (...) {
            for (auto& e : ev)
                if (e) (void)hipEventDestroy(e);
            throw;
        }
        for (auto& e : ev) HPUP_HIP(hipEventDestroy(e));
        return nk;
    }
    const char* profile_name(int i) const override {
        static const char* names[3] = {"k_rotmix", "k_axis_outer", "k_drift_slab"};
        if (fused()) {
            if (i == 0) return "k_axis3_mfma";
            return i == 1 ? drift_fused_name<T>(dr.view, den, bf16x3) : "";
        }
        return (i >= 0 && i < 3) ? names[i] : "";
    }
    // Algorithmic HBM bytes per launch of kernel i, averaged over the stages
    // of the plan's method (each input read once, each output written once).
    double profile_bytes(int i) const override {
        const double c = (double)sizeof(cT<T>) * (double)ncoef;
        const bool d3 = model.dim == 3;
        auto* self = const_cast<PlanT<T>*>(this);
        StageRef sr[4];
        int ns = 1;
        if (model.method == HPUP_EULER)
            sr[0].sc = StageCoef<T>{(T)model.dt, 0, 1, 0, 0, 0, 0, 0, 0};
        else
            ns = self->rk_stages(nullptr, sr);
        double sum = 0.0;
        for (int k = 0; k < ns; k++) {
            const StageCoef<T>& sc = sr[k].sc;
            // streams of the stage epilogue: F0, ACC read, ACC written
            const double epi = (sc.read_f0 ? 1 : 0) + (sc.read_acc ? 1 : 0) + (sc.write_acc ? 1 : 0);
            if (fused()) {
                if (i == 0) sum += d3 ? 2 * c : 0.0;         // read F; write T3 F
                else if (i == 1) sum += (d3 ? 3 + epi : 2 + epi) * c;  // read F (+ T3 F); write F
            } else {
                if (i == 0) sum += 4 * c;                    // read F; write W, -U1 F, -U2 F
                else if (i == 1) sum += d3 ? 2 * c : 0.0;    // read F; write T3 F
                else if (i == 2) sum += ((d3 ? 6 : 5) + epi) * c;  // read F, T3F, W, V1, V2; write F
            }
        }
        return sum / ns;
    }
    void full_to_compact(const void* a, void* b, hipStream_t st) override {
        launch_full_to_compact<T>(model.B, nx, a, b, st);
        HPUP_HIP(hipGetLastError());
    }
    void compact_to_full(const void* a, void* b, hipStream_t st) override {
        launch_compact_to_full<T>(model.B, nx, a, b, st);
        HPUP_HIP(hipGetLastError());
    }
    // mode 0: out = fin + scale df(fin) (out may be fin); mode 1: out = scale df(fin)
    void reset_grid(const void* fin, void* out, double scale, int mode, hipStream_t st) override {
        HPUP_REQUIRE(reset, HPUP_ERR_ARG, "the plan has no reset (hpup_plan_set_reset)");
        HPUP_REQUIRE(fin && out, HPUP_ERR_ARG, "null argument");
        HPUP_REQUIRE(mode == 0 || mode == 1, HPUP_ERR_ARG, "reset mode must be 0 or 1");
        HPUP_REQUIRE(mode == 0 || fin != out, HPUP_ERR_ARG, "reset mode 1 cannot run in place");
        if (reset->rows_fit(sizeof(T))) {  // LDS row-group kernel, in place on the region's rows
            if (mode == 0) {
                if (out != fin)
                    HPUP_HIP(hipMemcpyAsync(out, fin, sizeof(T) * (size_t)ngrid, hipMemcpyDeviceToDevice, st));
                reset->rows_pass<T>(out, out, scale, true, st);
            } else {
                reset->zero<T>(out, st);
                reset->rows_pass<T>(fin, out, scale, false, st);
            }
            return;
        }
        ensure(dfc, sizeof(T) * (size_t)std::max<int64_t>(1, (int64_t)reset->numR * nx));
        reset->df<T>(fin, dfc.p, st);
        if (mode == 0) {
            if (out != fin) HPUP_HIP(hipMemcpyAsync(out, fin, sizeof(T) * (size_t)ngrid, hipMemcpyDeviceToDevice, st));
            reset->apply<T>(dfc.p, out, out, scale, st);
        } else {
            reset->apply<T>(dfc.p, nullptr, out, scale, st);
        }
    }
    // out = g + scale R^H g (R the jump term; out may be g)
    void reset_adjoint(const void* g, void* out, double scale, hipStream_t st) override {
        HPUP_REQUIRE(reset, HPUP_ERR_ARG, "the plan has no reset (hpup_plan_set_reset)");
        HPUP_REQUIRE(g && out, HPUP_ERR_ARG, "null argument");
        if (out != g) HPUP_HIP(hipMemcpyAsync(out, g, sizeof(T) * (size_t)ngrid, hipMemcpyDeviceToDevice, st));
        if (reset->rows_fit(2 * sizeof(T))) {  // LDS row-group kernel (LDS atomics)
            reset->adjoint_rows<T>(out, out, scale, st);
            return;
        }
        ensure(dfc, sizeof(T) * (size_t)std::max<int64_t>(1, (int64_t)reset->numR * nx));
        reset->adjoint<T>(out, dfc.p, st);
        reset->apply<T>(dfc.p, out, out, scale, st);
    }
    // out[0] (device) += scale <g, dR/d eps f>
    void reset_eps_grad(const void* f, const void* g, double scale, double* out, hipStream_t st) override {
        HPUP_REQUIRE(reset, HPUP_ERR_ARG, "the plan has no reset (hpup_plan_set_reset)");
        HPUP_REQUIRE(f && g && out, HPUP_ERR_ARG, "null argument");
        ensure(epart, sizeof(double) * (size_t)reset->eps_grad_partials());
        reset->eps_grad<T>(f, g, epart.as<double>(), scale, out, st);
    }
    // pendulum_reduced_hybrid.m:224-234: continuous step on the coefficients,
    // inverse transform, f += dt df on the grid, forward transform; f holds
    // the last step's grid density after the reset
    void hybrid_step(void* F, void* f, int nsteps, hipStream_t st) override {
        HPUP_REQUIRE(reset, HPUP_ERR_ARG, "the plan has no reset (hpup_plan_set_reset)");
        HPUP_REQUIRE(F && f && nsteps >= 0, HPUP_ERR_ARG, "bad argument");
        for (int it = 0; it < nsteps; it++) {
            step(F, 1, st);
            backward(F, f, st);
            reset_grid(f, f, model.dt, 0, st);
            forward(f, F, st);
        }
    }
    void marginals(const void* f, const double* xg, std::vector<double>& hR, std::vector<double>& hG,
                   std::vector<double>& hx, hipStream_t st) override {
        grid_marginals<T>(model.B, model.dim, nx, f, w_dev.as<double>(), xg, st_work, st_out, hR, hG, hx, st);
    }
};

void PlanBase::get_stat(const void* f, hpup_stat& out, double* fx, bool fit, hipStream_t st) {
    const int B = model.B, dim = model.dim, N = std::max(model.N, 1);
    if (dim > 0 && !xg_dev.p) {
        xg_host.assign((size_t)dim * nx, 0.0);
        for (int64_t ix = 0; ix < nx; ix++)
            for (int a = 0; a < dim; a++)
                xg_host[(size_t)dim * ix + a] = -model.L / 2 + model.L / N * (double)((ix / ipow64(N, a)) % N);
        xg_dev = upload(xg_host);
    }
    std::vector<double> hR, hG, hx;
    marginals(f, xg_dev.as<double>(), hR, hG, hx, st);
    const double dv = dim > 0 ? std::pow(model.L / N, dim) : 1.0;
    const EulerGrid eg(B);
    mf::finish_stat(dim, (int64_t)hR.size(), nx, 2 * B, w.data(), hR.data(), hG.data(), hx.data(), xg_host.data(),
                    dv, eg, fit, out);
    if (fx) std::memcpy(fx, hx.data(), sizeof(double) * nx);
}

static PlanBase* make_plan(const hpup_model& m, const double* d, const double* w, const DriftSpec* spec) {
    HPUP_REQUIRE(m.B >= 1 && m.B <= 128, HPUP_ERR_ARG, "B out of range");
    HPUP_REQUIRE(m.dim == 0 || m.dim == 2 || m.dim == 3, HPUP_ERR_ARG, "dim must be 0, 2 or 3");
    HPUP_REQUIRE(m.dim == 0 || (m.N >= 2 && m.N % 2 == 0 && m.N <= 64), HPUP_ERR_ARG, "N must be even, 2..64");
    std::unique_ptr<PlanBase> p;
    if (m.precision == HPUP_FP64) {
        auto* q = new PlanT<double>();
        p.reset(q);
    } else if (m.precision == HPUP_FP32 || m.precision == HPUP_BF16X3) {
        auto* q = new PlanT<float>();
        q->bf16x3 = m.precision == HPUP_BF16X3;
        p.reset(q);
    } else {
        throw Error{HPUP_ERR_UNSUPPORTED, "precision mode not available"};
    }
    p->model = m;
    if (m.dim == 0) p->model.N = 1;
    p->nx = m.dim == 0 ? 1 : ipow64(m.N, m.dim);
    p->nrc = nrc_of(m.B);
    p->ncoef = p->nrc * p->nx;
    p->ngrid = 8LL * m.B * m.B * m.B * p->nx;
    p->w.assign(w, w + 2 * m.B);
    p->w_dev = upload(p->w);
    if (m.precision == HPUP_FP64) {
        auto* q = static_cast<PlanT<double>*>(p.get());
        q->init_xform(d, w);
        if (spec) q->init_drift(*spec);
    } else {
        auto* q = static_cast<PlanT<float>*>(p.get());
        q->init_xform(d, w);
        if (spec) q->init_drift(*spec);
    }
    return p.release();
}

}  // namespace hpup

using namespace hpup;

struct hpup_plan {
    std::unique_ptr<PlanBase> impl;
};

extern "C" {

const char* hpup_last_error(void) { return g_last_error.c_str(); }
int hpup_version(void) { return 100; }

int hpup_parse_method(const char* method) {
    if (!method) return -1;
    if (!strcasecmp(method, "euler")) return HPUP_EULER;
    if (!strcasecmp(method, "midpoint")) return HPUP_MIDPOINT;
    if (!strcasecmp(method, "RK2")) return HPUP_RK2;
    if (!strcasecmp(method, "RK4")) return HPUP_RK4;
    return -1;
}

int hpup_plan_create(const hpup_model* m, hpup_plan** out) {
    return guard([&] {
        HPUP_REQUIRE(m && out, HPUP_ERR_ARG, "null argument");
        *out = nullptr;
        HPUP_REQUIRE(m->method >= 0 && m->method <= 3, HPUP_ERR_METHOD, "unknown integrator");
        const int B = m->B, M = 2 * B - 1;
        std::vector<double> beta, w;
        host_grid(B, beta, w);
        std::vector<double> d((size_t)M * M * B * 2 * B);
        for (int j = 0; j < 2 * B; j++) host_wigner_d(beta[j], B - 1, d.data() + (size_t)j * M * M * B);
        std::unique_ptr<DriftSpec> spec;
        if (m->dim >= 2) spec.reset(new DriftSpec(build_drift_spec_from_model(*m)));
        auto* h = new hpup_plan();
        h->impl.reset(make_plan(*m, d.data(), w.data(), spec.get()));
        *out = h;
    });
}

int hpup_plan_create_operator(const hpup_model* m, int kind, hpup_plan** out) {
    return guard([&] {
        HPUP_REQUIRE(m && out, HPUP_ERR_ARG, "null argument");
        *out = nullptr;
        HPUP_REQUIRE(m->method >= 0 && m->method <= 3, HPUP_ERR_METHOD, "unknown integrator");
        HPUP_REQUIRE(m->dim >= 2, HPUP_ERR_ARG, "operator plans need an omega grid");
        const int B = m->B, M = 2 * B - 1;
        std::vector<double> beta, w;
        host_grid(B, beta, w);
        std::vector<double> d((size_t)M * M * B * 2 * B);
        for (int j = 0; j < 2 * B; j++) host_wigner_d(beta[j], B - 1, d.data() + (size_t)j * M * M * B);
        DriftSpec spec = build_drift_spec_from_model(*m);
        if (kind == 1) spec = adjoint_spec(spec);
        else if (kind >= 2) spec = parameter_spec(spec, kind);
        else HPUP_REQUIRE(kind == 0, HPUP_ERR_ARG, "unknown operator kind");
        auto* h = new hpup_plan();
        h->impl.reset(make_plan(*m, d.data(), w.data(), &spec));
        *out = h;
    });
}

int hpup_plan_destroy(hpup_plan* p) {
    return guard([&] { delete p; });
}

int64_t hpup_plan_num_coeffs(const hpup_plan* p) { return p ? p->impl->ncoef : -1; }
int64_t hpup_plan_num_grid(const hpup_plan* p) { return p ? p->impl->ngrid : -1; }

int hpup_plan_forward(hpup_plan* p, const void* f, void* F, void* stream) {
    return guard([&] { p->impl->forward(f, F, (hipStream_t)stream); });
}
int hpup_plan_backward(hpup_plan* p, const void* F, void* f, void* stream) {
    return guard([&] { p->impl->backward(F, f, (hipStream_t)stream); });
}
int hpup_plan_step(hpup_plan* p, void* F, int nsteps, void* stream) {
    return guard([&] { p->impl->step(F, nsteps, (hipStream_t)stream); });
}
int hpup_plan_derivative(hpup_plan* p, const void* F, void* dF, void* stream) {
    return guard([&] { p->impl->derivative(F, dF, (hipStream_t)stream); });
}
int hpup_plan_axpy_stage(hpup_plan* p, const void* in, const void* F0, double scale, void* out, void* stream) {
    return guard([&] { p->impl->axpy_stage(in, F0, scale, out, (hipStream_t)stream); });
}
int hpup_plan_adjoint_stage(hpup_plan* p, const void* ybar, const void* y, double scale, void* ybar_out,
                            void* fbar, int fbar_init, double* grad, void* stream) {
    return guard([&] { p->impl->adjoint_stage(ybar, y, scale, ybar_out, fbar, fbar_init, grad, (hipStream_t)stream); });
}
int hpup_plan_step_stages(hpup_plan* p, const void* F, void* const* ys, void* F_next, void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p, HPUP_ERR_ARG, "null argument");
        p->impl->step_stages(F, ys, F_next, (hipStream_t)stream);
    });
}
int hpup_plan_adjoint_step(hpup_plan* p, const void* F, const void* const* ys, const void* lam, void* Fbar,
                           double* grad, void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p, HPUP_ERR_ARG, "null argument");
        p->impl->adjoint_step(F, ys, lam, Fbar, grad, (hipStream_t)stream);
    });
}
int hpup_plan_adjoint_stage_noalloc(hpup_plan* p, const void* ybar, const void* y, double scale, void* ybar_out,
                                    void* fbar, int fbar_init, double* grad, void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p, HPUP_ERR_ARG, "null argument");
        p->impl->adjoint_stage(ybar, y, scale, ybar_out, fbar, fbar_init, grad, (hipStream_t)stream, true);
    });
}
int hpup_plan_full_to_compact(hpup_plan* p, const void* a, void* b, void* stream) {
    return guard([&] { p->impl->full_to_compact(a, b, (hipStream_t)stream); });
}
int hpup_plan_compact_to_full(hpup_plan* p, const void* a, void* b, void* stream) {
    return guard([&] { p->impl->compact_to_full(a, b, (hipStream_t)stream); });
}
int hpup_plan_profile(hpup_plan* p, void* F, int nreps, double* ms, int max_kernels, void* stream) {
    int n = 0;
    int rc = guard([&] {
        HPUP_REQUIRE(p && F && ms, HPUP_ERR_ARG, "null argument");
        n = p->impl->profile(F, nreps, ms, max_kernels, (hipStream_t)stream);
    });
    return rc == HPUP_OK ? n : rc;
}
const char* hpup_plan_profile_name(const hpup_plan* p, int i) { return p ? p->impl->profile_name(i) : ""; }
double hpup_plan_profile_bytes(const hpup_plan* p, int i) { return p ? p->impl->profile_bytes(i) : 0.0; }

int hpup_plan_stat_moments(hpup_plan* p, const void* f, double* ER, double* Ex, double* Varx, double* fx,
                           void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p && f, HPUP_ERR_ARG, "null argument");
        hpup_stat s;
        p->impl->get_stat(f, s, fx, false, (hipStream_t)stream);
        const int dim = p->impl->model.dim;
        if (ER) std::memcpy(ER, s.ER, sizeof(s.ER));
        if (Ex) std::memcpy(Ex, s.Ex, sizeof(double) * dim);
        if (Varx) std::memcpy(Varx, s.Varx, sizeof(double) * dim * dim);
    });
}

int hpup_plan_set_reset(hpup_plan* p, const hpup_reset_params* prm) {
    return guard([&] {
        HPUP_REQUIRE(p && prm, HPUP_ERR_ARG, "null argument");
        p->impl->set_reset(*prm);
    });
}
int hpup_plan_reset_region(const hpup_plan* p, int32_t* indR, int max_rows) {
    try {
        HPUP_REQUIRE(p, HPUP_ERR_ARG, "null argument");
        HPUP_REQUIRE(p->impl->reset, HPUP_ERR_ARG, "the plan has no reset (hpup_plan_set_reset)");
        const auto& reg = p->impl->reset->region;
        if (indR)
            for (int k = 0; k < (int)reg.size() && k < max_rows; k++) indR[k] = reg[k];
        return (int)reg.size();
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    }
}
int hpup_plan_reset(hpup_plan* p, const void* f_in, void* f_out, double scale, int mode, void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p, HPUP_ERR_ARG, "null argument");
        p->impl->reset_grid(f_in, f_out, scale, mode, (hipStream_t)stream);
    });
}
int hpup_plan_reset_adjoint(hpup_plan* p, const void* g, void* out, double scale, void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p, HPUP_ERR_ARG, "null argument");
        p->impl->reset_adjoint(g, out, scale, (hipStream_t)stream);
    });
}
int hpup_plan_reset_eps_grad(hpup_plan* p, const void* f, const void* g, double scale, double* out, void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p, HPUP_ERR_ARG, "null argument");
        p->impl->reset_eps_grad(f, g, scale, out, (hipStream_t)stream);
    });
}
int hpup_plan_hybrid_step(hpup_plan* p, void* F, void* f, int nsteps, void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p, HPUP_ERR_ARG, "null argument");
        p->impl->hybrid_step(F, f, nsteps, (hipStream_t)stream);
    });
}

int hpup_plan_get_stat(hpup_plan* p, const void* f, hpup_stat* out, double* fx, void* stream) {
    return guard([&] {
        HPUP_REQUIRE(p && f && out, HPUP_ERR_ARG, "null argument");
        HPUP_REQUIRE(p->impl->model.dim > 0, HPUP_ERR_ARG, "get_stat needs an omega grid (dim 2 or 3)");
        p->impl->get_stat(f, *out, fx, true, (hipStream_t)stream);
    });
}

// --------------------------------------------- gateways: get_stat, fit --

int hpup_get_stat(const double* f, const int64_t* dims, int ndims, const double* R, const double* x,
                  const double* w, hpup_stat* out, double* fx) {
    return guard([&] {
        HPUP_REQUIRE(f && dims && R && x && w && out, HPUP_ERR_ARG, "null argument");
        HPUP_REQUIRE(ndims == 5 || ndims == 6, HPUP_ERR_ARG, "f must have rank 5 or 6");
        HPUP_REQUIRE(dims[0] % 2 == 0 && dims[0] > 0 && dims[0] == dims[1] && dims[1] == dims[2], HPUP_ERR_ARG,
                     "f must be (2B,2B,2B,N,...)");
        const int B = (int)(dims[0] / 2), dim = ndims - 3, N = (int)dims[3];
        HPUP_REQUIRE(N >= 2, HPUP_ERR_ARG, "omega grid needs N >= 2");
        for (int i = 4; i < ndims; i++) HPUP_REQUIRE(dims[i] == N, HPUP_ERR_ARG, "omega axes must be equal");
        const int64_t nx = ipow64(N, dim), n2 = 2 * B, nR = n2 * n2 * n2;
        // L = x(1,end,1,1) + x(1,2,1,1) - 2 x(1,1,1,1), cell (L/N)^dim (pendulum.m:549-551)
        const double L = x[(size_t)dim * (N - 1)] + x[dim] - 2 * x[0];
        const double dv = std::pow(L / N, dim);
        DevBuf fd(sizeof(double) * nR * nx), wd, xd, work, ob;
        HPUP_HIP(hipMemcpy(fd.p, f, fd.bytes, hipMemcpyHostToDevice));
        wd = upload(std::vector<double>(w, w + n2));
        std::vector<double> xh(x, x + (size_t)dim * nx);
        xd = upload(xh);
        std::vector<double> hR, hG, hx;
        grid_marginals<double>(B, dim, nx, fd.p, wd.as<double>(), xd.as<double>(), work, ob, hR, hG, hx, nullptr);
        auto Rof = [R](int64_t iR, double* r) { std::memcpy(r, R + 9 * iR, 9 * sizeof(double)); };
        mf::finish_stat(dim, nR, nx, (int)n2, w, hR.data(), hG.data(), hx.data(), xh.data(), dv, Rof, true, *out);
        if (fx) std::memcpy(fx, hx.data(), sizeof(double) * nx);
    });
}

int hpup_pendulum_plot_getc(const double* f, const int64_t* dims, int ndims, const double* e, const int64_t* edims,
                            int nedims, double L, double* c) {
    return guard([&] {
        HPUP_REQUIRE(f && dims && e && edims && c, HPUP_ERR_ARG, "null argument");
        HPUP_REQUIRE(ndims == 5 || ndims == 6, HPUP_ERR_ARG, "f must have rank 5 or 6");
        HPUP_REQUIRE(dims[0] % 2 == 0 && dims[0] > 0 && dims[0] == dims[1] && dims[1] == dims[2], HPUP_ERR_ARG,
                     "f must be (2B,2B,2B,N,...)");
        HPUP_REQUIRE(nedims == 4 && edims[0] == 3 && edims[3] == 3 && edims[1] >= 0 && edims[2] >= 0, HPUP_ERR_ARG,
                     "e must be (3, ns, na, 3)");
        const int B = (int)(dims[0] / 2), dim = ndims - 3, N = (int)dims[3];
        HPUP_REQUIRE(B <= 64, HPUP_ERR_UNSUPPORTED, "bandlimit above 64");
        for (int i = 4; i < ndims; i++) HPUP_REQUIRE(dims[i] == N, HPUP_ERR_ARG, "omega axes must be equal");
        const int64_t nx = ipow64(N, dim), n2 = 2 * B, nR = n2 * n2 * n2, ns = edims[1], na = edims[2];
        const int64_t P = ns * na * 3;
        // rotation marginal, gridsize L / N per omega axis (getc.cpp:42-61)
        DevBuf fd(sizeof(double) * nR * nx), fR(sizeof(double) * nR);
        HPUP_HIP(hipMemcpy(fd.p, f, fd.bytes, hipMemcpyHostToDevice));
        launch_rot_marginal(fd.as<double>(), nR, nx, std::pow(L / N, dim), fR.as<double>(), nullptr);
        HPUP_HIP(hipGetLastError());
        // its SO(3) transform (nx = 1), tables of fftSO3R.m
        hpup_model m{};
        m.B = B;
        m.dim = 0;
        m.N = 1;
        m.precision = HPUP_FP64;
        const int M = 2 * B - 1, lmax = B - 1;
        std::vector<double> beta, w, d((size_t)M * M * B * n2);
        host_grid(B, beta, w);
        for (int j = 0; j < n2; j++) host_wigner_d(beta[j], lmax, d.data() + (size_t)j * M * M * B);
        std::unique_ptr<PlanBase> p(make_plan(m, d.data(), w.data(), nullptr));
        DevBuf Fc(sizeof(double) * 2 * p->ncoef);
        p->forward(fR.p, Fc.p, nullptr);
        // wedge seeds sqrt((2l)! / ((l+m)! (l-m)!)) (Wigner_d.m / getc.cpp:215-217)
        std::vector<double> sq((size_t)B * M, 0.0), fact(2 * B + 1, 1.0);
        for (int k = 1; k <= 2 * B; k++) fact[k] = fact[k - 1] * k;
        for (int l = 0; l < B; l++)
            for (int mm = -l; mm <= l; mm++) sq[(size_t)l * M + mm + l] = std::sqrt(fact[2 * l] / fact[l + mm] / fact[l - mm]);
        DevBuf sqd = upload(sq);
        std::vector<double> ph(P > 0 ? P : 1);
        if (P > 0) {
            DevBuf ed(sizeof(double) * 3 * P), od(sizeof(double) * P);
            HPUP_HIP(hipMemcpy(ed.p, e, ed.bytes, hipMemcpyHostToDevice));
            launch_point_eval(Fc.p, B, sqd.as<double>(), ed.as<double>(), P, od.as<double>(), nullptr);
            HPUP_HIP(hipGetLastError());
            HPUP_HIP(hipMemcpy(ph.data(), od.p, od.bytes, hipMemcpyDeviceToHost));
        }
        // c(s, k) = sum over the na angles (getc.cpp:24-35)
        for (int64_t k = 0; k < 3; k++)
            for (int64_t s = 0; s < ns; s++) {
                double acc = 0.0;
                for (int64_t a = 0; a < na; a++) acc += ph[s + ns * (a + na * k)];
                c[s + ns * k] = acc;
            }
    });
}

int hpup_psvd(const double* F, double* U, double* S, double* V) {
    return guard([&] {
        HPUP_REQUIRE(F && U && S && V, HPUP_ERR_ARG, "null argument");
        double s[3];
        mf::psvd(F, U, s, V);
        for (int i = 0; i < 9; i++) S[i] = 0.0;
        for (int i = 0; i < 3; i++) S[4 * i] = s[i];
    });
}

int hpup_pdf_MF_normal(const double* s, int bool_scaled, double* c, double* dc, double* ddc) {
    return guard([&] {
        HPUP_REQUIRE(s && c, HPUP_ERR_ARG, "null argument");
        const mf::Normalizer n = mf::normalizer(s);
        // c = C e^K, dc = D e^K, ddc = H e^K; the scaled form multiplies by
        // e^{-sum s} and differentiates that product
        const double sig = s[0] + s[1] + s[2];
        const double e = std::exp(n.K - (bool_scaled ? sig : 0.0));
        *c = n.C * e;
        if (dc)
            for (int i = 0; i < 3; i++) dc[i] = (bool_scaled ? n.D[i] - n.C : n.D[i]) * e;
        if (ddc)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    ddc[i + 3 * j] =
                        (bool_scaled ? n.H[i + 3 * j] - n.D[i] - n.D[j] + n.C : n.H[i + 3 * j]) * e;
    });
}

int hpup_pdf_MF_M2S(const double* d, const double* s0, double* s, double* fval, int* niter) {
    return guard([&] {
        HPUP_REQUIRE(d && s, HPUP_ERR_ARG, "null argument");
        const double z[3] = {0, 0, 0};
        double fv[3];
        const int it = mf::fit_s(d, s0 ? s0 : z, s, fv);
        if (fval) std::memcpy(fval, fv, sizeof(fv));
        if (niter) *niter = it;
    });
}

// ------------------------------------------------------- gateways: FFT --

int hpup_fftSO3R_forward(const double* f, const int64_t* dims, int ndims, const double* d, const double* w,
                         double* F) {
    return guard([&] {
        HPUP_REQUIRE(f && dims && d && w && F, HPUP_ERR_ARG, "null argument");
        HPUP_REQUIRE(ndims == 3 || ndims == 5 || ndims == 6, HPUP_ERR_ARG, "f must have rank 3, 5 or 6");
        HPUP_REQUIRE(dims[0] % 2 == 0 && dims[0] == dims[1] && dims[1] == dims[2], HPUP_ERR_ARG,
                     "f must be (2B,2B,2B,...)");
        hpup_model m{};
        m.B = (int)(dims[0] / 2);
        m.dim = ndims - 3;
        m.N = ndims > 3 ? (int)dims[3] : 1;
        for (int i = 4; i < ndims; i++) HPUP_REQUIRE(dims[i] == dims[3], HPUP_ERR_ARG, "omega axes must be equal");
        m.precision = HPUP_FP64;
        std::unique_ptr<PlanBase> p(make_plan(m, d, w, nullptr));
        const int64_t nfull = (int64_t)(2 * m.B - 1) * (2 * m.B - 1) * m.B * p->nx;
        DevBuf fd(sizeof(double) * p->ngrid), Fc(sizeof(double) * 2 * p->ncoef), Ff(sizeof(double) * 2 * nfull);
        HPUP_HIP(hipMemcpy(fd.p, f, fd.bytes, hipMemcpyHostToDevice));
        p->forward(fd.p, Fc.p, nullptr);
        p->compact_to_full(Fc.p, Ff.p, nullptr);
        HPUP_HIP(hipMemcpy(F, Ff.p, Ff.bytes, hipMemcpyDeviceToHost));
    });
}

int hpup_fftSO3R_backward(const double* F, const int64_t* dims, int ndims, const double* d, double* f) {
    return guard([&] {
        HPUP_REQUIRE(F && dims && d && f, HPUP_ERR_ARG, "null argument");
        HPUP_REQUIRE(ndims == 3 || ndims == 5 || ndims == 6, HPUP_ERR_ARG, "F must have rank 3, 5 or 6");
        const int B = (int)dims[2];
        HPUP_REQUIRE(dims[0] == 2 * B - 1 && dims[1] == 2 * B - 1, HPUP_ERR_ARG, "F must be (2B-1,2B-1,B,...)");
        hpup_model m{};
        m.B = B;
        m.dim = ndims - 3;
        m.N = ndims > 3 ? (int)dims[3] : 1;
        for (int i = 4; i < ndims; i++) HPUP_REQUIRE(dims[i] == dims[3], HPUP_ERR_ARG, "omega axes must be equal");
        m.precision = HPUP_FP64;
        std::vector<double> beta, w;
        host_grid(B, beta, w);  // weights are not used by the inverse
        std::unique_ptr<PlanBase> p(make_plan(m, d, w.data(), nullptr));
        const int64_t nfull = (int64_t)(2 * B - 1) * (2 * B - 1) * B * p->nx;
        DevBuf Ff(sizeof(double) * 2 * nfull), Fc(sizeof(double) * 2 * p->ncoef), fd(sizeof(double) * p->ngrid);
        HPUP_HIP(hipMemcpy(Ff.p, F, Ff.bytes, hipMemcpyHostToDevice));
        p->full_to_compact(Ff.p, Fc.p, nullptr);
        p->backward(Fc.p, fd.p, nullptr);
        HPUP_HIP(hipMemcpy(f, fd.p, fd.bytes, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"

// ------------------------------------------------ gateways: propagation --

namespace hpup {

template <class T>
static void propagate_gateway(int dim, const T* F, const int64_t* dims, const T* X, const T* OJO, const T* MR,
                              const T* bt, const T* G, double dtt, double L, const T* u, const T* const* CG,
                              const char* method, T* Fnew) {
    HPUP_REQUIRE(F && dims && X && MR && bt && G && u && CG && method && Fnew, HPUP_ERR_ARG, "null argument");
    HPUP_REQUIRE(dim == 2 || OJO, HPUP_ERR_ARG, "OJO required for the 3-D model");
    const int meth = hpup_parse_method(method);
    HPUP_REQUIRE(meth >= 0, HPUP_ERR_METHOD, std::string("unknown method '") + method + "'");
    const int B = (int)dims[2], M = 2 * B - 1, N = (int)dims[3];
    HPUP_REQUIRE(dims[0] == M && dims[1] == M, HPUP_ERR_ARG, "F must be (2B-1,2B-1,B,N,...)");
    for (int i = 4; i < 3 + dim; i++) HPUP_REQUIRE(dims[i] == N, HPUP_ERR_ARG, "omega axes must be equal");
    const int64_t nx = ipow64(N, dim), nrot = (int64_t)M * M * B;
    // promote the gateway arrays to FP64 for the table analysis
    auto to_d = [](const T* a, int64_t n) { return std::vector<double>(a, a + n); };
    std::vector<double> Xd = to_d(X, 2 * nx * dim), MRd = to_d(MR, 2 * nrot * dim), ud = to_d(u, 2 * nrot * 3);
    std::vector<double> Od = dim == 3 ? to_d(OJO, 2 * nx * 3) : std::vector<double>();
    double btd[3] = {0, 0, 0}, Gd[9] = {0};
    for (int i = 0; i < dim; i++) btd[i] = bt[i];
    for (int i = 0; i < dim * dim; i++) Gd[i] = G[i];
    std::vector<std::vector<double>> cells((size_t)B * B);
    std::vector<const double*> cptr((size_t)B * B, nullptr);
    for (int l1 = 0; l1 < B; l1++) {
        if (B < 2) break;
        const T* c = CG[l1 + (size_t)B * 1];
        HPUP_REQUIRE(c != nullptr, HPUP_ERR_ARG, "CG cell missing");
        const int64_t nc = (int64_t)(2 * l1 + 1) * 3;
        cells[l1 + (size_t)B] = to_d(c, nc * nc);
        cptr[l1 + (size_t)B] = cells[l1 + (size_t)B].data();
    }
    DriftSpec spec = build_drift_spec_from_arrays(B, dim, N, L, Xd.data(), dim == 3 ? Od.data() : nullptr,
                                                  MRd.data(), btd, Gd, ud.data(), cptr.data(),
                                                  sizeof(T) == 8 ? 1e-9 : 1e-4);
    hpup_model m{};
    m.B = B; m.dim = dim; m.N = N; m.L = L; m.dt = dtt; m.method = meth;
    m.precision = sizeof(T) == 8 ? HPUP_FP64 : HPUP_FP32;
    for (int i = 0; i < dim; i++) m.b[i] = btd[i];
    for (int i = 0; i < dim * dim; i++) m.G[i] = Gd[i];
    std::vector<double> beta, w;
    host_grid(B, beta, w);
    std::vector<double> d((size_t)M * M * B * 2 * B);
    for (int j = 0; j < 2 * B; j++) host_wigner_d(beta[j], B - 1, d.data() + (size_t)j * M * M * B);
    std::unique_ptr<PlanBase> p(make_plan(m, d.data(), w.data(), &spec));
    const int64_t nfull = nrot * nx;
    DevBuf Ff(sizeof(T) * 2 * nfull), Fc(sizeof(T) * 2 * p->ncoef);
    HPUP_HIP(hipMemcpy(Ff.p, F, Ff.bytes, hipMemcpyHostToDevice));
    p->full_to_compact(Ff.p, Fc.p, nullptr);
    p->step(Fc.p, 1, nullptr);
    p->compact_to_full(Fc.p, Ff.p, nullptr);
    HPUP_HIP(hipMemcpy(Fnew, Ff.p, Ff.bytes, hipMemcpyDeviceToHost));
}

// gyrobias_propagate (gyrobias_propagate/kernel.cu:8-120, integrate.cu:241-443; gyro_bias.m:235-298):
// the attitude kinematics with the gyro bias as the R^3 variable (the pendulum's -(F (*) X_p / nx) u_p^T
// term), the gyro random walk sum_ij G1_ij F u_i^T u_j^T and the bias random walk sum_ij G2_ij c_ij(k) F
// on the bias box L = 2; one Euler step.  The first and third terms are the pendulum drift with no
// damping, gyroscopic or gravity term; the second is a banded mixing of the n index of each l block,
// F M_l^T with M_l = sum_ij G1_ij u_j u_i, added from the step's input.
static void gyrobias_gateway(const double* F, const int64_t* dims, const double* X, double dt, const double* u,
                             const double* G1, const double* G2, double* Fnew) {
    HPUP_REQUIRE(F && dims && X && u && G1 && G2 && Fnew, HPUP_ERR_ARG, "null argument");
    const int B = (int)dims[2], M = 2 * B - 1, N = (int)dims[3], lmax = B - 1;
    HPUP_REQUIRE(dims[0] == M && dims[1] == M && dims[4] == N && dims[5] == N, HPUP_ERR_ARG,
                 "F must be (2B-1,2B-1,B,N,N,N)");
    const int64_t nx = ipow64(N, 3), nrot = (int64_t)M * M * B;
    std::vector<double> zO(2 * nx * 3, 0.0), zMR(2 * nrot * 3, 0.0);
    const double bt[3] = {0, 0, 0};
    std::vector<std::vector<double>> cells((size_t)B * B);
    std::vector<const double*> cptr((size_t)B * B, nullptr);
    for (int l1 = 0; l1 < B && B >= 2; l1++) {
        host_clebsch_gordan(l1, 1, cells[l1 + (size_t)B]);
        cptr[l1 + (size_t)B] = cells[l1 + (size_t)B].data();
    }
    const double L = 2.0;  // gyro_bias.m:64 (the bias box; c_ij of integrate.cu uses pi = 2 pi / L)
    DriftSpec spec = build_drift_spec_from_arrays(B, 3, N, L, X, zO.data(), zMR.data(), bt, G2, u, cptr.data(), 1e-9);
    hpup_model m{};
    m.B = B; m.dim = 3; m.N = N; m.L = L; m.dt = dt; m.method = 0;  // Euler
    m.precision = HPUP_FP64;
    for (int i = 0; i < 9; i++) m.G[i] = G2[i];
    std::vector<double> beta, w;
    host_grid(B, beta, w);
    std::vector<double> d((size_t)M * M * B * 2 * B);
    for (int j = 0; j < 2 * B; j++) host_wigner_d(beta[j], lmax, d.data() + (size_t)j * M * M * B);
    std::unique_ptr<PlanBase> p(make_plan(m, d.data(), w.data(), &spec));
    // M_l = sum_ij G1_ij u_j u_i on each l block, as a banded row table (compact m-major rows)
    auto U = [&](int i, int l, int a, int b) {  // u_i[a][b] of block l (a, b in [-l, l])
        const size_t k = (size_t)(a + lmax) + M * ((size_t)(b + lmax) + M * ((size_t)l + (size_t)B * i));
        return std::complex<double>(u[2 * k], u[2 * k + 1]);
    };
    const int64_t nrc = p->nrc;
    const int TS = 5;  // u_1, u_2 bidiagonal, u_3 diagonal: bandwidth 2
    std::vector<int32_t> src(nrc * TS, -1);
    std::vector<double> coef(2 * nrc * TS, 0.0);
    for (int l = 0; l < B; l++) {
        const int S = 2 * l + 1;
        std::vector<std::complex<double>> Ml((size_t)S * S, 0.0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double g = G1[i + 3 * j];
                if (g == 0.0) continue;
                for (int a = -l; a <= l; a++)
                    for (int c = -l; c <= l; c++) {
                        std::complex<double> s = 0.0;
                        for (int k = -l; k <= l; k++) s += U(j, l, a, k) * U(i, l, k, c);
                        Ml[(a + l) + (size_t)S * (c + l)] += g * s;
                    }
            }
        for (int mm = -l; mm <= l; mm++)
            for (int n = -l; n <= l; n++) {
                const int64_t r = ridx_B(B, l, mm, n);
                int t = 0;
                for (int np = -l; np <= l; np++) {
                    const std::complex<double> c = Ml[(n + l) + (size_t)S * (np + l)];
                    if (std::abs(c) == 0.0) continue;
                    HPUP_REQUIRE(t < TS, HPUP_ERR_UNSUPPORTED, "u_i u_j wider than the n +- 2 band");
                    src[r * TS + t] = (int32_t)ridx_B(B, l, mm, np);
                    coef[2 * (r * TS + t)] = c.real();
                    coef[2 * (r * TS + t) + 1] = c.imag();
                    t++;
                }
            }
    }
    DevBuf srcd = upload(src), coefd = upload(coef);
    const int64_t nfull = nrot * nx;
    DevBuf Ff(sizeof(double) * 2 * nfull), Fc(sizeof(double) * 2 * p->ncoef), F0(sizeof(double) * 2 * p->ncoef);
    HPUP_HIP(hipMemcpy(Ff.p, F, Ff.bytes, hipMemcpyHostToDevice));
    p->full_to_compact(Ff.p, Fc.p, nullptr);
    HPUP_HIP(hipMemcpy(F0.p, Fc.p, Fc.bytes, hipMemcpyDeviceToDevice));
    p->step(Fc.p, 1, nullptr);  // F + dt (kinematics + bias random walk)
    launch_row_mix_add<double>(F0.p, Fc.p, srcd.as<int32_t>(), coefd.p, TS, nrc, nx, dt, nullptr);
    HPUP_HIP(hipGetLastError());
    p->compact_to_full(Fc.p, Ff.p, nullptr);
    HPUP_HIP(hipMemcpy(Fnew, Ff.p, Ff.bytes, hipMemcpyDeviceToHost));
}

}  // namespace hpup

extern "C" {

int hpup_pendulum_propagate_reduced_den(const double* F, const int64_t* dims, const double* f, const double* X,
                                        const double* mR, const double* bt, const double* G, double dtt, double L,
                                        const double* u, const double* d, const double* w, const char* method,
                                        double* Fnew) {
    return guard([&] {
        HPUP_REQUIRE(F && dims && f && X && mR && bt && G && u && d && w && method && Fnew, HPUP_ERR_ARG,
                     "null argument");
        const int meth = hpup_parse_method(method);
        HPUP_REQUIRE(meth >= 0, HPUP_ERR_METHOD, std::string("unknown method '") + method + "'");
        const int B = (int)dims[2], M = 2 * B - 1, N = (int)dims[3];
        HPUP_REQUIRE(dims[0] == M && dims[1] == M && dims[4] == N, HPUP_ERR_ARG, "F must be (2B-1,2B-1,B,N,N)");
        const int64_t nx = (int64_t)N * N, nrot = (int64_t)M * M * B, nR = 8LL * B * B * B;
        // the operator without the Clebsch-Gordan gravity term (MR = 0)
        std::vector<double> MR0(2 * nrot * 2, 0.0);
        std::vector<const double*> cptr((size_t)B * B, nullptr);
        double btd[3] = {bt[0], bt[1], 0}, Gd[9] = {G[0], G[1], G[2], G[3]};
        DriftSpec spec = build_drift_spec_from_arrays(B, 2, N, L, X, nullptr, MR0.data(), btd, Gd, u, cptr.data());
        hpup_model m{};
        m.B = B; m.dim = 2; m.N = N; m.L = L; m.dt = dtt; m.method = meth; m.precision = HPUP_FP64;
        m.b[0] = btd[0]; m.b[1] = btd[1];
        for (int i = 0; i < 4; i++) m.G[i] = Gd[i];
        std::unique_ptr<PlanBase> p(make_plan(m, d, w, &spec));
        auto* q = static_cast<PlanT<double>*>(p.get());
        q->den = true;
        q->mR_dev = upload(std::vector<double>(mR, mR + 2 * nR));
        DevBuf fd(sizeof(double) * nR * nx), Ff(sizeof(double) * 2 * nrot * nx), Fc(sizeof(double) * 2 * p->ncoef);
        HPUP_HIP(hipMemcpy(fd.p, f, fd.bytes, hipMemcpyHostToDevice));
        HPUP_HIP(hipMemcpy(Ff.p, F, Ff.bytes, hipMemcpyHostToDevice));
        p->full_to_compact(Ff.p, Fc.p, nullptr);
        q->den_f = fd.p;
        p->step(Fc.p, 1, nullptr);
        p->compact_to_full(Fc.p, Ff.p, nullptr);
        HPUP_HIP(hipMemcpy(Fnew, Ff.p, Ff.bytes, hipMemcpyDeviceToHost));
    });
}

int hpup_pendulum_propagate_reduced(const double* F, const int64_t* dims, const double* X, const double* MR,
                                    const double* bt, const double* G, double dtt, double L, const double* u,
                                    const double* const* CG, const char* method, double* Fnew) {
    return guard([&] { propagate_gateway<double>(2, F, dims, X, nullptr, MR, bt, G, dtt, L, u, CG, method, Fnew); });
}
int hpup_pendulum_propagate_reduced32(const float* F, const int64_t* dims, const float* X, const float* MR,
                                      const float* bt, const float* G, float dtt, float L, const float* u,
                                      const float* const* CG, const char* method, float* Fnew) {
    return guard([&] { propagate_gateway<float>(2, F, dims, X, nullptr, MR, bt, G, dtt, L, u, CG, method, Fnew); });
}
int hpup_gyrobias_propagate(const double* F, const int64_t* dims, const double* X, double dt, const double* u,
                            const double* G1, const double* G2, double* Fnew) {
    return guard([&] { gyrobias_gateway(F, dims, X, dt, u, G1, G2, Fnew); });
}
int hpup_pendulum_propagate(const double* F, const int64_t* dims, const double* X, const double* OJO,
                            const double* MR, const double* bt, const double* G, double dtt, double L,
                            const double* u, const double* const* CG, const char* method, double* Fnew) {
    return guard([&] { propagate_gateway<double>(3, F, dims, X, OJO, MR, bt, G, dtt, L, u, CG, method, Fnew); });
}
int hpup_pendulum_propagate32(const float* F, const int64_t* dims, const float* X, const float* OJO,
                              const float* MR, const float* bt, const float* G, float dtt, float L, const float* u,
                              const float* const* CG, const char* method, float* Fnew) {
    return guard([&] { propagate_gateway<float>(3, F, dims, X, OJO, MR, bt, G, dtt, L, u, CG, method, Fnew); });
}

}  // extern "C"
