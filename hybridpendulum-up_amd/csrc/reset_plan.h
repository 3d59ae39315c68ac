// Device-resident hybrid reset of a plan (see reset_plan.hip).
#pragma once

#include <vector>

#include "hpup_internal.h"

namespace hpup {

constexpr int kResetMaxN = 64;

// Launch constants of the reset kernels (by value).
struct ResetConsts {
    int N;
    double xg[kResetMaxN];  // omega grid values along an axis
    double x0, hx, ihx;     // 3-D: grid origin, spacing and its reciprocal
    double xlo, xhi;        // 2-D: getOmega's box of valid pull-backs
    double Lg, dx2;         // 2-D: getIndRule's box length and cell area
    double kfac;            // (1 + eps) / eps
    double J;               // jump-term Jacobian factor (1/eps or 1)
    double ew[3];           // wall normal
};

struct ResetGeom {
    int dim = 0, N = 0, numR = 0;
    int64_t nR = 0, nx = 0;
    hpup_reset_params params{};
    ResetConsts consts{};
    DevBuf indR, geo;             // region rows and their geometry (FP64, structure of arrays)
    std::vector<int32_t> region;  // host copy of the region rows
    void build(int B, int dim, int N, double L, const hpup_reset_params& p);
    // dfc (numR, nx) = df on the region's rows
    template <class T> void df(const void* f, void* dfc, hipStream_t st) const;
    // out = base + scale dfc on the region's rows (base may be out); base
    // nullptr: out = scale dfc, zero elsewhere
    template <class T> void apply(const void* dfc, const void* base, void* out, double scale, hipStream_t st) const;
    // whether a row of `per` bytes per omega point fits the row-group kernels
    bool rows_fit(size_t per) const;
    template <class T> void zero(void* out, hipStream_t st) const;
    // Row-group forms (LDS-staged rows; false when one row does not fit):
    // fout's region rows = (accumulate ? fin : 0) + scale df(fin) (fout may be fin)
    template <class T> bool rows_pass(const void* fin, void* fout, double scale, bool accumulate, hipStream_t st) const;
    // out's region rows = g + scale R^H g (out may be g)
    template <class T> bool adjoint_rows(const void* g, void* out, double scale, hipStream_t st) const;
    // acc (numR, nx) = the transpose of the jump term applied to the grid g
    template <class T> void adjoint(const void* g, void* acc, hipStream_t st) const;
    // out[0] += scale <g, d(jump term)/d eps f> over the region (part: eps_grad_partials() doubles)
    template <class T>
    void eps_grad(const void* f, const void* g, double* part, double scale, double* out, hipStream_t st) const;
    int64_t eps_grad_partials() const;
};

}  // namespace hpup
