// Rotation marginal and point evaluation of the SO(3) inverse transform
// (pendulum_plot_getc).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hpup {

// fR[iR] = scale * sum_x f[iR + nR x]
void launch_rot_marginal(const double* f, int64_t nR, int64_t nx, double scale, double* fR, hipStream_t st);
// out[p] = sum_l (2l+1) sum_{m,n} Re(F^l_mn D^l_mn(e[3p], e[3p+1], e[3p+2])), F compact FP64 (m-major),
// sqb[l (2B-1) + m + l] = sqrt((2l)! / ((l+m)! (l-m)!))
void launch_point_eval(const void* F, int B, const double* sqb, const double* e, int64_t P, double* out,
                       hipStream_t st);

}  // namespace hpup
