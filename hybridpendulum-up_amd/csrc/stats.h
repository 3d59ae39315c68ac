#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hpup {

// Launch shape and work-space size of the one-pass grid statistics.
struct StatShape {
    int64_t nR, nx, ntile, nchunk, xper, nfx, work_doubles;
    int dim;
};
StatShape stat_shape(int B, int dim, int64_t nx);

// fR[nR], gR[dim][nR], fx[nx] (device, FP64) from the device grid f (plan
// precision T); xg[nx][dim] omega coordinates, w[2B] beta weights; work has
// work_doubles entries.
template <class T>
void launch_stat_pass(const StatShape& s, const void* f, int B, const double* w, const double* xg, double* work,
                      double* fR, double* gR, double* fx, hipStream_t st);

}  // namespace hpup
