#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hpup {
template <class T>
void launch_marginals(const void* f, int B, int64_t nx, const double* w_dev, double* fx, double* fR, hipStream_t st);
}
