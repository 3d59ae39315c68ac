// MFMA 16x16x4 wrappers (FP32 and FP64) shared by the drift and transform kernels.
#pragma once

#include <hip/hip_runtime.h>

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

}  // namespace hpup
