// Device-side helpers shared by the HIP kernels (complex arithmetic on
// interleaved re,im pairs in FP64 or FP32).
#pragma once

#include <hip/hip_runtime.h>

namespace hpup {

template <class T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };
template <class T> using cT = typename vec2<T>::type;

template <class T> __device__ __forceinline__ cT<T> cmk(T a, T b) { cT<T> r; r.x = a; r.y = b; return r; }
template <class T> __device__ __forceinline__ cT<T> czero() { return cmk<T>(T(0), T(0)); }
template <class T> __device__ __forceinline__ cT<T> cadd(cT<T> a, cT<T> b) { return cmk<T>(a.x + b.x, a.y + b.y); }
template <class T> __device__ __forceinline__ cT<T> csub(cT<T> a, cT<T> b) { return cmk<T>(a.x - b.x, a.y - b.y); }
template <class T> __device__ __forceinline__ cT<T> cscale(cT<T> a, T s) { return cmk<T>(a.x * s, a.y * s); }
template <class T> __device__ __forceinline__ cT<T> cmul(cT<T> a, cT<T> b) {
    return cmk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T> __device__ __forceinline__ cT<T> cconj(cT<T> a) { return cmk<T>(a.x, -a.y); }
// acc += a * b
template <class T> __device__ __forceinline__ void cfma(cT<T>& acc, cT<T> a, cT<T> b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
// i * s * a  (multiplication by the imaginary scalar i s)
template <class T> __device__ __forceinline__ cT<T> cimul(cT<T> a, T s) { return cmk<T>(-a.y * s, a.x * s); }

}  // namespace hpup
