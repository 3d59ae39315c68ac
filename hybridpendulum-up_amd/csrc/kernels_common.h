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
// Streaming (non-temporal) access for data touched once per kernel, so that
// it does not evict the rows other workgroups still re-read from L2.
template <class T> __device__ __forceinline__ cT<T> ld_stream(const cT<T>* p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return cmk<T>(v.x, v.y);
}
template <class T> __device__ __forceinline__ void st_stream(cT<T>* p, cT<T> a) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 v;
    v.x = a.x;
    v.y = a.y;
    __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
}
template <class T> __device__ __forceinline__ void cfma(cT<T>& acc, cT<T> a, cT<T> b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
// i * s * a  (multiplication by the imaginary scalar i s)
template <class T> __device__ __forceinline__ cT<T> cimul(cT<T> a, T s) { return cmk<T>(-a.y * s, a.x * s); }

// m-major compact rotation index (same formula as the host's ridx_B)
__device__ __forceinline__ int64_t d_sumsq(int64_t n) { return n * (n + 1) * (2 * n + 1) / 6; }
__device__ __forceinline__ int64_t d_moff(int B, int m) {
    const int lmax = B - 1;
    const int64_t s = m <= 0 ? d_sumsq(lmax) - d_sumsq(-m) : d_sumsq(lmax) + d_sumsq(m - 1);
    return (int64_t)(m + lmax) * B * B - s;
}
__device__ __forceinline__ int64_t d_ridx(int B, int l, int m, int n) {
    return d_moff(B, m) + (int64_t)l * l - (int64_t)m * m + (n + l);
}
// inverse of d_ridx
__device__ inline void d_rlmn(int B, int64_t r, int& l, int& m, int& n) {
    const int lmax = B - 1;
    int mm = -lmax;
    while (mm < lmax && d_moff(B, mm + 1) <= r) mm++;
    const int64_t o = r - d_moff(B, mm);
    int ll = mm < 0 ? -mm : mm;
    while (ll < lmax && (int64_t)(ll + 1) * (ll + 1) - (int64_t)mm * mm <= o) ll++;
    n = (int)(o - ((int64_t)ll * ll - (int64_t)mm * mm)) - ll;
    m = mm;
    l = ll;
}

}  // namespace hpup
