// Forward-mode dual numbers for the adjoint kernels: a value and its derivatives along N directions.  The
// adjoint sweep differentiates each particle's functions (deposit, deposit_S, the Adams-Bashforth forces, the
// beam kick) by evaluating them on Dual<N> over the particle's state or over the fields it gathers; the
// derivative of a sum lambda . f(state) along each direction is one component of the particle's adjoint.
#pragma once

#include <hip/hip_runtime.h>

namespace qz {

template <int N> struct Dual {
    double v;
    double d[N];
    __host__ __device__ __forceinline__ Dual() : v(0) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0;
    }
    __host__ __device__ __forceinline__ Dual(double a) : v(a) {   // constants promote
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0;
    }
    __host__ __device__ __forceinline__ static Dual var(double a, int k) {
        Dual r(a);
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = i == k ? 1.0 : 0.0;
        return r;
    }
};

#define QZ_DUAL_LOOP(expr)           \
    _Pragma("unroll") for (int i = 0; i < N; ++i) expr

template <int N> __host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v); QZ_DUAL_LOOP(r.d[i] = a.d[i] + b.d[i];) return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v); QZ_DUAL_LOOP(r.d[i] = a.d[i] - b.d[i];) return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v); QZ_DUAL_LOOP(r.d[i] = -a.d[i];) return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v); QZ_DUAL_LOOP(r.d[i] = a.d[i] * b.v + a.v * b.d[i];) return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v);
    const double ib = 1.0 / b.v;
    QZ_DUAL_LOOP(r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;)
    return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) {
    Dual<N> r = a; r.v += b; return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) { return b + a; }
template <int N> __host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) {
    Dual<N> r = a; r.v -= b; return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) {
    Dual<N> r(a - b.v); QZ_DUAL_LOOP(r.d[i] = -b.d[i];) return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r(a.v * b); QZ_DUAL_LOOP(r.d[i] = a.d[i] * b;) return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N> __host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
    return a * (1.0 / b);
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) {
    Dual<N> r(a / b.v);
    const double s = -r.v / b.v;
    QZ_DUAL_LOOP(r.d[i] = s * b.d[i];)
    return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> dsqrt_dual(const Dual<N>& a) {
    Dual<N> r(sqrt(a.v));
    const double h = r.v > 0 ? 0.5 / r.v : 0.0;
    QZ_DUAL_LOOP(r.d[i] = a.d[i] * h;)
    return r;
}
#undef QZ_DUAL_LOOP

__host__ __device__ __forceinline__ double vsqrt(double a) { return sqrt(a); }
template <int N> __host__ __device__ __forceinline__ Dual<N> vsqrt(const Dual<N>& a) { return dsqrt_dual(a); }
__host__ __device__ __forceinline__ double val(double a) { return a; }
template <int N> __host__ __device__ __forceinline__ double val(const Dual<N>& a) { return a.v; }

}  // namespace qz
