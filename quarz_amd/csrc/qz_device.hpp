// Device-side physics shared by the gfx950 kernels of the quasi-static step.
//
// Every expression keeps the operation order of the reference (file:line cited at each function) and
// the whole library is built with -ffp-contract=off: no multiply-add is fused, divisions and square roots
// are IEEE correctly rounded, so the kernels reproduce a serial CPU build of the same arithmetic bit for
// bit.  Component enums follow src/FieldStore.hpp:9-21 and src/Simulation.hpp:46.
#pragma once

#include <hip/hip_runtime.h>

namespace qz {

using Real = double;

enum FieldComp : int {
    F_EP1R = 0, F_EP1I, F_BP1R, F_BP1I, F_EZ0, F_BZ0,
    F_EP0R, F_EP0I, F_EP2R, F_EP2I, F_BP0R, F_BP0I, F_BP2R, F_BP2I,
    F_EZ1R, F_EZ1I, F_BZ1R, F_BZ1I,
    F_NC0 = 6, F_NC1 = 18
};
enum BeamComp : int {
    B_RT0 = 0, B_JZ0, B_JP1R, B_JP1I, B_RHO0,
    B_RT1R, B_RT1I, B_JZ1R, B_JZ1I, B_JP0R, B_JP0I, B_JP2R, B_JP2I, B_RHO1R, B_RHO1I,
    B_NC0 = 5, B_NC1 = 15
};
enum DiagComp : int { D_PSI0 = 0, D_NE0, D_NI0, D_PSI1R, D_PSI1I, D_NE1R, D_NE1I, D_NI1R, D_NI1I, D_NC0 = 3, D_NC1 = 9 };

// ------------------------------------------------------------------------------------ vectors (src/Pushers.hpp:24-32)
struct V3 { Real x, y, z; };
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(Real s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ Real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ Real dsqrt(Real x) { return __builtin_sqrt(x); }

// src/Pushers.hpp:36-39
__device__ __forceinline__ Real gamma2_root(Real x, Real c) {
    const Real s = dsqrt(x * x + c);
    return x >= Real(0) ? x + s : (s > -x ? c / (s - x) : Real(1));
}
// src/Pushers.hpp:44-52
__device__ __forceinline__ V3 solve_implicit(V3 a, V3 b) {
    const Real b2 = dot(b, b), ab = dot(a, b);
    const Real g2 = gamma2_root(Real(0.5) * (Real(1) + dot(a, a) - b2), b2 + ab * ab);
    const Real g = dsqrt(g2);
    const V3 axb = cross(a, b);
    const Real inv = Real(1) / (g2 + b2);
    return {inv * (g2 * a.x + g * axb.x + ab * b.x), inv * (g2 * a.y + g * axb.y + ab * b.y),
            inv * (g2 * a.z + g * axb.z + ab * b.z)};
}
// src/Pushers.hpp:54-61
__device__ __forceinline__ V3 boris(V3 u, V3 E, V3 B, Real eps) {
    const V3 um = u + eps * E;
    const Real g = dsqrt(Real(1) + dot(um, um));
    const V3 t = (eps / g) * B;
    const V3 s = (Real(2) / (Real(1) + dot(t, t))) * t;
    const V3 up = um + cross(um + cross(um, t), s);
    return up + eps * E;
}
// src/Pushers.hpp:63-68
__device__ __forceinline__ V3 vay(V3 u, V3 E, V3 B, Real eps) {
    const Real g0 = dsqrt(Real(1) + dot(u, u));
    const V3 up = u + (Real(2) * eps) * E + (eps / g0) * cross(u, B);
    return solve_implicit(up, eps * B);
}
// src/Pushers.hpp:70-80
__device__ __forceinline__ V3 higuera_cary(V3 u, V3 E, V3 B, Real eps) {
    const V3 um = u + eps * E;
    const V3 tau = eps * B;
    const Real ustar = dot(um, tau);
    const Real sigma = Real(1) + dot(um, um) - dot(tau, tau);
    const Real gp = dsqrt(Real(0.5) * (sigma + dsqrt(sigma * sigma + Real(4) * (dot(tau, tau) + ustar * ustar))));
    const V3 t = (Real(1) / gp) * tau;
    const Real s = Real(1) / (Real(1) + dot(t, t));
    const V3 upl = s * (um + dot(um, t) * t + cross(um, t));
    return upl + eps * E + cross(upl, t);
}
// src/Pushers.hpp:83-86
__device__ __forceinline__ V3 imp(V3 u, V3 E, V3 B, Real eps) {
    const V3 ubar = solve_implicit(u + eps * E, eps * B);
    return Real(2) * ubar - u;
}
// pusher selection as Beam::advance (src/Beam.cpp:1062-1071): 0 vay, 1 boris, 2 hc, 3 imp
__device__ __forceinline__ V3 push_u(int mode, V3 u, V3 E, V3 B, Real eps) {
    switch (mode) {
        case 1: return boris(u, E, B, eps);
        case 2: return higuera_cary(u, E, B, eps);
        case 3: return imp(u, E, B, eps);
        default: return vay(u, E, B, eps);
    }
}

// ------------------------------------------------------------------------------------ radial grid (src/RadialGrid.hpp:35-50)
__device__ __forceinline__ int grid_locate(const Real* __restrict__ r, int N, Real R, Real x) {
    if (x <= Real(0)) return 0;
    if (x >= R) return N - 1;
    int lo = 0, hi = N;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (r[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ void grid_weights(const Real* __restrict__ r, const Real* __restrict__ inv_h, int N, Real R,
                                             Real x, int& j, Real& t) {
    j = grid_locate(r, N, R, x);
    t = (x - r[j]) * inv_h[j];
    t = t < Real(0) ? Real(0) : (t > Real(1) ? Real(1) : t);
}

// ------------------------------------------------------------------------------------ mode sums (src/SliceData.hpp:70-83)
__device__ __forceinline__ Real eval_scalar(Real f0, Real fr, Real fi, Real c, Real s, bool m1) {
    return m1 ? f0 + Real(2) * (fr * c - fi * s) : f0;
}
__device__ __forceinline__ void eval_vector(Real r0, Real i0, Real r1, Real i1, Real r2, Real i2, Real c, Real s,
                                            bool m1, Real& vx, Real& vy) {
    vx = r1 * c - i1 * s;
    vy = r1 * s + i1 * c;
    if (m1) {
        const Real c2 = c * c - s * s, s2 = Real(2) * c * s;
        vx += r0 + r2 * c2 - i2 * s2;
        vy += i0 + r2 * s2 + i2 * c2;
    }
}

}  // namespace qz
