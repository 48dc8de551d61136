// Laser envelope of the quasi-static step on the device (src/Laser.hpp, src/Laser.cpp): the complex envelope
// a^ of one box, K slices x M nodes x (Re, Im) doubles in HBM, slice 0 at the head of the box, advanced by one
// time step per call.
//
//   - laser_pond_kernel (Laser::prepare_slice, src/Laser.cpp:292-315): <a^2>, d<a^2>/dr and d<a^2>/dxi of every
//     (slice, node) at time level n.  Each depends only on a^n, so all slices run at once: one thread per
//     (slice, node), HBM-bound (reads 48 B of envelope around the node, writes 24 B).
//   - laser_factor_kernel and laser_march_kernel (Laser::advance_slice, :317-419): the Crank-Nicolson step in t of
//     every slice, marched from the head of the box.  Slice k needs X = a^(n+1) - a^n of slices k-1 and k-2, so
//     the march is serial in xi and runs in one workgroup: per slice the workgroup forms the right-hand sides of
//     all nodes in LDS, one lane runs the right-hand side's Thomas recurrence and the back substitution (serial
//     chains, as in the reference), and the workgroup forms the carry and a^(n+1).  The pivots and c' of the
//     recurrence depend only on the slice's susceptibility, so the factor kernel forms them beforehand for all
//     slices at once, one thread per slice, and the march's serial chain holds one complex division per node
//     instead of two.  The carry and the right-hand sides never leave LDS.
//
// The arithmetic of each node follows the reference's expressions in the same order (no contraction), so the
// result equals the serial CPU restatement bit for bit.  Scope: laser.l = 0 (no vortex), no ionization losses,
// one rank that owns the whole box (the guard rows of the head of the box are zero).
#include "qz_internal.hpp"
#include "host/simulation.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>

namespace {

using qz::Real;

struct Cx { double re, im; };
// src/Laser.cpp:13-21
__device__ __forceinline__ Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx csub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cx cscale(double s, Cx a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ Cx cdiv(Cx a, Cx b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

constexpr int kPondThreads = 256;
constexpr int kFactorThreads = 64;
constexpr int kMarchThreads = 512;
// LDS doubles per node of the march: c' (2), pivot (2), rhs / d' / X (2), the carry (4), -a_j (1)
constexpr int kMarchLdsPerNode = 11;
constexpr long kMaxLds = 160 * 1024;
constexpr int kMaxNodes = 1861;   // 11 doubles per node in 160 KB

// the diagonal's constant part, cd 2/(dt dxi) - 4 i k0 / dt (src/Laser.cpp:352-361)
__device__ __forceinline__ Cx march_alpha(double dxi, double dt, double k0, int scheme) {
    const double cd = scheme == 0 ? 4.0 : (scheme == 1 ? 3.0 : 2.0);
    const double s0 = 2.0 / (dt * dxi);
    const double s1 = cd * s0, s2 = 4.0 / dt;
    return Cx{s1, -s2 * k0};
}

__device__ __forceinline__ void pond_node(long i, const double* __restrict__ a, double* __restrict__ pond,
                                          const double* __restrict__ gm, const double* __restrict__ g0,
                                          const double* __restrict__ gp, int M, double fac, double idxi) {
    const int k = static_cast<int>(i / M), j = static_cast<int>(i - static_cast<long>(k) * M);
    const int N = M - 1;
    auto mag = [&](int l, int jj) {   // <a^2> / fac at (slice l, node jj)
        const double* p = a + (static_cast<long>(l) * M + jj) * 2;
        return p[0] * p[0] + p[1] * p[1];
    };
    const double m0 = fac * mag(k, j);
    const double m1 = k >= 1 ? fac * mag(k - 1, j) : 0.0;
    const double m2 = k >= 2 ? fac * mag(k - 2, j) : 0.0;
    // FieldSolver::gradient of <a^2> (src/FieldSolver.cpp:193-201): one-sided at the wall, zero on the axis
    double d;
    if (j == 0) d = 0.0;
    else if (j == N) d = g0[N] * m0 + gm[N] * (fac * mag(k, N - 1)) + gp[N] * (fac * mag(k, N - 2));
    else d = gm[j] * (fac * mag(k, j - 1)) + g0[j] * m0 + gp[j] * (fac * mag(k, j + 1));
    double* q = pond + i * 3;
    q[0] = m0;
    q[1] = d;
    q[2] = (3.0 * m0 - 4.0 * m1 + m2) * 0.5 * idxi;
}

// The coefficient half of each slice's Thomas recurrence, for every slice at once (one thread per slice): the
// pivots m_j = diag_j + a_j c'_{j-1} and c'_j = -c_j / m_j of src/Laser.cpp:385-393 depend on the slice's
// susceptibility and the operator, not on the carry, so only the right-hand side's recurrence stays in the
// march.  fac: (K, N, 4) doubles, m.re, m.im, c'.re, c'.im of (slice k, node j): slice-major, so that the march's
// workgroup reads a slice's rows with neighbouring addresses.
__device__ __forceinline__ void factor_slice(int k, const double* __restrict__ chi, const double* __restrict__ la,
                                             const double* __restrict__ lb, const double* __restrict__ lc,
                                             double* __restrict__ fac, int M, double dxi, double dt, double k0,
                                             int scheme) {
    const int N = M - 1;
    const Cx alpha = march_alpha(dxi, dt, k0, scheme);
    const double* ch = chi ? chi + static_cast<long>(k) * M : nullptr;
    Cx cprev{0.0, 0.0};
    for (int j = 0; j < N; ++j) {
        const double aj = j > 0 ? la[j] : 0.0, bj = lb[j], cj = lc[j];
        const Cx hd{bj - (ch ? ch[j] : 0.0), 0.0};
        const Cx diag = csub(alpha, hd);
        const double lo = -aj, up = -cj;
        const Cx m = csub(diag, cscale(lo, cprev));
        const Cx cpj = cdiv(Cx{up, 0.0}, m);
        double* f = fac + (static_cast<long>(k) * N + j) * 4;
        f[0] = m.re; f[1] = m.im; f[2] = cpj.re; f[3] = cpj.im;
        cprev = cpj;
    }
}

// One workgroup marches the K slices of one box.  Every barrier is reached by all threads in every slice.
__device__ __forceinline__ void march_box(const double* __restrict__ a, const double* __restrict__ chi,
                                          double* __restrict__ an, const double* __restrict__ la,
                                          const double* __restrict__ lb, const double* __restrict__ lc,
                                          const double* __restrict__ fac, int K, int M, double dxi, double dt,
                                          double k0, int scheme) {
    extern __shared__ double lds[];
    Cx* cpl = reinterpret_cast<Cx*>(lds);   // c' of the slice
    Cx* ml = cpl + M;                       // pivots m of the slice
    Cx* rhs = ml + M;                       // rhs, then d', then X
    Cx* car0 = rhs + M;                     // X of the previous slice
    Cx* car1 = car0 + M;                    // G of the previous slice (trapezoid) or X two slices back
    double* low = reinterpret_cast<double*>(car1 + M);   // -a_j (0 at j = 0)
    const int tid = threadIdx.x, N = M - 1;
    // src/Laser.cpp:352-361
    const double c1 = scheme == 0 ? 4.0 : (scheme == 1 ? 4.0 : 2.0);
    const double c2 = scheme == 1 ? -1.0 : 0.0;
    const double s0 = 2.0 / (dt * dxi);
    const double s2 = 4.0 / dt;
    for (int j = tid; j < M; j += kMarchThreads) {
        car0[j] = Cx{0.0, 0.0};
        car1[j] = Cx{0.0, 0.0};
        const double aj = j > 0 ? la[j] : 0.0;
        low[j] = -aj;
    }
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        const double* ak = a + static_cast<long>(k) * M * 2;
        const double* ch = chi ? chi + static_cast<long>(k) * M : nullptr;
        double* ank = an + static_cast<long>(k) * M * 2;
        auto A = [&](int j) { return Cx{ak[2 * j], ak[2 * j + 1]}; };
        // the right-hand side of node j: hist + 2 H a^n (src/Laser.cpp:371-384); the slice's pivots and c' into LDS
        const double* fk = fac + static_cast<long>(k) * N * 4;
        for (int j = tid; j < N; j += kMarchThreads) {
            const double* f = fk + static_cast<long>(j) * 4;
            ml[j] = Cx{f[0], f[1]};
            cpl[j] = Cx{f[2], f[3]};
            const double aj = j > 0 ? la[j] : 0.0, bj = lb[j], cj = lc[j];
            const Cx hd{bj - (ch ? ch[j] : 0.0), 0.0};
            Cx Ha = cmul(hd, A(j));
            if (j > 0) Ha = cadd(Ha, cscale(aj, A(j - 1)));
            Ha = cadd(Ha, cscale(cj, A(j + 1)));
            const Cx xp = car0[j], gp = car1[j];
            const Cx hist = scheme == 0 ? cadd(cscale(c1 * s0, xp), cscale(s2, gp))
                                        : cadd(cscale(c1 * s0, xp), cscale(c2 * s0, gp));
            rhs[j] = cadd(hist, cscale(2.0, Ha));
        }
        __syncthreads();
        // d'_j = (rhs_j + a_j d'_{j-1}) / m_j with the slice's pivots, then back substitution (src/Laser.cpp:385-399),
        // X_N = 0.  Node j +- 1's operands are read from LDS before node j is formed, so the chain does not wait
        // for LDS.
        if (tid == 0) {
            Cx dprev{0.0, 0.0};
            Cx m_n = ml[0], r_n = rhs[0];
            double lo_n = low[0];
            for (int j = 0; j < N; ++j) {
                const Cx m = m_n, r = r_n;
                const double lo = lo_n;
                if (j + 1 < N) {
                    m_n = ml[j + 1];
                    r_n = rhs[j + 1];
                    lo_n = low[j + 1];
                }
                const Cx dpj = cdiv(csub(r, cscale(lo, dprev)), m);
                rhs[j] = dpj;
                dprev = dpj;
            }
            rhs[N] = Cx{0.0, 0.0};
            Cx xn{0.0, 0.0};
            Cx c_n = cpl[N - 1], d_n = rhs[N - 1];
            for (int j = N - 1; j >= 0; --j) {
                const Cx cp = c_n, dp = d_n;
                if (j > 0) {
                    c_n = cpl[j - 1];
                    d_n = rhs[j - 1];
                }
                const Cx x = csub(dp, cmul(cp, xn));
                rhs[j] = x;
                xn = x;
            }
        }
        __syncthreads();
        // the carry for the next slice and a^(n+1) = a^n + X (src/Laser.cpp:400-418)
        for (int j = tid; j <= N; j += kMarchThreads) {
            const Cx X = rhs[j];
            const Cx an_j = A(j);
            if (scheme != 0) {
                car1[j] = car0[j];
                car0[j] = X;
            } else {
                Cx G{0.0, 0.0};
                if (j < N) {
                    auto Y = [&](int jj) {
                        const Cx x = rhs[jj];
                        return Cx{x.re + 2.0 * ak[2 * jj], x.im + 2.0 * ak[2 * jj + 1]};
                    };
                    const double aj = j > 0 ? la[j] : 0.0, bj = lb[j], cj = lc[j];
                    const Cx hd{bj - (ch ? ch[j] : 0.0), 0.0};
                    Cx HY = cmul(hd, Y(j));
                    if (j > 0) HY = cadd(HY, cscale(aj, Y(j - 1)));
                    HY = cadd(HY, cscale(cj, Y(j + 1)));
                    G = cadd(cmul(Cx{0.0, k0}, X), cscale(0.25 * dt, HY));
                }
                car0[j] = X;
                car1[j] = G;
            }
            ank[2 * j] = X.re + an_j.re;
            ank[2 * j + 1] = X.im + an_j.im;
        }
        __syncthreads();
    }
}

// One box of a step, the kernels' argument: its arrays, operator rows and constants.  active = 0: a box whose
// envelope is not advanced (dt <= 0); only its ponderomotive arrays are formed.
struct LaserBox {
    const double* a;
    const double* chi;
    double *an, *pond, *fac;
    const double *la, *lb, *lc, *gm, *g0, *gp;
    int K, M, scheme, active;
    double dxi, dt, k0, fac2, idxi;   // fac2: <a^2> / |a^|^2 (1 circular, 1/2 linear); idxi = 1 / dxi
};

// kBatch = false: one box, passed by value; true: box blockIdx.y (pond, factor) or blockIdx.x (march) of `many`
template <bool kBatch>
__global__ __launch_bounds__(kPondThreads) void laser_pond_kernel(LaserBox one, const LaserBox* __restrict__ many) {
    const LaserBox B = kBatch ? many[blockIdx.y] : one;
    const long i = static_cast<long>(blockIdx.x) * kPondThreads + threadIdx.x;
    if (i >= static_cast<long>(B.K) * B.M) return;
    pond_node(i, B.a, B.pond, B.gm, B.g0, B.gp, B.M, B.fac2, B.idxi);
}

template <bool kBatch>
__global__ __launch_bounds__(kFactorThreads) void laser_factor_kernel(LaserBox one,
                                                                      const LaserBox* __restrict__ many) {
    const LaserBox B = kBatch ? many[blockIdx.y] : one;
    const int k = blockIdx.x * kFactorThreads + threadIdx.x;
    if (!B.active || k >= B.K) return;
    factor_slice(k, B.chi, B.la, B.lb, B.lc, B.fac, B.M, B.dxi, B.dt, B.k0, B.scheme);
}

template <bool kBatch>
__global__ __launch_bounds__(kMarchThreads) void laser_march_kernel(LaserBox one, const LaserBox* __restrict__ many) {
    const LaserBox B = kBatch ? many[blockIdx.x] : one;
    if (!B.active) return;   // the whole workgroup, before any barrier
    march_box(B.a, B.chi, B.an, B.la, B.lb, B.lc, B.fac, B.K, B.M, B.dxi, B.dt, B.k0, B.scheme);
}

long march_lds(int M) { return static_cast<long>(kMarchLdsPerNode) * M * static_cast<long>(sizeof(double)); }

// glibc exp / cos / sin called through pointers, so that no compiler folds or replaces them: the pulse is
// formed from the C library's values at run time, as the CPU restatement forms it
double (*volatile host_exp)(double) = ::exp;
double (*volatile host_cos)(double) = ::cos;
double (*volatile host_sin)(double) = ::sin;

}  // namespace

struct qz_laser {
    qz::RadialGrid grid;
    qz::FieldSolver solver;
    int K = 0, M = 0, scheme = 1;
    bool circular = false;
    double xi_min = 0, dxi = 0, dt = 0, k0 = 0;
    qz::DevBuf<Real> a, an, pond, fac;
    hipStream_t own = nullptr;
    // a batched step led by this handle: the boxes' descriptors (kept until the next such step) and their copy
    std::vector<LaserBox> boxes;
    qz::DevBuf<char> dboxes;
    ~qz_laser() {
        if (own) (void)hipStreamDestroy(own);
    }
};

#define QZ_LASER_TRY(...)                                   \
    try {                                                   \
        __VA_ARGS__;                                        \
        return QZ_OK;                                       \
    } catch (const std::exception& e) {                     \
        return qz_fail(QZ_EINVAL, e.what());                \
    }

static void laser_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

extern "C" int qz_laser_create(const double* nodes, int M, int K, double xi_min, double dxi, double dt, double k0,
                               int scheme, int circular, qz_laser** out) {
    if (!nodes || !out) return qz_fail(QZ_EINVAL, "qz_laser_create: null argument");
    if (M < 3 || K < 1) return qz_fail(QZ_EINVAL, "qz_laser_create: need M >= 3 nodes and K >= 1 slices");
    if (scheme < 0 || scheme > 2)
        return qz_fail(QZ_EINVAL, "qz_laser_create: scheme 0 (trapezoid), 1 (bdf2) or 2 (euler)");
    if (!(dxi > 0)) return qz_fail(QZ_EINVAL, "qz_laser_create: dxi must be positive");
    if (!(k0 > 1)) return qz_fail(QZ_EINVAL, "qz_laser_create: k0 = omega0/omega_p must be > 1 (underdense plasma)");
    if (M > kMaxNodes || march_lds(M) > kMaxLds)
        return qz_fail(QZ_ERANGE, "qz_laser_create: the march holds at most 1861 nodes in LDS");
    QZ_LASER_TRY({
        auto L = std::make_unique<qz_laser>();
        laser_check(hipStreamCreateWithFlags(&L->own, hipStreamNonBlocking), "hipStreamCreate");
        L->grid.finalize(std::vector<Real>(nodes, nodes + M), L->own);
        L->solver.init(L->grid, L->own);
        L->K = K; L->M = M; L->scheme = scheme; L->circular = circular != 0;
        L->xi_min = xi_min; L->dxi = dxi; L->dt = dt; L->k0 = k0;
        const size_t n = static_cast<size_t>(K) * M;
        L->a.alloc(2 * n); L->an.alloc(2 * n); L->pond.alloc(3 * n);
        if (dt > 0) L->fac.alloc(4 * static_cast<size_t>(K) * (M - 1));
        laser_check(hipMemsetAsync(L->a.p, 0, sizeof(Real) * 2 * n, L->own), "hipMemsetAsync");
        laser_check(hipMemsetAsync(L->an.p, 0, sizeof(Real) * 2 * n, L->own), "hipMemsetAsync");
        laser_check(hipMemsetAsync(L->pond.p, 0, sizeof(Real) * 3 * n, L->own), "hipMemsetAsync");
        laser_check(hipStreamSynchronize(L->own), "hipStreamSynchronize");
        *out = L.release();
    })
}

extern "C" void qz_laser_destroy(qz_laser* L) { delete L; }

// Laser::Laser's Gaussian pulse (src/Laser.cpp:195-236), laser.l = 0: a^ = a0 / (1 + i s)
// exp(-r^2 / (w0^2 (1 + i s))) exp(-(xi - xi0)^2 / L0^2), s = -focus / Z_R, Z_R = k0 w0^2 / 2, zero at the wall
extern "C" int qz_laser_gaussian(qz_laser* L, double a0, double w0, double L0, double xi0, double focus) {
    if (!L) return qz_fail(QZ_EINVAL, "qz_laser_gaussian: null argument");
    if (!(w0 > 0) || !(L0 > 0)) return qz_fail(QZ_EINVAL, "qz_laser_gaussian: w0 and L0 must be positive");
    QZ_LASER_TRY({
        const int K = L->K, M = L->M;
        std::vector<Real> h(static_cast<size_t>(K) * M * 2);
        const Real zR = 0.5 * L->k0 * w0 * w0;
        const Real s = -focus / zR;
        for (int kl = 0; kl < K; ++kl) {
            const Real xi = L->xi_min + kl * L->dxi;
            const Real lon = host_exp(-(xi - xi0) * (xi - xi0) / (L0 * L0));
            for (int j = 0; j < M; ++j) {
                const Real r = L->grid.r[j];
                const Real den = 1 + s * s;
                const Real pre_re = 1 / den, pre_im = -s / den;
                const Real ex = host_exp(-r * r / (w0 * w0 * den)), ph = r * r * s / (w0 * w0 * den);
                const Real e_re = ex * host_cos(ph), e_im = ex * host_sin(ph);
                const Real amp = (j == M - 1) ? 0.0 : a0 * lon;
                Real* p = h.data() + (static_cast<size_t>(kl) * M + j) * 2;
                p[0] = amp * (pre_re * e_re - pre_im * e_im);
                p[1] = amp * (pre_re * e_im + pre_im * e_re);
            }
        }
        L->a.upload(h, L->own);
        laser_check(hipStreamSynchronize(L->own), "hipStreamSynchronize");
    })
}

extern "C" int qz_laser_set_envelope(qz_laser* L, const double* a_host) {
    if (!L || !a_host) return qz_fail(QZ_EINVAL, "qz_laser_set_envelope: null argument");
    QZ_LASER_TRY({
        laser_check(hipMemcpyAsync(L->a.p, a_host, sizeof(Real) * L->a.n, hipMemcpyHostToDevice, L->own),
                    "hipMemcpyAsync");
        laser_check(hipStreamSynchronize(L->own), "hipStreamSynchronize");
    })
}

extern "C" int qz_laser_get(qz_laser* L, int what, double* out_host) {
    if (!L || !out_host) return qz_fail(QZ_EINVAL, "qz_laser_get: null argument");
    if (what != 0 && what != 1) return qz_fail(QZ_EINVAL, "qz_laser_get: what is 0 (envelope) or 1 (ponderomotive)");
    const qz::DevBuf<Real>& b = what == 0 ? L->a : L->pond;
    QZ_LASER_TRY({
        laser_check(hipMemcpyAsync(out_host, b.p, sizeof(Real) * b.n, hipMemcpyDeviceToHost, L->own),
                    "hipMemcpyAsync");
        laser_check(hipStreamSynchronize(L->own), "hipStreamSynchronize");
    })
}

extern "C" int qz_laser_device(qz_laser* L, int what, double** ptr) {
    if (!L || !ptr) return qz_fail(QZ_EINVAL, "qz_laser_device: null argument");
    if (what != 0 && what != 1) return qz_fail(QZ_EINVAL, "qz_laser_device: what is 0 (envelope) or 1 (ponderomotive)");
    *ptr = what == 0 ? L->a.p : L->pond.p;
    return QZ_OK;
}

static LaserBox box_of(qz_laser* L, const double* chi) {
    LaserBox B{};
    B.a = L->a.p; B.chi = chi; B.an = L->an.p; B.pond = L->pond.p; B.fac = L->fac.p;
    B.la = L->solver.d_a[qz::FieldSolver::L0].p;
    B.lb = L->solver.d_b[qz::FieldSolver::L0].p;
    B.lc = L->solver.d_c[qz::FieldSolver::L0].p;
    B.gm = L->solver.d_gm.p; B.g0 = L->solver.d_g0.p; B.gp = L->solver.d_gp.p;
    B.K = L->K; B.M = L->M; B.scheme = L->scheme; B.active = L->dt > 0 ? 1 : 0;
    B.dxi = L->dxi; B.dt = L->dt; B.k0 = L->k0;
    B.fac2 = L->circular ? 1.0 : 0.5;
    B.idxi = 1.0 / L->dxi;
    return B;
}

static int set_march_lds(long lds) {
    QZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(laser_march_kernel<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    QZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(laser_march_kernel<true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    return QZ_OK;
}

// One time step of the envelope: prepare_slice of every slice (the ponderomotive arrays at level n), then the
// march of advance_slice from the head of the box; the advanced envelope becomes the current one
// (Laser::begin_step of the next step, src/Laser.cpp:258-262).  chi: (K, M) device array, or null for vacuum.
extern "C" int qz_laser_step(qz_laser* L, const double* chi, void* stream) {
    if (!L) return qz_fail(QZ_EINVAL, "qz_laser_step: null argument");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : L->own;
    const LaserBox B = box_of(L, chi);
    const long n = static_cast<long>(B.K) * B.M;
    hipLaunchKernelGGL(laser_pond_kernel<false>, dim3(static_cast<unsigned>((n + kPondThreads - 1) / kPondThreads)),
                       dim3(kPondThreads), 0, st, B, nullptr);
    QZ_HIP(hipGetLastError());
    if (B.active) {
        const int rc = set_march_lds(march_lds(B.M));
        if (rc != QZ_OK) return rc;
        hipLaunchKernelGGL(laser_factor_kernel<false>, dim3((B.K + kFactorThreads - 1) / kFactorThreads),
                           dim3(kFactorThreads), 0, st, B, nullptr);
        QZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(laser_march_kernel<false>, dim3(1), dim3(kMarchThreads), march_lds(B.M), st, B, nullptr);
        QZ_HIP(hipGetLastError());
        std::swap(L->a, L->an);
    }
    if (!stream) QZ_HIP(hipStreamSynchronize(st));
    return QZ_OK;
}

// Many boxes, one step each, in one launch per kernel: every box's march is one workgroup, so the boxes march
// at once.  Each box's results are those of qz_laser_step.
extern "C" int qz_laser_step_batched(qz_laser* const* Ls, int n, const double* const* chis, void* stream) {
    if (!Ls || n < 1) return qz_fail(QZ_EINVAL, "qz_laser_step_batched: no boxes");
    if (n > 65535) return qz_fail(QZ_ERANGE, "qz_laser_step_batched: at most 65535 boxes per call");
    for (int i = 0; i < n; ++i) {
        if (!Ls[i]) return qz_fail(QZ_EINVAL, "qz_laser_step_batched: null box");
        for (int j = 0; j < i; ++j)
            if (Ls[j] == Ls[i]) return qz_fail(QZ_EINVAL, "qz_laser_step_batched: a box appears twice");
    }
    qz_laser* L0 = Ls[0];
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : L0->own;
    QZ_LASER_TRY({
        // the descriptors of an earlier batched step on this stream may still be read: wait for it
        laser_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        L0->boxes.resize(static_cast<size_t>(n));
        long maxn = 1, lds = 0;
        int maxk = 1;
        for (int i = 0; i < n; ++i) {
            L0->boxes[i] = box_of(Ls[i], chis ? chis[i] : nullptr);
            const LaserBox& B = L0->boxes[i];
            maxn = std::max(maxn, static_cast<long>(B.K) * B.M);
            maxk = std::max(maxk, B.K);
            if (B.active) lds = std::max(lds, march_lds(B.M));
        }
        const size_t bytes = sizeof(LaserBox) * static_cast<size_t>(n);
        if (L0->dboxes.n < bytes) L0->dboxes.alloc(bytes);
        laser_check(hipMemcpyAsync(L0->dboxes.p, L0->boxes.data(), bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        const LaserBox* dev = reinterpret_cast<const LaserBox*>(L0->dboxes.p);
        hipLaunchKernelGGL(laser_pond_kernel<true>,
                           dim3(static_cast<unsigned>((maxn + kPondThreads - 1) / kPondThreads), n),
                           dim3(kPondThreads), 0, st, LaserBox{}, dev);
        laser_check(hipGetLastError(), "laser_pond_kernel");
        if (lds > 0) {
            if (set_march_lds(lds) != QZ_OK) throw std::runtime_error(qz_last_error());
            hipLaunchKernelGGL(laser_factor_kernel<true>, dim3((maxk + kFactorThreads - 1) / kFactorThreads, n),
                               dim3(kFactorThreads), 0, st, LaserBox{}, dev);
            laser_check(hipGetLastError(), "laser_factor_kernel");
            hipLaunchKernelGGL(laser_march_kernel<true>, dim3(n), dim3(kMarchThreads), lds, st, LaserBox{}, dev);
            laser_check(hipGetLastError(), "laser_march_kernel");
            for (int i = 0; i < n; ++i)
                if (L0->boxes[i].active) std::swap(Ls[i]->a, Ls[i]->an);
        }
        if (!stream) laser_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    })
}

// Laser::sums (src/Laser.cpp:421-438) of the current envelope, on the host in slice and node order:
// {sum |a|^2 dV, sum |a|^2 xi dV (the user's xi), sum |a|^2 r^2 dV, max |a|, sum |a|^2 xi^2 dV}
extern "C" int qz_laser_sums(qz_laser* L, double* out5) {
    if (!L || !out5) return qz_fail(QZ_EINVAL, "qz_laser_sums: null argument");
    QZ_LASER_TRY({
        std::vector<Real> h;
        L->a.download(h, L->own);
        const int K = L->K, M = L->M;
        double S[5] = {0, 0, 0, 0, 0};
        for (int kl = 0; kl < K; ++kl) {
            const Real xi = L->xi_min + kl * L->dxi;
            for (int j = 0; j < M; ++j) {
                const Real* p = h.data() + (static_cast<size_t>(kl) * M + j) * 2;
                const Real m2 = p[0] * p[0] + p[1] * p[1];
                const Real dv = L->grid.V[j] * L->dxi;
                S[0] += m2 * dv;
                S[1] += m2 * (Real(0) - xi) * dv;
                S[2] += m2 * L->grid.r[j] * L->grid.r[j] * dv;
                S[3] = std::max(S[3], std::sqrt(m2));
                S[4] += m2 * xi * xi * dv;
            }
        }
        std::memcpy(out5, S, sizeof(S));
    })
}
