// The adjoint of one m = 0 time step, for the differentiable witness profile: the reverse xi sweep of the
// plasma (qz_sweep_adjoint), and the adjoints of the beam kick, the beam deposit and the energy-spread objective.
// The reference has no adjoint; its forward is Simulation::step_fields (src/Simulation.cpp:636-795) and
// Beam::advance (src/Beam.cpp:948-1131), and these kernels are reverse-mode differentiation of that forward as
// the sweep kernel computes it:
//
//   - the sweep walks the slices backwards, one workgroup per run.  Each slice's plasma state comes from a
//     recomputation of its checkpoint segment (qz_sweep with a trajectory buffer); the slice's fields from the
//     forward's field array.  The node adjoints of the slice live in LDS;
//   - the per-particle functions (deposit, deposit_S, the Adams-Bashforth forces) are differentiated in forward
//     mode with dual numbers (qz_dual.hpp): once over the particle's five state variables (the gather follows
//     the particle's position), once over the fields it gathers, whose adjoints scatter to the cell's two nodes
//     by LDS atomics;
//   - the node-level linear algebra by its transpose: transposed Thomas solves (the chi-dependent diagonal of
//     the B+ operators from the slice's recomputed chi), transposed gradient stencils, the transposed dxi
//     difference of the beam sources;
//   - the recurrence x_{k+1} = x_k + dxi sum_m c_m f_{k-m} (the first push fills every history slot with f_0)
//     by a ring of the state adjoints of the next `order` slices, per particle.
//
// Removed (QSA) and wall-reflected plasma particles are fixed events; node sums are atomic (no fixed order):
// the adjoint is a floating-point derivative, not a bit-exact restatement.
#include <hip/hip_runtime.h>

#include "qz_device.hpp"
#include "qz_dual.hpp"
#include "qz_internal.hpp"
#include "qz_sweep_common.hpp"

namespace qz {
namespace {

constexpr int kAdjThreads = 512;
constexpr int kAdjLdsBytes = 160 * 1024;

// LDS arrays (Mp doubles each): grid, the slice's fields, the node adjoints, the slice's chi
enum AdjLds : int {
    L_R = 0, L_IH, F_WR, F_EZ, F_BZ, F_BR, F_BI,
    A_WR, A_EZ, A_BZ, A_BR, A_BI, A_SR, A_SI, A_CHI, A_JZ, A_RHOT, A_DRHO, A_JPI,
    S_CHI, L_NARR
};
// the solves' arrays (M doubles each): two buffers of the reduced system (sub-diagonal, diagonal,
// super-diagonal, right-hand side) of the parallel cyclic reduction, the solutions of the slice's five solves
// (two at a time) and the psi right-hand side.  In LDS after the arrays above when the grid leaves room, else in
// the global workspace; then (global) the particles' partial adjoints (5 x NP)
enum AdjSolve : int {
    P_A0 = 0, P_B0, P_C0, P_D0, P_A1, P_B1, P_C1, P_D1,
    T_MU0, T_MU1, T_MU2, T_RHS, S_NARR
};

// ---------------------------------------------------------------- particle functions over any number type
template <class T> struct AGeo {
    T r, c, s, W0, W1, Wor0, Wor1;
    double dW0, dW1;
    int j;
};
// PlasmaSpecies geometry (src/PlasmaSpecies.cpp:52-89)
template <class T>
__device__ __forceinline__ AGeo<T> ageo(const double* r_, const double* ih_, int N, double R, const T& x, const T& y) {
    AGeo<T> p;
    p.r = vsqrt(x * x + y * y);
    const double rv = val(p.r);
    if (rv > 0) { p.c = x / p.r; p.s = y / p.r; } else { p.c = T(1.0); p.s = T(0.0); }
    int j;
    double t0;
    grid_weights(r_, ih_, N, R, rv, j, t0);
    const double ih = ih_[j];
    T t = (p.r - r_[j]) * ih;
    if (val(t) < 0) t = T(0.0);
    else if (val(t) > 1) t = T(1.0);
    p.j = j;
    p.W0 = 1.0 - t;
    p.W1 = t;
    p.dW0 = -ih;
    p.dW1 = ih;
    const T ir = rv > 0 ? 1.0 / p.r : T(0.0);
    p.Wor0 = j == 0 ? T(0.0) : p.W0 * ir;
    p.Wor1 = j == 0 ? T(ih) : p.W1 * ir;
    return p;
}
template <class T, class G> __device__ __forceinline__ AGeo<T> promote(const AGeo<G>& g) {
    AGeo<T> p;
    p.r = T(val(g.r)); p.c = T(val(g.c)); p.s = T(val(g.s)); p.W0 = T(val(g.W0)); p.W1 = T(val(g.W1));
    p.Wor0 = T(val(g.Wor0)); p.Wor1 = T(val(g.Wor1)); p.dW0 = g.dW0; p.dW1 = g.dW1; p.j = g.j;
    return p;
}
template <class T> __device__ __forceinline__ T alin(const double* f, const AGeo<T>& p) {
    return p.W0 * f[p.j] + p.W1 * f[p.j + 1];
}
template <class T> struct AGath { T Wx, Wy, Ez, Bz, Bx, By; };
// the m = 0 gather of PlasmaSpecies (src/SliceData.hpp:70-83): W+ and B+ mode 1 (W+'s imaginary part is zero)
template <class T>
__device__ __forceinline__ AGath<T> agather(const double* wr, const double* ez, const double* bz, const double* br,
                                            const double* bi, const AGeo<T>& p, bool withB) {
    AGath<T> o;
    const T lw = alin(wr, p);
    o.Wx = lw * p.c;
    o.Wy = lw * p.s;
    o.Ez = alin(ez, p);
    o.Bz = alin(bz, p);
    if (withB) {
        const T lr = alin(br, p), li = alin(bi, p);
        o.Bx = lr * p.c - li * p.s;
        o.By = lr * p.s + li * p.c;
    } else {
        o.Bx = T(0.0);
        o.By = T(0.0);
    }
    return o;
}
template <class T> struct AMom { T gam, invd, pz, ux, uy; };
template <class T> __device__ __forceinline__ AMom<T> amom(const T& PX, const T& PY, const T& D) {
    AMom<T> m;
    m.gam = (1.0 + PX * PX + PY * PY + D * D) / (2.0 * D);
    m.invd = 1.0 / D;
    m.pz = m.gam - D;
    m.ux = PX * m.invd;
    m.uy = PY * m.invd;
    return m;
}
// lambda . (the m = 0 deposit terms the fields read), PlasmaSpecies::deposit (src/PlasmaSpecies.cpp:409-478)
template <class T>
__device__ __forceinline__ T deposit_dot(const AGeo<T>& p, const T& PX, const T& PY, const T& D, double wq, double qm,
                                         const double* inv_V, const double* Lrhot, const double* Ldrho,
                                         const double* Ljz, const double* Lchi, const double* Ljpi) {
    const AMom<T> m = amom(PX, PY, D);
    const T ur = p.c * m.ux + p.s * m.uy, uth = p.c * m.uy - p.s * m.ux;
    T acc(0.0);
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
        const int jj = p.j + cc;
        const double iv = inv_V[jj];
        const T a = wq * (cc ? p.W1 : p.W0) * iv;
        const T dr = wq * ur * (cc ? p.dW1 : p.dW0) * iv;
        const T vjz = a * m.pz * m.invd, vchi = a * qm * m.invd;
        acc = acc + Lrhot[jj] * a + Ldrho[jj] * dr + Ljz[jj] * vjz + Lchi[jj] * vchi + Ljpi[jj] * (a * uth);
    }
    return acc;
}
// lambda . (S_r1, S_i1 contributions), PlasmaSpecies::deposit_S mode 1 (src/PlasmaSpecies.cpp:508-559)
template <class T>
__device__ __forceinline__ T deposit_S_dot(const AGeo<T>& p, const AGath<T>& F, const T& PX, const T& PY, const T& D,
                                           double wq, double qm, const double* inv_V, const double* LSr,
                                           const double* LSi) {
    const AMom<T> m = amom(PX, PY, D);
    const T ur = p.c * m.ux + p.s * m.uy, uth = p.c * m.uy - p.s * m.ux;
    const T dD = qm * ((PX * F.Wx + PY * F.Wy) * m.invd - F.Ez);
    const T ax = (qm * (m.gam * F.Wx + PY * F.Bz) * m.invd) * m.invd - PX * dD * m.invd * m.invd;
    const T ay = (qm * (m.gam * F.Wy - PX * F.Bz) * m.invd) * m.invd - PY * dD * m.invd * m.invd;
    T acc(0.0);
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
        const int jj = p.j + cc;
        const double iv = inv_V[jj] * wq;
        const T f2 = uth * (cc ? p.Wor1 : p.Wor0);
        const T W = cc ? p.W1 : p.W0;
        const double dW = cc ? p.dW1 : p.dW0;
        const T Ar = ax * W + m.ux * ur * dW + f2 * m.uy;
        const T Ai = ay * W + m.uy * ur * dW - f2 * m.ux;
        acc = acc + LSr[jj] * (iv * (Ar * p.c + Ai * p.s)) + LSi[jj] * (iv * (Ai * p.c - Ar * p.s));
    }
    return acc;
}
// lambda . f, the forces of PlasmaSpecies::push (src/PlasmaSpecies.cpp:595-610)
template <class T>
__device__ __forceinline__ T forces_dot(const AGath<T>& F, const T& PX, const T& PY, const T& D, double qm,
                                        const double* lf) {
    const AMom<T> m = amom(PX, PY, D);
    const T f2 = qm * (m.gam * F.Wx * m.invd + F.By + PY * F.Bz * m.invd);
    const T f3 = qm * (m.gam * F.Wy * m.invd - F.Bx - PX * F.Bz * m.invd);
    const T f4 = qm * ((PX * F.Wx + PY * F.Wy) * m.invd - F.Ez);
    return lf[0] * (PX * m.invd) + lf[1] * (PY * m.invd) + lf[2] * f2 + lf[3] * f3 + lf[4] * f4;
}

// scatter the adjoints of gathered values to the cell's two nodes (LDS atomics)
__device__ __forceinline__ void scatter_gathered(const AGeo<double>& p, double lWx, double lWy, double lEz, double lBz,
                                                 double lBx, double lBy, bool withB, double* Awr, double* Aez,
                                                 double* Abz, double* Abr, double* Abi) {
    const double lw = lWx * p.c + lWy * p.s;
    const double lr = lBx * p.c + lBy * p.s, li = -lBx * p.s + lBy * p.c;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
        const int jj = p.j + cc;
        const double W = cc ? p.W1 : p.W0;
        atomicAdd(&Awr[jj], W * lw);
        atomicAdd(&Aez[jj], W * lEz);
        atomicAdd(&Abz[jj], W * lBz);
        if (withB) {
            atomicAdd(&Abr[jj], W * lr);
            atomicAdd(&Abi[jj], W * li);
        }
    }
}

// T^T mu = rhs for an operator of FieldSolver::solve (src/FieldSolver.cpp:242-261): Dirichlet rows (row 0 unless
// axis_regular, row N unless flux) are identity rows, the other diagonals b - chi (chi may be null).  T^T has the
// diagonals c[i-1], diag, a[i+1].  Parallel cyclic reduction over the whole workgroup: log2(M) steps, each
// eliminating every row's neighbours at the current stride (the unknowns' coupling doubles in reach); the
// adjoint needs the solution, not Thomas's rounding (src/Tridiag.cpp:111-126 is the forward's serial form).
__device__ void pcr_transposed(const double* __restrict__ a, const double* __restrict__ b,
                               const double* __restrict__ c, const double* chi, bool axis_regular, bool flux,
                               const double* rhs, double* mu, double* const (&buf)[8], int N, int tid, int nth) {
    const int n = N + 1;
    double *A0 = buf[0], *B0 = buf[1], *C0 = buf[2], *D0 = buf[3];
    double *A1 = buf[4], *B1 = buf[5], *C1 = buf[6], *D1 = buf[7];
    for (int i = tid; i < n; i += nth) {
        const bool dir = (i == N && !flux) || (i == 0 && !axis_regular);
        A0[i] = i > 0 ? c[i - 1] : 0.0;
        B0[i] = (dir || !chi) ? b[i] : b[i] - chi[i];
        C0[i] = i < N ? a[i + 1] : 0.0;
        D0[i] = rhs[i];
    }
    __syncthreads();
    for (int st = 1; st < n; st <<= 1) {
        for (int i = tid; i < n; i += nth) {
            const int im = i - st, ip = i + st;
            double an = 0.0, bn = B0[i], cn = 0.0, dn = D0[i];
            if (im >= 0) {
                const double k1 = A0[i] / B0[im];
                an = -A0[im] * k1;
                bn -= C0[im] * k1;
                dn -= D0[im] * k1;
            }
            if (ip < n) {
                const double k2 = C0[i] / B0[ip];
                cn = -C0[ip] * k2;
                bn -= A0[ip] * k2;
                dn -= D0[ip] * k2;
            }
            A1[i] = an;
            B1[i] = bn;
            C1[i] = cn;
            D1[i] = dn;
        }
        __syncthreads();
        double* t;
        t = A0; A0 = A1; A1 = t;
        t = B0; B0 = B1; B1 = t;
        t = C0; C0 = C1; C1 = t;
        t = D0; D0 = D1; D1 = t;
    }
    for (int i = tid; i < n; i += nth) mu[i] = D0[i] / B0[i];
    __syncthreads();
}

// pcr_transposed for an operator that is the same every slice (no chi): the reduction of the matrix is done
// once, keeping each level's two elimination factors and the last level's diagonal (plan: kPcrLevels x 2 rows of
// n factors, then n diagonals), so that a slice's solves only apply them to their right-hand sides -- the same
// operations in the same order as pcr_transposed, so the same results.
constexpr int kPcrLevels = 11;   // ceil(log2(n)) for n <= 2048 rows
__device__ void pcr_plan(const double* __restrict__ a, const double* __restrict__ b, const double* __restrict__ c,
                         double* __restrict__ plan, double* const (&buf)[8], int N, int tid, int nth) {
    const int n = N + 1;
    double *A0 = buf[0], *B0 = buf[1], *C0 = buf[2];
    double *A1 = buf[4], *B1 = buf[5], *C1 = buf[6];
    for (int i = tid; i < n; i += nth) {
        A0[i] = i > 0 ? c[i - 1] : 0.0;
        B0[i] = b[i];
        C0[i] = i < N ? a[i + 1] : 0.0;
    }
    __syncthreads();   // a Dirichlet row's diagonal is b itself, as every row's without chi
    int lev = 0;
    for (int st = 1; st < n; st <<= 1, ++lev) {
        double* k1v = plan + static_cast<size_t>(2 * lev) * n;
        double* k2v = k1v + n;
        for (int i = tid; i < n; i += nth) {
            const int im = i - st, ip = i + st;
            double an = 0.0, bn = B0[i], cn = 0.0, k1 = 0.0, k2 = 0.0;
            if (im >= 0) {
                k1 = A0[i] / B0[im];
                an = -A0[im] * k1;
                bn -= C0[im] * k1;
            }
            if (ip < n) {
                k2 = C0[i] / B0[ip];
                cn = -C0[ip] * k2;
                bn -= A0[ip] * k2;
            }
            k1v[i] = k1;
            k2v[i] = k2;
            A1[i] = an;
            B1[i] = bn;
            C1[i] = cn;
        }
        __syncthreads();
        double* t;
        t = A0; A0 = A1; A1 = t;
        t = B0; B0 = B1; B1 = t;
        t = C0; C0 = C1; C1 = t;
    }
    double* bl = plan + static_cast<size_t>(2 * kPcrLevels) * n;
    for (int i = tid; i < n; i += nth) bl[i] = B0[i];
    __syncthreads();
}

// three right-hand sides through one plan: mu_r = T^-T rhs_r; buf holds two n-row buffers per side
__device__ void pcr_apply3(const double* __restrict__ plan, const double* const (&rhs)[3], double* const (&mu)[3],
                           double* const (&buf)[8], int N, int tid, int nth) {
    const int n = N + 1;
    double* D0[3] = {buf[0], buf[1], buf[2]};
    double* D1[3] = {buf[3], buf[4], buf[5]};
    for (int i = tid; i < n; i += nth)
        for (int r = 0; r < 3; ++r) D0[r][i] = rhs[r][i];
    __syncthreads();
    int lev = 0;
    for (int st = 1; st < n; st <<= 1, ++lev) {
        const double* k1v = plan + static_cast<size_t>(2 * lev) * n;
        const double* k2v = k1v + n;
        for (int i = tid; i < n; i += nth) {
            const int im = i - st, ip = i + st;
            const double k1 = k1v[i], k2 = k2v[i];
            for (int r = 0; r < 3; ++r) {
                double dn = D0[r][i];
                if (im >= 0) dn -= D0[r][im] * k1;
                if (ip < n) dn -= D0[r][ip] * k2;
                D1[r][i] = dn;
            }
        }
        __syncthreads();
        for (int r = 0; r < 3; ++r) {
            double* t = D0[r];
            D0[r] = D1[r];
            D1[r] = t;
        }
    }
    const double* bl = plan + static_cast<size_t>(2 * kPcrLevels) * n;
    for (int i = tid; i < n; i += nth)
        for (int r = 0; r < 3; ++r) mu[r][i] = D0[r][i] / bl[i];
    __syncthreads();
}

// out_j of D^T l for the radial gradient stencils (FieldSolver::gradient / gradient_odd, src/FieldSolver.cpp:263-281)
__device__ __forceinline__ double gradient_T(const double* gm, const double* g0, const double* gp, const double* ir,
                                             const double* l, int j, int N, bool odd) {
    double o = 0;
    if (j + 1 >= 1 && j + 1 <= N - 1) o += gm[j + 1] * l[j + 1];
    if (j >= 1 && j <= N - 1) o += g0[j] * l[j];
    if (j - 1 >= 1 && j - 1 <= N - 1) o += gp[j - 1] * l[j - 1];
    if (j == N) o += g0[N] * l[N];
    if (j == N - 1) o += gm[N] * l[N];
    if (j == N - 2) o += gp[N] * l[N];
    if (odd && j == 1) o += ir[0] * l[0];
    return o;
}
// FieldSolver::over_r (src/FieldSolver.cpp:283-289) transposed
__device__ __forceinline__ double over_r_T(const double* ir, const double* l, int j) {
    double o = j >= 1 ? ir[j] * l[j] : 0.0;
    if (j == 1) o += ir[0] * l[0];
    return o;
}

__global__ void __launch_bounds__(kAdjThreads) sweep_adjoint_kernel(qz_sweep_args A, qz_adjoint_args J) {
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nth = blockDim.x;
    unsigned long long t_last = wall_clock64();
    // phase clocks (lane 0, after the phase's closing barrier)
    auto phase = [&](int i) {
        if (J.ticks && tid == 0) {
            const unsigned long long now = wall_clock64();
            atomicAdd(&J.ticks[i], now - t_last);
            t_last = now;
        }
    };
    const int N = A.grid.N, M = N + 1;
    const int Mp = (M + 1) & ~1;
    auto LA = [&](int a) { return lds + static_cast<size_t>(a) * Mp; };
    // the solves' arrays: LDS when they fit after the node arrays, else global
    const bool solve_lds = static_cast<size_t>(L_NARR + S_NARR) * Mp * 8 <= static_cast<size_t>(kAdjLdsBytes);
    double* sbase = solve_lds ? lds + static_cast<size_t>(L_NARR) * Mp : J.work;
    const size_t sstride = solve_lds ? static_cast<size_t>(Mp) : static_cast<size_t>(M);
    auto SA = [&](int a) { return sbase + static_cast<size_t>(a) * sstride; };
    const int np0 = A.nspecies > 0 ? A.sp[0].np : 0, np1 = A.nspecies > 1 ? A.sp[1].np : 0;
    const int NP = np0 + np1;
    double* part = J.work + static_cast<size_t>(S_NARR) * M;   // 5 x NP
    const int K = A.nslices, ord = A.ab_order;
    const double dxi = A.dxi, idxi = 1.0 / dxi, R = A.grid.R;
    const double* inv_V = A.grid.inv_V;
    const double *gm = A.grad_m, *g0 = A.grad_0, *gp = A.grad_p, *ir = A.inv_r;
    const double* r_ = LA(L_R);
    const double* ih_ = LA(L_IH);
    for (int j = tid; j < M; j += nth) {
        LA(L_R)[j] = A.grid.r[j];
        LA(L_IH)[j] = j < N ? A.grid.inv_h[j] : 0.0;
    }
    double* const pbuf[8] = {SA(P_A0), SA(P_B0), SA(P_C0), SA(P_D0), SA(P_A1), SA(P_B1), SA(P_C1), SA(P_D1)};
    __syncthreads();
    // L0 is the same every slice: its reduction once per launch
    double* l0plan = part + 5 * static_cast<size_t>(NP);
    pcr_plan(A.op_a[OP_L0], A.op_b[OP_L0], A.op_c[OP_L0], l0plan, pbuf, N, tid, nth);
    for (int k = J.k_hi - 1; k >= J.k_lo; --k) {
        const double* st = J.traj + static_cast<size_t>(k - J.k_lo) * 6 * NP;
        // ---- the slice's fields; the objective's field adjoints through store_slice (src/Simulation.cpp:594-633)
        for (int j = tid; j < M; j += nth) {
            const double* F = A.fld + (static_cast<size_t>(k) * M + j) * A.ncf;
            const double* lf = J.lam_fld + (static_cast<size_t>(k) * M + j) * A.ncf;
            LA(F_WR)[j] = F[F_EP1R] - F[F_BP1I];
            LA(F_EZ)[j] = F[F_EZ0];
            LA(F_BZ)[j] = F[F_BZ0];
            LA(F_BR)[j] = F[F_BP1R];
            LA(F_BI)[j] = F[F_BP1I];
            LA(A_WR)[j] = lf[F_EP1R];
            LA(A_BI)[j] = lf[F_EP1R] + lf[F_BP1I];
            LA(A_BR)[j] = -lf[F_EP1I] + lf[F_BP1R];
            LA(A_EZ)[j] = lf[F_EZ0];
            LA(A_BZ)[j] = lf[F_BZ0];
            for (int a = A_SR; a <= S_CHI; ++a) LA(a)[j] = 0.0;
        }
        __syncthreads();
        phase(0);
        // ---- chi of the slice (the B+ operators' diagonal) and the adjoint of the slice's forces
        const bool pushed = k < K - 1;
        for (int gi = tid; gi < NP; gi += nth) {
            double* pp = part + gi;
            pp[0] = pp[NP] = pp[2 * NP] = pp[3 * NP] = pp[4 * NP] = 0.0;
            const double wi = st[5 * NP + gi];
            if (wi == 0.0) continue;
            const bool s1 = gi >= np0;
            const double q = s1 ? A.sp[1].q : A.sp[0].q, qm = q / (s1 ? A.sp[1].m : A.sp[0].m), wq = wi * q;
            const double X = st[gi], Y = st[NP + gi], PX = st[2 * NP + gi], PY = st[3 * NP + gi], D = st[4 * NP + gi];
            const AGeo<double> p = ageo<double>(r_, ih_, N, R, X, Y);
            const double invd = 1.0 / D;
            atomicAdd(&LA(S_CHI)[p.j], wq * p.W0 * inv_V[p.j] * qm * invd);
            atomicAdd(&LA(S_CHI)[p.j + 1], wq * p.W1 * inv_V[p.j + 1] * qm * invd);
            if (!pushed) continue;
            double lfv[5] = {0, 0, 0, 0, 0};
            for (int m = 0; m < ord; ++m) {
                double c = A.ab_c[m];
                if (k == 0) {   // the first push fills every history slot with f_0
                    c = 0;
                    for (int mm = m; mm < ord; ++mm) c += A.ab_c[mm];
                }
                const double* rg = J.ring + static_cast<size_t>(m) * 5 * NP + gi;
#pragma unroll
                for (int v = 0; v < 5; ++v) lfv[v] += dxi * c * rg[static_cast<size_t>(v) * NP];
            }
            {   // over the state (the gather follows the position)
                using D5 = Dual<5>;
                const D5 x = D5::var(X, 0), y = D5::var(Y, 1);
                const AGeo<D5> pd = ageo<D5>(r_, ih_, N, R, x, y);
                const AGath<D5> g = agather(LA(F_WR), LA(F_EZ), LA(F_BZ), LA(F_BR), LA(F_BI), pd, true);
                const D5 L = forces_dot(g, D5::var(PX, 2), D5::var(PY, 3), D5::var(D, 4), qm, lfv);
#pragma unroll
                for (int v = 0; v < 5; ++v) pp[static_cast<size_t>(v) * NP] += L.d[v];
            }
            {   // over the gathered fields
                using D6 = Dual<6>;
                const AGath<double> g0v = agather(LA(F_WR), LA(F_EZ), LA(F_BZ), LA(F_BR), LA(F_BI), p, true);
                const AGath<D6> g{D6::var(g0v.Wx, 0), D6::var(g0v.Wy, 1), D6::var(g0v.Ez, 2), D6::var(g0v.Bz, 3),
                                  D6::var(g0v.Bx, 4), D6::var(g0v.By, 5)};
                const D6 L = forces_dot(g, D6(PX), D6(PY), D6(D), qm, lfv);
                scatter_gathered(p, L.d[0], L.d[1], L.d[2], L.d[3], L.d[4], L.d[5], true, LA(A_WR), LA(A_EZ),
                                 LA(A_BZ), LA(A_BR), LA(A_BI));
            }
        }
        __syncthreads();
        phase(1);
        // ---- B+ (src/Simulation.cpp:535-591, m = 0: one pass): br = L1D(chi)^-1 (-S_i1),
        //      bi = L1F(chi, jz)^-1 (grad jz + S_r1)
        pcr_transposed(A.op_a[OP_L1D], A.op_b[OP_L1D], A.op_c[OP_L1D], LA(S_CHI), false, false, LA(A_BR), SA(T_MU0),
                       pbuf, N, tid, nth);
        pcr_transposed(A.op_a[OP_L1F], A.op_b[OP_L1F], A.op_c[OP_L1F], LA(S_CHI), false, true, LA(A_BI), SA(T_MU1),
                       pbuf, N, tid, nth);
        phase(2);
        {
            const double* mr = SA(T_MU0);
            double* mi = SA(T_MU1);
            for (int j = tid; j < M; j += nth) {
                const double vr = (j == 0 || j == N) ? 0.0 : mr[j];   // L1D rows 0 and N are Dirichlet rows
                const double vi = j == 0 ? 0.0 : mi[j];                // L1F row 0 is
                LA(A_SI)[j] = -vr;
                LA(A_SR)[j] = vi;
                LA(A_CHI)[j] = vr * LA(F_BR)[j] + vi * LA(F_BI)[j];
            }
            __syncthreads();
            if (tid == 0) mi[0] = 0.0;
            __syncthreads();
            for (int j = tid; j < M; j += nth)
                LA(A_JZ)[j] = gradient_T(gm, g0, gp, ir, mi, j, N, false) + (j == N ? -A.flux_coef * mi[N] : 0.0);
        }
        __syncthreads();
        phase(3);
        // ---- deposit_S: the state and the gathered W+, Ez, Bz
        for (int gi = tid; gi < NP; gi += nth) {
            const double wi = st[5 * NP + gi];
            if (wi == 0.0) continue;
            double* pp = part + gi;
            const bool s1 = gi >= np0;
            const double q = s1 ? A.sp[1].q : A.sp[0].q, qm = q / (s1 ? A.sp[1].m : A.sp[0].m), wq = wi * q;
            const double X = st[gi], Y = st[NP + gi], PX = st[2 * NP + gi], PY = st[3 * NP + gi], D = st[4 * NP + gi];
            {
                using D5 = Dual<5>;
                const AGeo<D5> pd = ageo<D5>(r_, ih_, N, R, D5::var(X, 0), D5::var(Y, 1));
                const AGath<D5> g = agather(LA(F_WR), LA(F_EZ), LA(F_BZ), LA(F_BR), LA(F_BI), pd, false);
                const D5 L = deposit_S_dot(pd, g, D5::var(PX, 2), D5::var(PY, 3), D5::var(D, 4), wq, qm, inv_V,
                                           LA(A_SR), LA(A_SI));
#pragma unroll
                for (int v = 0; v < 5; ++v) pp[static_cast<size_t>(v) * NP] += L.d[v];
            }
            {
                using D4 = Dual<4>;
                const AGeo<double> p = ageo<double>(r_, ih_, N, R, X, Y);
                const AGath<double> g0v = agather(LA(F_WR), LA(F_EZ), LA(F_BZ), LA(F_BR), LA(F_BI), p, false);
                const AGath<D4> g{D4::var(g0v.Wx, 0), D4::var(g0v.Wy, 1), D4::var(g0v.Ez, 2), D4::var(g0v.Bz, 3),
                                  D4(0.0), D4(0.0)};
                const D4 L = deposit_S_dot(promote<D4>(p), g, D4(PX), D4(PY), D4(D), wq, qm, inv_V, LA(A_SR),
                                           LA(A_SI));
                scatter_gathered(p, L.d[0], L.d[1], L.d[2], L.d[3], 0.0, 0.0, false, LA(A_WR), LA(A_EZ), LA(A_BZ),
                                 nullptr, nullptr);
            }
        }
        __syncthreads();
        phase(4);
        // ---- Ez, Bz and psi (src/Simulation.cpp:477-532, 797-819): bz = L0^-1 t3, t3 = -(grad_odd + over_r) jp_i1;
        //      ez = L0^-1 (-drho); wr = -grad psi, psi = L0^-1 (-rhot)
        for (int j = tid; j < M; j += nth) SA(T_RHS)[j] = -gradient_T(gm, g0, gp, ir, LA(A_WR), j, N, false);
        __syncthreads();
        {
            const double* const rhs[3] = {LA(A_BZ), LA(A_EZ), SA(T_RHS)};
            double* const mu[3] = {SA(T_MU0), SA(T_MU1), SA(T_MU2)};
            pcr_apply3(l0plan, rhs, mu, pbuf, N, tid, nth);
        }
        phase(5);
        if (tid == 0) SA(T_MU0)[N] = 0.0;   // L0's row N is a Dirichlet row
        __syncthreads();
        for (int j = tid; j < M; j += nth) {
            LA(A_DRHO)[j] = j < N ? -SA(T_MU1)[j] : 0.0;
            LA(A_RHOT)[j] = j < N ? -SA(T_MU2)[j] : 0.0;
            const double* lt = SA(T_MU0);
            LA(A_JPI)[j] = -(gradient_T(gm, g0, gp, ir, lt, j, N, true) + over_r_T(ir, lt, j));
        }
        __syncthreads();
        phase(6);
        // ---- the beam sources (combine_sources, src/Simulation.cpp:408-457) and the plasma deposit
        for (int j = tid; j < M; j += nth) {
            auto LB = [&](int kk, int comp) -> double& {
                return J.lam_bsrc[(static_cast<size_t>(kk) * M + j) * A.ncb + comp];
            };
            LB(k, B_RT0) += LA(A_RHOT)[j];
            LB(k, B_JZ0) += LA(A_JZ)[j];
            LB(k, B_JP1I) += LA(A_JPI)[j];
            const double d = LA(A_DRHO)[j] * idxi;
            if (A.ngp) {
                LB(k, B_RT0) += d;
                if (k > 0) LB(k - 1, B_RT0) -= d;
            } else if (K > 1) {
                if (k == 0) { LB(1, B_RT0) += d; LB(0, B_RT0) -= d; }
                else if (k == K - 1) { LB(K - 1, B_RT0) += d; LB(K - 2, B_RT0) -= d; }
                else { LB(k + 1, B_RT0) += 0.5 * d; LB(k - 1, B_RT0) -= 0.5 * d; }
            }
        }
        for (int gi = tid; gi < NP; gi += nth) {
            const double wi = st[5 * NP + gi];
            double Lk[5] = {0, 0, 0, 0, 0};
            if (wi != 0.0) {
                const bool s1 = gi >= np0;
                const double q = s1 ? A.sp[1].q : A.sp[0].q, qm = q / (s1 ? A.sp[1].m : A.sp[0].m), wq = wi * q;
                using D5 = Dual<5>;
                const AGeo<D5> pd = ageo<D5>(r_, ih_, N, R, D5::var(st[gi], 0), D5::var(st[NP + gi], 1));
                const D5 L = deposit_dot(pd, D5::var(st[2 * NP + gi], 2), D5::var(st[3 * NP + gi], 3),
                                         D5::var(st[4 * NP + gi], 4), wq, qm, inv_V, LA(A_RHOT), LA(A_DRHO), LA(A_JZ),
                                         LA(A_CHI), LA(A_JPI));
                // Lambda_k = Lambda_{k+1} (the update's identity part) + this slice's partials
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    Lk[v] = J.ring[static_cast<size_t>(v) * NP + gi] + part[static_cast<size_t>(v) * NP + gi] + L.d[v];
            }
            for (int m = ord - 1; m > 0; --m)
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    J.ring[(static_cast<size_t>(m) * 5 + v) * NP + gi] = J.ring[(static_cast<size_t>(m - 1) * 5 + v) * NP + gi];
#pragma unroll
            for (int v = 0; v < 5; ++v) J.ring[static_cast<size_t>(v) * NP + gi] = Lk[v];
        }
        __syncthreads();
        phase(7);
    }
}

// ---------------------------------------------------------------- beam side
// dJ/dgamma and the direct dJ/dw of J = sigma / mu over the live particles (w != 0):
// mu = sum w gamma / W, sigma^2 = sum w (gamma - mu)^2 / W
__global__ void spread_adjoint_kernel(int np, qz_beam_soa P, double mu, double sig, double W,
                                      double* __restrict__ lam_gamma, double* __restrict__ lam_w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double w = P.w[i];
    if (w == 0.0) {
        lam_gamma[i] = 0.0;
        lam_w[i] = 0.0;
        return;
    }
    const double g = sqrt(1.0 + P.px[i] * P.px[i] + P.py[i] * P.py[i] + P.pz[i] * P.pz[i]);
    const double dg = g - mu;
    lam_gamma[i] = w / W * (dg / (sig * mu) - sig / (mu * mu));
    lam_w[i] = (dg * dg - sig * sig) / (2.0 * sig * W * mu) - sig * dg / (mu * mu * W);
}

template <class T> struct V3T { T x, y, z; };
template <class T> __device__ __forceinline__ V3T<T> vadd(const V3T<T>& a, const V3T<T>& b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T> __device__ __forceinline__ V3T<T> vscale(const T& s, const V3T<T>& a) {
    return {s * a.x, s * a.y, s * a.z};
}
template <class T> __device__ __forceinline__ T vdot(const V3T<T>& a, const V3T<T>& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T> __device__ __forceinline__ V3T<T> vcross(const V3T<T>& a, const V3T<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// gamma after the kick of Beam::advance with the Vay pusher (src/Pushers.hpp:36-68), with the half back-kick of the
// first push (src/Beam.cpp:900-905)
template <class T>
__device__ __forceinline__ T kicked_gamma(const V3T<T>& u0, const V3T<T>& E, const V3T<T>& B, double qm,
                                          double dt_kick, bool started) {
    V3T<T> u = u0;
    if (!started) {
        const double h = 0.5 * qm * (-0.5 * dt_kick);
        const T ig = 1.0 / vsqrt(1.0 + vdot(u, u));
        u = vadd(u, vscale(T(2.0 * h), vadd(E, vcross(vscale(ig, u), B))));
    }
    const double eps = 0.5 * qm * dt_kick;
    const T g0 = vsqrt(1.0 + vdot(u, u));
    const V3T<T> a = vadd(vadd(u, vscale(T(2.0 * eps), E)), vscale(eps / g0, vcross(u, B)));
    const V3T<T> b = vscale(T(eps), B);
    const T b2 = vdot(b, b), ab = vdot(a, b);
    const T x = 0.5 * (1.0 + vdot(a, a) - b2), c = b2 + ab * ab;
    const T sq = vsqrt(x * x + c);
    const T g2 = val(x) >= 0 ? x + sq : (val(sq) > -val(x) ? c / (sq - x) : T(1.0));
    const T g = vsqrt(g2);
    const V3T<T> axb = vcross(a, b);
    const T inv = 1.0 / (g2 + b2);
    const V3T<T> un{inv * (g2 * a.x + g * axb.x + ab * b.x), inv * (g2 * a.y + g * axb.y + ab * b.y),
                    inv * (g2 * a.z + g * axb.z + ab * b.z)};
    return vsqrt(1.0 + vdot(un, un));
}

// lam_fld += dJ/dfld of the beam's kick (the gather of Beam::advance, m = 0, transposed)
__global__ void kick_adjoint_kernel(qz_radial_grid g, qz_beam_grid bg, double qm, int np, qz_beam_soa P,
                                    const double* __restrict__ fld, int nc, double dt_kick, int started,
                                    const double* __restrict__ lam_gamma, double* __restrict__ lam_fld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np || P.w[i] == 0.0 || lam_gamma[i] == 0.0) return;
    const double xmin = bg.xi_min, idx = 1.0 / bg.dxi;
    const int k0 = bg.k0, nloc = bg.nloc, M = g.N + 1;
    const bool ngp = bg.ngp != 0;
    const double X = P.x[i], Y = P.y[i];
    const double r = sqrt(X * X + Y * Y);
    const double c = r > 0 ? X / r : 1.0, sn = r > 0 ? Y / r : 0.0;
    double s = (P.xi[i] - xmin) * idx - k0;
    int k;
    double tk;
    if (ngp) {
        k = static_cast<int>(floor(s + 0.5));
        k = k < 0 ? 0 : (k > nloc - 1 ? nloc - 1 : k);
        tk = 0;
    } else {
        k = static_cast<int>(floor(s));
        if (k < 0) { k = 0; s = 0; }
        if (k > nloc - 2) { k = nloc - 2; s = nloc - 1; }
        tk = s - k;
    }
    const int k1 = ngp ? k : k + 1;
    int j;
    double t;
    grid_weights(g.r, g.inv_h, g.N, g.R, r, j, t);
    const double* F0 = fld + (static_cast<size_t>(k) * M + j) * nc;
    const double* F1 = fld + (static_cast<size_t>(k1) * M + j) * nc;
    auto bil = [&](int comp) {
        return (1.0 - tk) * ((1.0 - t) * F0[comp] + t * F0[nc + comp]) + tk * ((1.0 - t) * F1[comp] + t * F1[nc + comp]);
    };
    const double er = bil(F_EP1R), ei = bil(F_EP1I), br = bil(F_BP1R), bi = bil(F_BP1I);
    using D6 = Dual<6>;
    const V3T<D6> E{D6::var(er * c - ei * sn, 0), D6::var(er * sn + ei * c, 1), D6::var(bil(F_EZ0), 2)};
    const V3T<D6> B{D6::var(br * c - bi * sn, 3), D6::var(br * sn + bi * c, 4), D6::var(bil(F_BZ0), 5)};
    const V3T<D6> u0{D6(P.px[i]), D6(P.py[i]), D6(P.pz[i])};
    const D6 g1 = kicked_gamma(u0, E, B, qm, dt_kick, started != 0);
    const double lg = lam_gamma[i];
    double lE[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) lE[q] = lg * g1.d[q];
    const double l_er = lE[0] * c + lE[1] * sn, l_ei = -lE[0] * sn + lE[1] * c;
    const double l_br = lE[3] * c + lE[4] * sn, l_bi = -lE[3] * sn + lE[4] * c;
    const double wts[4] = {(1 - tk) * (1 - t), (1 - tk) * t, tk * (1 - t), tk * t};
    const int kks[4] = {k, k, k1, k1}, jjs[4] = {j, j + 1, j, j + 1};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (wts[q] == 0.0) continue;
        double* o = lam_fld + (static_cast<size_t>(kks[q]) * M + jjs[q]) * nc;
        atomicAdd(&o[F_EP1R], wts[q] * l_er);
        atomicAdd(&o[F_EP1I], wts[q] * l_ei);
        atomicAdd(&o[F_BP1R], wts[q] * l_br);
        atomicAdd(&o[F_BP1I], wts[q] * l_bi);
        atomicAdd(&o[F_EZ0], wts[q] * lE[2]);
        atomicAdd(&o[F_BZ0], wts[q] * lE[5]);
    }
}

// lam_w += the beam deposit's transpose (Beam::deposit, m = 0, per unit weight) against lam_bsrc
__global__ void deposit_adjoint_kernel(qz_radial_grid g, qz_beam_grid bg, double q, int np, qz_beam_soa P,
                                       const double* __restrict__ lam_bsrc, int nc, double* __restrict__ lam_w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np || P.w[i] == 0.0) return;
    const double X = P.x[i], Y = P.y[i];
    const double r = sqrt(X * X + Y * Y);
    if (r > g.R) return;
    const double xmin = bg.xi_min, idx = 1.0 / bg.dxi;
    const int k0 = bg.k0, nloc = bg.nloc, M = g.N + 1;
    const bool ngp = bg.ngp != 0;
    const double c = r > 0 ? X / r : 1.0, sn = r > 0 ? Y / r : 0.0;
    const double sx = (P.xi[i] - xmin) * idx;
    const int k = ngp ? static_cast<int>(floor(sx + 0.5)) : static_cast<int>(floor(sx));
    if (k < k0 - 1 || k >= k0 + nloc) return;
    const double tk = ngp ? 0.0 : sx - k;
    const double px = P.px[i], py = P.py[i], pz = P.pz[i];
    const double p2 = px * px + py * py;
    const double gam = sqrt(1.0 + p2 + pz * pz);
    const double omvz = (1.0 + p2) / (gam * (gam + pz));
    const double vz = pz / gam, vx = px / gam, vy = py / gam;
    const double vr = c * vx + sn * vy, vt = c * vy - sn * vx;
    int j;
    double t;
    grid_weights(g.r, g.inv_h, g.N, g.R, r, j, t);
    const double Wr[2] = {1.0 - t, t}, Wk[2] = {1.0 - tk, tk};
    double acc = 0;
    for (int a = 0; a < (ngp ? 1 : 2); ++a) {
        const int kk = k + a - k0;
        if (kk < 0 || kk >= nloc) continue;
        for (int b = 0; b < 2; ++b) {
            const int jj = j + b;
            const double d = q * Wk[a] * Wr[b] * g.inv_V[jj] * idx;
            const double* l = lam_bsrc + (static_cast<size_t>(kk) * M + jj) * nc;
            acc += d * (l[B_RT0] * omvz + l[B_JZ0] * vz + l[B_JP1R] * vr + l[B_JP1I] * vt + l[B_RHO0]);
        }
    }
    lam_w[i] += acc;
}

// grad_s += sum_i lam_w_i wmac d profile(u_i) / d v_s: the profile is linear between the centres of nseg equal
// segments of [u0, u1] and flat beyond the first and the last centre
__global__ void profile_gradient_kernel(int np, const double* __restrict__ u, const double* __restrict__ lam_w,
                                        int nseg, double u0, double u1, double wmac, double* __restrict__ grad) {
    extern __shared__ double sg[];
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) sg[s] = 0.0;
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        const double l = lam_w[i] * wmac;
        if (l == 0.0) continue;
        const double tt = (u[i] - u0) / (u1 - u0) * nseg - 0.5;
        if (!(tt > 0)) atomicAdd(&sg[0], l);
        else if (tt >= nseg - 1) atomicAdd(&sg[nseg - 1], l);
        else {
            const int s0 = static_cast<int>(tt);
            const double f = tt - s0;
            atomicAdd(&sg[s0], l * (1.0 - f));
            atomicAdd(&sg[s0 + 1], l * f);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) atomicAdd(&grad[s], sg[s]);
}

// LDS of a launch: the node arrays, and the solves' arrays when they fit beside them
size_t adjoint_lds_bytes(int M) {
    const size_t Mp = static_cast<size_t>((M + 1) & ~1);
    const size_t all = static_cast<size_t>(L_NARR + S_NARR) * Mp * 8;
    return all <= static_cast<size_t>(kAdjLdsBytes) ? all : static_cast<size_t>(L_NARR) * Mp * 8;
}

}  // namespace
}  // namespace qz

using namespace qz;

extern "C" size_t qz_sweep_adjoint_workspace_doubles(int M, int np) {
    const size_t m = M > 0 ? static_cast<size_t>(M) : 0;
    return static_cast<size_t>(S_NARR) * m + 5 * static_cast<size_t>(np > 0 ? np : 0) +
           static_cast<size_t>(2 * kPcrLevels + 1) * m;   // the solve arrays, the particles' adjoints, L0's plan
}

extern "C" int qz_sweep_adjoint(const qz_sweep_args* a, const qz_adjoint_args* j, void* stream) {
    if (!a || !j) return qz_fail(QZ_EINVAL, "qz_sweep_adjoint: null argument");
    if (a->modes != 0) return qz_fail(QZ_EINVAL, "qz_sweep_adjoint: modes = 0 only");
    if (a->groups > 1) return qz_fail(QZ_EINVAL, "qz_sweep_adjoint: one workgroup per run (groups <= 1)");
    if (!j->traj || !j->lam_fld || !j->lam_bsrc || !j->ring || !j->work || !a->fld)
        return qz_fail(QZ_EINVAL, "qz_sweep_adjoint: null array");
    if (a->nspecies < 0 || a->nspecies > QZ_MAX_SPECIES || a->ab_order < 1 || a->ab_order > 5)
        return qz_fail(QZ_EINVAL, "qz_sweep_adjoint: species or order out of range");
    if (j->k_lo < 0 || j->k_hi > a->nslices || j->k_lo >= j->k_hi)
        return qz_fail(QZ_EINVAL, "qz_sweep_adjoint: slice range outside the box");
    if (a->grid.N < 3) return qz_fail(QZ_EINVAL, "qz_sweep_adjoint: grid too small");
    const size_t lds = adjoint_lds_bytes(a->grid.N + 1);
    if (lds > static_cast<size_t>(kAdjLdsBytes))
        return qz_fail(QZ_ERANGE, "qz_sweep_adjoint: the slice's node arrays exceed LDS (grids up to 1023 cells)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    QZ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(sweep_adjoint_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, kAdjLdsBytes));
    hipLaunchKernelGGL(sweep_adjoint_kernel, dim3(1), dim3(kAdjThreads), lds, st, *a, *j);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

extern "C" int qz_beam_spread_adjoint(int np, const qz_beam_soa* p, double mu, double sigma, double wsum,
                                      double* lam_gamma, double* lam_w, void* stream) {
    if (!p || !lam_gamma || !lam_w || np < 0) return qz_fail(QZ_EINVAL, "qz_beam_spread_adjoint: bad argument");
    if (!(mu > 0) || !(sigma > 0) || !(wsum > 0))
        return qz_fail(QZ_EINVAL, "qz_beam_spread_adjoint: needs a beam with a positive mean and spread");
    if (np == 0) return QZ_OK;
    hipLaunchKernelGGL(spread_adjoint_kernel, dim3((np + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       np, *p, mu, sigma, wsum, lam_gamma, lam_w);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

extern "C" int qz_beam_kick_adjoint(const qz_radial_grid* grid, const qz_beam_grid* box, double qm, int pusher, int np,
                                    const qz_beam_soa* before, const double* fld, int ncf, double dt_kick, int started,
                                    const double* lam_gamma, double* lam_fld, void* stream) {
    if (!grid || !box || !before || !fld || !lam_gamma || !lam_fld || np < 0)
        return qz_fail(QZ_EINVAL, "qz_beam_kick_adjoint: bad argument");
    if (pusher != 0) return qz_fail(QZ_EINVAL, "qz_beam_kick_adjoint: the Vay pusher only");
    if (ncf != F_NC0) return qz_fail(QZ_EINVAL, "qz_beam_kick_adjoint: modes = 0 only");
    if (np == 0) return QZ_OK;
    hipLaunchKernelGGL(kick_adjoint_kernel, dim3((np + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       *grid, *box, qm, np, *before, fld, ncf, dt_kick, started, lam_gamma, lam_fld);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

extern "C" int qz_beam_deposit_adjoint(const qz_radial_grid* grid, const qz_beam_grid* box, double q, int np,
                                       const qz_beam_soa* before, const double* lam_bsrc, int ncb, double* lam_w,
                                       void* stream) {
    if (!grid || !box || !before || !lam_bsrc || !lam_w || np < 0)
        return qz_fail(QZ_EINVAL, "qz_beam_deposit_adjoint: bad argument");
    if (ncb != B_NC0) return qz_fail(QZ_EINVAL, "qz_beam_deposit_adjoint: modes = 0 only");
    if (np == 0) return QZ_OK;
    hipLaunchKernelGGL(deposit_adjoint_kernel, dim3((np + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       *grid, *box, q, np, *before, lam_bsrc, ncb, lam_w);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

extern "C" int qz_profile_gradient(int np, const double* u, const double* lam_w, int nseg, double u0, double u1,
                                   double wmac, double* grad, void* stream) {
    if (!u || !lam_w || !grad || np < 0 || nseg < 1 || nseg > 4096 || !(u1 != u0))
        return qz_fail(QZ_EINVAL, "qz_profile_gradient: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int rc = qz_zero_async(grad, sizeof(double) * nseg, st);
    if (rc) return rc;
    if (np == 0) return QZ_OK;
    const int blocks = np < 256 * 512 ? (np + 255) / 256 : 512;
    hipLaunchKernelGGL(profile_gradient_kernel, dim3(blocks), dim3(256), sizeof(double) * nseg, st, np, u, lam_w, nseg,
                       u0, u1, wmac, grad);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}
