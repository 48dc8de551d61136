// Device helpers of the quasi-static sweep kernels (one workgroup per run, qz_sweep.hip; several per run,
// qz_sweep_multi.hip): the workspace layout of a run, the deposit terms, the Thomas recurrences, the sequential
// node sums and the radial stencils.  Every expression keeps the reference's operation order (file:line at
// each function), so both kernels reproduce a serial CPU build bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "qz_device.hpp"
#include "../../include/quarz_amd.h"

namespace qz {

// workspace arrays (M doubles each)
enum WsArr : int {
    // node sums of the plasma deposit (the first 24 ids are also the deposit component ids)
    W_RHOT0 = 0, W_DRHO0, W_JZ0, W_CHI0, W_NE0, W_NI0, W_JPR1, W_JPI1,
    W_RHOT1R, W_RHOT1I, W_DRHO1R, W_DRHO1I, W_JZ1R, W_JZ1I, W_CHI1R, W_CHI1I, W_NE1R, W_NE1I, W_NI1R, W_NI1I,
    W_JPR0, W_JPI0, W_JPR2, W_JPI2,
    // S+ node sums, order r1 i1 r0 i0 r2 i2
    W_SR1, W_SI1, W_SR0, W_SI0, W_SR2, W_SI2,
    // fields of the slice
    W_EZ0, W_EZ1R, W_EZ1I, W_BZ0, W_BZ1R, W_BZ1I,
    W_WPR1, W_WPI1, W_WPR0, W_WPI0, W_WPR2, W_WPI2,
    W_BPR1, W_BPI1, W_BPR0, W_BPI0, W_BPR2, W_BPI2,
    // Thomas factors (m, c') of this slice's chi-dependent operators, in the order L1D, L1F, L0, L2D
    W_FAC0,
    W_NARR = W_FAC0 + 8
};

enum OpId : int { OP_L0 = 0, OP_L1D = 1, OP_L1F = 2, OP_L2D = 3 };

template <int V> using IC = std::integral_constant<int, V>;

// value of deposit component `comp` for the node on side cc (0: j, 1: j+1) of one particle
// (PlasmaSpecies::deposit, src/PlasmaSpecies.cpp:420-477; rho is not read on this path and not formed)
__device__ __forceinline__ Real dep_value(int comp, int cc, int j, Real t, Real c, Real sn, Real ir, Real invd,
                                          Real gam, Real PX, Real PY, Real pz, Real wi, Real q, Real qm, Real iv,
                                          Real ih) {
    const Real W = cc == 0 ? Real(1) - t : t;
    const Real dW = cc == 0 ? -ih : ih;
    const Real ux = PX * invd, uy = PY * invd;
    const Real ur = c * ux + sn * uy, uth = c * uy - sn * ux;
    const Real wq = wi * q;
    const Real chif = Real(1);
    const Real a = wq * W * iv;
    switch (comp) {
        case W_RHOT0: return a;
        case W_DRHO0: return wq * ur * dW * iv;
        case W_JZ0: return a * pz * invd;
        case W_CHI0: return chif * a * qm * invd;
        case W_NE0: return q < 0 ? wi * W * iv * gam * invd : Real(0);
        case W_NI0: return q < 0 ? Real(0) : wi * W * iv * gam * invd;
        case W_JPR1: return a * ur;
        case W_JPI1: return a * uth;
        case W_RHOT1R: return a * c;
        case W_RHOT1I: return -a * sn;
        case W_DRHO1R:
        case W_DRHO1I: {
            const Real dr = wq * ur * dW * iv;
            const Real Wor = (j == 0) ? (cc == 0 ? Real(0) : ih) : W * ir;
            const Real th = wq * uth * Wor * iv;
            return comp == W_DRHO1R ? dr * c - th * sn : -dr * sn - th * c;
        }
        case W_JZ1R: return a * pz * invd * c;
        case W_JZ1I: return -(a * pz * invd) * sn;
        case W_CHI1R: return chif * a * qm * invd * c;
        case W_CHI1I: return -(chif * a * qm * invd) * sn;
        case W_NE1R: return q < 0 ? wi * W * iv * gam * invd * c : Real(0);
        case W_NE1I: return q < 0 ? -(wi * W * iv * gam * invd) * sn : Real(0);
        case W_NI1R: return q < 0 ? Real(0) : wi * W * iv * gam * invd * c;
        case W_NI1I: return q < 0 ? Real(0) : -(wi * W * iv * gam * invd) * sn;
        case W_JPR0: return a * ux;
        case W_JPI0: return a * uy;
        case W_JPR2: {
            const Real c2 = c * c - sn * sn, s2 = Real(2) * c * sn;
            return a * (ux * c2 + uy * s2);
        }
        case W_JPI2: {
            const Real c2 = c * c - sn * sn, s2 = Real(2) * c * sn;
            return a * (uy * c2 - ux * s2);
        }
        default: return Real(0);
    }
}

// Thomas solve with factors formed beforehand, every array in LDS (Tridiag.cpp:114-125 with c', m precomputed):
//   dp_0 = d_0 / b_0,  dp_i = (d_i - a_i dp_{i-1}) m_i,  x_{M-1} = dp_{M-1},  x_i = dp_i - c'_i x_{i+1}
// d_i is the prepared right-hand side of FieldSolver::solve (src/FieldSolver.cpp:253-259): 0 on Dirichlet
// rows, scale * rhs otherwise, minus fcw on the flux row.  w holds rhs on entry and dp afterwards; the
// solution goes to x (which may be w).  Rows go in groups of four through two register sets: one group is
// loaded while the other is computed, so the LDS latency stays off the recurrence.
struct FwdRows { Real d0, d1, d2, d3, a0, a1, a2, a3, m0, m1, m2, m3; };
struct BwdRows { Real p0, p1, p2, p3, c0, c1, c2, c3; };

__device__ __forceinline__ void thomas_lds(const Real* a, const Real* m, const Real* cp, Real b0, Real* w, Real* x,
                                           Real scale, bool dir0, bool dirN, bool fluxN, Real fcw, int M) {
    const int N = M - 1;
    Real prev = (dir0 ? Real(0) : scale * w[0]) / b0;
    w[0] = prev;
    auto ldf = [&](FwdRows& g, int r) {
        g.d0 = w[r]; g.d1 = w[r + 1]; g.d2 = w[r + 2]; g.d3 = w[r + 3];
        g.a0 = a[r]; g.a1 = a[r + 1]; g.a2 = a[r + 2]; g.a3 = a[r + 3];
        g.m0 = m[r]; g.m1 = m[r + 1]; g.m2 = m[r + 2]; g.m3 = m[r + 3];
    };
    auto stepf = [&](const FwdRows& g, int r) {
        prev = (scale * g.d0 - g.a0 * prev) * g.m0; w[r] = prev;
        prev = (scale * g.d1 - g.a1 * prev) * g.m1; w[r + 1] = prev;
        prev = (scale * g.d2 - g.a2 * prev) * g.m2; w[r + 2] = prev;
        prev = (scale * g.d3 - g.a3 * prev) * g.m3; w[r + 3] = prev;
    };
    const int nq = (N - 1) >> 2;   // full groups of four rows in 1 .. N-1
    int i = 1;
    if (nq > 0) {
        FwdRows A, B;
        ldf(A, 1);
        int q = 0;
        for (; q + 2 <= nq; q += 2, i += 8) {
            ldf(B, i + 4);
            stepf(A, i);
            ldf(A, (q + 2 < nq) ? i + 8 : i + 4);   // the last pass reloads rows it already holds
            stepf(B, i + 4);
        }
        if (q < nq) {
            stepf(A, i);
            i += 4;
        }
    }
    for (; i < N; ++i) {
        prev = (scale * w[i] - a[i] * prev) * m[i];
        w[i] = prev;
    }
    {
        Real dN;
        if (dirN) dN = Real(0);
        else {
            dN = scale * w[N];
            if (fluxN) dN -= fcw;
        }
        prev = (dN - a[N] * prev) * m[N];
    }
    Real xn = prev;
    x[N] = xn;
    // backward: rows N-1 .. 0, groups of four from the top
    auto ldb = [&](BwdRows& g, int r) {
        g.p0 = w[r]; g.p1 = w[r - 1]; g.p2 = w[r - 2]; g.p3 = w[r - 3];
        g.c0 = cp[r]; g.c1 = cp[r - 1]; g.c2 = cp[r - 2]; g.c3 = cp[r - 3];
    };
    auto stepb = [&](const BwdRows& g, int r) {
        xn = g.p0 - g.c0 * xn; x[r] = xn;
        xn = g.p1 - g.c1 * xn; x[r - 1] = xn;
        xn = g.p2 - g.c2 * xn; x[r - 2] = xn;
        xn = g.p3 - g.c3 * xn; x[r - 3] = xn;
    };
    i = N - 1;
    const int nb = N >> 2;
    if (nb > 0) {
        BwdRows A, B;
        ldb(A, i);
        int q = 0;
        for (; q + 2 <= nb; q += 2, i -= 8) {
            ldb(B, i - 4);
            stepb(A, i);
            ldb(A, (q + 2 < nb) ? i - 8 : i - 4);
            stepb(B, i - 4);
        }
        if (q < nb) {
            stepb(A, i);
            i -= 4;
        }
    }
    for (; i >= 0; --i) {
        xn = w[i] - cp[i] * xn;
        x[i] = xn;
    }
}

// Thomas factors of (L - chi) with the Dirichlet rows of FieldSolver::solve (src/FieldSolver.cpp:253-256):
//   bw_i = b_i (Dirichlet row) or b_i - chi_i;  c'_0 = c_0 / bw_0;  m_i = 1 / (bw_i - a_i c'_{i-1});  c'_i = c_i m_i
// Rows 0 .. N-1 from a, b, c (LDS); the last row's a_N, b_N given (the flux and Dirichlet variants of L_1
// share every other row).  m, c' go to global memory.  Returns bw_0.
__device__ __forceinline__ Real thomas_factor(const Real* a, const Real* b, const Real* c, const Real* chi, Real aN, Real bN,
                              bool dir0, bool dirN, Real* __restrict__ m, Real* __restrict__ cp, int M) {
    const int N = M - 1;
    const Real bw0 = dir0 ? b[0] : b[0] - chi[0];
    Real cprev = c[0] / bw0;
    cp[0] = cprev;
    m[0] = Real(0);
    // rows 1 .. N-1 in groups of four through two register sets, as thomas_lds
    struct Rows { Real w0, w1, w2, w3, a0, a1, a2, a3, c0, c1, c2, c3; };
    auto ld = [&](Rows& g, int r) {
        g.w0 = b[r] - chi[r]; g.w1 = b[r + 1] - chi[r + 1]; g.w2 = b[r + 2] - chi[r + 2]; g.w3 = b[r + 3] - chi[r + 3];
        g.a0 = a[r]; g.a1 = a[r + 1]; g.a2 = a[r + 2]; g.a3 = a[r + 3];
        g.c0 = c[r]; g.c1 = c[r + 1]; g.c2 = c[r + 2]; g.c3 = c[r + 3];
    };
    auto step = [&](const Rows& g, int r) {
        Real mi;
        mi = Real(1) / (g.w0 - g.a0 * cprev); cprev = g.c0 * mi; m[r] = mi; cp[r] = cprev;
        mi = Real(1) / (g.w1 - g.a1 * cprev); cprev = g.c1 * mi; m[r + 1] = mi; cp[r + 1] = cprev;
        mi = Real(1) / (g.w2 - g.a2 * cprev); cprev = g.c2 * mi; m[r + 2] = mi; cp[r + 2] = cprev;
        mi = Real(1) / (g.w3 - g.a3 * cprev); cprev = g.c3 * mi; m[r + 3] = mi; cp[r + 3] = cprev;
    };
    const int nq = (N - 1) >> 2;
    int i = 1;
    if (nq > 0) {
        Rows A, B;
        ld(A, 1);
        int q = 0;
        for (; q + 2 <= nq; q += 2, i += 8) {
            ld(B, i + 4);
            step(A, i);
            ld(A, (q + 2 < nq) ? i + 8 : i + 4);
            step(B, i + 4);
        }
        if (q < nq) {
            step(A, i);
            i += 4;
        }
    }
    for (; i < N; ++i) {
        const Real mi = Real(1) / ((b[i] - chi[i]) - a[i] * cprev);
        cprev = c[i] * mi;
        m[i] = mi;
        cp[i] = cprev;
    }
    const Real bwN = dirN ? bN : bN - chi[N];
    const Real mN = Real(1) / (bwN - aN * cprev);
    m[N] = mN;
    cp[N] = c[N] * mN;
    return bw0;
}

// exclusive scan of n ints in place by one wave
__device__ __forceinline__ void wave_exclusive_scan(int* __restrict__ v, int n, int lane) {
    const int per = (n + 63) / 64;
    const int lo = lane * per, hi = min(lo + per, n);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += v[i];
    int incl = s;
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    int run = incl - s;
    for (int i = lo; i < hi; ++i) {
        const int t = v[i];
        v[i] = run;
        run += t;
    }
}

// ((0 + v[e0]) + v[e0+1]) + ... + v[e1-1]: the sequential sum of one node list.  Lists are long where the
// plasma is compressed (hundreds of entries in a blowout sheath), so the loads run one batch of eight ahead of
// the dependent adds: a long list costs about one add latency per entry.
struct Batch8 {
    Real a0, a1, a2, a3, a4, a5, a6, a7;
    __device__ __forceinline__ void load(const Real* v) {
        a0 = v[0]; a1 = v[1]; a2 = v[2]; a3 = v[3]; a4 = v[4]; a5 = v[5]; a6 = v[6]; a7 = v[7];
    }
    __device__ __forceinline__ void add_to(Real& acc) const {
        acc += a0; acc += a1; acc += a2; acc += a3; acc += a4; acc += a5; acc += a6; acc += a7;
    }
};

__device__ __forceinline__ Real sum_list(const Real* v, int e0, int e1) {
    Real acc = Real(0);
    int e = e0;
    // 24 entries per iteration, all loads issued before the adds: the adds of the first batch cover the
    // latency of the other two, leaving one exposed LDS round trip per 24 entries
    for (; e + 24 <= e1; e += 24) {
        Batch8 x, y, z;
        x.load(v + e);
        y.load(v + e + 8);
        z.load(v + e + 16);
        x.add_to(acc);
        y.add_to(acc);
        z.add_to(acc);
    }
    if (e + 16 <= e1) {
        Batch8 x, y;
        x.load(v + e);
        y.load(v + e + 8);
        x.add_to(acc);
        y.add_to(acc);
        e += 16;
    } else if (e + 8 <= e1) {
        Batch8 x;
        x.load(v + e);
        x.add_to(acc);
        e += 8;
    }
    if (e + 4 <= e1) {
        const Real x0 = v[e], x1 = v[e + 1], x2 = v[e + 2], x3 = v[e + 3];
        acc += x0; acc += x1; acc += x2; acc += x3;
        e += 4;
    }
    for (; e < e1; ++e) acc += v[e];
    return acc;
}

// entries < key in the index-sorted range idx[lo, hi)
__device__ __forceinline__ int count_less(const unsigned short* __restrict__ idx, int lo, int hi, int key) {
    int a = lo, b = hi;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (static_cast<int>(idx[mid]) < key) a = mid + 1; else b = mid;
    }
    return a - lo;
}

// Opaque register fence: the compiler must assume the values changed here, so nothing derived from them is
// hoisted out of the enclosing loop (per-particle live ranges stay at the cache itself).
template <class T, int PT> __device__ __forceinline__ void fence_regs(T (&a)[PT]) {
#pragma unroll
    for (int i = 0; i < PT; ++i) asm volatile("" : "+v"(a[i]));
}

__device__ __forceinline__ Real lin(const Real* f, int j, Real t) { return (Real(1) - t) * f[j] + t * f[j + 1]; }

// 3-point gradients and f/r (FieldSolver::gradient / gradient_odd / over_r, src/FieldSolver.cpp:263-289)
struct Stencil {
    const Real *gm, *g0, *gp, *ir;
    int N;
    __device__ __forceinline__ Real grad(const Real* f, int j) const {
        if (j == 0) return Real(0);
        if (j == N) return g0[N] * f[N] + gm[N] * f[N - 1] + gp[N] * f[N - 2];
        return gm[j] * f[j - 1] + g0[j] * f[j] + gp[j] * f[j + 1];
    }
    __device__ __forceinline__ Real grad_odd(const Real* f, int j) const {
        if (j == 0) return f[1] * ir[0];
        if (j == N) return g0[N] * f[N] + gm[N] * f[N - 1] + gp[N] * f[N - 2];
        return gm[j] * f[j - 1] + g0[j] * f[j] + gp[j] * f[j + 1];
    }
    __device__ __forceinline__ Real over_r(const Real* f, int j) const { return (j == 0 ? f[1] : f[j]) * ir[j]; }
};

// The cold rings of one species at the head of the box (PlasmaSpecies::load, src/PlasmaSpecies.cpp:202-295, with
// the host load's expression order): particles [i0, i1) of `st`, node radii r (LDS), and a fresh AB history.
__device__ __forceinline__ void load_rings(const qz_plasma_rings& rg, const qz_species_state& st, int ord,
                                           const Real* r, int i0, int i1, int stride) {
    constexpr Real PI = 3.14159265358979323846;
    const int nt = rg.ntheta, ppc = rg.ppc;
    for (int i = i0; i < i1; i += stride) {
        const int ring = i / nt, l = i - ring * nt;
        const int j = ring / ppc, sub = ring - j * ppc;
        const Real hc = r[j + 1] - r[j];
        const Real ra = r[j] + Real(sub) * hc / Real(ppc), rb = r[j] + Real(sub + 1) * hc / Real(ppc);
        const Real rc = (Real(2.0) / Real(3.0)) * (rb * rb * rb - ra * ra * ra) / (rb * rb - ra * ra);
        const Real volume = (Real(1.0) * rg.density) * PI * (rb * rb - ra * ra);
        const int a = (ring % 2) * nt + l;
        st.x[i] = nt == 1 ? rc : rc * rg.cs[a];
        st.y[i] = nt == 1 ? Real(0) : rc * rg.cs[2 * nt + a];
        st.px[i] = Real(0);
        st.py[i] = Real(0);
        st.dl[i] = Real(1);   // gamma - p_z of a particle at rest
        st.w[i] = Real(1.0) * volume / Real(nt);
        for (int v = 0; v < 5 * ord; ++v) st.hist[static_cast<size_t>(v) * st.cap + i] = Real(0);
    }
}

}  // namespace qz
