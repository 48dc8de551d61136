// Persistent gfx950 kernel of one quasi-static sweep: the whole slice loop of Simulation::step_fields
// (src/Simulation.cpp:727-783) for one run in one workgroup, walking xi from the head of the box.
//
// Per slice:
//   1. geometry of every live plasma particle (kept in registers for the slice);
//   2. a stable counting sort of the particles by radial cell in LDS (ballot ranks per wave), from which
//      every particle gets its position in the contribution lists of its two nodes, in particle-index
//      order;
//   3. the mode deposit (PlasmaSpecies::deposit, src/PlasmaSpecies.cpp:409-478): particles write their two
//      contributions per array into LDS, one thread per node then sums its list sequentially -- the
//      reference's atomic_add performed in index order, i.e. the serial result bit for bit;
//   4. combine with the background and the beam sources (Simulation::combine_sources);
//   5. field solves: psi, E_z, B_z (chi-free Thomas with factors formed once per run) while another wave
//      forms the per-slice Thomas factors of the chi-dependent B+ operators; W+, the S+ deposit, and the
//      Picard solves for B+ (Simulation::solve_slice_fields / compute_wplus / compute_bz / solve_bplus);
//   6. the field store (Simulation::store_slice) and the Adams-Bashforth push (PlasmaSpecies::push).
// Per-slice arrays live in the run's workspace (L2-resident); the deposit lists and the sort in LDS.
#include <hip/hip_runtime.h>

#include "qz_device.hpp"
#include "qz_internal.hpp"

namespace qz {

// workspace arrays (M doubles each)
enum WsArr : int {
    // deposit components (the first 30 are also the deposit component ids)
    W_RHOT0 = 0, W_DRHO0, W_JZ0, W_CHI0, W_NE0, W_NI0, W_JPR1, W_JPI1,
    W_RHOT1R, W_RHOT1I, W_DRHO1R, W_DRHO1I, W_JZ1R, W_JZ1I, W_CHI1R, W_CHI1I, W_NE1R, W_NE1I, W_NI1R, W_NI1I,
    W_JPR0, W_JPI0, W_JPR2, W_JPI2,
    W_SR1, W_SI1, W_SR0, W_SI0, W_SR2, W_SI2,
    // fields of the slice
    W_PSI0, W_PSI1R, W_PSI1I, W_EZ0, W_EZ1R, W_EZ1I, W_BZ0, W_BZ1R, W_BZ1I,
    W_WPR0, W_WPI0, W_WPR1, W_WPI1, W_WPR2, W_WPI2,
    W_BPR0, W_BPI0, W_BPR1, W_BPI1, W_BPR2, W_BPI2,
    // right-hand sides
    W_T3, W_T4, W_RHS,
    W_GR0, W_GI0, W_GR1, W_GI1, W_GR2, W_GI2,
    W_RBR0, W_RBI0, W_RBR1, W_RBI1, W_RBR2, W_RBI2,
    // Thomas factors (m, c') of the chi-dependent operators of this slice
    W_FM_L0, W_FCP_L0, W_FM_L1D, W_FCP_L1D, W_FM_L1F, W_FCP_L1F, W_FM_L2D, W_FCP_L2D,
    // forward-sweep scratch of concurrent solves
    W_DP0, W_DP_END = W_DP0 + 9,
    W_NARR = W_DP_END
};

enum OpId : int { OP_L0 = 0, OP_L1D = 1, OP_L1F = 2, OP_L2D = 3 };

constexpr int kBS = 512;           // threads per run
constexpr int kNW = kBS / 64;      // waves per run
constexpr int kLdsBytes = 160 * 1024;

struct SpecInfo {
    const qz_species_state* st;
    int off, np;
    Real q, qm;
};

// One Thomas solve with factors formed beforehand (Tridiag.cpp:114-125 with cp, m precomputed):
//   dp_0 = d_0 / b_0,  dp_i = (d_i - a_i dp_{i-1}) m_i,  x_{M-1} = dp_{M-1},  x_i = dp_i - cp_i x_{i+1}
// d_i is the prepared right-hand side of FieldSolver::solve (src/FieldSolver.cpp:253-259).
__device__ void thomas_apply(const Real* __restrict__ a, const Real* __restrict__ m, const Real* __restrict__ cp,
                             Real b0, const Real* __restrict__ rhs, Real scale, bool dir0, bool dirN, bool fluxN,
                             Real fcw, Real* __restrict__ dp, Real* __restrict__ x, int M) {
    const int N = M - 1;
    Real d0 = dir0 ? Real(0) : scale * rhs[0];
    Real prev = d0 / b0;
    dp[0] = prev;
    for (int i = 1; i < N; ++i) {
        const Real di = scale * rhs[i];
        prev = (di - a[i] * prev) * m[i];
        dp[i] = prev;
    }
    {
        Real dN;
        if (dirN) dN = Real(0);
        else {
            dN = scale * rhs[N];
            if (fluxN) dN -= fcw;
        }
        prev = (dN - a[N] * prev) * m[N];
        dp[N] = prev;
    }
    Real xn = prev;
    x[N] = xn;
    for (int i = N - 1; i >= 0; --i) {
        xn = dp[i] - cp[i] * xn;
        x[i] = xn;
    }
}

// Thomas factors of (L - chi) with the Dirichlet rows of FieldSolver::solve (src/FieldSolver.cpp:253-256):
//   bw_i = b_i (Dirichlet row) or b_i - chi_i;  cp_0 = c_0 / bw_0;  m_i = 1 / (bw_i - a_i cp_{i-1});  cp_i = c_i m_i
// Returns bw_0 for the first row of the solve.
__device__ Real thomas_factor(const Real* __restrict__ a, const Real* __restrict__ b, const Real* __restrict__ c,
                              const Real* __restrict__ chi, bool dir0, bool dirN, Real* __restrict__ m,
                              Real* __restrict__ cp, int M) {
    const int N = M - 1;
    const Real bw0 = dir0 ? b[0] : b[0] - chi[0];
    Real cprev = c[0] / bw0;
    cp[0] = cprev;
    m[0] = Real(0);
    for (int i = 1; i < M; ++i) {
        const Real bwi = (i == N && dirN) ? b[i] : b[i] - chi[i];
        const Real mi = Real(1) / (bwi - a[i] * cprev);
        cprev = c[i] * mi;
        m[i] = mi;
        cp[i] = cprev;
    }
    return bw0;
}

// block-wide exclusive scan of n ints in place (one wave does it; the others wait at the caller's barrier)
__device__ void wave_exclusive_scan(int* __restrict__ v, int n, int lane) {
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

// number of entries < key in the sorted range idx[lo, hi)
__device__ __forceinline__ int count_less(const unsigned short* __restrict__ idx, int lo, int hi, int key) {
    int a = lo, b = hi;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (static_cast<int>(idx[mid]) < key) a = mid + 1; else b = mid;
    }
    return a - lo;
}

template <bool M1, int PT>
struct Sweep {
    // ---------------------------------------------------------------- per-thread particle cache (one slice)
    int cj[PT], cpl[PT], cpr[PT], cspec[PT], cli[PT];
    Real ct[PT], cc[PT], cs[PT], cir[PT], cinvd[PT], cgam[PT], cPX[PT], cPY[PT], cD[PT], cw[PT];

    __device__ __forceinline__ int slot(int p, int w, int lane) const { return w * (PT * 64) + p * 64 + lane; }
};

// value of deposit component `comp` for the node on side cc (0: j, 1: j+1) of a particle
// (PlasmaSpecies::deposit, src/PlasmaSpecies.cpp:420-477; rho is not used on this path and not formed)
__device__ __forceinline__ Real dep_value(int comp, int cc, int j, Real t, Real c, Real sn, Real ir, Real invd,
                                          Real gam, Real PX, Real PY, Real D, Real wi, Real q, Real qm,
                                          const Real* __restrict__ inv_V, const Real* __restrict__ inv_h) {
    const int jj = j + cc;
    const Real iv = inv_V[jj];
    const Real ih = inv_h[j];
    const Real W = cc == 0 ? Real(1) - t : t;
    const Real dW = cc == 0 ? -ih : ih;
    const Real pz = gam - D;
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

__device__ __forceinline__ Real lin(const Real* __restrict__ f, int j, Real t) {
    return (Real(1) - t) * f[j] + t * f[j + 1];
}

template <bool M1, int PT>
__global__ __launch_bounds__(kBS) void sweep_kernel(const qz_sweep_args* __restrict__ runs) {
    const qz_sweep_args& A = runs[blockIdx.x];
    extern __shared__ double smem_d[];
    char* smem = reinterpret_cast<char*>(smem_d);

    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int N = A.grid.N, M = N + 1;
    const Real R = A.grid.R;
    const Real* __restrict__ gr = A.grid.r;
    const Real* __restrict__ gih = A.grid.inv_h;
    const Real* __restrict__ giv = A.grid.inv_V;
    Real* __restrict__ ws = A.ws;
    auto WS = [&](int arr) -> Real* { return ws + static_cast<size_t>(arr) * M; };
    const int K = A.nslices;
    const Real dxi = A.dxi;
    const Real idxi = Real(1) / dxi;

    // ---- species
    const int nsp = A.nspecies;
    SpecInfo sp[QZ_MAX_SPECIES];
    int NP = 0;
    bool has_neg = false, has_pos = false;
    for (int s = 0; s < QZ_MAX_SPECIES; ++s) {
        sp[s].st = &A.sp[s];
        sp[s].off = NP;
        sp[s].np = s < nsp ? A.sp[s].np : 0;
        sp[s].q = s < nsp ? A.sp[s].q : Real(0);
        sp[s].qm = s < nsp ? A.sp[s].q / A.sp[s].m : Real(0);
        if (s < nsp) {
            NP += sp[s].np;
            if (sp[s].q < 0) has_neg = true; else has_pos = true;
        }
    }

    // ---- LDS carve-up: wave counters, cell starts, sorted indices, contribution lists
    unsigned short* s_wcnt = reinterpret_cast<unsigned short*>(smem);          // kNW * N
    int off = ((kNW * N * 2) + 15) & ~15;
    int* s_cs = reinterpret_cast<int*>(smem + off);                           // N + 3 (cell starts shifted by one)
    off += (((N + 3) * 4) + 15) & ~15;
    unsigned short* s_sidx = reinterpret_cast<unsigned short*>(smem + off);   // kBS * PT
    off += ((kBS * PT * 2) + 15) & ~15;
    int* s_misc = reinterpret_cast<int*>(smem + off);                         // 16 ints
    off += 64;
    Real* s_vbuf = reinterpret_cast<Real*>(smem + off);
    const int vbuf_len = (kLdsBytes - off) / 8;

    // ---- deposit component lists
    int comps[24];
    int ncomp = 0;
    comps[ncomp++] = W_RHOT0; comps[ncomp++] = W_DRHO0; comps[ncomp++] = W_JZ0; comps[ncomp++] = W_CHI0;
    if (has_neg) comps[ncomp++] = W_NE0;
    if (has_pos) comps[ncomp++] = W_NI0;
    comps[ncomp++] = W_JPR1; comps[ncomp++] = W_JPI1;
    if (M1) {
        comps[ncomp++] = W_RHOT1R; comps[ncomp++] = W_RHOT1I; comps[ncomp++] = W_DRHO1R; comps[ncomp++] = W_DRHO1I;
        comps[ncomp++] = W_JZ1R; comps[ncomp++] = W_JZ1I; comps[ncomp++] = W_CHI1R; comps[ncomp++] = W_CHI1I;
        if (has_neg) { comps[ncomp++] = W_NE1R; comps[ncomp++] = W_NE1I; }
        if (has_pos) { comps[ncomp++] = W_NI1R; comps[ncomp++] = W_NI1I; }
        comps[ncomp++] = W_JPR0; comps[ncomp++] = W_JPI0; comps[ncomp++] = W_JPR2; comps[ncomp++] = W_JPI2;
    }
    const int nS = M1 ? 6 : 2;   // S components W_SR1.. in order r1 i1 r0 i0 r2 i2

    // ---- B+ := 0 at the head of the box (src/Simulation.cpp:705-708); absent density components are 0
    for (int n = tid; n < M; n += kBS) {
        for (int a = W_BPR0; a <= W_BPI2; ++a) WS(a)[n] = Real(0);
    }
    __syncthreads();

    Sweep<M1, PT> P;
    const Real dmin = A.delta_min, qsamax = A.max_qsa;
    const int ord = A.ab_order;

    for (int k = 0; k < K; ++k) {
        // ================================================================ 1. geometry (PlasmaSpecies.cpp:52-70)
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int gi = P.slot(p, w, lane);
            P.cj[p] = -1;
            if (gi < NP) {
                const int s = (nsp > 1 && gi >= sp[1].off) ? 1 : 0;
                const int li = gi - sp[s].off;
                P.cspec[p] = s;
                P.cli[p] = li;
                const qz_species_state* st = sp[s].st;
                const Real wi = st->w[li];
                P.cw[p] = wi;
                if (wi != Real(0)) {
                    const Real X = st->x[li], Y = st->y[li];
                    const Real PX = st->px[li], PY = st->py[li], D = st->dl[li];
                    const Real r = dsqrt(X * X + Y * Y);
                    Real c, s_;
                    if (r > Real(0)) { c = X / r; s_ = Y / r; } else { c = Real(1); s_ = Real(0); }
                    int j; Real t;
                    grid_weights(gr, gih, N, R, r, j, t);
                    P.cj[p] = j;
                    P.ct[p] = t;
                    P.cc[p] = c;
                    P.cs[p] = s_;
                    P.cir[p] = r > Real(0) ? Real(1) / r : Real(0);
                    P.cgam[p] = (Real(1) + PX * PX + PY * PY + D * D) / (Real(2) * D);
                    P.cinvd[p] = Real(1) / D;
                    P.cPX[p] = PX; P.cPY[p] = PY; P.cD[p] = D;
                }
            }
        }
        // ================================================================ 2. stable sort by cell
        for (int i = tid; i < kNW * N; i += kBS) s_wcnt[i] = 0;
        __syncthreads();
        int lrank[PT];
        const unsigned long long lt_mask = (1ull << lane) - 1ull;
        int nbits = 1;
        while ((1 << nbits) < N) ++nbits;
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int key = P.cj[p];
            unsigned long long mask = __ballot(key >= 0);
            for (int b = 0; b < nbits; ++b) {
                const bool bit = key >= 0 && ((key >> b) & 1);
                const unsigned long long bb = __ballot(bit);
                mask &= bit ? bb : ~bb;
            }
            int old = 0;
            if (key >= 0) old = s_wcnt[w * N + key];
            const int rk = __popcll(mask & lt_mask);
            lrank[p] = old + rk;
            __builtin_amdgcn_wave_barrier();
            if (key >= 0 && rk == 0) s_wcnt[w * N + key] = static_cast<unsigned short>(old + __popcll(mask));
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        // per cell: exclusive prefix over waves (in place) and the cell total
        for (int c = tid; c < N; c += kBS) {
            int run = 0;
            for (int ww = 0; ww < kNW; ++ww) {
                const int v = s_wcnt[ww * N + c];
                s_wcnt[ww * N + c] = static_cast<unsigned short>(run);
                run += v;
            }
            s_cs[c + 1] = run;
        }
        __syncthreads();
        if (w == 0) wave_exclusive_scan(s_cs + 1, N, lane);   // s_cs[c + 1] = start of cell c
        if (tid == 0) s_cs[0] = 0;                             // empty cell -1
        __syncthreads();
        // place the sorted particle indices: (cell, index) order
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int j = P.cj[p];
            if (j >= 0) {
                const int pos = s_cs[j + 1] + s_wcnt[w * N + j] + lrank[p];
                s_sidx[pos] = static_cast<unsigned short>(P.slot(p, w, lane));
                lrank[p] = pos - s_cs[j + 1];   // rank within the cell (index order)
            }
        }
        // live total L = number of depositing particles
        if (tid == 0) s_misc[0] = 0;
        __syncthreads();
        {
            int cnt = 0;
#pragma unroll
            for (int p = 0; p < PT; ++p) cnt += P.cj[p] >= 0 ? 1 : 0;
            for (int d = 32; d >= 1; d >>= 1) cnt += __shfl_xor(cnt, d, 64);
            if (lane == 0) atomicAdd(&s_misc[0], cnt);
        }
        __syncthreads();
        const int L = s_misc[0];
        if (tid == 0) {
            s_cs[N + 1] = L;
            s_cs[N + 2] = L;
            if (A.stats) atomicAdd(A.stats, static_cast<unsigned long long>(L));
        }
        __syncthreads();
        // positions in the node lists: node n holds cells n-1 (right entries) and n (left entries), index order
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int j = P.cj[p];
            if (j >= 0) {
                const int gi = P.slot(p, w, lane);
                const int less_m = count_less(s_sidx, s_cs[j], s_cs[j + 1], gi);        // cell j-1
                const int less_p = count_less(s_sidx, s_cs[j + 2], s_cs[j + 3], gi);   // cell j+1
                P.cpl[p] = (s_cs[j] + s_cs[j + 1]) + lrank[p] + less_m;
                P.cpr[p] = (s_cs[j + 1] + s_cs[j + 2]) + lrank[p] + less_p;
            }
        }
        const int G = max(1, min(vbuf_len / max(2 * L, 1), 8));

        // ================================================================ 3. deposit (node sums in index order)
        for (int c0 = 0; c0 < ncomp; c0 += G) {
            const int g1 = min(G, ncomp - c0);
            __syncthreads();
#pragma unroll
            for (int p = 0; p < PT; ++p) {
                const int j = P.cj[p];
                if (j >= 0) {
                    const SpecInfo& S = sp[P.cspec[p]];
                    for (int g = 0; g < g1; ++g) {
                        const int comp = comps[c0 + g];
                        s_vbuf[g * 2 * L + P.cpl[p]] = dep_value(comp, 0, j, P.ct[p], P.cc[p], P.cs[p], P.cir[p], P.cinvd[p], P.cgam[p],
                                                                 P.cPX[p], P.cPY[p], P.cD[p], P.cw[p], S.q, S.qm, giv, gih);
                        s_vbuf[g * 2 * L + P.cpr[p]] = dep_value(comp, 1, j, P.ct[p], P.cc[p], P.cs[p], P.cir[p], P.cinvd[p], P.cgam[p],
                                                                 P.cPX[p], P.cPY[p], P.cD[p], P.cw[p], S.q, S.qm, giv, gih);
                    }
                }
            }
            __syncthreads();
            for (int n = tid; n < M; n += kBS) {
                const int e0 = s_cs[n] + s_cs[n + 1], e1 = s_cs[n + 1] + s_cs[n + 2];
                for (int g = 0; g < g1; ++g) {
                    Real acc = Real(0);
                    const Real* vb = s_vbuf + g * 2 * L;
                    for (int e = e0; e < e1; ++e) acc += vb[e];
                    WS(comps[c0 + g])[n] = acc;
                }
            }
        }
        __syncthreads();

        // absent density components stay zero (zero_sources, Simulation.cpp:394-406)
        for (int n = tid; n < M; n += kBS) {
            if (!has_neg) { WS(W_NE0)[n] = Real(0); WS(W_NE1R)[n] = Real(0); WS(W_NE1I)[n] = Real(0); }
            if (!has_pos) { WS(W_NI0)[n] = Real(0); WS(W_NI1R)[n] = Real(0); WS(W_NI1I)[n] = Real(0); }
        }
        __syncthreads();
        // ================================================================ 4. combine_sources (Simulation.cpp:408-457)
        {
            const int KL = K;
            const int ncb = A.ncb;
            const Real* __restrict__ bs = A.bsrc;
            const Real* __restrict__ bg = A.bg;
            const bool ngp = A.ngp != 0;
            auto b = [&](int kk, int j, int comp) { return bs[(static_cast<size_t>(kk) * M + j) * ncb + comp]; };
            for (int j = tid; j < M; j += kBS) {
                auto dxi_b = [&](int comp) {
                    if (ngp) {
                        const Real prev = (k > 0) ? b(k - 1, j, comp) : Real(0);
                        return (b(k, j, comp) - prev) * idxi;
                    }
                    if (KL == 1) return Real(0);
                    if (k == 0) return (b(1, j, comp) - b(0, j, comp)) * idxi;
                    if (k == KL - 1) return (b(KL - 1, j, comp) - b(KL - 2, j, comp)) * idxi;
                    return Real(0.5) * (b(k + 1, j, comp) - b(k - 1, j, comp)) * idxi;
                };
                const Real bgt0 = bg[j], bgr0 = bg[3 * M + j];
                WS(W_RHOT0)[j] += bgt0 + b(k, j, B_RT0);
                WS(W_DRHO0)[j] += dxi_b(B_RT0);
                WS(W_JZ0)[j] += b(k, j, B_JZ0);
                WS(W_JPR1)[j] += b(k, j, B_JP1R);
                WS(W_JPI1)[j] += b(k, j, B_JP1I);
                if (bgr0 > Real(0)) WS(W_NI0)[j] += bgr0;
                if (M1) {
                    WS(W_RHOT1R)[j] += bg[M + j] + b(k, j, B_RT1R);
                    WS(W_RHOT1I)[j] += bg[2 * M + j] + b(k, j, B_RT1I);
                    WS(W_DRHO1R)[j] += dxi_b(B_RT1R);
                    WS(W_DRHO1I)[j] += dxi_b(B_RT1I);
                    WS(W_JZ1R)[j] += b(k, j, B_JZ1R);
                    WS(W_JZ1I)[j] += b(k, j, B_JZ1I);
                    WS(W_JPR0)[j] += b(k, j, B_JP0R);
                    WS(W_JPI0)[j] += b(k, j, B_JP0I);
                    WS(W_JPR2)[j] += b(k, j, B_JP2R);
                    WS(W_JPI2)[j] += b(k, j, B_JP2I);
                }
            }
        }
        __syncthreads();

        // ================================================================ 5a. B_z right-hand side (compute_bz, Simulation.cpp:506-526)
        {
            const Real* __restrict__ gm = A.grad_m;
            const Real* __restrict__ g0 = A.grad_0;
            const Real* __restrict__ gp = A.grad_p;
            const Real* __restrict__ ir = A.inv_r;
            auto grad = [&](const Real* f, int j) -> Real {
                if (j == 0) return Real(0);
                if (j == N) return g0[N] * f[N] + gm[N] * f[N - 1] + gp[N] * f[N - 2];
                return gm[j] * f[j - 1] + g0[j] * f[j] + gp[j] * f[j + 1];
            };
            auto grad_odd = [&](const Real* f, int j) -> Real {
                if (j == 0) return f[1] * ir[0];
                if (j == N) return g0[N] * f[N] + gm[N] * f[N - 1] + gp[N] * f[N - 2];
                return gm[j] * f[j - 1] + g0[j] * f[j] + gp[j] * f[j + 1];
            };
            auto over_r = [&](const Real* f, int j) -> Real { return (j == 0 ? f[1] : f[j]) * ir[j]; };
            for (int j = tid; j < M; j += kBS) {
                const Real t1 = grad_odd(WS(W_JPI1), j), t2 = over_r(WS(W_JPI1), j);
                WS(W_T3)[j] = -(t1 + t2);
                if (M1) {
                    const Real gr0 = grad_odd(WS(W_JPR2), j), gi0 = over_r(WS(W_JPR2), j);
                    const Real gr1 = grad_odd(WS(W_JPI2), j), gi1 = over_r(WS(W_JPI2), j);
                    const Real gr2 = grad(WS(W_JPR0), j), gi2 = grad(WS(W_JPI0), j);
                    const Real D1r = gr0 + Real(2) * gi0, D1i = gr1 + Real(2) * gi1;
                    const Real Dmr = gr2, Dmi = gi2;
                    WS(W_T4)[j] = -Real(0.5) * (D1i + Dmi);
                    WS(W_RHS)[j] = Real(0.5) * (D1r - Dmr);
                }
            }
        }
        __syncthreads();

        // ================================================================ 5b. psi, E_z, B_z solves || chi factors
        if (w == 0) {
            // lanes: 0 psi0, 1 ez0, 2 bz0 (L0);  m1: 3 psi1r, 4 psi1i, 5 ez1r, 6 ez1i, 7 bz1r, 8 bz1i (L1D)
            const int njobs = M1 ? 9 : 3;
            if (lane < njobs) {
                int op, rhs, out; Real scale;
                switch (lane) {
                    case 0: op = OP_L0; rhs = W_RHOT0; out = W_PSI0; scale = Real(-1); break;
                    case 1: op = OP_L0; rhs = W_DRHO0; out = W_EZ0; scale = Real(-1); break;
                    case 2: op = OP_L0; rhs = W_T3; out = W_BZ0; scale = Real(1); break;
                    case 3: op = OP_L1D; rhs = W_RHOT1R; out = W_PSI1R; scale = Real(-1); break;
                    case 4: op = OP_L1D; rhs = W_RHOT1I; out = W_PSI1I; scale = Real(-1); break;
                    case 5: op = OP_L1D; rhs = W_DRHO1R; out = W_EZ1R; scale = Real(-1); break;
                    case 6: op = OP_L1D; rhs = W_DRHO1I; out = W_EZ1I; scale = Real(-1); break;
                    case 7: op = OP_L1D; rhs = W_T4; out = W_BZ1R; scale = Real(1); break;
                    default: op = OP_L1D; rhs = W_RHS; out = W_BZ1I; scale = Real(1); break;
                }
                const bool l0 = op == OP_L0;
                thomas_apply(A.op_a[op], l0 ? A.fac_l0_m : A.fac_l1_m, l0 ? A.fac_l0_cp : A.fac_l1_cp, A.op_b[op][0],
                             WS(rhs), scale, !l0, true, false, Real(0), WS(W_DP0 + lane), WS(out), M);
            }
        } else if (w == 1) {
            // per-slice factors of (L - chi0): lanes 0 L1D, 1 L1F, m1: 2 L0, 3 L2D
            const int nf = M1 ? 4 : 2;
            if (lane < nf) {
                const int op = lane == 0 ? OP_L1D : lane == 1 ? OP_L1F : lane == 2 ? OP_L0 : OP_L2D;
                const int fm = lane == 0 ? W_FM_L1D : lane == 1 ? W_FM_L1F : lane == 2 ? W_FM_L0 : W_FM_L2D;
                const bool dir0 = op != OP_L0;
                const bool dirN = op != OP_L1F;
                const Real bw0 = thomas_factor(A.op_a[op], A.op_b[op], A.op_c[op], WS(W_CHI0), dir0, dirN, WS(fm), WS(fm + 1), M);
                // bw0 of each operator is kept in LDS for the Picard solves
                reinterpret_cast<Real*>(s_misc + 8)[lane] = bw0;
            }
        }
        __syncthreads();

        // ================================================================ 5c. W+ (compute_wplus, Simulation.cpp:477-498)
        {
            const Real* __restrict__ gm = A.grad_m;
            const Real* __restrict__ g0 = A.grad_0;
            const Real* __restrict__ gp = A.grad_p;
            const Real* __restrict__ ir = A.inv_r;
            for (int j = tid; j < M; j += kBS) {
                const Real* f = WS(W_PSI0);
                Real t1;
                if (j == 0) t1 = Real(0);
                else if (j == N) t1 = g0[N] * f[N] + gm[N] * f[N - 1] + gp[N] * f[N - 2];
                else t1 = gm[j] * f[j - 1] + g0[j] * f[j] + gp[j] * f[j + 1];
                WS(W_WPR1)[j] = -t1;
                WS(W_WPI1)[j] = Real(0);
                if (M1) {
                    const Real* fr = WS(W_PSI1R);
                    const Real* fi = WS(W_PSI1I);
                    Real t2, t4;
                    if (j == 0) { t2 = fr[1] * ir[0]; t4 = fi[1] * ir[0]; }
                    else if (j == N) {
                        t2 = g0[N] * fr[N] + gm[N] * fr[N - 1] + gp[N] * fr[N - 2];
                        t4 = g0[N] * fi[N] + gm[N] * fi[N - 1] + gp[N] * fi[N - 2];
                    } else {
                        t2 = gm[j] * fr[j - 1] + g0[j] * fr[j] + gp[j] * fr[j + 1];
                        t4 = gm[j] * fi[j - 1] + g0[j] * fi[j] + gp[j] * fi[j + 1];
                    }
                    const Real t3 = (j == 0 ? fr[1] : fr[j]) * ir[j];
                    const Real rh = (j == 0 ? fi[1] : fi[j]) * ir[j];
                    WS(W_WPR2)[j] = -(t2 - t3);
                    WS(W_WPI2)[j] = -(t4 - rh);
                    WS(W_WPR0)[j] = -(t2 + t3);
                    WS(W_WPI0)[j] = (t4 + rh);
                }
            }
        }
        __syncthreads();

        // ================================================================ 5d. S+ deposit (PlasmaSpecies::deposit_S, PlasmaSpecies.cpp:516-558)
        {
            const int GS = max(1, min(G, nS));
            for (int c0 = 0; c0 < nS; c0 += GS) {
                const int g1 = min(GS, nS - c0);
                __syncthreads();
#pragma unroll
                for (int p = 0; p < PT; ++p) {
                    const int j = P.cj[p];
                    if (j < 0) continue;
                    const SpecInfo& S = sp[P.cspec[p]];
                    const Real t = P.ct[p], c = P.cc[p], sn = P.cs[p];
                    const Real qq = S.q, qm = S.qm;
                    const Real PX = P.cPX[p], PY = P.cPY[p];
                    const Real wi = P.cw[p];
                    // gather (PlasmaSpecies.cpp:75-89) without B+
                    Real Wx, Wy;
                    eval_vector(lin(WS(W_WPR0), j, t), lin(WS(W_WPI0), j, t), lin(WS(W_WPR1), j, t), lin(WS(W_WPI1), j, t),
                                lin(WS(W_WPR2), j, t), lin(WS(W_WPI2), j, t), c, sn, M1, Wx, Wy);
                    const Real Ez = eval_scalar(lin(WS(W_EZ0), j, t), M1 ? lin(WS(W_EZ1R), j, t) : Real(0),
                                                M1 ? lin(WS(W_EZ1I), j, t) : Real(0), c, sn, M1);
                    const Real Bz = eval_scalar(lin(WS(W_BZ0), j, t), M1 ? lin(WS(W_BZ1R), j, t) : Real(0),
                                                M1 ? lin(WS(W_BZ1I), j, t) : Real(0), c, sn, M1);
                    const Real dA = Real(0);
                    const Real gam = P.cgam[p];
                    const Real invd = P.cinvd[p];
                    const Real fp = -Real(0.5) * qm * qm * dA * invd;
                    const Real ux = PX * invd, uy = PY * invd;
                    const Real ur = c * ux + sn * uy, uth = c * uy - sn * ux;
                    const Real dD = qm * ((PX * Wx + PY * Wy) * invd - Ez);
                    const Real ax = (qm * (gam * Wx + PY * Bz) * invd + fp * c) * invd - PX * dD * invd * invd;
                    const Real ay = (qm * (gam * Wy - PX * Bz) * invd + fp * sn) * invd - PY * dD * invd * invd;
                    const Real wq = wi * qq;
                    const Real ih = gih[j];
                    for (int g = 0; g < g1; ++g) {
                        const int sc = c0 + g;              // 0 r1, 1 i1, 2 r0, 3 i0, 4 r2, 5 i2
                        const int mode = sc < 2 ? 1 : (sc < 4 ? 0 : 2);
                        const Real ck = mode == 1 ? c : (mode == 0 ? Real(1) : c * c - sn * sn);
                        const Real sk = mode == 1 ? sn : (mode == 0 ? Real(0) : Real(2) * c * sn);
                        for (int cc = 0; cc < 2; ++cc) {
                            const int jj = j + cc;
                            const Real iv = giv[jj] * wq;
                            const Real W = cc == 0 ? Real(1) - t : t;
                            const Real dW = cc == 0 ? -ih : ih;
                            Real Ar = ax * W + ux * ur * dW;
                            Real Ai = ay * W + uy * ur * dW;
                            if (mode != 0) {
                                const Real Wor = (j == 0) ? (cc == 0 ? Real(0) : ih) : W * P.cir[p];
                                const Real f2 = mode * uth * Wor;
                                Ar += f2 * uy;
                                Ai -= f2 * ux;
                            }
                            const Real v = (sc & 1) == 0 ? iv * (Ar * ck + Ai * sk) : iv * (Ai * ck - Ar * sk);
                            s_vbuf[g * 2 * L + (cc == 0 ? P.cpl[p] : P.cpr[p])] = v;
                        }
                    }
                }
                __syncthreads();
                for (int n = tid; n < M; n += kBS) {
                    const int e0 = s_cs[n] + s_cs[n + 1], e1 = s_cs[n + 1] + s_cs[n + 2];
                    for (int g = 0; g < g1; ++g) {
                        Real acc = Real(0);
                        const Real* vb = s_vbuf + g * 2 * L;
                        for (int e = e0; e < e1; ++e) acc += vb[e];
                        WS(W_SR1 + c0 + g)[n] = acc;
                    }
                }
            }
        }
        __syncthreads();

        // ================================================================ 5e. gz (solve_bplus, Simulation.cpp:542-558)
        {
            const Real* __restrict__ gm = A.grad_m;
            const Real* __restrict__ g0 = A.grad_0;
            const Real* __restrict__ gp = A.grad_p;
            const Real* __restrict__ ir = A.inv_r;
            for (int j = tid; j < M; j += kBS) {
                const Real* f = WS(W_JZ0);
                Real t1;
                if (j == 0) t1 = Real(0);
                else if (j == N) t1 = g0[N] * f[N] + gm[N] * f[N - 1] + gp[N] * f[N - 2];
                else t1 = gm[j] * f[j - 1] + g0[j] * f[j] + gp[j] * f[j + 1];
                WS(W_GR1)[j] = t1 + WS(W_SR1)[j];
                WS(W_GI1)[j] = WS(W_SI1)[j];
                if (M1) {
                    const Real* fr = WS(W_JZ1R);
                    const Real* fi = WS(W_JZ1I);
                    Real t2, t4;
                    if (j == 0) { t2 = fr[1] * ir[0]; t4 = fi[1] * ir[0]; }
                    else if (j == N) {
                        t2 = g0[N] * fr[N] + gm[N] * fr[N - 1] + gp[N] * fr[N - 2];
                        t4 = g0[N] * fi[N] + gm[N] * fi[N - 1] + gp[N] * fi[N - 2];
                    } else {
                        t2 = gm[j] * fr[j - 1] + g0[j] * fr[j] + gp[j] * fr[j + 1];
                        t4 = gm[j] * fi[j - 1] + g0[j] * fi[j] + gp[j] * fi[j + 1];
                    }
                    const Real t3 = (j == 0 ? fr[1] : fr[j]) * ir[j];
                    const Real rh = (j == 0 ? fi[1] : fi[j]) * ir[j];
                    WS(W_GR2)[j] = (t2 - t3) + WS(W_SR2)[j];
                    WS(W_GI2)[j] = (t4 - rh) + WS(W_SI2)[j];
                    WS(W_GR0)[j] = (t2 + t3) + WS(W_SR0)[j];
                    WS(W_GI0)[j] = -(t4 + rh) + WS(W_SI0)[j];
                }
            }
        }
        const Real* bw0s = reinterpret_cast<const Real*>(s_misc + 8);
        const int niter = M1 ? A.picard : 1;
        for (int it = 0; it < niter; ++it) {
            __syncthreads();
            // ---- Picard right-hand sides (Simulation.cpp:563-581)
            for (int j = tid; j < M; j += kBS) {
                Real r1 = -WS(W_GI1)[j], i1 = WS(W_GR1)[j];
                if (M1) {
                    Real r0 = -WS(W_GI0)[j], i0 = WS(W_GR0)[j], r2 = -WS(W_GI2)[j], i2 = WS(W_GR2)[j];
                    const Real cr = WS(W_CHI1R)[j], ci = WS(W_CHI1I)[j];
                    const Real b0r = WS(W_BPR0)[j], b0i = WS(W_BPI0)[j], b1r = WS(W_BPR1)[j], b1i = WS(W_BPI1)[j];
                    const Real b2r = WS(W_BPR2)[j], b2i = WS(W_BPI2)[j];
                    r0 += cr * b1r + ci * b1i; i0 += cr * b1i - ci * b1r;
                    r1 += cr * b0r - ci * b0i + cr * b2r + ci * b2i;
                    i1 += cr * b0i + ci * b0r + cr * b2i - ci * b2r;
                    r2 += cr * b1r - ci * b1i; i2 += cr * b1i + ci * b1r;
                    WS(W_RBR0)[j] = r0; WS(W_RBI0)[j] = i0; WS(W_RBR2)[j] = r2; WS(W_RBI2)[j] = i2;
                }
                WS(W_RBR1)[j] = r1; WS(W_RBI1)[j] = i1;
            }
            __syncthreads();
            // ---- solves (Simulation.cpp:582-589): 0 B_r0 L1D, 1 B_theta0 L1F (flux wall J_z(R)),
            //      m1: 2 B0r L0, 3 B0i L0, 4 B2r L2D, 5 B2i L2D; all with chi0
            if (w == 0) {
                const int nj = M1 ? 6 : 2;
                if (lane < nj) {
                    int op, fm, rhs, out;
                    switch (lane) {
                        case 0: op = OP_L1D; fm = W_FM_L1D; rhs = W_RBR1; out = W_BPR1; break;
                        case 1: op = OP_L1F; fm = W_FM_L1F; rhs = W_RBI1; out = W_BPI1; break;
                        case 2: op = OP_L0; fm = W_FM_L0; rhs = W_RBR0; out = W_BPR0; break;
                        case 3: op = OP_L0; fm = W_FM_L0; rhs = W_RBI0; out = W_BPI0; break;
                        case 4: op = OP_L2D; fm = W_FM_L2D; rhs = W_RBR2; out = W_BPR2; break;
                        default: op = OP_L2D; fm = W_FM_L2D; rhs = W_RBI2; out = W_BPI2; break;
                    }
                    const int fidx = op == OP_L1D ? 0 : op == OP_L1F ? 1 : op == OP_L0 ? 2 : 3;
                    const bool flux = op == OP_L1F;
                    const Real fcw = flux ? A.flux_coef * WS(W_JZ0)[N] : Real(0);
                    thomas_apply(A.op_a[op], WS(fm), WS(fm + 1), bw0s[fidx], WS(rhs), Real(1), op != OP_L0, !flux, flux,
                                 fcw, WS(W_DP0 + lane), WS(out), M);
                }
            }
        }
        __syncthreads();

        // ================================================================ 6. store (Simulation.cpp:594-633) + push
        {
            Real* __restrict__ F = A.fld + static_cast<size_t>(k) * M * A.ncf;
            Real* __restrict__ Dg = A.diag + static_cast<size_t>(k) * M * A.ncd;
            const int ncf = A.ncf, ncd = A.ncd;
            for (int j = tid; j < M; j += kBS) {
                Real* Fj = F + static_cast<size_t>(j) * ncf;
                Real* Dj = Dg + static_cast<size_t>(j) * ncd;
                Fj[F_EP1R] = WS(W_WPR1)[j] + WS(W_BPI1)[j];
                Fj[F_EP1I] = WS(W_WPI1)[j] - WS(W_BPR1)[j];
                Fj[F_BP1R] = WS(W_BPR1)[j];
                Fj[F_BP1I] = WS(W_BPI1)[j];
                Fj[F_EZ0] = WS(W_EZ0)[j];
                Fj[F_BZ0] = WS(W_BZ0)[j];
                Dj[D_PSI0] = WS(W_PSI0)[j];
                Dj[D_NE0] = WS(W_NE0)[j];
                Dj[D_NI0] = WS(W_NI0)[j];
                if (M1) {
                    Fj[F_EP0R] = WS(W_WPR0)[j] + WS(W_BPI0)[j];
                    Fj[F_EP0I] = WS(W_WPI0)[j] - WS(W_BPR0)[j];
                    Fj[F_EP2R] = WS(W_WPR2)[j] + WS(W_BPI2)[j];
                    Fj[F_EP2I] = WS(W_WPI2)[j] - WS(W_BPR2)[j];
                    Fj[F_BP0R] = WS(W_BPR0)[j];
                    Fj[F_BP0I] = WS(W_BPI0)[j];
                    Fj[F_BP2R] = WS(W_BPR2)[j];
                    Fj[F_BP2I] = WS(W_BPI2)[j];
                    Fj[F_EZ1R] = WS(W_EZ1R)[j];
                    Fj[F_EZ1I] = WS(W_EZ1I)[j];
                    Fj[F_BZ1R] = WS(W_BZ1R)[j];
                    Fj[F_BZ1I] = WS(W_BZ1I)[j];
                    Dj[D_PSI1R] = WS(W_PSI1R)[j];
                    Dj[D_PSI1I] = WS(W_PSI1I)[j];
                    Dj[D_NE1R] = WS(W_NE1R)[j];
                    Dj[D_NE1I] = WS(W_NE1I)[j];
                    Dj[D_NI1R] = WS(W_NI1R)[j];
                    Dj[D_NI1I] = WS(W_NI1I)[j];
                }
            }
        }
        if (k < K - 1) {
            // ---- Adams-Bashforth push (PlasmaSpecies::push, src/PlasmaSpecies.cpp:600-667)
            const int slot = k % ord;
            const bool first = (k == 0);
#pragma unroll
            for (int p = 0; p < PT; ++p) {
                const int j = P.cj[p];
                if (j < 0) continue;
                const SpecInfo& S = sp[P.cspec[p]];
                const qz_species_state* st = S.st;
                const int li = P.cli[p];
                const int cap = st->cap;
                Real* __restrict__ hist = st->hist;
                const Real X = st->x[li], Y = st->y[li];
                const Real PX = P.cPX[p], PY = P.cPY[p], D = P.cD[p];
                const Real qm = S.qm;
                const Real t = P.ct[p], c = P.cc[p], sn = P.cs[p];
                Real Wx, Wy, Bx, By;
                eval_vector(lin(WS(W_WPR0), j, t), lin(WS(W_WPI0), j, t), lin(WS(W_WPR1), j, t), lin(WS(W_WPI1), j, t),
                            lin(WS(W_WPR2), j, t), lin(WS(W_WPI2), j, t), c, sn, M1, Wx, Wy);
                const Real Ez = eval_scalar(lin(WS(W_EZ0), j, t), M1 ? lin(WS(W_EZ1R), j, t) : Real(0),
                                            M1 ? lin(WS(W_EZ1I), j, t) : Real(0), c, sn, M1);
                const Real Bz = eval_scalar(lin(WS(W_BZ0), j, t), M1 ? lin(WS(W_BZ1R), j, t) : Real(0),
                                            M1 ? lin(WS(W_BZ1I), j, t) : Real(0), c, sn, M1);
                eval_vector(lin(WS(W_BPR0), j, t), lin(WS(W_BPI0), j, t), lin(WS(W_BPR1), j, t), lin(WS(W_BPI1), j, t),
                            lin(WS(W_BPR2), j, t), lin(WS(W_BPI2), j, t), c, sn, M1, Bx, By);
                const Real dA = Real(0);
                const Real gam = P.cgam[p];
                const Real invd = P.cinvd[p];
                const Real fp = -Real(0.5) * qm * qm * dA * invd;
                Real fv[5];
                fv[0] = PX * invd;
                fv[1] = PY * invd;
                fv[2] = qm * (gam * Wx * invd + By + PY * Bz * invd) + fp * c;
                fv[3] = qm * (gam * Wy * invd - Bx - PX * Bz * invd) + fp * sn;
                fv[4] = qm * ((PX * Wx + PY * Wy) * invd - Ez);
                auto H = [&](int sl, int v) -> Real& { return hist[(static_cast<size_t>(sl) * 5 + v) * cap + li]; };
                if (first) {
                    for (int sl = 0; sl < ord; ++sl)
                        for (int v = 0; v < 5; ++v) H(sl, v) = fv[v];
                } else {
                    for (int v = 0; v < 5; ++v) H(slot, v) = fv[v];
                }
                Real sum[5] = {0, 0, 0, 0, 0};
                for (int mm = 0; mm < ord; ++mm) {
                    const int sl = (slot - mm + ord) % ord;
                    for (int v = 0; v < 5; ++v) sum[v] += A.ab_c[mm] * (first ? fv[v] : (sl == slot ? fv[v] : H(sl, v)));
                }
                Real xn = X + dxi * sum[0], yn = Y + dxi * sum[1];
                Real pxn = PX + dxi * sum[2], pyn = PY + dxi * sum[3];
                const Real dn = D + dxi * sum[4];
                const Real gn = (Real(1) + pxn * pxn + pyn * pyn + dn * dn) / (Real(2) * dn);
                if (!(dn > dmin) || gn > qsamax * dn) {
                    st->w[li] = Real(0);
                    atomicAdd(&A.lost[P.cspec[p]], 1);
                    continue;
                }
                const Real rn = dsqrt(xn * xn + yn * yn);
                if (rn > R) {
                    const Real nx = xn / rn, ny = yn / rn;
                    const Real rr = Real(2) * R - rn;
                    if (rr < Real(0)) {
                        st->w[li] = Real(0);
                        atomicAdd(&A.lost[P.cspec[p]], 1);
                        continue;
                    }
                    xn = nx * rr; yn = ny * rr;
                    const Real pn = pxn * nx + pyn * ny;
                    pxn -= Real(2) * pn * nx; pyn -= Real(2) * pn * ny;
                    for (int sl = 0; sl < ord; ++sl) {
                        const Real aa = H(sl, 0) * nx + H(sl, 1) * ny;
                        H(sl, 0) -= Real(2) * aa * nx; H(sl, 1) -= Real(2) * aa * ny;
                        const Real bb = H(sl, 2) * nx + H(sl, 3) * ny;
                        H(sl, 2) -= Real(2) * bb * nx; H(sl, 3) -= Real(2) * bb * ny;
                    }
                }
                st->x[li] = xn; st->y[li] = yn; st->px[li] = pxn; st->py[li] = pyn; st->dl[li] = dn;
            }
        }
        __syncthreads();
    }
}

}  // namespace qz

// ============================================================================ host launch
using namespace qz;

size_t qz_sweep_workspace_doubles_impl(int M) { return static_cast<size_t>(W_NARR) * M; }

int qz_sweep_launch(const qz_sweep_args* dev_runs, const qz_sweep_args* host_runs, int nruns, hipStream_t stream) {
    int maxnp = 0, maxN = 0;
    bool m1 = host_runs[0].modes == 1;
    for (int i = 0; i < nruns; ++i) {
        int np = 0;
        for (int s = 0; s < host_runs[i].nspecies; ++s) np += host_runs[i].sp[s].np;
        maxnp = std::max(maxnp, np);
        maxN = std::max(maxN, host_runs[i].grid.N);
        if ((host_runs[i].modes == 1) != m1) return qz_fail(QZ_EINVAL, "qz_sweep_batched: all runs need the same modes");
    }
    if (maxnp > 65535) return qz_fail(QZ_ERANGE, "sweep: more than 65535 plasma particles per slice");
    const int shmem = kLdsBytes;
    {   // the LDS carve-up of the kernel must leave room for one contribution list of 2 L doubles
        const int pt_cap = ((maxnp + kBS - 1) / kBS <= 2) ? 2 : ((maxnp + kBS - 1) / kBS <= 4 ? 4 : 8);
        int off = ((kNW * maxN * 2) + 15) & ~15;
        off += (((maxN + 3) * 4) + 15) & ~15;
        off += ((kBS * pt_cap * 2) + 15) & ~15;
        off += 64;
        if (off + 16 * kBS * pt_cap > kLdsBytes) return qz_fail(QZ_ERANGE, "sweep: LDS too small for this grid");
    }
    auto launch = [&](auto kern) -> int {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, shmem);
        if (e != hipSuccess) return qz_fail_hip(e, "hipFuncSetAttribute");
        hipLaunchKernelGGL(kern, dim3(nruns), dim3(kBS), shmem, stream, dev_runs);
        e = hipGetLastError();
        if (e != hipSuccess) return qz_fail_hip(e, "sweep launch");
        return QZ_OK;
    };
    const int pt = (maxnp + kBS - 1) / kBS;
    if (pt <= 2) return m1 ? launch(sweep_kernel<true, 2>) : launch(sweep_kernel<false, 2>);
    if (pt <= 4) return m1 ? launch(sweep_kernel<true, 4>) : launch(sweep_kernel<false, 4>);
    if (pt <= 8) return m1 ? launch(sweep_kernel<true, 8>) : launch(sweep_kernel<false, 8>);
    return qz_fail(QZ_ERANGE, "sweep: more than 4096 plasma particles per slice (one workgroup per run)");
}
