// Body of the persistent gfx950 kernel of one quasi-static sweep (sweep_run): the whole slice loop of
// Simulation::step_fields (src/Simulation.cpp:727-783) for one run in one workgroup, walking xi from the head of
// the box.
//
// Per slice:
//   1. geometry of every live plasma particle (kept in registers for the slice);
//   2. a stable counting sort of the particles by radial cell in LDS (ballot ranks per wave); every
//      particle gets its position in the contribution lists of its two nodes, in particle-index order;
//   3. the mode deposit (PlasmaSpecies::deposit, src/PlasmaSpecies.cpp:409-478): particles write their two
//      contributions per array into LDS, one thread per (node, array) sums the list sequentially -- the
//      reference's atomic_add performed in index order, i.e. the serial result bit for bit;
//   4. combine with the background and the beam sources (Simulation::combine_sources);
//   5. psi, E_z, B_z solves (chi-free Thomas, factors formed once per run) on one wave while another wave
//      forms this slice's Thomas factors of the chi-dependent B+ operators; W+; the S+ deposit; the
//      Picard solves for B+ (Simulation::solve_slice_fields, compute_wplus, compute_bz, solve_bplus);
//   6. the field store (Simulation::store_slice) and the Adams-Bashforth push (PlasmaSpecies::push).
// The Thomas chains and the particle gathers read LDS; the run's workspace (L2-resident) carries arrays
// between phases.  LDS holds the grid, the sort and one region that every phase reuses for its own arrays.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "qz_device.hpp"
#include "qz_internal.hpp"
#include "qz_sweep_common.hpp"

namespace qz {

constexpr int kBS = 512;           // threads per run
constexpr int kNW = kBS / 64;      // waves per run
constexpr int kLdsBytes = 160 * 1024;
constexpr int kMiscBytes = 512;    // counters, bw0 of the chi factors, deposit component list, phase clocks

// byte offsets of the LDS carve-up: r | 1/h | 1/V | wave counters | cell starts | sorted ids | misc | region
struct LdsPlan {
    int Mp, wcnt, cs, sidx, misc, region;
    __host__ __device__ LdsPlan(int N, int pt) {
        const int M = N + 1;
        Mp = (M + 1) & ~1;
        wcnt = 3 * Mp * 8;
        cs = wcnt + (((kNW * N * 2) + 15) & ~15);
        sidx = cs + ((((N + 3) * 4) + 15) & ~15);
        misc = sidx + (((kBS * pt * 2) + 15) & ~15);
        region = misc + kMiscBytes;
    }
};

// array slots (Mp doubles each) of the shared LDS region in each phase
template <bool M1> struct Lay {
    // solves: chi-free operators and factors, right-hand sides (solved in place), J+, chi factor inputs
    static constexpr int L0A = 0, L0M = 1, L0CP = 2, L1A = 3, L1M = 4, L1CP = 5;
    static constexpr int WPSI0 = M1 ? 6 : 4, WEZ0 = WPSI0 + 1, WBZ0 = WPSI0 + 2;
    static constexpr int WPSI1R = 9, WPSI1I = 10, WEZ1R = 11, WEZ1I = 12, WBZ1R = 13, WBZ1I = 14;
    static constexpr int JP = M1 ? 15 : 7;          // J+ r1 i1 r0 i0 r2 i2
    static constexpr int L1B = M1 ? 21 : 9, L1C = L1B + 1, CHI0 = L1B + 2;
    static constexpr int L0B = 24, L0C = 25, L2A = 26, L2B = 27, L2C = 28;
    static constexpr int NSOLVE = M1 ? 29 : 12;
    // S+ deposit: W+ (r1 i1 r0 i0 r2 i2), E_z, B_z modes; the contribution lists follow
    static constexpr int SW = 0, SEZ = M1 ? 6 : 2, SBZ = SEZ + (M1 ? 3 : 1), NSDEP = SBZ + (M1 ? 3 : 1);
    // Picard: a of each operator, factors (m, c') per operator, gz, right-hand sides, B+, chi_1
    static constexpr int PA_L1D = 0, PA_L1F = 1, PA_L0 = 2, PA_L2D = 3;
    static constexpr int PF = M1 ? 4 : 2;             // (m, c') of L1D, L1F, [L0, L2D]
    static constexpr int PGZ = PF + (M1 ? 8 : 4);     // r1 i1 r0 i0 r2 i2
    static constexpr int PRB = PGZ + (M1 ? 6 : 2);
    static constexpr int PBP = PRB + (M1 ? 6 : 2);
    static constexpr int PCHI = PBP + (M1 ? 6 : 2);   // chi_1 r, i
    static constexpr int NPICARD = PCHI + (M1 ? 2 : 0);
    // push: W+ (r1 i1 r0 i0 r2 i2), E_z, B_z, B+ (r1 i1 r0 i0 r2 i2)
    static constexpr int GW = 0, GEZ = M1 ? 6 : 2, GBZ = GEZ + (M1 ? 3 : 1), GBP = GBZ + (M1 ? 3 : 1);
    static constexpr int NPUSH = GBP + (M1 ? 6 : 2);
    static constexpr int NMAX = NSOLVE > NPICARD ? (NSOLVE > NPUSH ? NSOLVE : NPUSH) : (NPICARD > NPUSH ? NPICARD : NPUSH);
};

// one run's sweep by the calling workgroup
template <bool M1, int PT>
__device__ __forceinline__ void sweep_run(const qz_sweep_args& A) {
    using LY = Lay<M1>;
    extern __shared__ double smem_d[];
    char* smem = reinterpret_cast<char*>(smem_d);

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int N = A.grid.N, M = N + 1;
    const Real R = A.grid.R;
    Real* __restrict__ ws = A.ws;
    auto WS = [&](int arr) -> Real* { return ws + static_cast<size_t>(arr) * M; };
    const int K = A.nslices;
    const int kb = A.k_begin, ke = A.k_end > 0 ? A.k_end : K;   // slices of this launch
    const Real dxi = A.dxi;
    const Real idxi = Real(1) / dxi;
    const Stencil stn{A.grad_m, A.grad_0, A.grad_p, A.inv_r, N};
    unsigned long long* __restrict__ stats = A.stats;

    // ---- species (at most two; scalars, no indexed register arrays)
    const int nsp = A.nspecies;
    const int np0 = nsp > 0 ? A.sp[0].np : 0, np1 = nsp > 1 ? A.sp[1].np : 0;
    const int NP = np0 + np1;
    // species arrays read once: a particle's array is a per-lane choice between two uniform pointers
    const Real* const px0 = nsp > 0 ? A.sp[0].px : nullptr;
    const Real* const py0 = nsp > 0 ? A.sp[0].py : nullptr;
    const Real* const dl0 = nsp > 0 ? A.sp[0].dl : nullptr;
    const Real* const w0 = nsp > 0 ? A.sp[0].w : nullptr;
    const Real* const px1 = nsp > 1 ? A.sp[1].px : px0;
    const Real* const py1 = nsp > 1 ? A.sp[1].py : py0;
    const Real* const dl1 = nsp > 1 ? A.sp[1].dl : dl0;
    const Real* const w1 = nsp > 1 ? A.sp[1].w : w0;
    const Real q0 = nsp > 0 ? A.sp[0].q : Real(0), q1 = nsp > 1 ? A.sp[1].q : Real(0);
    const Real qm0 = nsp > 0 ? A.sp[0].q / A.sp[0].m : Real(0), qm1 = nsp > 1 ? A.sp[1].q / A.sp[1].m : Real(0);
    const bool has_neg = (nsp > 0 && q0 < 0) || (nsp > 1 && q1 < 0);
    const bool has_pos = (nsp > 0 && !(q0 < 0)) || (nsp > 1 && !(q1 < 0));

    // ---- LDS carve-up
    const LdsPlan plan(N, PT);
    const int Mp = plan.Mp;
    Real* s_r = smem_d;
    Real* s_ih = s_r + Mp;
    Real* s_iv = s_ih + Mp;
    unsigned short* s_wcnt = reinterpret_cast<unsigned short*>(smem + plan.wcnt);
    int* s_cs = reinterpret_cast<int*>(smem + plan.cs);          // N + 3 (cell starts shifted by one)
    unsigned short* s_sidx = reinterpret_cast<unsigned short*>(smem + plan.sidx);
    int* s_misc = reinterpret_cast<int*>(smem + plan.misc);      // [0] live count
    Real* s_bw0 = reinterpret_cast<Real*>(smem + plan.misc + 16);                        // 4 doubles
    int* s_comp = reinterpret_cast<int*>(smem + plan.misc + 48);                         // 24 ints
    unsigned long long* s_ph = reinterpret_cast<unsigned long long*>(smem + plan.misc + 160);   // phase clocks + last
    // order of the per-node sums (in the wave-counter area once the sort is done): M node ids
    unsigned short* s_norder = reinterpret_cast<unsigned short*>(smem + plan.wcnt);
    Real* B = reinterpret_cast<Real*>(smem + plan.region);
    const int Bcap = (kLdsBytes - plan.region) / 8;
    auto BA = [&](int slot) -> Real* { return B + static_cast<size_t>(slot) * Mp; };

    // deposit components in pass order (chi_0 first: it feeds the B+ factors); the sequence of the
    // deposit phase below follows this list
    {
        int n = 0;
        auto put = [&](int c) { if (tid == 0) s_comp[n] = c; ++n; };
        put(W_CHI0); put(W_RHOT0); put(W_DRHO0); put(W_JZ0);
        if (has_neg) put(W_NE0);
        if (has_pos) put(W_NI0);
        put(W_JPR1); put(W_JPI1);
        if (M1) {
            put(W_CHI1R); put(W_CHI1I); put(W_RHOT1R); put(W_RHOT1I); put(W_DRHO1R); put(W_DRHO1I);
            put(W_JZ1R); put(W_JZ1I);
            if (has_neg) { put(W_NE1R); put(W_NE1I); }
            if (has_pos) { put(W_NI1R); put(W_NI1I); }
            put(W_JPR0); put(W_JPI0); put(W_JPR2); put(W_JPI2);
        }
    }
    for (int j = tid; j < M; j += kBS) {
        s_r[j] = A.grid.r[j];
        s_iv[j] = A.grid.inv_V[j];
        if (j < N) s_ih[j] = A.grid.inv_h[j];
        if (kb == 0)
            for (int a = W_BPR1; a <= W_BPI2; ++a) WS(a)[j] = Real(0);   // B+ := 0 at the head (Simulation.cpp:705-708)
    }
    if (tid == 0) {
        for (int i = 0; i < QZ_NPHASE; ++i) s_ph[i] = 0;
        s_ph[QZ_NPHASE] = wall_clock64();
    }
    const int nS = M1 ? 6 : 2;
    const Real dmin = A.delta_min, qsamax = A.max_qsa;
    const int ord = A.ab_order;
#define QZ_PHASE(i)                                          \
    do {                                                     \
        if (tid == 0 && stats) {                             \
            const unsigned long long tnow = wall_clock64();  \
            s_ph[i] += tnow - s_ph[QZ_NPHASE];               \
            s_ph[QZ_NPHASE] = tnow;                          \
        }                                                    \
    } while (0)
    __syncthreads();
    if (kb == 0 && A.load) {   // the plasma at the head of the box, made here instead of copied in
        for (int s = 0; s < nsp; ++s) load_rings(A.rings[s], A.sp[s], ord, s_r, tid, A.sp[s].np, kBS);
        __syncthreads();
    }

    // per-thread particle cache of the slice (pp is a compile-time index in every loop): cell, node-list
    // positions, and the geometry that costs divisions; momenta and weights are re-read where needed
    int cj[PT], cpos[PT];   // cell; node-list positions left | right << 16
    Real ct[PT], cc[PT], cs[PT], cir[PT], cinvd[PT], cgam[PT];
    auto fence_cache = [&]() {
        fence_regs(cj); fence_regs(cpos); fence_regs(ct); fence_regs(cc); fence_regs(cs); fence_regs(cir);
        fence_regs(cinvd); fence_regs(cgam);
    };

    for (int k = kb; k < ke; ++k) {
        // every per-thread value (addresses, lane roles) is formed inside the slice, not kept live across it
        int tid_k = threadIdx.x;
        asm volatile("" : "+v"(tid_k));
        const int tid = tid_k;
        const int lane = tid & 63;
        auto slot_of = [&](int pp) { return w * (PT * 64) + pp * 64 + lane; };
        // momenta, Delta and weight of cached particle pp (species arrays, L2-resident); the slot is made
        // opaque so the addresses are formed at each use instead of being held in registers
        auto load_mom = [&](int pp, Real& PX, Real& PY, Real& D, Real& wi) {
            int gi = slot_of(pp);
            asm volatile("" : "+v"(gi));
            const bool s1 = gi >= np0;
            const int li = s1 ? gi - np0 : gi;
            PX = (s1 ? px1 : px0)[li]; PY = (s1 ? py1 : py0)[li]; D = (s1 ? dl1 : dl0)[li]; wi = (s1 ? w1 : w0)[li];
        };
        // the same for PB consecutive cached particles from p0, all loads issued before any is used
        constexpr int PB = PT < 4 ? PT : 4;
        auto load_batch = [&](int p0, Real (&PX)[PB], Real (&PY)[PB], Real (&D)[PB], Real (&wi)[PB]) {
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                PX[u] = PY[u] = D[u] = wi[u] = Real(0);
                if (cj[p0 + u] >= 0) load_mom(p0 + u, PX[u], PY[u], D[u], wi[u]);
            }
        };
        // ================================================================ 1. geometry (PlasmaSpecies.cpp:52-70)
#pragma unroll
        for (int pp = 0; pp < PT; ++pp) {
            const int gi = slot_of(pp);
            cj[pp] = -1;
            if (gi < NP) {
                const bool s1 = gi >= np0;
                const qz_species_state& st = s1 ? A.sp[1] : A.sp[0];
                const int li = s1 ? gi - np0 : gi;
                const Real wi = st.w[li];
                // the state at the start of the slice, for the adjoint's recomputation
                Real* tr = A.traj ? A.traj + static_cast<size_t>(k - kb) * 6 * NP + gi : nullptr;
                if (tr) tr[5 * static_cast<size_t>(NP)] = wi;
                if (wi != Real(0)) {
                    const Real X = st.x[li], Y = st.y[li];
                    const Real PX = st.px[li], PY = st.py[li], D = st.dl[li];
                    if (tr) {
                        tr[0] = X;
                        tr[NP] = Y;
                        tr[2 * static_cast<size_t>(NP)] = PX;
                        tr[3 * static_cast<size_t>(NP)] = PY;
                        tr[4 * static_cast<size_t>(NP)] = D;
                    }
                    const Real r = dsqrt(X * X + Y * Y);
                    Real c, s_;
                    if (r > Real(0)) { c = X / r; s_ = Y / r; } else { c = Real(1); s_ = Real(0); }
                    int j; Real t;
                    grid_weights(s_r, s_ih, N, R, r, j, t);
                    cj[pp] = j;
                    ct[pp] = t;
                    cc[pp] = c;
                    cs[pp] = s_;
                    cir[pp] = r > Real(0) ? Real(1) / r : Real(0);
                    const Real gam = (Real(1) + PX * PX + PY * PY + D * D) / (Real(2) * D);
                    cgam[pp] = gam;
                    cinvd[pp] = Real(1) / D;
                }
            }
        }
        QZ_PHASE(0);
        // ================================================================ 2. stable sort by cell
        for (int i = tid; i < kNW * N; i += kBS) s_wcnt[i] = 0;
        if (tid == 0) s_misc[0] = 0;
        __syncthreads();
        int lrank[PT];
        {
            const unsigned long long lt_mask = (1ull << lane) - 1ull;
            int nbits = 1;
            while ((1 << nbits) < N) ++nbits;
            int live = 0;
#pragma unroll
            for (int pp = 0; pp < PT; ++pp) {
                const int key = cj[pp];
                unsigned long long mask = __ballot(key >= 0);
                for (int b = 0; b < nbits; ++b) {
                    const bool bit = key >= 0 && ((key >> b) & 1);
                    const unsigned long long bb = __ballot(bit);
                    mask &= bit ? bb : ~bb;
                }
                int old = 0;
                if (key >= 0) old = s_wcnt[w * N + key];
                const int rk = __popcll(mask & lt_mask);
                lrank[pp] = old + rk;
                live += key >= 0 ? 1 : 0;
                __builtin_amdgcn_wave_barrier();
                if (key >= 0 && rk == 0) s_wcnt[w * N + key] = static_cast<unsigned short>(old + __popcll(mask));
                __builtin_amdgcn_wave_barrier();
            }
            for (int d = 32; d >= 1; d >>= 1) live += __shfl_xor(live, d, 64);
            if (lane == 0) atomicAdd(&s_misc[0], live);
        }
        __syncthreads();
        for (int c = tid; c < N; c += kBS) {   // exclusive prefix over the waves per cell; cell totals
            int run = 0;
            for (int ww = 0; ww < kNW; ++ww) {
                const int v = s_wcnt[ww * N + c];
                s_wcnt[ww * N + c] = static_cast<unsigned short>(run);
                run += v;
            }
            s_cs[c + 1] = run;
        }
        __syncthreads();
        const int L = s_misc[0];
        if (w == 0) wave_exclusive_scan(s_cs + 1, N, lane);   // s_cs[c + 1] = start of cell c
        if (tid == 0) {
            s_cs[0] = 0;
            s_cs[N + 1] = L;
            s_cs[N + 2] = L;
            if (stats) atomicAdd(stats, static_cast<unsigned long long>(L));
        }
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < PT; ++pp) {
            const int j = cj[pp];
            if (j >= 0) {
                const int pos = s_cs[j + 1] + s_wcnt[w * N + j] + lrank[pp];
                s_sidx[pos] = static_cast<unsigned short>(slot_of(pp));
                lrank[pp] = pos - s_cs[j + 1];   // rank within the cell (index order)
            }
        }
        __syncthreads();
        // Order of the per-node sums (m = 1): nodes whose lists exceed twice the mean length (a compressed sheath
        // holds hundreds of entries) first, then the others, each group in node order.  A wave's item round costs
        // its longest list, so a thread then sums at most one long list per pass: 3.4% off the c2 sweep (eleven
        // passes of two components).  In m = 0 (four components, passes of up to eight) it costs more than it
        // saves (+2.9% on the c5 batch), and the sums keep node order.  Wave 0 builds it while the other waves
        // search their list positions; it is read after the deposit's first barrier (s_misc[1] = number of long
        // nodes).
        if (M1 && w == 0) {
            const unsigned long long lt_mask = (1ull << lane) - 1ull;
            const int thr = max(8, (4 * L) / M);
            int nlong = 0;
            for (int t0 = 0; t0 < M; t0 += 64) {
                const int n = t0 + lane;
                nlong += __popcll(__ballot(n < M && s_cs[n + 2] - s_cs[n] > thr));
            }
            int cl = 0, csh = nlong;
            for (int t0 = 0; nlong > 0 && t0 < M; t0 += 64) {   // no long list: the sums take plain node order
                const int n = t0 + lane;
                const bool valid = n < M;
                const bool lg = valid && s_cs[n + 2] - s_cs[n] > thr;
                const unsigned long long bl = __ballot(lg), bs = __ballot(valid && !lg);
                if (lg) s_norder[cl + __popcll(bl & lt_mask)] = static_cast<unsigned short>(n);
                else if (valid) s_norder[csh + __popcll(bs & lt_mask)] = static_cast<unsigned short>(n);
                cl += __popcll(bl);
                csh += __popcll(bs);
            }
            if (lane == 0) s_misc[1] = nlong;
        }
        // node n's list holds cell n-1 (right entries) and cell n (left entries), merged in particle-index order
#pragma unroll
        for (int pp = 0; pp < PT; ++pp) {
            const int j = cj[pp];
            cpos[pp] = 0;
            if (j >= 0) {
                const int gi = slot_of(pp);
                const int less_m = count_less(s_sidx, s_cs[j], s_cs[j + 1], gi);
                const int less_p = count_less(s_sidx, s_cs[j + 2], s_cs[j + 3], gi);
                const int pl = (s_cs[j] + s_cs[j + 1]) + lrank[pp] + less_m;
                const int pr = (s_cs[j + 1] + s_cs[j + 2]) + lrank[pp] + less_p;
                cpos[pp] = pl | (pr << 16);
            }
        }
        const int G = max(1, min(Bcap / max(2 * L, 1), 8));
        QZ_PHASE(1);

        // ================================================================ 3. deposit (node sums in index order)
        // Components run in a fixed compile-time sequence (the order of s_comp); every G of them share one
        // pass: contributions into LDS lists, a barrier, the per-node sums, a barrier.
        Real* __restrict__ Dg = A.diag + static_cast<size_t>(k) * M * A.ncd;
        const int ncd = A.ncd;
        // (component, node) of sum item `item` of a pass over gcnt <= 8 components, in the order of s_norder
        // (plain node order when no list is long); the component count by subtraction, not division
        auto item_node = [&](int item, int gcnt, int nl, int& g, int& n) {
            int r = item;
            g = 0;
            if (r < gcnt * nl) {
                while (r >= nl) { r -= nl; ++g; }
                n = s_norder[r];
            } else {
                r -= gcnt * nl;
                const int ns = M - nl;
                while (r >= ns) { r -= ns; ++g; }
                n = nl > 0 ? s_norder[nl + r] : r;
            }
        };
        {
            int gcur = 0, c0 = 0;
            auto reduce = [&]() {
                __syncthreads();
                const int nl = M1 ? s_misc[1] : 0;
                for (int item = tid; item < M * gcur; item += kBS) {
                    int g, n;
                    item_node(item, gcur, nl, g, n);
                    const int e0 = s_cs[n] + s_cs[n + 1], e1 = s_cs[n + 1] + s_cs[n + 2];
                    const Real* vb = B + g * 2 * L;
                    const Real acc = sum_list(vb, e0, e1);
                    const int comp = s_comp[c0 + g];
                    Real* Dn = Dg + static_cast<size_t>(n) * ncd;
                    if (comp == W_NE0) Dn[D_NE0] = acc;
                    else if (comp == W_NE1R) Dn[D_NE1R] = acc;
                    else if (comp == W_NE1I) Dn[D_NE1I] = acc;
                    else if (comp == W_NI1R) Dn[D_NI1R] = acc;
                    else if (comp == W_NI1I) Dn[D_NI1I] = acc;
                    else if (comp == W_NI0) {   // combine_sources adds the positive background (Simulation.cpp:441)
                        const Real bgr0 = A.bg[3 * M + n];
                        Dn[D_NI0] = bgr0 > Real(0) ? acc + bgr0 : acc;
                    } else WS(comp)[n] = acc;
                }
                __syncthreads();
                QZ_PHASE(3);
                c0 += gcur;
                gcur = 0;
            };
            // contributions of every cached particle to component C (straight-line over the particles)
            auto comp_pass = [&](auto cst) {
                constexpr int C = decltype(cst)::value;
                Real* vb = B + gcur * 2 * L;
                fence_cache();
#pragma unroll
                for (int p0 = 0; p0 < PT; p0 += PB) {
                Real PXb[PB], PYb[PB], Db[PB], wb[PB];
                load_batch(p0, PXb, PYb, Db, wb);
#pragma unroll
                for (int u = 0; u < PB; ++u) {
                    const int pp = p0 + u;
                    const int j = cj[pp];
                    if (j < 0) continue;
                    const Real PX = PXb[u], PY = PYb[u], D = Db[u], wi = wb[u];
                    const bool s1 = slot_of(pp) >= np0;
                    const Real q = s1 ? q1 : q0, qm = s1 ? qm1 : qm0;
                    const int pl = cpos[pp] & 0xffff, pr = cpos[pp] >> 16;
                    const Real ih = s_ih[j];
                    const Real pz = cgam[pp] - D;
                    vb[pl] = dep_value(C, 0, j, ct[pp], cc[pp], cs[pp], cir[pp], cinvd[pp], cgam[pp], PX, PY, pz, wi, q,
                                       qm, s_iv[j], ih);
                    vb[pr] = dep_value(C, 1, j, ct[pp], cc[pp], cs[pp], cir[pp], cinvd[pp], cgam[pp], PX, PY, pz, wi, q,
                                       qm, s_iv[j + 1], ih);
                }
                }
                QZ_PHASE(2);
                if (++gcur == G) reduce();
            };
            comp_pass(IC<W_CHI0>{});
            comp_pass(IC<W_RHOT0>{});
            comp_pass(IC<W_DRHO0>{});
            comp_pass(IC<W_JZ0>{});
            if (has_neg) comp_pass(IC<W_NE0>{});
            if (has_pos) comp_pass(IC<W_NI0>{});
            comp_pass(IC<W_JPR1>{});
            comp_pass(IC<W_JPI1>{});
            if (M1) {
                comp_pass(IC<W_CHI1R>{});
                comp_pass(IC<W_CHI1I>{});
                comp_pass(IC<W_RHOT1R>{});
                comp_pass(IC<W_RHOT1I>{});
                comp_pass(IC<W_DRHO1R>{});
                comp_pass(IC<W_DRHO1I>{});
                comp_pass(IC<W_JZ1R>{});
                comp_pass(IC<W_JZ1I>{});
                if (has_neg) { comp_pass(IC<W_NE1R>{}); comp_pass(IC<W_NE1I>{}); }
                if (has_pos) { comp_pass(IC<W_NI1R>{}); comp_pass(IC<W_NI1I>{}); }
                comp_pass(IC<W_JPR0>{});
                comp_pass(IC<W_JPI0>{});
                comp_pass(IC<W_JPR2>{});
                comp_pass(IC<W_JPI2>{});
            }
            if (gcur > 0) reduce();
        }
        // densities of absent species are zero, plus the positive background for n_i
        for (int n = tid; n < M; n += kBS) {
            Real* Dn = Dg + static_cast<size_t>(n) * ncd;
            if (!has_neg) {
                Dn[D_NE0] = Real(0);
                if (M1) { Dn[D_NE1R] = Real(0); Dn[D_NE1I] = Real(0); }
            }
            if (!has_pos) {
                const Real bgr0 = A.bg[3 * M + n];
                Dn[D_NI0] = bgr0 > Real(0) ? Real(0) + bgr0 : Real(0);
                if (M1) { Dn[D_NI1R] = Real(0); Dn[D_NI1I] = Real(0); }
            }
        }
        __syncthreads();
        QZ_PHASE(3);

        // ================================================================ 4. combine_sources (Simulation.cpp:408-457)
        {
            const int KL = K;
            const int ncb = A.ncb;
            const Real* __restrict__ bs = A.bsrc;
            const Real* __restrict__ bg = A.bg;
            const bool ngp = A.ngp != 0;
            auto bb = [&](int kk, int j, int comp) { return bs[(static_cast<size_t>(kk) * M + j) * ncb + comp]; };
            for (int j = tid; j < M; j += kBS) {
                auto dxi_b = [&](int comp) {
                    if (ngp) {
                        const Real prev = (k > 0) ? bb(k - 1, j, comp) : Real(0);
                        return (bb(k, j, comp) - prev) * idxi;
                    }
                    if (KL == 1) return Real(0);
                    if (k == 0) return (bb(1, j, comp) - bb(0, j, comp)) * idxi;
                    if (k == KL - 1) return (bb(KL - 1, j, comp) - bb(KL - 2, j, comp)) * idxi;
                    return Real(0.5) * (bb(k + 1, j, comp) - bb(k - 1, j, comp)) * idxi;
                };
                BA(LY::WPSI0)[j] = WS(W_RHOT0)[j] + (bg[j] + bb(k, j, B_RT0));
                BA(LY::WEZ0)[j] = WS(W_DRHO0)[j] + dxi_b(B_RT0);
                WS(W_JZ0)[j] += bb(k, j, B_JZ0);
                BA(LY::JP + 0)[j] = WS(W_JPR1)[j] + bb(k, j, B_JP1R);
                BA(LY::JP + 1)[j] = WS(W_JPI1)[j] + bb(k, j, B_JP1I);
                if (M1) {
                    BA(LY::WPSI1R)[j] = WS(W_RHOT1R)[j] + (bg[M + j] + bb(k, j, B_RT1R));
                    BA(LY::WPSI1I)[j] = WS(W_RHOT1I)[j] + (bg[2 * M + j] + bb(k, j, B_RT1I));
                    BA(LY::WEZ1R)[j] = WS(W_DRHO1R)[j] + dxi_b(B_RT1R);
                    BA(LY::WEZ1I)[j] = WS(W_DRHO1I)[j] + dxi_b(B_RT1I);
                    WS(W_JZ1R)[j] += bb(k, j, B_JZ1R);
                    WS(W_JZ1I)[j] += bb(k, j, B_JZ1I);
                    BA(LY::JP + 2)[j] = WS(W_JPR0)[j] + bb(k, j, B_JP0R);
                    BA(LY::JP + 3)[j] = WS(W_JPI0)[j] + bb(k, j, B_JP0I);
                    BA(LY::JP + 4)[j] = WS(W_JPR2)[j] + bb(k, j, B_JP2R);
                    BA(LY::JP + 5)[j] = WS(W_JPI2)[j] + bb(k, j, B_JP2I);
                }
                // operators of the solves and of the chi factors
                BA(LY::L0A)[j] = A.op_a[OP_L0][j];
                BA(LY::L0M)[j] = A.fac_l0_m[j];
                BA(LY::L0CP)[j] = A.fac_l0_cp[j];
                BA(LY::L1A)[j] = A.op_a[OP_L1D][j];
                BA(LY::L1B)[j] = A.op_b[OP_L1D][j];
                BA(LY::L1C)[j] = A.op_c[OP_L1D][j];
                BA(LY::CHI0)[j] = WS(W_CHI0)[j];
                if (A.chi_out) A.chi_out[static_cast<size_t>(k) * M + j] = WS(W_CHI0)[j];   // for the laser march
                if (M1) {
                    BA(LY::L1M)[j] = A.fac_l1_m[j];
                    BA(LY::L1CP)[j] = A.fac_l1_cp[j];
                    BA(LY::L0B)[j] = A.op_b[OP_L0][j];
                    BA(LY::L0C)[j] = A.op_c[OP_L0][j];
                    BA(LY::L2A)[j] = A.op_a[OP_L2D][j];
                    BA(LY::L2B)[j] = A.op_b[OP_L2D][j];
                    BA(LY::L2C)[j] = A.op_c[OP_L2D][j];
                }
            }
        }
        __syncthreads();
        QZ_PHASE(4);
        // B_z right-hand side (compute_bz, Simulation.cpp:506-526)
        for (int j = tid; j < M; j += kBS) {
            const Real t1 = stn.grad_odd(BA(LY::JP + 1), j), t2 = stn.over_r(BA(LY::JP + 1), j);
            BA(LY::WBZ0)[j] = -(t1 + t2);
            if (M1) {
                const Real gr0 = stn.grad_odd(BA(LY::JP + 4), j), gi0 = stn.over_r(BA(LY::JP + 4), j);
                const Real gr1 = stn.grad_odd(BA(LY::JP + 5), j), gi1 = stn.over_r(BA(LY::JP + 5), j);
                const Real gr2 = stn.grad(BA(LY::JP + 2), j), gi2 = stn.grad(BA(LY::JP + 3), j);
                const Real D1r = gr0 + Real(2) * gi0, D1i = gr1 + Real(2) * gi1;
                const Real Dmr = gr2, Dmi = gi2;
                BA(LY::WBZ1R)[j] = -Real(0.5) * (D1i + Dmi);
                BA(LY::WBZ1I)[j] = Real(0.5) * (D1r - Dmr);
            }
        }
        __syncthreads();
        QZ_PHASE(5);

        // ================================================================ 5a. psi, E_z, B_z solves || chi factors
        if (w == 0) {
            // lanes: 0 psi0, 1 ez0, 2 bz0 (L0);  m1: 3 psi1r, 4 psi1i, 5 ez1r, 6 ez1i, 7 bz1r, 8 bz1i (L1D)
            const int njobs = M1 ? 9 : 3;
            if (lane < njobs) {
                const bool l0 = lane < 3;
                const int slot = l0 ? LY::WPSI0 + lane : LY::WPSI1R + (lane - 3);
                const Real scale = (lane == 2 || lane >= 7) ? Real(1) : Real(-1);
                const Real b0 = l0 ? A.op_b[OP_L0][0] : A.op_b[OP_L1D][0];
                thomas_lds(BA(l0 ? LY::L0A : LY::L1A), BA(l0 ? LY::L0M : LY::L1M), BA(l0 ? LY::L0CP : LY::L1CP), b0,
                           BA(slot), BA(slot), scale, !l0, true, false, Real(0), M);
            }
        } else if (w == 1) {
            // this slice's factors of (L - chi0): lanes 0 L1D, 1 L1F, m1: 2 L0, 3 L2D
            const int nf = M1 ? 4 : 2;
            if (lane < nf) {
                const Real* a = lane < 2 ? BA(LY::L1A) : lane == 2 ? BA(LY::L0A) : BA(LY::L2A);
                const Real* b = lane < 2 ? BA(LY::L1B) : lane == 2 ? BA(LY::L0B) : BA(LY::L2B);
                const Real* c = lane < 2 ? BA(LY::L1C) : lane == 2 ? BA(LY::L0C) : BA(LY::L2C);
                const int op = lane == 0 ? OP_L1D : lane == 1 ? OP_L1F : lane == 2 ? OP_L0 : OP_L2D;
                const Real aN = A.op_a[op][N], bN = A.op_b[op][N];
                const bool dir0 = op != OP_L0, dirN = op != OP_L1F;
                s_bw0[lane] = thomas_factor(a, b, c, BA(LY::CHI0), aN, bN, dir0, dirN, WS(W_FAC0 + 2 * lane),
                                            WS(W_FAC0 + 2 * lane + 1), M);
            }
        }
        __syncthreads();
        QZ_PHASE(6);

        // ================================================================ 5b. field outputs, W+ (compute_wplus, Simulation.cpp:477-498)
        {
            Real* __restrict__ F = A.fld + static_cast<size_t>(k) * M * A.ncf;
            const int ncf = A.ncf;
            for (int j = tid; j < M; j += kBS) {
                Real* Fj = F + static_cast<size_t>(j) * ncf;
                Real* Dj = Dg + static_cast<size_t>(j) * ncd;
                const Real ez0 = BA(LY::WEZ0)[j], bz0 = BA(LY::WBZ0)[j];
                Dj[D_PSI0] = BA(LY::WPSI0)[j];
                Fj[F_EZ0] = ez0; Fj[F_BZ0] = bz0;
                WS(W_EZ0)[j] = ez0; WS(W_BZ0)[j] = bz0;
                const Real t1 = stn.grad(BA(LY::WPSI0), j);
                WS(W_WPR1)[j] = -t1;
                WS(W_WPI1)[j] = Real(0);
                if (M1) {
                    const Real* fr = BA(LY::WPSI1R);
                    const Real* fi = BA(LY::WPSI1I);
                    Dj[D_PSI1R] = fr[j]; Dj[D_PSI1I] = fi[j];
                    const Real e1r = BA(LY::WEZ1R)[j], e1i = BA(LY::WEZ1I)[j];
                    const Real z1r = BA(LY::WBZ1R)[j], z1i = BA(LY::WBZ1I)[j];
                    Fj[F_EZ1R] = e1r; Fj[F_EZ1I] = e1i; Fj[F_BZ1R] = z1r; Fj[F_BZ1I] = z1i;
                    WS(W_EZ1R)[j] = e1r; WS(W_EZ1I)[j] = e1i; WS(W_BZ1R)[j] = z1r; WS(W_BZ1I)[j] = z1i;
                    const Real t2 = stn.grad_odd(fr, j), t3 = stn.over_r(fr, j);
                    const Real t4 = stn.grad_odd(fi, j), rh = stn.over_r(fi, j);
                    WS(W_WPR2)[j] = -(t2 - t3);
                    WS(W_WPI2)[j] = -(t4 - rh);
                    WS(W_WPR0)[j] = -(t2 + t3);
                    WS(W_WPI0)[j] = (t4 + rh);
                }
            }
        }
        __syncthreads();
        // W+, E_z, B_z into LDS for the S+ gather
        for (int j = tid; j < M; j += kBS) {
            BA(LY::SW + 0)[j] = WS(W_WPR1)[j];
            BA(LY::SW + 1)[j] = WS(W_WPI1)[j];
            BA(LY::SEZ)[j] = WS(W_EZ0)[j];
            BA(LY::SBZ)[j] = WS(W_BZ0)[j];
            if (M1) {
                BA(LY::SW + 2)[j] = WS(W_WPR0)[j]; BA(LY::SW + 3)[j] = WS(W_WPI0)[j];
                BA(LY::SW + 4)[j] = WS(W_WPR2)[j]; BA(LY::SW + 5)[j] = WS(W_WPI2)[j];
                BA(LY::SEZ + 1)[j] = WS(W_EZ1R)[j]; BA(LY::SEZ + 2)[j] = WS(W_EZ1I)[j];
                BA(LY::SBZ + 1)[j] = WS(W_BZ1R)[j]; BA(LY::SBZ + 2)[j] = WS(W_BZ1I)[j];
            }
        }
        QZ_PHASE(7);

        // ================================================================ 5c. S+ deposit (PlasmaSpecies::deposit_S, PlasmaSpecies.cpp:516-558)
        {
            Real* VB = BA(LY::NSDEP);
            const int vcap = Bcap - LY::NSDEP * Mp;
            const int GS = max(1, min(vcap / max(2 * L, 1), nS));
            __syncthreads();
            // gather W+, E_z, B_z (PlasmaSpecies.cpp:75-89, without B+) and form the particle's a_x, a_y
            Real cax[PT], cay[PT];
            fence_cache();
#pragma unroll
            for (int p0 = 0; p0 < PT; p0 += PB) {
            Real PXb[PB], PYb[PB], Db[PB], wb[PB];
            load_batch(p0, PXb, PYb, Db, wb);
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                const int pp = p0 + u;
                const int j = cj[pp];
                if (j < 0) continue;
                const Real PX = PXb[u], PY = PYb[u];
                const bool s1 = slot_of(pp) >= np0;
                const Real qm = s1 ? qm1 : qm0;
                const Real t = ct[pp], c = cc[pp], sn = cs[pp];
                const Real invd = cinvd[pp];
                Real Wx, Wy;
                if (M1)
                    eval_vector(lin(BA(LY::SW + 2), j, t), lin(BA(LY::SW + 3), j, t), lin(BA(LY::SW + 0), j, t),
                                lin(BA(LY::SW + 1), j, t), lin(BA(LY::SW + 4), j, t), lin(BA(LY::SW + 5), j, t), c, sn, true,
                                Wx, Wy);
                else
                    eval_vector(Real(0), Real(0), lin(BA(LY::SW + 0), j, t), lin(BA(LY::SW + 1), j, t), Real(0), Real(0), c,
                                sn, false, Wx, Wy);
                const Real Ez = eval_scalar(lin(BA(LY::SEZ), j, t), M1 ? lin(BA(LY::SEZ + 1), j, t) : Real(0),
                                            M1 ? lin(BA(LY::SEZ + 2), j, t) : Real(0), c, sn, M1);
                const Real Bz = eval_scalar(lin(BA(LY::SBZ), j, t), M1 ? lin(BA(LY::SBZ + 1), j, t) : Real(0),
                                            M1 ? lin(BA(LY::SBZ + 2), j, t) : Real(0), c, sn, M1);
                const Real dA = Real(0);
                const Real gam = cgam[pp];
                const Real fp = -Real(0.5) * qm * qm * dA * invd;
                const Real dD = qm * ((PX * Wx + PY * Wy) * invd - Ez);
                cax[pp] = (qm * (gam * Wx + PY * Bz) * invd + fp * c) * invd - PX * dD * invd * invd;
                cay[pp] = (qm * (gam * Wy - PX * Bz) * invd + fp * sn) * invd - PY * dD * invd * invd;
            }
            }
            QZ_PHASE(8);
            int gcur = 0, c0 = 0;
            auto reduce = [&]() {
                __syncthreads();
                const int nl = M1 ? s_misc[1] : 0;
                for (int item = tid; item < M * gcur; item += kBS) {
                    int g, n;
                    item_node(item, gcur, nl, g, n);
                    const int e0 = s_cs[n] + s_cs[n + 1], e1 = s_cs[n + 1] + s_cs[n + 2];
                    const Real* vb = VB + g * 2 * L;
                    const Real acc = sum_list(vb, e0, e1);
                    WS(W_SR1 + c0 + g)[n] = acc;
                }
                __syncthreads();
                QZ_PHASE(10);
                c0 += gcur;
                gcur = 0;
            };
            // component SC of S+ (0 r1, 1 i1, 2 r0, 3 i0, 4 r2, 5 i2) for every cached particle
            auto s_pass = [&](auto cst) {
                constexpr int SC = decltype(cst)::value;
                constexpr int mode = SC < 2 ? 1 : (SC < 4 ? 0 : 2);
                Real* vb = VB + gcur * 2 * L;
                fence_cache();
                fence_regs(cax); fence_regs(cay);
#pragma unroll
                for (int p0 = 0; p0 < PT; p0 += PB) {
                Real PXb[PB], PYb[PB], Db[PB], wb[PB];
                load_batch(p0, PXb, PYb, Db, wb);
#pragma unroll
                for (int u = 0; u < PB; ++u) {
                    const int pp = p0 + u;
                    const int j = cj[pp];
                    if (j < 0) continue;
                    const Real PX = PXb[u], PY = PYb[u], wi = wb[u];
                    const bool s1 = slot_of(pp) >= np0;
                    const Real qq = s1 ? q1 : q0;
                    const Real t = ct[pp], c = cc[pp], sn = cs[pp];
                    const Real invd = cinvd[pp];
                    const Real ux = PX * invd, uy = PY * invd;
                    const Real ur = c * ux + sn * uy, uth = c * uy - sn * ux;
                    const Real ck = mode == 1 ? c : (mode == 0 ? Real(1) : c * c - sn * sn);
                    const Real sk = mode == 1 ? sn : (mode == 0 ? Real(0) : Real(2) * c * sn);
                    const Real ax = cax[pp], ay = cay[pp];
                    const Real wq = wi * qq;
                    const Real ih = s_ih[j];
                    const int pl = cpos[pp] & 0xffff, pr = cpos[pp] >> 16;
#pragma unroll
                    for (int side = 0; side < 2; ++side) {
                        const Real iv = s_iv[j + side] * wq;
                        const Real W = side == 0 ? Real(1) - t : t;
                        const Real dW = side == 0 ? -ih : ih;
                        Real Ar = ax * W + ux * ur * dW;
                        Real Ai = ay * W + uy * ur * dW;
                        if (mode != 0) {
                            const Real Wor = (j == 0) ? (side == 0 ? Real(0) : ih) : W * cir[pp];
                            const Real f2 = mode * uth * Wor;
                            Ar += f2 * uy;
                            Ai -= f2 * ux;
                        }
                        const Real v = (SC & 1) == 0 ? iv * (Ar * ck + Ai * sk) : iv * (Ai * ck - Ar * sk);
                        vb[side == 0 ? pl : pr] = v;
                    }
                }
                }
                QZ_PHASE(9);
                if (++gcur == GS) reduce();
            };
            s_pass(IC<0>{});
            s_pass(IC<1>{});
            if (M1) {
                s_pass(IC<2>{});
                s_pass(IC<3>{});
                s_pass(IC<4>{});
                s_pass(IC<5>{});
            }
            if (gcur > 0) reduce();
        }
        __syncthreads();
        QZ_PHASE(10);

        // ================================================================ 5d. gz and the Picard iterations (Simulation.cpp:535-591)
        for (int j = tid; j < M; j += kBS) {
            // ((d_x + i d_y) J_z)_k + S_k
            const Real t1 = stn.grad(WS(W_JZ0), j);
            BA(LY::PGZ + 0)[j] = t1 + WS(W_SR1)[j];
            BA(LY::PGZ + 1)[j] = WS(W_SI1)[j];
            if (M1) {
                const Real* fr = WS(W_JZ1R);
                const Real* fi = WS(W_JZ1I);
                const Real t2 = stn.grad_odd(fr, j), t3 = stn.over_r(fr, j);
                const Real t4 = stn.grad_odd(fi, j), rh = stn.over_r(fi, j);
                BA(LY::PGZ + 4)[j] = (t2 - t3) + WS(W_SR2)[j];
                BA(LY::PGZ + 5)[j] = (t4 - rh) + WS(W_SI2)[j];
                BA(LY::PGZ + 2)[j] = (t2 + t3) + WS(W_SR0)[j];
                BA(LY::PGZ + 3)[j] = -(t4 + rh) + WS(W_SI0)[j];
                BA(LY::PCHI + 0)[j] = WS(W_CHI1R)[j];
                BA(LY::PCHI + 1)[j] = WS(W_CHI1I)[j];
            }
            BA(LY::PA_L1D)[j] = A.op_a[OP_L1D][j];
            BA(LY::PA_L1F)[j] = A.op_a[OP_L1F][j];
            for (int f = 0; f < (M1 ? 8 : 4); ++f) BA(LY::PF + f)[j] = WS(W_FAC0 + f)[j];
            for (int f = 0; f < (M1 ? 6 : 2); ++f) BA(LY::PBP + f)[j] = WS(W_BPR1 + f)[j];
            if (M1) {
                BA(LY::PA_L0)[j] = A.op_a[OP_L0][j];
                BA(LY::PA_L2D)[j] = A.op_a[OP_L2D][j];
            }
        }
        {
            const Real fcw = A.flux_coef * WS(W_JZ0)[N];
            const int niter = M1 ? A.picard : 1;
            __syncthreads();
            QZ_PHASE(11);
            for (int it = 0; it < niter; ++it) {
                if (it > 0) {
                    __syncthreads();
                    QZ_PHASE(13);
                }
                for (int j = tid; j < M; j += kBS) {
                    Real r1 = -BA(LY::PGZ + 1)[j], i1 = BA(LY::PGZ + 0)[j];
                    if (M1) {
                        Real r0 = -BA(LY::PGZ + 3)[j], i0 = BA(LY::PGZ + 2)[j];
                        Real r2 = -BA(LY::PGZ + 5)[j], i2 = BA(LY::PGZ + 4)[j];
                        const Real cr = BA(LY::PCHI)[j], ci = BA(LY::PCHI + 1)[j];
                        const Real b1r = BA(LY::PBP + 0)[j], b1i = BA(LY::PBP + 1)[j];
                        const Real b0r = BA(LY::PBP + 2)[j], b0i = BA(LY::PBP + 3)[j];
                        const Real b2r = BA(LY::PBP + 4)[j], b2i = BA(LY::PBP + 5)[j];
                        r0 += cr * b1r + ci * b1i; i0 += cr * b1i - ci * b1r;
                        r1 += cr * b0r - ci * b0i + cr * b2r + ci * b2i;
                        i1 += cr * b0i + ci * b0r + cr * b2i - ci * b2r;
                        r2 += cr * b1r - ci * b1i; i2 += cr * b1i + ci * b1r;
                        BA(LY::PRB + 2)[j] = r0; BA(LY::PRB + 3)[j] = i0;
                        BA(LY::PRB + 4)[j] = r2; BA(LY::PRB + 5)[j] = i2;
                    }
                    BA(LY::PRB + 0)[j] = r1; BA(LY::PRB + 1)[j] = i1;
                }
                __syncthreads();
                QZ_PHASE(12);
                // lanes: 0 B_r1 (L1D), 1 B_i1 (L1F, flux wall J_z(R)); m1: 2 B_r0, 3 B_i0 (L0), 4 B_r2, 5 B_i2 (L2D)
                if (w == 0) {
                    const int nj = M1 ? 6 : 2;
                    if (lane < nj) {
                        const int fidx = lane < 2 ? lane : (lane < 4 ? 2 : 3);   // factor set: L1D, L1F, L0, L2D
                        const Real* a = BA(fidx == 0 ? LY::PA_L1D : fidx == 1 ? LY::PA_L1F : fidx == 2 ? LY::PA_L0 : LY::PA_L2D);
                        const bool flux = fidx == 1;
                        thomas_lds(a, BA(LY::PF + 2 * fidx), BA(LY::PF + 2 * fidx + 1), s_bw0[fidx], BA(LY::PRB + lane),
                                   BA(LY::PBP + lane), Real(1), fidx != 2, !flux, flux, flux ? fcw : Real(0), M);
                    }
                }
            }
        }
        __syncthreads();
        QZ_PHASE(13);

        // ================================================================ 6. store (Simulation.cpp:594-633) + push
        for (int j = tid; j < M; j += kBS)
            for (int f = 0; f < (M1 ? 6 : 2); ++f) WS(W_BPR1 + f)[j] = BA(LY::PBP + f)[j];   // warm start of the next slice
        __syncthreads();
        for (int j = tid; j < M; j += kBS) {
            BA(LY::GW + 0)[j] = WS(W_WPR1)[j];
            BA(LY::GW + 1)[j] = WS(W_WPI1)[j];
            BA(LY::GEZ)[j] = WS(W_EZ0)[j];
            BA(LY::GBZ)[j] = WS(W_BZ0)[j];
            BA(LY::GBP + 0)[j] = WS(W_BPR1)[j];
            BA(LY::GBP + 1)[j] = WS(W_BPI1)[j];
            if (M1) {
                BA(LY::GW + 2)[j] = WS(W_WPR0)[j]; BA(LY::GW + 3)[j] = WS(W_WPI0)[j];
                BA(LY::GW + 4)[j] = WS(W_WPR2)[j]; BA(LY::GW + 5)[j] = WS(W_WPI2)[j];
                BA(LY::GEZ + 1)[j] = WS(W_EZ1R)[j]; BA(LY::GEZ + 2)[j] = WS(W_EZ1I)[j];
                BA(LY::GBZ + 1)[j] = WS(W_BZ1R)[j]; BA(LY::GBZ + 2)[j] = WS(W_BZ1I)[j];
                BA(LY::GBP + 2)[j] = WS(W_BPR0)[j]; BA(LY::GBP + 3)[j] = WS(W_BPI0)[j];
                BA(LY::GBP + 4)[j] = WS(W_BPR2)[j]; BA(LY::GBP + 5)[j] = WS(W_BPI2)[j];
            }
        }
        __syncthreads();
        {
            Real* __restrict__ F = A.fld + static_cast<size_t>(k) * M * A.ncf;
            const int ncf = A.ncf;
            for (int j = tid; j < M; j += kBS) {
                Real* Fj = F + static_cast<size_t>(j) * ncf;
                const Real b1r = BA(LY::GBP + 0)[j], b1i = BA(LY::GBP + 1)[j];
                Fj[F_EP1R] = BA(LY::GW + 0)[j] + b1i;
                Fj[F_EP1I] = BA(LY::GW + 1)[j] - b1r;
                Fj[F_BP1R] = b1r;
                Fj[F_BP1I] = b1i;
                if (M1) {
                    const Real b0r = BA(LY::GBP + 2)[j], b0i = BA(LY::GBP + 3)[j];
                    const Real b2r = BA(LY::GBP + 4)[j], b2i = BA(LY::GBP + 5)[j];
                    Fj[F_EP0R] = BA(LY::GW + 2)[j] + b0i;
                    Fj[F_EP0I] = BA(LY::GW + 3)[j] - b0r;
                    Fj[F_EP2R] = BA(LY::GW + 4)[j] + b2i;
                    Fj[F_EP2I] = BA(LY::GW + 5)[j] - b2r;
                    Fj[F_BP0R] = b0r; Fj[F_BP0I] = b0i; Fj[F_BP2R] = b2r; Fj[F_BP2I] = b2i;
                }
            }
        }
        QZ_PHASE(14);
        if (k < K - 1) {
            // ---- Adams-Bashforth push (PlasmaSpecies::push, src/PlasmaSpecies.cpp:600-667)
            const int slot = k % ord;
            const bool first = (k == 0);
            fence_cache();
#pragma unroll
            for (int pp = 0; pp < PT; ++pp) {
                const int j = cj[pp];
                if (j < 0) continue;
                const int gi = slot_of(pp);
                const bool s1 = gi >= np0;
                const qz_species_state& st = s1 ? A.sp[1] : A.sp[0];
                const int li = s1 ? gi - np0 : gi;
                const int cap = st.cap;
                Real* __restrict__ hist = st.hist;
                const Real X = st.x[li], Y = st.y[li], D = st.dl[li];
                const Real PX = st.px[li], PY = st.py[li];
                const Real qm = s1 ? qm1 : qm0;
                const Real t = ct[pp], c = cc[pp], sn = cs[pp];
                Real Wx, Wy, Bx, By;
                if (M1) {
                    eval_vector(lin(BA(LY::GW + 2), j, t), lin(BA(LY::GW + 3), j, t), lin(BA(LY::GW + 0), j, t),
                                lin(BA(LY::GW + 1), j, t), lin(BA(LY::GW + 4), j, t), lin(BA(LY::GW + 5), j, t), c, sn, true,
                                Wx, Wy);
                    eval_vector(lin(BA(LY::GBP + 2), j, t), lin(BA(LY::GBP + 3), j, t), lin(BA(LY::GBP + 0), j, t),
                                lin(BA(LY::GBP + 1), j, t), lin(BA(LY::GBP + 4), j, t), lin(BA(LY::GBP + 5), j, t), c, sn,
                                true, Bx, By);
                } else {
                    eval_vector(Real(0), Real(0), lin(BA(LY::GW + 0), j, t), lin(BA(LY::GW + 1), j, t), Real(0), Real(0),
                                c, sn, false, Wx, Wy);
                    eval_vector(Real(0), Real(0), lin(BA(LY::GBP + 0), j, t), lin(BA(LY::GBP + 1), j, t), Real(0),
                                Real(0), c, sn, false, Bx, By);
                }
                const Real Ez = eval_scalar(lin(BA(LY::GEZ), j, t), M1 ? lin(BA(LY::GEZ + 1), j, t) : Real(0),
                                            M1 ? lin(BA(LY::GEZ + 2), j, t) : Real(0), c, sn, M1);
                const Real Bz = eval_scalar(lin(BA(LY::GBZ), j, t), M1 ? lin(BA(LY::GBZ + 1), j, t) : Real(0),
                                            M1 ? lin(BA(LY::GBZ + 2), j, t) : Real(0), c, sn, M1);
                const Real dA = Real(0);
                const Real gam = cgam[pp];
                const Real invd = cinvd[pp];
                const Real fp = -Real(0.5) * qm * qm * dA * invd;
                const Real f0 = PX * invd;
                const Real f1 = PY * invd;
                const Real f2 = qm * (gam * Wx * invd + By + PY * Bz * invd) + fp * c;
                const Real f3 = qm * (gam * Wy * invd - Bx - PX * Bz * invd) + fp * sn;
                const Real f4 = qm * ((PX * Wx + PY * Wy) * invd - Ez);
                auto H = [&](int sl, int v) -> Real& { return hist[(static_cast<size_t>(sl) * 5 + v) * cap + li]; };
                Real s0 = 0, sm1 = 0, sm2 = 0, sm3 = 0, sm4 = 0;
                if (first) {
                    for (int sl = 0; sl < ord; ++sl) {
                        H(sl, 0) = f0; H(sl, 1) = f1; H(sl, 2) = f2; H(sl, 3) = f3; H(sl, 4) = f4;
                    }
                    for (int mm = 0; mm < ord; ++mm) {
                        const Real cm = A.ab_c[mm];
                        s0 += cm * f0; sm1 += cm * f1; sm2 += cm * f2; sm3 += cm * f3; sm4 += cm * f4;
                    }
                } else {
                    H(slot, 0) = f0; H(slot, 1) = f1; H(slot, 2) = f2; H(slot, 3) = f3; H(slot, 4) = f4;
                    for (int mm = 0; mm < ord; ++mm) {
                        const int sl = (slot - mm + ord) % ord;
                        const Real cm = A.ab_c[mm];
                        if (mm == 0) {
                            s0 += cm * f0; sm1 += cm * f1; sm2 += cm * f2; sm3 += cm * f3; sm4 += cm * f4;
                        } else {
                            s0 += cm * H(sl, 0); sm1 += cm * H(sl, 1); sm2 += cm * H(sl, 2); sm3 += cm * H(sl, 3);
                            sm4 += cm * H(sl, 4);
                        }
                    }
                }
                Real xn = X + dxi * s0, yn = Y + dxi * sm1;
                Real pxn = PX + dxi * sm2, pyn = PY + dxi * sm3;
                const Real dn = D + dxi * sm4;
                const Real gn = (Real(1) + pxn * pxn + pyn * pyn + dn * dn) / (Real(2) * dn);
                if (!(dn > dmin) || gn > qsamax * dn) {
                    st.w[li] = Real(0);
                    atomicAdd(&A.lost[s1 ? 1 : 0], 1);
                    continue;
                }
                const Real rn = dsqrt(xn * xn + yn * yn);
                if (rn > R) {
                    const Real nx = xn / rn, ny = yn / rn;
                    const Real rr = Real(2) * R - rn;
                    if (rr < Real(0)) {
                        st.w[li] = Real(0);
                        atomicAdd(&A.lost[s1 ? 1 : 0], 1);
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
                st.x[li] = xn; st.y[li] = yn; st.px[li] = pxn; st.py[li] = pyn; st.dl[li] = dn;
            }
        }
        __syncthreads();
        QZ_PHASE(15);
    }
    if (tid == 0 && stats)
        for (int i = 0; i < QZ_NPHASE; ++i) atomicAdd(&stats[1 + i], s_ph[i]);
#undef QZ_PHASE
}

}  // namespace qz
