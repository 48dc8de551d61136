// Body of the persistent gfx950 kernel of one quasi-static sweep (the slice loop of Simulation::step_fields,
// src/Simulation.cpp:727-783) run by G cooperating workgroups.  It serves runs that one workgroup cannot hold:
// more than 4096 plasma particles per slice, or grids whose solve sets do not fit one compute unit's LDS (the
// m = 0 + 1 long-box workload: 32768 particles, 1025 nodes).
//
// Workgroup g owns the particles [g P, (g + 1) P) of the run (P = 512 x PT) for the whole sweep.  Per slice:
//   1. geometry of the owned particles (registers) and a stable counting sort by cell in LDS;
//   2. the per-cell counts of every workgroup meet in global memory (grid barrier); each workgroup reads them
//      all and forms the global cell starts and its own offsets, so that every contribution gets its position in
//      its node's list in global particle-index order: workgroups hold consecutive index ranges, so the particles
//      of earlier workgroups come first, and within a workgroup the local sort gives the order;
//   3. the contributions of every deposit array go to these global lists (grid barrier); one thread per
//      (node, array), over all workgroups, sums a list sequentially -- the serial atomic_add order -- and adds
//      the background and beam sources (combine_sources) (grid barrier);
//   4. the solves: the psi / E_z / B_z systems of L0, those of L1D (m = 1) and the chi-dependent Thomas factors
//      of the B+ operators are three jobs for three workgroups, each staging its operators in its own LDS and
//      running its chains one lane per system (grid barrier);
//   5. every workgroup forms W+ from psi and stages W+, E_z, B_z in LDS; the owned particles gather them and
//      write their S+ contributions (grid barrier); the node sums of S+ as in 3 (grid barrier);
//   6. the Picard iterations of B+: the systems of each operator pair are one job per workgroup, reading the
//      previous iterate of the coupled modes from global memory and writing the next one to the other of two
//      buffers (one grid barrier per iteration);
//   7. the field store (node ranges per workgroup) and the Adams-Bashforth push of the owned particles.
// The grid barrier is a counter in the run's workspace that only grows during a launch: barrier n of the launch
// completes when it reaches n G; the last workgroup to leave the launch puts it back to zero.  Every wait is bounded in time: a barrier not reached within two seconds marks the run's error
// flag and ends the launch in every workgroup, and the host reports it.  A launch is refused unless all its
// workgroups (and those of the launches the caller keeps in flight beside it) fit the device together, one per
// compute unit, so every workgroup is resident.  Arithmetic per term is that of the one-workgroup kernel
// (qz_sweep.hip), so the result is the serial oracle's bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "qz_device.hpp"
#include "qz_internal.hpp"
#include "qz_sweep_common.hpp"

namespace qz {

namespace {
constexpr int mBS = 512;              // threads per workgroup
constexpr int mNW = mBS / 64;         // waves per workgroup
constexpr int mLds = 160 * 1024;      // LDS bytes per workgroup
constexpr int kDepComps = 24;         // deposit arrays (WsArr ids 0 .. 23)
constexpr unsigned long long kBarrierTicks = 200000000ull;   // 2 s of the 100 MHz wall clock

// layout of the multi-workgroup workspace of a run (doubles): misc | counts | lists | Picard second buffer
struct MwsLayout {
    size_t misc = 0, cnt, lists, bpalt, total;
    __host__ __device__ MwsLayout(int M, int N, int np, int G, bool m1) {
        cnt = 64;   // words 0 barrier, 1 exit count, 2 error; doubles [8..12) bw0
        lists = cnt + (static_cast<size_t>(G) * (N + 1) + 1) / 2 + 1;     // G x N ints
        bpalt = lists + static_cast<size_t>(kDepComps) * 2 * (np > 1 ? np : 1);
        total = bpalt + static_cast<size_t>(m1 ? 6 : 2) * M;
    }
};

// LDS carve-up (bytes): r | 1/h | 1/V | wave counters | local cell starts | global cell starts | own offsets |
// local sorted slots | misc (flags, counts, the deposit array list) | region (Mp-double slots)
struct MLds {
    int Mp, wcnt, cs, gcs, goff, sidx, misc, region;
    __host__ __device__ MLds(int N, int pt) {
        const int M = N + 1;
        Mp = (M + 1) & ~1;
        wcnt = 3 * Mp * 8;
        cs = wcnt + (((mNW * N * 2) + 15) & ~15);
        gcs = cs + ((((N + 3) * 4) + 15) & ~15);
        goff = gcs + ((((N + 3) * 4) + 15) & ~15);
        sidx = goff + ((((N + 2) * 4) + 15) & ~15);
        misc = sidx + (((mBS * pt * 2) + 15) & ~15);
        region = misc + 256;
    }
};

// region slots (Mp doubles each)
template <bool M1> struct MSlots {
    // fields staged for the S+ gather and the push: W+ (r1 i1 r0 i0 r2 i2), E_z modes, B_z modes
    static constexpr int W = 0, EZ = M1 ? 6 : 2, BZ = EZ + (M1 ? 3 : 1), NFIELD = BZ + (M1 ? 3 : 1);
    // the largest job: m = 1 the chi factors (three operators and chi0), m = 0 the Picard pair (two a, two factor
    // pairs, two right-hand sides)
    static constexpr int NJOB = M1 ? 10 : 8;
    static constexpr int NREGION = NFIELD > NJOB ? NFIELD : NJOB;
};
}  // namespace

// Grid barrier of the G workgroups of a run.  `target` counts this workgroup's barriers; the shared counter only
// grows, so barrier n completes when it reaches n G.  Returns false (and sets the run's error flag) when the
// others do not arrive in time, or when another workgroup has already given up.
// One release fence before the arrival and one acquire fence after the wait, with the counter itself relaxed
// (fence-to-fence synchronisation through the counter's release sequence): on this GPU an agent-scope release
// writes back the die's whole L2 and an acquire invalidates it, operations every other kernel on the die waits
// behind, so the barrier issues one of each and none per poll.
__device__ __forceinline__ bool grid_sync(unsigned int* __restrict__ ctr, unsigned int* __restrict__ err, int G,
                                          unsigned int& target, int* s_ok) {
    __syncthreads();
    target += 1;
    if (threadIdx.x == 0) {
        int ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // this workgroup's writes before the arrival
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned int want = target * static_cast<unsigned int>(G);
        const unsigned long long t0 = wall_clock64();
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                wall_clock64() - t0 > kBarrierTicks) {
                __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the other workgroups' writes after it
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// The sweep of workgroup `block` of a launch: block_map[2 block] is its run, block_map[2 block + 1] its index in the
// run's G workgroups.
template <bool M1, int PT>
__device__ __forceinline__ void sweep_multi_run(const qz_sweep_args* __restrict__ runs,
                                                const int* __restrict__ block_map, int block) {
    using SL = MSlots<M1>;
    const int run = block_map[2 * block], grp = block_map[2 * block + 1];
    const qz_sweep_args& A = runs[run];
    const int G = A.groups;
    extern __shared__ double smem_d[];
    char* smem = reinterpret_cast<char*>(smem_d);

    const int N = A.grid.N, M = N + 1;
    const Real R = A.grid.R;
    Real* __restrict__ ws = A.ws;
    auto WS = [&](int arr) -> Real* { return ws + static_cast<size_t>(arr) * M; };
    const int K = A.nslices;
    const int kb = A.k_begin, ke = A.k_end > 0 ? A.k_end : K;
    const Real dxi = A.dxi;
    const Real idxi = Real(1) / dxi;
    const Stencil stn{A.grad_m, A.grad_0, A.grad_p, A.inv_r, N};

    // species, as the one-workgroup kernel: particle gi of the run is slot gi of species 0, then species 1
    const int nsp = A.nspecies;
    const int np0 = nsp > 0 ? A.sp[0].np : 0, np1 = nsp > 1 ? A.sp[1].np : 0;
    const int NP = np0 + np1;
    const Real q0 = nsp > 0 ? A.sp[0].q : Real(0), q1 = nsp > 1 ? A.sp[1].q : Real(0);
    const Real qm0 = nsp > 0 ? A.sp[0].q / A.sp[0].m : Real(0), qm1 = nsp > 1 ? A.sp[1].q / A.sp[1].m : Real(0);
    const bool has_neg = (nsp > 0 && q0 < 0) || (nsp > 1 && q1 < 0);
    const bool has_pos = (nsp > 0 && !(q0 < 0)) || (nsp > 1 && !(q1 < 0));
    const int own0 = grp * mBS * PT;   // first particle of this workgroup

    // the run's multi-workgroup workspace
    const MwsLayout mw(M, N, NP, G, M1);
    Real* __restrict__ mws = A.mws;
    unsigned int* bar = reinterpret_cast<unsigned int*>(mws);
    unsigned int* err = bar + 2;
    Real* g_bw0 = mws + 8;
    int* g_cnt = reinterpret_cast<int*>(mws + mw.cnt);
    Real* __restrict__ g_list = mws + mw.lists;
    const size_t lstride = 2 * static_cast<size_t>(NP > 1 ? NP : 1);
    Real* bpbuf[2] = {WS(W_BPR1), mws + mw.bpalt};   // Picard iterates: W_BPR1.. of ws, then the second buffer
    unsigned int nbar = 0;

    // LDS
    const MLds plan(N, PT);
    const int Mp = plan.Mp;
    Real* s_r = smem_d;
    Real* s_ih = s_r + Mp;
    Real* s_iv = s_ih + Mp;
    unsigned short* s_wcnt = reinterpret_cast<unsigned short*>(smem + plan.wcnt);
    int* s_cs = reinterpret_cast<int*>(smem + plan.cs);
    int* s_gcs = reinterpret_cast<int*>(smem + plan.gcs);
    int* s_goff = reinterpret_cast<int*>(smem + plan.goff);
    unsigned short* s_sidx = reinterpret_cast<unsigned short*>(smem + plan.sidx);
    int* s_misc = reinterpret_cast<int*>(smem + plan.misc);   // [0] barrier verdict [1] live [2] last cell count
    int* s_comps = s_misc + 8;                                 // the deposit arrays of this run, in pass order
    int& s_ok = s_misc[0];
    int& s_live = s_misc[1];
    Real* B = reinterpret_cast<Real*>(smem + plan.region);
    auto BA = [&](int slot) -> Real* { return B + static_cast<size_t>(slot) * Mp; };

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int gthreads = G * mBS, gtid = grp * mBS + tid;
    for (int j = tid; j < M; j += mBS) {
        s_r[j] = A.grid.r[j];
        s_iv[j] = A.grid.inv_V[j];
        if (j < N) s_ih[j] = A.grid.inv_h[j];
    }
    if (kb == 0)   // B+ := 0 at the head (Simulation.cpp:705-708); every workgroup its share of the nodes
        for (int j = gtid; j < M; j += gthreads)
            for (int a = 0; a < (M1 ? 6 : 2); ++a) WS(W_BPR1 + a)[j] = Real(0);
    // deposit arrays of this run in a fixed order
    int ncomp = 0;
    {
        auto put = [&](int c) {
            if (tid == 0) s_comps[ncomp] = c;
            ++ncomp;
        };
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
    const int nS = M1 ? 6 : 2;
    const Real dmin = A.delta_min, qsamax = A.max_qsa;
    const int ord = A.ab_order;
    const int niter = M1 ? A.picard : 1;
    __syncthreads();
    if (kb == 0 && A.load) {   // this workgroup's particles of the plasma at the head of the box
        const int own1 = own0 + mBS * PT;
        if (nsp > 0) load_rings(A.rings[0], A.sp[0], ord, s_r, own0 + tid, min(own1, np0), mBS);
        if (nsp > 1) load_rings(A.rings[1], A.sp[1], ord, s_r, max(own0, np0) - np0 + tid, min(own1 - np0, np1), mBS);
        __syncthreads();
    }

    int cj[PT], lrank[PT], cpl[PT], cpr[PT];
    Real ct[PT], cc[PT], cs[PT], cir[PT], cinvd[PT], cgam[PT];
    auto slot_of = [&](int pp) { return w * (PT * 64) + pp * 64 + lane; };   // local slot; particle own0 + slot
    auto species_of = [&](int gi, int& li) -> const qz_species_state& {
        const bool s1 = gi >= np0;
        li = s1 ? gi - np0 : gi;
        return s1 ? A.sp[1] : A.sp[0];
    };

    bool alive = true;
    for (int k = kb; k < ke && alive; ++k) {
        // ================================================================ 1. geometry and the local sort
#pragma unroll
        for (int pp = 0; pp < PT; ++pp) {
            const int gi = own0 + slot_of(pp);
            cj[pp] = -1;
            if (gi < NP) {
                int li;
                const qz_species_state& st = species_of(gi, li);
                const Real wi = st.w[li];
                if (wi != Real(0)) {
                    const Real X = st.x[li], Y = st.y[li];
                    const Real PX = st.px[li], PY = st.py[li], D = st.dl[li];
                    const Real r = dsqrt(X * X + Y * Y);
                    Real c, s_;
                    if (r > Real(0)) { c = X / r; s_ = Y / r; } else { c = Real(1); s_ = Real(0); }
                    int j; Real t;
                    grid_weights(s_r, s_ih, N, R, r, j, t);
                    cj[pp] = j; ct[pp] = t; cc[pp] = c; cs[pp] = s_;
                    cir[pp] = r > Real(0) ? Real(1) / r : Real(0);
                    cgam[pp] = (Real(1) + PX * PX + PY * PY + D * D) / (Real(2) * D);
                    cinvd[pp] = Real(1) / D;
                }
            }
        }
        for (int i = tid; i < mNW * N; i += mBS) s_wcnt[i] = 0;
        if (tid == 0) s_live = 0;
        __syncthreads();
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
            if (lane == 0) atomicAdd(&s_live, live);
        }
        __syncthreads();
        for (int c = tid; c < N; c += mBS) {   // per-cell prefix over the waves; cell totals to global memory
            int run_ = 0;
            for (int ww = 0; ww < mNW; ++ww) {
                const int v = s_wcnt[ww * N + c];
                s_wcnt[ww * N + c] = static_cast<unsigned short>(run_);
                run_ += v;
            }
            s_cs[c + 1] = run_;
            g_cnt[static_cast<size_t>(grp) * N + c] = run_;
        }
        __syncthreads();
        const int Lloc = s_live;
        if (w == 0) wave_exclusive_scan(s_cs + 1, N, lane);
        if (tid == 0) { s_cs[0] = 0; s_cs[N + 1] = Lloc; s_cs[N + 2] = Lloc; }
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < PT; ++pp) {
            const int j = cj[pp];
            if (j >= 0) {
                const int pos = s_cs[j + 1] + s_wcnt[w * N + j] + lrank[pp];
                s_sidx[pos] = static_cast<unsigned short>(slot_of(pp));
                lrank[pp] = pos - s_cs[j + 1];   // rank within the cell, index order
            }
        }
        if (!(alive = grid_sync(bar, err, G, nbar, &s_ok))) break;

        // ================================================================ 2. global cell starts, own offsets
        // s_gcs[c + 1] = start of cell c over all workgroups; s_goff[c + 1] = particles of cell c in the
        // workgroups before this one (s_goff[0] = s_goff[N + 1] = 0: no cells beyond the grid)
        for (int c = tid; c < N; c += mBS) {
            int before = 0, total = 0;
            for (int g = 0; g < G; ++g) {
                const int v = g_cnt[static_cast<size_t>(g) * N + c];
                before += g < grp ? v : 0;
                total += v;
            }
            s_goff[c + 1] = before;
            s_gcs[c + 1] = total;
            if (c == N - 1) s_misc[2] = total;
        }
        __syncthreads();
        if (w == 0) wave_exclusive_scan(s_gcs + 1, N, lane);
        __syncthreads();
        const int Ltot = s_gcs[N] + s_misc[2];   // start of the last cell and its particles
        if (tid == 0) {
            s_gcs[0] = 0; s_gcs[N + 1] = Ltot; s_gcs[N + 2] = Ltot;
            s_goff[0] = 0; s_goff[N + 1] = 0;
            if (grp == 0 && A.stats) atomicAdd(A.stats, static_cast<unsigned long long>(Ltot));
        }
        __syncthreads();
        // positions of the two contributions of every owned particle in the global node lists: node n holds
        // cell n-1 (right entries) and cell n (left entries), merged in particle-index order
#pragma unroll
        for (int pp = 0; pp < PT; ++pp) {
            const int j = cj[pp];
            if (j >= 0) {
                const int sl = slot_of(pp);
                const int rank = s_goff[j + 1] + lrank[pp];
                const int less_m = s_goff[j] + count_less(s_sidx, s_cs[j], s_cs[j + 1], sl);
                const int less_p = s_goff[j + 2] + count_less(s_sidx, s_cs[j + 2], s_cs[j + 3], sl);
                cpl[pp] = (s_gcs[j] + s_gcs[j + 1]) + rank + less_m;
                cpr[pp] = (s_gcs[j + 1] + s_gcs[j + 2]) + rank + less_p;
            }
        }

        // ================================================================ 3. deposit: contributions, node sums
#pragma unroll
        for (int pp = 0; pp < PT; ++pp) {
            const int j = cj[pp];
            if (j < 0) continue;
            int li;
            const int gi = own0 + slot_of(pp);
            const qz_species_state& st = species_of(gi, li);
            const bool s1 = gi >= np0;
            const Real PX = st.px[li], PY = st.py[li], D = st.dl[li], wi = st.w[li];
            const Real q = s1 ? q1 : q0, qm = s1 ? qm1 : qm0;
            const Real ih = s_ih[j];
            const Real pz = cgam[pp] - D;
            for (int ci = 0; ci < ncomp; ++ci) {
                Real* L = g_list + static_cast<size_t>(ci) * lstride;
                const int C = s_comps[ci];
                L[cpl[pp]] = dep_value(C, 0, j, ct[pp], cc[pp], cs[pp], cir[pp], cinvd[pp], cgam[pp], PX, PY, pz, wi,
                                       q, qm, s_iv[j], ih);
                L[cpr[pp]] = dep_value(C, 1, j, ct[pp], cc[pp], cs[pp], cir[pp], cinvd[pp], cgam[pp], PX, PY, pz, wi,
                                       q, qm, s_iv[j + 1], ih);
            }
        }
        if (!(alive = grid_sync(bar, err, G, nbar, &s_ok))) break;
        {
            Real* __restrict__ Dg = A.diag + static_cast<size_t>(k) * M * A.ncd;
            const int ncd = A.ncd, ncb = A.ncb;
            const Real* __restrict__ bs = A.bsrc;
            const Real* __restrict__ bg = A.bg;
            const bool ngp = A.ngp != 0;
            auto bb = [&](int kk, int j, int comp) { return bs[(static_cast<size_t>(kk) * M + j) * ncb + comp]; };
            auto dxi_b = [&](int j, int comp) {   // combine_sources' xi derivative of the beam source
                if (ngp) {
                    const Real prev = (k > 0) ? bb(k - 1, j, comp) : Real(0);
                    return (bb(k, j, comp) - prev) * idxi;
                }
                if (K == 1) return Real(0);
                if (k == 0) return (bb(1, j, comp) - bb(0, j, comp)) * idxi;
                if (k == K - 1) return (bb(K - 1, j, comp) - bb(K - 2, j, comp)) * idxi;
                return Real(0.5) * (bb(k + 1, j, comp) - bb(k - 1, j, comp)) * idxi;
            };
            for (int item = gtid; item < ncomp * M; item += gthreads) {
                const int ci = item / M, n = item - ci * M;
                const int e0 = s_gcs[n] + s_gcs[n + 1], e1 = s_gcs[n + 1] + s_gcs[n + 2];
                const Real acc = sum_list(g_list + static_cast<size_t>(ci) * lstride, e0, e1);
                const int comp = s_comps[ci];
                Real* Dn = Dg + static_cast<size_t>(n) * ncd;
                // the node sum combined with the background and the beam sources (Simulation.cpp:408-457)
                switch (comp) {
                    case W_RHOT0: WS(comp)[n] = acc + (bg[n] + bb(k, n, B_RT0)); break;
                    case W_DRHO0: WS(comp)[n] = acc + dxi_b(n, B_RT0); break;
                    case W_JZ0: WS(comp)[n] = acc + bb(k, n, B_JZ0); break;
                    case W_JPR1: WS(comp)[n] = acc + bb(k, n, B_JP1R); break;
                    case W_JPI1: WS(comp)[n] = acc + bb(k, n, B_JP1I); break;
                    case W_RHOT1R: WS(comp)[n] = acc + (bg[M + n] + bb(k, n, B_RT1R)); break;
                    case W_RHOT1I: WS(comp)[n] = acc + (bg[2 * M + n] + bb(k, n, B_RT1I)); break;
                    case W_DRHO1R: WS(comp)[n] = acc + dxi_b(n, B_RT1R); break;
                    case W_DRHO1I: WS(comp)[n] = acc + dxi_b(n, B_RT1I); break;
                    case W_JZ1R: WS(comp)[n] = acc + bb(k, n, B_JZ1R); break;
                    case W_JZ1I: WS(comp)[n] = acc + bb(k, n, B_JZ1I); break;
                    case W_JPR0: WS(comp)[n] = acc + bb(k, n, B_JP0R); break;
                    case W_JPI0: WS(comp)[n] = acc + bb(k, n, B_JP0I); break;
                    case W_JPR2: WS(comp)[n] = acc + bb(k, n, B_JP2R); break;
                    case W_JPI2: WS(comp)[n] = acc + bb(k, n, B_JP2I); break;
                    case W_NE0: Dn[D_NE0] = acc; break;
                    case W_NE1R: Dn[D_NE1R] = acc; break;
                    case W_NE1I: Dn[D_NE1I] = acc; break;
                    case W_NI1R: Dn[D_NI1R] = acc; break;
                    case W_NI1I: Dn[D_NI1I] = acc; break;
                    case W_NI0: {   // combine_sources adds the positive background (Simulation.cpp:441)
                        const Real bgr0 = bg[3 * M + n];
                        Dn[D_NI0] = bgr0 > Real(0) ? acc + bgr0 : acc;
                        break;
                    }
                    default: WS(comp)[n] = acc;   // chi
                }
            }
            // densities of absent species are zero, plus the positive background for n_i
            for (int n = gtid; n < M; n += gthreads) {
                Real* Dn = Dg + static_cast<size_t>(n) * ncd;
                if (!has_neg) {
                    Dn[D_NE0] = Real(0);
                    if (M1) { Dn[D_NE1R] = Real(0); Dn[D_NE1I] = Real(0); }
                }
                if (!has_pos) {
                    const Real bgr0 = bg[3 * M + n];
                    Dn[D_NI0] = bgr0 > Real(0) ? Real(0) + bgr0 : Real(0);
                    if (M1) { Dn[D_NI1R] = Real(0); Dn[D_NI1I] = Real(0); }
                }
            }
        }
        if (!(alive = grid_sync(bar, err, G, nbar, &s_ok))) break;

        // ================================================================ 4. the solves (three jobs)
        // job 0: psi0, E_z0, B_z0 with L0;  job 1 (m = 1): psi1, E_z1, B_z1 (r, i) with L1D;  job 2: this
        // slice's Thomas factors of (L - chi0) for L1D, L1F (and L0, L2D in m = 1).  Job j runs on workgroup j % G.
        for (int job = grp; job < 3; job += G) {
            if (job == 1 && !M1) continue;
            if (job < 2) {
                const bool l0 = job == 0;
                const int op = l0 ? OP_L0 : OP_L1D;
                Real* a = BA(0);
                Real* m = BA(1);
                Real* cp = BA(2);
                const int nsys = l0 ? 3 : 6;
                for (int j = tid; j < M; j += mBS) {
                    a[j] = A.op_a[op][j];
                    m[j] = l0 ? A.fac_l0_m[j] : A.fac_l1_m[j];
                    cp[j] = l0 ? A.fac_l0_cp[j] : A.fac_l1_cp[j];
                    if (l0) {
                        BA(3)[j] = WS(W_RHOT0)[j];
                        BA(4)[j] = WS(W_DRHO0)[j];
                        // B_z right-hand side (compute_bz, Simulation.cpp:506-526)
                        const Real t1 = stn.grad_odd(WS(W_JPI1), j), t2 = stn.over_r(WS(W_JPI1), j);
                        BA(5)[j] = -(t1 + t2);
                    } else {
                        BA(3)[j] = WS(W_RHOT1R)[j];
                        BA(4)[j] = WS(W_RHOT1I)[j];
                        BA(5)[j] = WS(W_DRHO1R)[j];
                        BA(6)[j] = WS(W_DRHO1I)[j];
                        const Real gr0 = stn.grad_odd(WS(W_JPR2), j), gi0 = stn.over_r(WS(W_JPR2), j);
                        const Real gr1 = stn.grad_odd(WS(W_JPI2), j), gi1 = stn.over_r(WS(W_JPI2), j);
                        const Real gr2 = stn.grad(WS(W_JPR0), j), gi2 = stn.grad(WS(W_JPI0), j);
                        const Real D1r = gr0 + Real(2) * gi0, D1i = gr1 + Real(2) * gi1;
                        const Real Dmr = gr2, Dmi = gi2;
                        BA(7)[j] = -Real(0.5) * (D1i + Dmi);
                        BA(8)[j] = Real(0.5) * (D1r - Dmr);
                    }
                }
                __syncthreads();
                if (w == 0 && lane < nsys) {
                    // psi and E_z: -rhs; B_z: +rhs (FieldSolver::solve scale); L0 keeps the axis row regular
                    const bool bz = l0 ? lane == 2 : lane >= 4;
                    thomas_lds(a, m, cp, A.op_b[op][0], BA(3 + lane), BA(3 + lane), bz ? Real(1) : Real(-1), !l0, true,
                               false, Real(0), M);
                }
                __syncthreads();
                // results: psi over its right-hand side in ws, E_z and B_z to their arrays
                for (int j = tid; j < M; j += mBS) {
                    if (l0) {
                        WS(W_RHOT0)[j] = BA(3)[j];
                        WS(W_EZ0)[j] = BA(4)[j];
                        WS(W_BZ0)[j] = BA(5)[j];
                    } else {
                        WS(W_RHOT1R)[j] = BA(3)[j];
                        WS(W_RHOT1I)[j] = BA(4)[j];
                        WS(W_EZ1R)[j] = BA(5)[j];
                        WS(W_EZ1I)[j] = BA(6)[j];
                        WS(W_BZ1R)[j] = BA(7)[j];
                        WS(W_BZ1I)[j] = BA(8)[j];
                    }
                }
                __syncthreads();
            } else {
                // rows of L1 (shared by L1D and L1F up to the last row), L0 and L2D, and chi0
                for (int j = tid; j < M; j += mBS) {
                    BA(0)[j] = A.op_a[OP_L1D][j]; BA(1)[j] = A.op_b[OP_L1D][j]; BA(2)[j] = A.op_c[OP_L1D][j];
                    BA(3)[j] = WS(W_CHI0)[j];
                    if (A.chi_out) A.chi_out[static_cast<size_t>(k) * M + j] = WS(W_CHI0)[j];   // for the laser march
                    if (M1) {
                        BA(4)[j] = A.op_a[OP_L0][j]; BA(5)[j] = A.op_b[OP_L0][j]; BA(6)[j] = A.op_c[OP_L0][j];
                        BA(7)[j] = A.op_a[OP_L2D][j]; BA(8)[j] = A.op_b[OP_L2D][j]; BA(9)[j] = A.op_c[OP_L2D][j];
                    }
                }
                __syncthreads();
                const int nf = M1 ? 4 : 2;
                if (w == 0 && lane < nf) {
                    const int base = lane < 2 ? 0 : (lane == 2 ? 4 : 7);
                    const int op = lane == 0 ? OP_L1D : lane == 1 ? OP_L1F : lane == 2 ? OP_L0 : OP_L2D;
                    const Real aN = A.op_a[op][N], bN = A.op_b[op][N];
                    const bool dir0 = op != OP_L0, dirN = op != OP_L1F;
                    g_bw0[lane] = thomas_factor(BA(base), BA(base + 1), BA(base + 2), BA(3), aN, bN, dir0, dirN,
                                                WS(W_FAC0 + 2 * lane), WS(W_FAC0 + 2 * lane + 1), M);
                }
                __syncthreads();
            }
        }
        if (!(alive = grid_sync(bar, err, G, nbar, &s_ok))) break;

        // ================================================================ 5. W+, field outputs, S+ deposit
        {
            Real* __restrict__ F = A.fld + static_cast<size_t>(k) * M * A.ncf;
            Real* __restrict__ Dg = A.diag + static_cast<size_t>(k) * M * A.ncd;
            const int ncf = A.ncf, ncd = A.ncd;
            const Real* psi0 = WS(W_RHOT0);
            const Real* psr = WS(W_RHOT1R);
            const Real* psi_ = WS(W_RHOT1I);
            for (int j = tid; j < M; j += mBS) {   // compute_wplus (Simulation.cpp:477-498) into LDS
                const Real t1 = stn.grad(psi0, j);
                BA(SL::W + 0)[j] = -t1;
                BA(SL::W + 1)[j] = Real(0);
                BA(SL::EZ)[j] = WS(W_EZ0)[j];
                BA(SL::BZ)[j] = WS(W_BZ0)[j];
                if (M1) {
                    const Real t2 = stn.grad_odd(psr, j), t3 = stn.over_r(psr, j);
                    const Real t4 = stn.grad_odd(psi_, j), rh = stn.over_r(psi_, j);
                    BA(SL::W + 4)[j] = -(t2 - t3);
                    BA(SL::W + 5)[j] = -(t4 - rh);
                    BA(SL::W + 2)[j] = -(t2 + t3);
                    BA(SL::W + 3)[j] = (t4 + rh);
                    BA(SL::EZ + 1)[j] = WS(W_EZ1R)[j]; BA(SL::EZ + 2)[j] = WS(W_EZ1I)[j];
                    BA(SL::BZ + 1)[j] = WS(W_BZ1R)[j]; BA(SL::BZ + 2)[j] = WS(W_BZ1I)[j];
                }
            }
            __syncthreads();
            // this workgroup's share of the nodes: psi, E_z, B_z to the stores, W+ to ws for the field store
            for (int j = gtid; j < M; j += gthreads) {
                Real* Fj = F + static_cast<size_t>(j) * ncf;
                Real* Dj = Dg + static_cast<size_t>(j) * ncd;
                Dj[D_PSI0] = psi0[j];
                Fj[F_EZ0] = BA(SL::EZ)[j];
                Fj[F_BZ0] = BA(SL::BZ)[j];
                for (int a = 0; a < (M1 ? 6 : 2); ++a) WS(W_WPR1 + a)[j] = BA(SL::W + a)[j];
                if (M1) {
                    Dj[D_PSI1R] = psr[j]; Dj[D_PSI1I] = psi_[j];
                    Fj[F_EZ1R] = BA(SL::EZ + 1)[j]; Fj[F_EZ1I] = BA(SL::EZ + 2)[j];
                    Fj[F_BZ1R] = BA(SL::BZ + 1)[j]; Fj[F_BZ1I] = BA(SL::BZ + 2)[j];
                }
            }
            // S+ contributions of the owned particles (PlasmaSpecies::deposit_S, PlasmaSpecies.cpp:516-558)
#pragma unroll
            for (int pp = 0; pp < PT; ++pp) {
                const int j = cj[pp];
                if (j < 0) continue;
                int li;
                const int gi = own0 + slot_of(pp);
                const qz_species_state& st = species_of(gi, li);
                const bool s1 = gi >= np0;
                const Real PX = st.px[li], PY = st.py[li], wi = st.w[li];
                const Real qm = s1 ? qm1 : qm0, qq = s1 ? q1 : q0;
                const Real t = ct[pp], c = cc[pp], sn = cs[pp];
                const Real invd = cinvd[pp];
                Real Wx, Wy;
                if (M1)
                    eval_vector(lin(BA(SL::W + 2), j, t), lin(BA(SL::W + 3), j, t), lin(BA(SL::W + 0), j, t),
                                lin(BA(SL::W + 1), j, t), lin(BA(SL::W + 4), j, t), lin(BA(SL::W + 5), j, t), c, sn, true,
                                Wx, Wy);
                else
                    eval_vector(Real(0), Real(0), lin(BA(SL::W + 0), j, t), lin(BA(SL::W + 1), j, t), Real(0), Real(0), c,
                                sn, false, Wx, Wy);
                const Real Ez = eval_scalar(lin(BA(SL::EZ), j, t), M1 ? lin(BA(SL::EZ + 1), j, t) : Real(0),
                                            M1 ? lin(BA(SL::EZ + 2), j, t) : Real(0), c, sn, M1);
                const Real Bz = eval_scalar(lin(BA(SL::BZ), j, t), M1 ? lin(BA(SL::BZ + 1), j, t) : Real(0),
                                            M1 ? lin(BA(SL::BZ + 2), j, t) : Real(0), c, sn, M1);
                const Real dA = Real(0);
                const Real gam = cgam[pp];
                const Real fp = -Real(0.5) * qm * qm * dA * invd;
                const Real dD = qm * ((PX * Wx + PY * Wy) * invd - Ez);
                const Real ax = (qm * (gam * Wx + PY * Bz) * invd + fp * c) * invd - PX * dD * invd * invd;
                const Real ay = (qm * (gam * Wy - PX * Bz) * invd + fp * sn) * invd - PY * dD * invd * invd;
                const Real ux = PX * invd, uy = PY * invd;
                const Real ur = c * ux + sn * uy, uth = c * uy - sn * ux;
                const Real wq = wi * qq;
                const Real ih = s_ih[j];
                for (int SC = 0; SC < nS; ++SC) {
                    const int mode = SC < 2 ? 1 : (SC < 4 ? 0 : 2);
                    const Real ck = mode == 1 ? c : (mode == 0 ? Real(1) : c * c - sn * sn);
                    const Real sk = mode == 1 ? sn : (mode == 0 ? Real(0) : Real(2) * c * sn);
                    Real* L = g_list + static_cast<size_t>(SC) * lstride;
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
                        L[side == 0 ? cpl[pp] : cpr[pp]] = v;
                    }
                }
            }
        }
        if (!(alive = grid_sync(bar, err, G, nbar, &s_ok))) break;
        for (int item = gtid; item < nS * M; item += gthreads) {
            const int sc = item / M, n = item - sc * M;
            const int e0 = s_gcs[n] + s_gcs[n + 1], e1 = s_gcs[n + 1] + s_gcs[n + 2];
            WS(W_SR1 + sc)[n] = sum_list(g_list + static_cast<size_t>(sc) * lstride, e0, e1);
        }
        if (!(alive = grid_sync(bar, err, G, nbar, &s_ok))) break;

        // ================================================================ 6. Picard iterations of B+ (Simulation.cpp:535-591)
        // jobs: 0 (B_r1 with L1D, B_i1 with L1F), m = 1: 1 (B_r0, B_i0 with L0), 2 (B_r2, B_i2 with L2D).  The
        // iterate `it` reads buffer it % 2 (buffer 0 is ws' B+: the warm start) and writes buffer (it + 1) % 2.
        {
            const int njob = M1 ? 3 : 1;
            const Real fcw = A.flux_coef * WS(W_JZ0)[N];
            for (int it = 0; it < niter; ++it) {
                const Real* bp = bpbuf[it & 1];
                Real* bn = bpbuf[(it + 1) & 1];
                auto BP = [&](int f, int j) { return bp[static_cast<size_t>(f) * M + j]; };
                for (int job = grp; job < njob; job += G) {
                    // slots: 0 a (first system), 1 a (second), 2,3 factors of the first, 4,5 of the second,
                    // 6,7 right-hand sides, then solutions in place
                    const int sys0 = 2 * job;   // B+ arrays: 0 r1, 1 i1, 2 r0, 3 i0, 4 r2, 5 i2
                    const int fa = job == 0 ? 0 : (job == 1 ? 2 : 3);                      // factor set: L1D L1F L0 L2D
                    const int fb = job == 0 ? 1 : fa;
                    const int opa = job == 0 ? OP_L1D : (job == 1 ? OP_L0 : OP_L2D);
                    const int opb = job == 0 ? OP_L1F : opa;
                    for (int j = tid; j < M; j += mBS) {
                        BA(0)[j] = A.op_a[opa][j];
                        BA(1)[j] = A.op_a[opb][j];
                        BA(2)[j] = WS(W_FAC0 + 2 * fa)[j]; BA(3)[j] = WS(W_FAC0 + 2 * fa + 1)[j];
                        BA(4)[j] = WS(W_FAC0 + 2 * fb)[j]; BA(5)[j] = WS(W_FAC0 + 2 * fb + 1)[j];
                        // gz: ((d_x + i d_y) J_z)_k + S_k, and the chi_1 coupling with the previous iterate
                        Real gr, gi;
                        if (job == 0) {
                            const Real t1 = stn.grad(WS(W_JZ0), j);
                            gr = t1 + WS(W_SR1)[j];
                            gi = WS(W_SI1)[j];
                        } else {
                            const Real* fr = WS(W_JZ1R);
                            const Real* fi = WS(W_JZ1I);
                            const Real t2 = stn.grad_odd(fr, j), t3 = stn.over_r(fr, j);
                            const Real t4 = stn.grad_odd(fi, j), rh = stn.over_r(fi, j);
                            if (job == 1) {
                                gr = (t2 + t3) + WS(W_SR0)[j];
                                gi = -(t4 + rh) + WS(W_SI0)[j];
                            } else {
                                gr = (t2 - t3) + WS(W_SR2)[j];
                                gi = (t4 - rh) + WS(W_SI2)[j];
                            }
                        }
                        Real rr = -gi, ri = gr;
                        if (M1) {
                            const Real cr = WS(W_CHI1R)[j], ci = WS(W_CHI1I)[j];
                            const Real b1r = BP(0, j), b1i = BP(1, j), b0r = BP(2, j), b0i = BP(3, j);
                            const Real b2r = BP(4, j), b2i = BP(5, j);
                            if (job == 0) {
                                rr += cr * b0r - ci * b0i + cr * b2r + ci * b2i;
                                ri += cr * b0i + ci * b0r + cr * b2i - ci * b2r;
                            } else if (job == 1) {
                                rr += cr * b1r + ci * b1i; ri += cr * b1i - ci * b1r;
                            } else {
                                rr += cr * b1r - ci * b1i; ri += cr * b1i + ci * b1r;
                            }
                        }
                        BA(6)[j] = rr;
                        BA(7)[j] = ri;
                    }
                    __syncthreads();
                    if (w == 0 && lane < 2) {
                        const int f = lane == 0 ? fa : fb;
                        const bool flux = f == 1;
                        thomas_lds(BA(lane), BA(2 + 2 * lane), BA(3 + 2 * lane), g_bw0[f], BA(6 + lane), BA(6 + lane),
                                   Real(1), f != 2, !flux, flux, flux ? fcw : Real(0), M);
                    }
                    __syncthreads();
                    for (int j = tid; j < M; j += mBS) {
                        bn[static_cast<size_t>(sys0) * M + j] = BA(6)[j];
                        bn[static_cast<size_t>(sys0 + 1) * M + j] = BA(7)[j];
                    }
                    __syncthreads();
                }
                if (!(alive = grid_sync(bar, err, G, nbar, &s_ok))) break;
            }
            if (!alive) break;
        }

        // ================================================================ 7. store (Simulation.cpp:594-633) and push
        // the Picard jobs reused the region: W+, E_z and B_z are staged again for the push gather
        for (int j = tid; j < M; j += mBS) {
            for (int a = 0; a < (M1 ? 6 : 2); ++a) BA(SL::W + a)[j] = WS(W_WPR1 + a)[j];
            BA(SL::EZ)[j] = WS(W_EZ0)[j];
            BA(SL::BZ)[j] = WS(W_BZ0)[j];
            if (M1) {
                BA(SL::EZ + 1)[j] = WS(W_EZ1R)[j]; BA(SL::EZ + 2)[j] = WS(W_EZ1I)[j];
                BA(SL::BZ + 1)[j] = WS(W_BZ1R)[j]; BA(SL::BZ + 2)[j] = WS(W_BZ1I)[j];
            }
        }
        __syncthreads();
        {
            const Real* bfin = bpbuf[niter & 1];
            auto BF = [&](int f, int j) { return bfin[static_cast<size_t>(f) * M + j]; };
            Real* __restrict__ F = A.fld + static_cast<size_t>(k) * M * A.ncf;
            const int ncf = A.ncf;
            for (int j = gtid; j < M; j += gthreads) {
                Real* Fj = F + static_cast<size_t>(j) * ncf;
                const Real b1r = BF(0, j), b1i = BF(1, j);
                Fj[F_EP1R] = WS(W_WPR1)[j] + b1i;
                Fj[F_EP1I] = WS(W_WPI1)[j] - b1r;
                Fj[F_BP1R] = b1r;
                Fj[F_BP1I] = b1i;
                if (M1) {
                    const Real b0r = BF(2, j), b0i = BF(3, j), b2r = BF(4, j), b2i = BF(5, j);
                    Fj[F_EP0R] = WS(W_WPR0)[j] + b0i;
                    Fj[F_EP0I] = WS(W_WPI0)[j] - b0r;
                    Fj[F_EP2R] = WS(W_WPR2)[j] + b2i;
                    Fj[F_EP2I] = WS(W_WPI2)[j] - b2r;
                    Fj[F_BP0R] = b0r; Fj[F_BP0I] = b0i; Fj[F_BP2R] = b2r; Fj[F_BP2I] = b2i;
                }
                if (niter & 1)   // the warm start of the next slice goes back to ws
                    for (int f = 0; f < (M1 ? 6 : 2); ++f) WS(W_BPR1 + f)[j] = BF(f, j);
            }
            if (k < K - 1) {
                // ---- Adams-Bashforth push (PlasmaSpecies::push, src/PlasmaSpecies.cpp:600-667)
                const int slot = k % ord;
                const bool first = (k == 0);
                auto glin = [&](int f, int j, Real t) { return (Real(1) - t) * BF(f, j) + t * BF(f, j + 1); };
#pragma unroll
                for (int pp = 0; pp < PT; ++pp) {
                    const int j = cj[pp];
                    if (j < 0) continue;
                    int li;
                    const int gi = own0 + slot_of(pp);
                    const bool s1 = gi >= np0;
                    const qz_species_state& st = species_of(gi, li);
                    const int cap = st.cap;
                    Real* __restrict__ hist = st.hist;
                    const Real X = st.x[li], Y = st.y[li], D = st.dl[li];
                    const Real PX = st.px[li], PY = st.py[li];
                    const Real qm = s1 ? qm1 : qm0;
                    const Real t = ct[pp], c = cc[pp], sn = cs[pp];
                    Real Wx, Wy, Bx, By;
                    if (M1) {
                        eval_vector(lin(BA(SL::W + 2), j, t), lin(BA(SL::W + 3), j, t), lin(BA(SL::W + 0), j, t),
                                    lin(BA(SL::W + 1), j, t), lin(BA(SL::W + 4), j, t), lin(BA(SL::W + 5), j, t), c, sn,
                                    true, Wx, Wy);
                        eval_vector(glin(2, j, t), glin(3, j, t), glin(0, j, t), glin(1, j, t), glin(4, j, t),
                                    glin(5, j, t), c, sn, true, Bx, By);
                    } else {
                        eval_vector(Real(0), Real(0), lin(BA(SL::W + 0), j, t), lin(BA(SL::W + 1), j, t), Real(0),
                                    Real(0), c, sn, false, Wx, Wy);
                        eval_vector(Real(0), Real(0), glin(0, j, t), glin(1, j, t), Real(0), Real(0), c, sn, false, Bx,
                                    By);
                    }
                    const Real Ez = eval_scalar(lin(BA(SL::EZ), j, t), M1 ? lin(BA(SL::EZ + 1), j, t) : Real(0),
                                                M1 ? lin(BA(SL::EZ + 2), j, t) : Real(0), c, sn, M1);
                    const Real Bz = eval_scalar(lin(BA(SL::BZ), j, t), M1 ? lin(BA(SL::BZ + 1), j, t) : Real(0),
                                                M1 ? lin(BA(SL::BZ + 2), j, t) : Real(0), c, sn, M1);
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
                            const Real bb2 = H(sl, 2) * nx + H(sl, 3) * ny;
                            H(sl, 2) -= Real(2) * bb2 * nx; H(sl, 3) -= Real(2) * bb2 * ny;
                        }
                    }
                    st.x[li] = xn; st.y[li] = yn; st.px[li] = pxn; st.py[li] = pyn; st.dl[li] = dn;
                }
            }
        }
        __syncthreads();
    }
    // every workgroup has passed its last barrier: the last one out puts the barrier counter (word 0) and the
    // exit counter (word 1) back to zero, so the next launch on this workspace needs no reset of its own (the
    // error flag, word 2, stays)
    if (threadIdx.x == 0) {
        const unsigned int out = __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (out + 1 == static_cast<unsigned int>(G)) {
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(bar + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace qz
