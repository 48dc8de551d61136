// Beam kernels of the quasi-static step on gfx950: the push against the stored (xi, r, mode) fields
// (Beam::advance, src/Beam.cpp:948-1131) and the deposit of rho - J_z, J_z, J+ and rho
// (Beam::deposit, src/Beam.cpp:725-785).
//
// The deposit is deterministic and equals the reference's serial atomic_add order: particles are
// bucketed by their (xi, r) cell (counting sort, each bucket then ordered by particle index), and one
// thread per (slice, node) merges the index-ordered buckets of the four cells that touch its node and
// adds their contributions sequentially, in particle-index order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "qz_device.hpp"
#include "qz_internal.hpp"

namespace qz {

// per-particle record of the deposit (the node-independent factors of src/Beam.cpp:738-757)
enum BRec : int { R_WQ = 0, R_TK, R_T, R_OMVZ, R_VZ, R_VR, R_VT, R_C, R_S, R_VX, R_VY, R_N };

// count: every active lane adds `add` (0 or 1) to *ctr, with one atomic per wave instead of one per lane
__device__ __forceinline__ void wave_count(unsigned long long* __restrict__ ctr, bool add) {
    const unsigned long long m = __ballot(add ? 1 : 0);
    const unsigned long long act = __ballot(1);
    if (static_cast<int>(__lane_id()) == __ffsll(static_cast<long long>(act)) - 1 && m)
        atomicAdd(ctr, static_cast<unsigned long long>(__popcll(m)));
}

// With pstep, only the particles whose gather slice k (after clamping) lies in [s_lo, s_hi) and that are not
// yet marked with `step` are pushed, and they are marked: a step pushed slice range by slice range pushes
// every particle once, also those that slip into a later range.
__global__ void beam_advance_kernel(qz_radial_grid g, qz_beam_grid bg, int m1, Real qm, int mode, int i0, int np,
                                    qz_beam_soa P, const Real* __restrict__ fld, int nc, Real dt_kick, Real dt,
                                    int move, int s_lo, int s_hi, int step, int* __restrict__ pstep,
                                    unsigned long long* __restrict__ nmarked) {
    const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < np && P.w[i] != Real(0);
    if (!live) return;
    const Real xmin = bg.xi_min, idxi = Real(1) / bg.dxi;
    const Real xmax = bg.xi_min + (bg.nxi - 1) * bg.dxi;
    const int k0 = bg.k0, nloc = bg.nloc;
    const bool ngp = bg.ngp != 0;
    const Real R = g.R;
    const int M = g.N + 1;
    const Real X = P.x[i], Y = P.y[i];
    const Real r = dsqrt(X * X + Y * Y);
    const Real c = r > Real(0) ? X / r : Real(1), sn = r > Real(0) ? Y / r : Real(0);
    Real s = (P.xi[i] - xmin) * idxi - k0;
    int k;
    Real tk;
    if (ngp) {
        k = static_cast<int>(floor(s + Real(0.5)));
        k = k < 0 ? 0 : (k > nloc - 1 ? nloc - 1 : k);
        tk = Real(0);
    } else {
        k = static_cast<int>(floor(s));
        if (k < 0) { k = 0; s = 0; }
        if (k > nloc - 2) { k = nloc - 2; s = nloc - 1; }
        tk = s - k;
    }
    if (pstep) {
        // markers grow with the step: a particle a later step has already pushed (a slow one can slip past
        // the ranges this step has still to push) is not pushed again
        if (k < s_lo || k >= s_hi || pstep[i] >= step) return;
        pstep[i] = step;
        // the live particles this step pushes, each once: Simulation::step's count before the push
        if (nmarked) wave_count(nmarked, true);
    }
    const int k1 = ngp ? k : k + 1;
    int j; Real t;
    grid_weights(g.r, g.inv_h, g.N, R, r, j, t);
    const Real* F0 = fld + (static_cast<size_t>(k) * M + j) * nc;
    const Real* F1 = fld + (static_cast<size_t>(k1) * M + j) * nc;
    auto bil = [&](int comp) {
        return (Real(1) - tk) * ((Real(1) - t) * F0[comp] + t * F0[nc + comp]) +
               tk * ((Real(1) - t) * F1[comp] + t * F1[nc + comp]);
    };
    Real Ex, Ey, Bx, By, Ez, Bz;
    if (m1) {
        eval_vector(bil(F_EP0R), bil(F_EP0I), bil(F_EP1R), bil(F_EP1I), bil(F_EP2R), bil(F_EP2I), c, sn, true, Ex, Ey);
        eval_vector(bil(F_BP0R), bil(F_BP0I), bil(F_BP1R), bil(F_BP1I), bil(F_BP2R), bil(F_BP2I), c, sn, true, Bx, By);
        Ez = eval_scalar(bil(F_EZ0), bil(F_EZ1R), bil(F_EZ1I), c, sn, true);
        Bz = eval_scalar(bil(F_BZ0), bil(F_BZ1R), bil(F_BZ1I), c, sn, true);
    } else {
        eval_vector(Real(0), Real(0), bil(F_EP1R), bil(F_EP1I), Real(0), Real(0), c, sn, false, Ex, Ey);
        eval_vector(Real(0), Real(0), bil(F_BP1R), bil(F_BP1I), Real(0), Real(0), c, sn, false, Bx, By);
        Ez = bil(F_EZ0);
        Bz = bil(F_BZ0);
    }
    const V3 E{Ex, Ey, Ez}, B{Bx, By, Bz};
    V3 u{P.px[i], P.py[i], P.pz[i]};
    const Real h = Real(0.5) * qm * dt_kick;
    const V3 Fp{Real(0), Real(0), Real(0)};   // no laser on this path
    if (!move) {
        const Real g0 = dsqrt(Real(1) + dot(u, u));
        u = u + (Real(2) * h) * (E + cross((Real(1) / g0) * u, B)) + dt_kick * Fp;
        P.px[i] = u.x; P.py[i] = u.y; P.pz[i] = u.z;
        return;
    }
    u = u + (Real(0.5) * dt_kick) * Fp;
    u = push_u(mode, u, E, B, h);
    u = u + (Real(0.5) * dt_kick) * Fp;
    const Real p2 = u.x * u.x + u.y * u.y;
    const Real gam = dsqrt(Real(1) + p2 + u.z * u.z);
    const Real omvz = (Real(1) + p2) / (gam * (gam + u.z));
    const Real xn = X + u.x / gam * dt, yn = Y + u.y / gam * dt;
    P.x[i] = xn; P.y[i] = yn;
    P.px[i] = u.x; P.py[i] = u.y; P.pz[i] = u.z;
    const Real xin = P.xi[i] + omvz * dt;
    P.xi[i] = xin;
    if (xn * xn + yn * yn >= R * R || xin < xmin || xin > xmax) P.w[i] = Real(0);
}

// live particles (w != 0) of a beam, added to *count
// zero fill: 16-byte stores over n16 chunks, then 4-byte stores over the ntail words after them
__global__ void __launch_bounds__(256) zero_kernel(uint4* __restrict__ p, size_t n16, unsigned int* __restrict__ tail,
                                                   int ntail) {
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && static_cast<int>(threadIdx.x) < ntail) tail[threadIdx.x] = 0u;
}

// live particles (w != 0): a grid-stride count per thread, one atomic per block of 256 threads
__global__ void __launch_bounds__(256) beam_count_live_kernel(int np, const Real* __restrict__ w,
                                                              unsigned long long* __restrict__ count) {
    __shared__ unsigned int part[4];
    unsigned int c = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < np; i += gridDim.x * 256) c += w[i] != Real(0) ? 1u : 0u;
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long s = static_cast<unsigned long long>(part[0]) + part[1] + part[2] + part[3];
        if (s) atomicAdd(count, s);
    }
}

// deposit step 1: cell key, per-particle record and bucket counts
// i0: the particle's index in the deposit's common index space (beams after one another).  pstep (optional):
// particles whose marker is `marker` or later are left out (a pipelined step's deposit skips the particles
// that step, or a later one, has already pushed)
// live (optional): live particles whose lower slice lies in the window's rows [k0, k0 + nloc), those before the
// box in the box's first window and those after it in its last -- each live particle once over a step's windows,
// the count Simulation::step takes before its deposit.
__global__ void beam_key_kernel(qz_radial_grid g, qz_beam_grid bg, Real q, int np, qz_beam_soa P, int i0,
                                const int* __restrict__ pstep, int marker, int* __restrict__ key,
                                Real* __restrict__ rec, int* __restrict__ counts, unsigned long long* __restrict__ live) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int gi = i0 + i;
    key[gi] = -1;
    // the slice window first: a deposit restricted to a few rows reads only xi of the other particles
    const Real xmin = bg.xi_min, idxi = Real(1) / bg.dxi;
    const int k0 = bg.k0, nloc = bg.nloc;
    const bool ngp = bg.ngp != 0;
    const Real sx = (P.xi[i] - xmin) * idxi;
    const int k = ngp ? static_cast<int>(floor(sx + Real(0.5))) : static_cast<int>(floor(sx));
    if (live) {   // the window's share of the live count: particles before (after) the box go to the first (last)
        const bool first = k0 == 0, last = k0 + nloc >= bg.nxi;
        wave_count(live, (k >= k0 || first) && (k < k0 + nloc || last) && !(pstep && pstep[i] >= marker) &&
                             P.w[i] != Real(0));
    }
    if (k < k0 - 1 || k >= k0 + nloc) return;
    if (pstep && pstep[i] >= marker) return;
    const Real wi = P.w[i];
    if (wi == Real(0)) return;
    const Real X = P.x[i], Y = P.y[i];
    const Real r = dsqrt(X * X + Y * Y);
    if (r > g.R) return;
    const Real c = r > Real(0) ? X / r : Real(1), s = r > Real(0) ? Y / r : Real(0);
    const Real tk = ngp ? Real(0) : sx - k;
    const Real px = P.px[i], py = P.py[i], pz = P.pz[i];
    const Real p2 = px * px + py * py;
    const Real gam = dsqrt(Real(1) + p2 + pz * pz);
    const Real omvz = (Real(1) + p2) / (gam * (gam + pz));
    const Real vz = pz / gam, vx = px / gam, vy = py / gam;
    const Real vr = c * vx + s * vy, vt = c * vy - s * vx;
    int j; Real t;
    grid_weights(g.r, g.inv_h, g.N, g.R, r, j, t);
    const int kl = k - k0;   // local slice, -1 .. nloc-1
    const int kk = (kl + 1) * g.N + j;
    key[gi] = kk;
    Real* rr = rec + static_cast<size_t>(gi) * R_N;
    rr[R_WQ] = wi * q; rr[R_TK] = tk; rr[R_T] = t; rr[R_OMVZ] = omvz; rr[R_VZ] = vz; rr[R_VR] = vr; rr[R_VT] = vt;
    rr[R_C] = c; rr[R_S] = s; rr[R_VX] = vx; rr[R_VY] = vy;
    atomicAdd(&counts[kk], 1);
}

// deposit steps 2-3: the deposited particles' indices grouped by bucket, each bucket in particle-index order --
// the order of the serial deposit's sums.  A stable least-significant-digit radix sort of the (bucket key,
// particle index) pairs: the pairs are first compacted in index order, then every pass is a stable counting
// sort of one 8-bit digit of the key, so pairs with equal keys keep their index order by construction.
// Blocks of kRThreads threads own tiles of kRTile consecutive pairs; within a tile, ranks come from wave
// ballots over the digit's bits and a prefix over the waves of each 256-pair chunk.
constexpr int kRThreads = 256;
constexpr int kRTile = 4096;
constexpr int kRWaves = kRThreads / 64;

// rank of this lane among the lanes below it with the same `digit` (valid lanes only), and how many lanes of the
// wave share it; nbits: the digit's width
__device__ __forceinline__ void wave_match_rank(int digit, bool valid, int nbits, int& rank, int& count) {
    const int lane = threadIdx.x & 63;
    unsigned long long mask = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
        const bool bit = valid && ((digit >> b) & 1);
        const unsigned long long bb = __ballot(bit);
        mask &= bit ? bb : ~bb;
    }
    const unsigned long long lt = (1ull << lane) - 1ull;
    rank = valid ? __popcll(mask & lt) : 0;
    count = valid ? __popcll(mask) : 0;
}

// per tile: the number of deposited particles (key >= 0)
__global__ void __launch_bounds__(kRThreads) compact_count_kernel(int np, const int* __restrict__ key,
                                                                  int* __restrict__ tile_count) {
    __shared__ int sh[kRWaves];
    const int t = threadIdx.x, w = t >> 6;
    const long base = static_cast<long>(blockIdx.x) * kRTile;
    int c = 0;
    for (int i = t; i < kRTile; i += kRThreads) {
        const long e = base + i;
        c += (e < np && key[e] >= 0) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if ((t & 63) == 0) sh[w] = c;
    __syncthreads();
    if (t == 0) {
        int tot = 0;
        for (int q = 0; q < kRWaves; ++q) tot += sh[q];
        tile_count[blockIdx.x] = tot;
    }
}

// the (key, index) pairs of the deposited particles, in index order, at tile_off[tile] + rank in the tile
__global__ void __launch_bounds__(kRThreads) compact_write_kernel(int np, const int* __restrict__ key,
                                                                  const int* __restrict__ tile_off,
                                                                  int* __restrict__ kout, int* __restrict__ vout) {
    __shared__ int wc[kRWaves];
    __shared__ int run;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const long base = static_cast<long>(blockIdx.x) * kRTile;
    if (t == 0) run = tile_off[blockIdx.x];
    __syncthreads();
    for (int c0 = 0; c0 < kRTile; c0 += kRThreads) {
        const long e = base + c0 + t;
        const int kk = e < np ? key[e] : -1;
        const bool v = kk >= 0;
        const unsigned long long m = __ballot(v);
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wc[w] = __popcll(m);
        __syncthreads();
        int pre = run;
        for (int q = 0; q < w; ++q) pre += wc[q];
        if (v) {
            kout[pre + rank] = kk;
            vout[pre + rank] = static_cast<int>(e);
        }
        __syncthreads();
        if (t == 0) {
            int tot = 0;
            for (int q = 0; q < kRWaves; ++q) tot += wc[q];
            run += tot;
        }
        __syncthreads();
    }
}

// one radix pass, step 1: digit counts per tile, stored digit-major (hist[d * ntiles + tile])
__global__ void __launch_bounds__(kRThreads) radix_hist_kernel(const int* __restrict__ nvalid,
                                                               const int* __restrict__ kin, int shift, int ntiles,
                                                               int* __restrict__ hist) {
    __shared__ int h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int n = *nvalid;
    const long base = static_cast<long>(blockIdx.x) * kRTile;
    for (int i = t; i < kRTile; i += kRThreads) {
        const long e = base + i;
        if (e < n) atomicAdd(&h[(kin[e] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[static_cast<size_t>(t) * ntiles + blockIdx.x] = h[t];
}

// one radix pass, step 2 (after the exclusive scan of hist into off): the stable scatter
__global__ void __launch_bounds__(kRThreads) radix_scatter_kernel(const int* __restrict__ nvalid,
                                                                  const int* __restrict__ kin,
                                                                  const int* __restrict__ vin, int shift, int ntiles,
                                                                  const int* __restrict__ off,
                                                                  int* __restrict__ kout, int* __restrict__ vout) {
    __shared__ int running[256];
    __shared__ int wc[kRWaves][256];
    const int t = threadIdx.x, w = t >> 6;
    running[t] = off[static_cast<size_t>(t) * ntiles + blockIdx.x];
    for (int q = 0; q < kRWaves; ++q) wc[q][t] = 0;
    __syncthreads();
    const int n = *nvalid;
    const long base = static_cast<long>(blockIdx.x) * kRTile;
    for (int c0 = 0; c0 < kRTile; c0 += kRThreads) {
        if (base + c0 >= n) break;   // uniform over the block
        const long e = base + c0 + t;
        const bool v = e < n;
        const int kk = v ? kin[e] : 0;
        const int d = (kk >> shift) & 255;
        int rank, cnt;
        wave_match_rank(d, v, 8, rank, cnt);
        if (v && rank == 0) wc[w][d] = cnt;
        __syncthreads();
        if (v) {
            int pos = running[d] + rank;
            for (int q = 0; q < w; ++q) pos += wc[q][d];
            kout[pos] = kk;
            vout[pos] = vin[e];
        }
        __syncthreads();
        int add = 0;
        for (int q = 0; q < kRWaves; ++q) {
            add += wc[q][t];
            wc[q][t] = 0;
        }
        running[t] += add;
        __syncthreads();
    }
}

// deposit step 4: one thread per (local slice, node): the four buckets touching the node, merged by index
__global__ void beam_node_kernel(qz_radial_grid g, qz_beam_grid bg, int m1, const int* __restrict__ start,
                                 const int* __restrict__ list, const Real* __restrict__ rec, Real* __restrict__ src, int nc,
                                 int overwrite) {
    const int M = g.N + 1, N = g.N;
    const long node = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= static_cast<long>(bg.nloc) * M) return;
    const int kk = static_cast<int>(node / M), jj = static_cast<int>(node % M);
    const bool ngp = bg.ngp != 0;
    const Real idxi = Real(1) / bg.dxi;
    // buckets q = 2 a + b (a = slice offset, b = node offset of the particle's corner): cell (kk - a, jj - b).  A
    // particle sits in one of them at most, so the merge by index has no ties.  The head of each bucket is kept
    // in a register, and the advanced bucket's next head is loaded while the chosen particle's record is read.
    constexpr int kEnd = 0x7fffffff;
    int lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0, lo2 = 0, hi2 = 0, lo3 = 0, hi3 = 0;
    auto bucket = [&](int a, int b, int& lo, int& hi) {
        const int kc = kk - a, jc = jj - b;
        if ((a == 1 && ngp) || jc < 0 || jc >= N) return;
        const int key = (kc + 1) * N + jc;   // kc >= -1
        lo = start[key];
        hi = start[key + 1];
    };
    bucket(0, 0, lo0, hi0);
    bucket(0, 1, lo1, hi1);
    bucket(1, 0, lo2, hi2);
    bucket(1, 1, lo3, hi3);
    Real* S = src + static_cast<size_t>(node) * nc;
    if (lo0 >= hi0 && lo1 >= hi1 && lo2 >= hi2 && lo3 >= hi3) {   // no particle: the node keeps what src holds,
        if (overwrite)                                               // or is zero when the deposit overwrites
            for (int c = 0; c < nc; ++c) S[c] = Real(0);
        return;
    }
    int h0 = lo0 < hi0 ? list[lo0] : kEnd, h1 = lo1 < hi1 ? list[lo1] : kEnd;
    int h2 = lo2 < hi2 ? list[lo2] : kEnd, h3 = lo3 < hi3 ? list[lo3] : kEnd;
    Real acc[15];   // from zero when overwriting: +0 + x is x, as the zero-filled source plus x
    for (int c = 0; c < nc; ++c) acc[c] = overwrite ? Real(0) : S[c];
    const Real iv = g.inv_V[jj];
    while (true) {
        // next particle in index order among the buckets
        int best = 0, bi = h0;
        if (h1 < bi) { bi = h1; best = 1; }
        if (h2 < bi) { bi = h2; best = 2; }
        if (h3 < bi) { bi = h3; best = 3; }
        if (bi == kEnd) break;
        switch (best) {
            case 0: ++lo0; h0 = lo0 < hi0 ? list[lo0] : kEnd; break;
            case 1: ++lo1; h1 = lo1 < hi1 ? list[lo1] : kEnd; break;
            case 2: ++lo2; h2 = lo2 < hi2 ? list[lo2] : kEnd; break;
            default: ++lo3; h3 = lo3 < hi3 ? list[lo3] : kEnd; break;
        }
        const Real* rr = rec + static_cast<size_t>(bi) * R_N;
        const Real tk = rr[R_TK], t = rr[R_T];
        const Real Wk = (best >> 1) == 0 ? Real(1) - tk : tk;
        const Real Wr = (best & 1) == 0 ? Real(1) - t : t;
        const Real d = rr[R_WQ] * Wk * Wr * iv * idxi;
        const Real omvz = rr[R_OMVZ], vz = rr[R_VZ];
        acc[B_RT0] += d * omvz;
        acc[B_JZ0] += d * vz;
        acc[B_JP1R] += d * rr[R_VR];
        acc[B_JP1I] += d * rr[R_VT];
        acc[B_RHO0] += d;
        if (m1) {
            const Real c = rr[R_C], s = rr[R_S], vx = rr[R_VX], vy = rr[R_VY];
            acc[B_RT1R] += d * omvz * c;
            acc[B_RT1I] += -d * omvz * s;
            acc[B_JZ1R] += d * vz * c;
            acc[B_JZ1I] += -d * vz * s;
            acc[B_RHO1R] += d * c;
            acc[B_RHO1I] += -d * s;
            acc[B_JP0R] += d * vx;
            acc[B_JP0I] += d * vy;
            const Real c2 = c * c - s * s, s2 = Real(2) * c * s;
            acc[B_JP2R] += d * (vx * c2 + vy * s2);
            acc[B_JP2I] += d * (vy * c2 - vx * s2);
        }
    }
    for (int c = 0; c < nc; ++c) S[c] = acc[c];
}

// ---- exclusive scan of n ints (counts -> starts, start[n] = total): per-block sums, one-block scan, add
__global__ void scan_block_sums(const int* __restrict__ in, int n, int* __restrict__ bsum) {
    __shared__ int sh[1024];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    sh[threadIdx.x] = i < n ? in[i] : 0;
    __syncthreads();
    for (int d = 512; d > 0; d >>= 1) {
        if (threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}
__global__ void scan_sums_serial(int* __restrict__ bsum, int nb, int* __restrict__ total) {
    // one block: exclusive scan of the block sums by 1024 threads in chunks
    __shared__ int sh[1024];
    int carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int t = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
        const int blocktotal = sh[1023];
        __syncthreads();
        carry += blocktotal;
    }
    if (threadIdx.x == 0) *total = carry;
}
__global__ void scan_finish(const int* __restrict__ in, int n, const int* __restrict__ bsum, int* __restrict__ out) {
    __shared__ int sh[1024];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int t = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = bsum[blockIdx.x] + sh[threadIdx.x] - v;
}

// ---- the same exclusive scan in one block when n <= 1024 * kScanOne (a deposit restricted to a few rows): each
// thread sums a contiguous chunk, the block scans the chunk sums, each thread writes its chunk's prefixes
constexpr int kScanOne = 64;
__global__ void __launch_bounds__(1024) scan_one_block(const int* __restrict__ in, int n, int* __restrict__ out) {
    __shared__ int sh[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(b + per, n);
    int s = 0;
    for (int i = b; i < e; ++i) s += in[i];
    sh[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int run = sh[t] - s;
    for (int i = b; i < e; ++i) {
        const int v = in[i];
        out[i] = run;
        run += v;
    }
}

}  // namespace qz

using namespace qz;

namespace {
struct DepWork {   // layout of the deposit workspace
    size_t key, rec, counts, start, list, bsum, total, kc, kt, vt, tcount, toff, hist, hoff, hbsum, htotal, bytes;
    int ntiles;
};
DepWork dep_layout(int np, int nloc, int N) {
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t nc = static_cast<size_t>(nloc + 1) * N;
    const size_t nb = (nc + 1 + 1023) / 1024;
    DepWork d;
    d.ntiles = static_cast<int>((static_cast<size_t>(np) + kRTile - 1) / kRTile);
    const size_t nt = static_cast<size_t>(d.ntiles > 0 ? d.ntiles : 1);
    const size_t nh = 256 * nt + 1;
    size_t o = 0;
    d.key = o; o = al(o + 4 * static_cast<size_t>(np));
    d.rec = o; o = al(o + 8 * static_cast<size_t>(R_N) * np);
    d.counts = o; o = al(o + 4 * (nc + 1));
    d.start = o; o = al(o + 4 * (nc + 1));
    d.list = o; o = al(o + 4 * static_cast<size_t>(np));
    d.bsum = o; o = al(o + 4 * nb);
    d.total = o; o = al(o + 4);
    d.kc = o; o = al(o + 4 * static_cast<size_t>(np));    // radix sort: keys, two buffers
    d.kt = o; o = al(o + 4 * static_cast<size_t>(np));
    d.vt = o; o = al(o + 4 * static_cast<size_t>(np));    // the second value buffer (the first is `list`)
    d.tcount = o; o = al(o + 4 * (nt + 1));
    d.toff = o; o = al(o + 4 * (nt + 1));
    d.hist = o; o = al(o + 4 * nh);
    d.hoff = o; o = al(o + 4 * nh);
    d.hbsum = o; o = al(o + 4 * ((std::max(nh, nt + 1) + 1023) / 1024 + 1));
    d.htotal = o; o = al(o + 4);
    d.bytes = o;
    return d;
}

// exclusive scan of n ints (one block for short arrays, three launches otherwise); bsum: (n + 1023) / 1024 ints
void scan_exclusive(const int* in, int n, int* out, int* bsum, int* total, hipStream_t st) {
    const int nb = (n + 1023) / 1024;
    if (nb <= kScanOne) {
        hipLaunchKernelGGL(scan_one_block, dim3(1), dim3(1024), 0, st, in, n, out);
    } else {
        hipLaunchKernelGGL(scan_block_sums, dim3(nb), dim3(1024), 0, st, in, n, bsum);
        hipLaunchKernelGGL(scan_sums_serial, dim3(1), dim3(1024), 0, st, bsum, nb, total);
        hipLaunchKernelGGL(scan_finish, dim3(nb), dim3(1024), 0, st, in, n, bsum, out);
    }
}
}  // namespace

// Threads per block of the beam kernels (QZ_K_*), chosen by the autotuner (qz_sim_autotune) or set with
// qz_set_block; 256 until then.  One table for the process.
static int g_block[QZ_K_COUNT] = {256, 256, 256, 256};

extern "C" int qz_set_block(int kernel, int threads) {
    if (kernel < 0 || kernel >= QZ_K_COUNT) return qz_fail(QZ_EINVAL, "qz_set_block: unknown kernel");
    if (threads < 64 || threads > 1024 || threads % 64 != 0)
        return qz_fail(QZ_EINVAL, "qz_set_block: threads must be a multiple of 64 in [64, 1024]");
    g_block[kernel] = threads;
    return QZ_OK;
}

extern "C" int qz_get_block(int kernel) { return (kernel < 0 || kernel >= QZ_K_COUNT) ? -1 : g_block[kernel]; }

int qz_zero_async(void* p, size_t bytes, hipStream_t st) {
    if (bytes == 0) return QZ_OK;
    if (!p || bytes % 4 != 0) return qz_fail(QZ_EINVAL, "zero fill: null pointer or size not a multiple of 4");
    char* c = static_cast<char*>(p);
    size_t head = (16 - reinterpret_cast<uintptr_t>(c) % 16) % 16;   // bytes before the first 16-byte boundary
    if (head > bytes) head = bytes;
    if (head) {   // at most 3 words: a separate tiny launch keeps the main kernel aligned
        hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, st, static_cast<uint4*>(nullptr), size_t(0),
                           reinterpret_cast<unsigned int*>(c), static_cast<int>(head / 4));
        QZ_HIP(hipGetLastError());
    }
    const size_t body = bytes - head;
    const size_t n16 = body / 16;
    const int ntail = static_cast<int>((body % 16) / 4);
    const size_t nb = std::min<size_t>(std::max<size_t>((n16 + 255) / 256, 1), 2048);
    hipLaunchKernelGGL(zero_kernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, st, reinterpret_cast<uint4*>(c + head),
                       n16, reinterpret_cast<unsigned int*>(c + head + 16 * n16), ntail);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

// One output component of the fields_*.bin writer: out[r][j] = s0 a[k][j][c0] + s1 a[k][j][c1] with k = K-1-r
// (the file lists xi increasing, the head last).  s0, s1 are 0, +-1 or +-2, so every product is exact and the
// value equals the host expression (one term, 2 x, -2 x, x + y or x - y) bit for bit.
__global__ void field_component_kernel(const Real* __restrict__ a, int nc, int K, int M, int c0, Real s0, int c1,
                                       Real s1, Real* __restrict__ out) {
    const long n = static_cast<long>(K) * M;
    for (long e = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += static_cast<long>(gridDim.x) * blockDim.x) {
        const int r = static_cast<int>(e / M), j = static_cast<int>(e % M);
        const Real* A = a + (static_cast<size_t>(K - 1 - r) * M + j) * nc;
        Real v = s0 * A[c0];
        if (s1 != Real(0)) v = v + s1 * A[c1];
        out[e] = v;
    }
}

int qz_field_component(const double* a, int nc, int K, int M, int c0, double s0, int c1, double s1, double* out,
                       hipStream_t st) {
    if (!a || !out || nc < 1 || K < 1 || M < 1 || c0 < 0 || c0 >= nc || c1 < 0 || c1 >= nc)
        return qz_fail(QZ_EINVAL, "field component: bad argument");
    const long n = static_cast<long>(K) * M;
    const unsigned nb = static_cast<unsigned>(std::min<long>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(field_component_kernel, dim3(nb), dim3(256), 0, st, a, nc, K, M, c0, s0, c1, s1, out);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

int qz_beam_count_live(int np, const double* w, unsigned long long* count, hipStream_t st) {
    if (np <= 0) return QZ_OK;
    const int nb = std::min((np + 255) / 256, 512);
    hipLaunchKernelGGL(beam_count_live_kernel, dim3(nb), dim3(256), 0, st, np, w, count);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

extern "C" size_t qz_beam_deposit_workspace_bytes(int np, int nloc, int N) {
    return dep_layout(np, nloc, N).bytes;
}

// The deposit of nbeams beams into one bsrc, their particles indexed one beam after another: every node sums
// the first beam's contributions in particle order, then the next beam's, as nbeams deposits one after another do.
static int deposit_beams(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, int nbeams, const double* q,
                         const int* np, const qz_beam_soa* p, double* bsrc, int nc, void* work,
                         const int* const* pstep, int marker, unsigned long long* live, hipStream_t st,
                         bool overwrite = false) {
    long ntot = 0;
    for (int b = 0; b < nbeams; ++b) {
        if (np[b] < 0) return qz_fail(QZ_EINVAL, "qz_beam_deposit: bad extent");
        ntot += np[b];
    }
    if (ntot > 0x7fffffffL) return qz_fail(QZ_ERANGE, "qz_beam_deposit: more than 2^31 - 1 particles");
    if (!grid || !box || !bsrc || (ntot > 0 && !work)) return qz_fail(QZ_EINVAL, "qz_beam_deposit: null argument");
    if (grid->N < 2 || box->nloc < 1) return qz_fail(QZ_EINVAL, "qz_beam_deposit: bad extent");
    if (nc != (modes == 1 ? B_NC1 : B_NC0)) return qz_fail(QZ_EINVAL, "qz_beam_deposit: nc must be 5 (m = 0) or 15 (m = 1)");
    if (ntot == 0) return overwrite ? qz_zero_async(bsrc, sizeof(Real) * static_cast<size_t>(box->nloc) *
                                                          (grid->N + 1) * nc, st)
                                    : QZ_OK;
    const int npt = static_cast<int>(ntot);
    const DepWork L = dep_layout(npt, box->nloc, grid->N);
    char* w = static_cast<char*>(work);
    int* key = reinterpret_cast<int*>(w + L.key);
    Real* rec = reinterpret_cast<Real*>(w + L.rec);
    int* counts = reinterpret_cast<int*>(w + L.counts);
    int* start = reinterpret_cast<int*>(w + L.start);
    int* list = reinterpret_cast<int*>(w + L.list);
    int* bsum = reinterpret_cast<int*>(w + L.bsum);
    int* total = reinterpret_cast<int*>(w + L.total);
    const int ncells = (box->nloc + 1) * grid->N;
    {
        const int rc = qz_zero_async(counts, 4 * (static_cast<size_t>(ncells) + 1), st);
        if (rc) return rc;
    }
    const int tb = g_block[QZ_K_BEAM_PARTICLES], tnode = g_block[QZ_K_BEAM_NODES];
    for (int b = 0, i0 = 0; b < nbeams; i0 += np[b], ++b)
        if (np[b] > 0)
            hipLaunchKernelGGL(beam_key_kernel, dim3((np[b] + tb - 1) / tb), dim3(tb), 0, st, *grid, *box, q[b], np[b], p[b],
                               i0, pstep ? pstep[b] : nullptr, marker, key, rec, counts, live);
    const int n1 = ncells + 1;   // counts[ncells] = 0 -> start[ncells] = total
    scan_exclusive(counts, n1, start, bsum, total, st);
    // the bucket lists, each in particle-index order: compaction, then LSD radix passes over the key's digits
    {
        const int nt = L.ntiles;
        int* kc = reinterpret_cast<int*>(w + L.kc);
        int* kt = reinterpret_cast<int*>(w + L.kt);
        int* vt = reinterpret_cast<int*>(w + L.vt);
        int* tcount = reinterpret_cast<int*>(w + L.tcount);
        int* toff = reinterpret_cast<int*>(w + L.toff);
        int* hist = reinterpret_cast<int*>(w + L.hist);
        int* hoff = reinterpret_cast<int*>(w + L.hoff);
        int* hbsum = reinterpret_cast<int*>(w + L.hbsum);
        int* htotal = reinterpret_cast<int*>(w + L.htotal);
        int bits = 0;
        while ((1 << bits) < ncells) ++bits;
        const int passes = (bits + 7) / 8;
        // the values end in `list` after the last pass
        int* vA = passes % 2 == 0 ? list : vt;
        int* vB = passes % 2 == 0 ? vt : list;
        hipLaunchKernelGGL(compact_count_kernel, dim3(nt), dim3(kRThreads), 0, st, npt, key, tcount);
        scan_exclusive(tcount, nt, toff, hbsum, htotal, st);
        hipLaunchKernelGGL(compact_write_kernel, dim3(nt), dim3(kRThreads), 0, st, npt, key, toff, kc, vA);
        const int* nvalid = start + ncells;
        int *kin = kc, *kout = kt, *vin = vA, *vout = vB;
        for (int p = 0; p < passes; ++p) {
            hipLaunchKernelGGL(radix_hist_kernel, dim3(nt), dim3(kRThreads), 0, st, nvalid, kin, 8 * p, nt, hist);
            scan_exclusive(hist, 256 * nt, hoff, hbsum, htotal, st);
            hipLaunchKernelGGL(radix_scatter_kernel, dim3(nt), dim3(kRThreads), 0, st, nvalid, kin, vin, 8 * p, nt,
                               hoff, kout, vout);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
    }
    const long nodes = static_cast<long>(box->nloc) * (grid->N + 1);
    hipLaunchKernelGGL(beam_node_kernel, dim3(static_cast<unsigned>((nodes + tnode - 1) / tnode)), dim3(tnode), 0, st, *grid, *box,
                       modes == 1 ? 1 : 0, start, list, rec, bsrc, nc, overwrite ? 1 : 0);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

int qz_beam_deposit_window(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, int nbeams,
                           const double* q, const int* np, const qz_beam_soa* p, double* bsrc, int nc, void* work,
                           const int* const* pstep, int marker, unsigned long long* live_count, hipStream_t st) {
    if (nbeams < 1 || !q || !np || !p) return qz_fail(QZ_EINVAL, "qz_beam_deposit_window: null argument or no beam");
    return deposit_beams(grid, box, modes, nbeams, q, np, p, bsrc, nc, work, pstep, marker, live_count, st, true);
}

extern "C" int qz_beam_deposit(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, double q, int np,
                               const qz_beam_soa* p, double* bsrc, int nc, void* work, void* stream) {
    if (!p) return qz_fail(QZ_EINVAL, "qz_beam_deposit: null argument");
    return deposit_beams(grid, box, modes, 1, &q, &np, p, bsrc, nc, work, nullptr, 0, nullptr,
                         static_cast<hipStream_t>(stream));
}

extern "C" int qz_beam_deposit_beams(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, int nbeams,
                                     const double* q, const int* np, const qz_beam_soa* p, double* bsrc, int nc,
                                     void* work, const int* const* pstep, int marker, unsigned long long* live_count,
                                     void* stream) {
    if (nbeams < 1 || !q || !np || !p) return qz_fail(QZ_EINVAL, "qz_beam_deposit_beams: null argument or no beam");
    return deposit_beams(grid, box, modes, nbeams, q, np, p, bsrc, nc, work, pstep, marker, live_count,
                         static_cast<hipStream_t>(stream));
}

extern "C" int qz_beam_advance(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, double qm, int pusher,
                               int i0, int np, const qz_beam_soa* p, const double* fld, int nc, double dt_kick,
                               double dt, int move, void* stream) {
    if (!grid || !box || !p || !fld) return qz_fail(QZ_EINVAL, "qz_beam_advance: null argument");
    if (np < 0 || i0 < 0 || grid->N < 2) return qz_fail(QZ_EINVAL, "qz_beam_advance: bad extent");
    if (!box->ngp && box->nloc < 2) return qz_fail(QZ_EINVAL, "qz_beam_advance: linear shape needs >= 2 slices");
    if (nc != (modes == 1 ? F_NC1 : F_NC0)) return qz_fail(QZ_EINVAL, "qz_beam_advance: nc must be 6 (m = 0) or 18 (m = 1)");
    if (pusher < 0 || pusher > 3) return qz_fail(QZ_EINVAL, "qz_beam_advance: pusher must be 0 vay, 1 boris, 2 hc, 3 imp");
    if (np - i0 <= 0) return QZ_OK;
    const int tb = g_block[QZ_K_BEAM_ADVANCE];
    hipLaunchKernelGGL(beam_advance_kernel, dim3((np - i0 + tb - 1) / tb), dim3(tb), 0, static_cast<hipStream_t>(stream),
                       *grid, *box, modes == 1 ? 1 : 0, qm, pusher, i0, np, *p, fld, nc, dt_kick, dt, move, 0, 0, 0,
                       static_cast<int*>(nullptr), static_cast<unsigned long long*>(nullptr));
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

extern "C" int qz_beam_advance_slices(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, double qm,
                                      int pusher, int np, const qz_beam_soa* p, const double* fld, int nc,
                                      double dt_kick, double dt, int move, int s_lo, int s_hi, int step, int* pstep,
                                      unsigned long long* pushed_count, void* stream) {
    if (!grid || !box || !p || !fld || !pstep) return qz_fail(QZ_EINVAL, "qz_beam_advance_slices: null argument");
    if (np < 0 || grid->N < 2) return qz_fail(QZ_EINVAL, "qz_beam_advance_slices: bad extent");
    if (!box->ngp && box->nloc < 2) return qz_fail(QZ_EINVAL, "qz_beam_advance_slices: linear shape needs >= 2 slices");
    if (nc != (modes == 1 ? F_NC1 : F_NC0)) return qz_fail(QZ_EINVAL, "qz_beam_advance_slices: nc must be 6 or 18");
    if (pusher < 0 || pusher > 3) return qz_fail(QZ_EINVAL, "qz_beam_advance_slices: pusher must be 0..3");
    if (np == 0 || s_lo >= s_hi) return QZ_OK;
    const int tb = g_block[QZ_K_BEAM_ADVANCE];
    hipLaunchKernelGGL(beam_advance_kernel, dim3((np + tb - 1) / tb), dim3(tb), 0, static_cast<hipStream_t>(stream),
                       *grid, *box, modes == 1 ? 1 : 0, qm, pusher, 0, np, *p, fld, nc, dt_kick, dt, move, s_lo, s_hi,
                       step, pstep, pushed_count);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

// Beam::sums (src/Beam.cpp:1133-1155): copied to the host and summed in index order like the reference
// ---------------------------------------------------------------------------- on-device beam sums
// The 17 sums of Beam::sums (src/Beam.cpp:1133-1155) as a fixed-shape reduction: block b sums particles
// [b * kSumChunk, (b + 1) * kSumChunk) with thread t taking every kSumBlock-th particle from t, then a tree in
// LDS; a second kernel adds the block partials in block order.  The order depends only on np, so the result
// is reproducible run to run (it differs from the sequential host loop in the last bits).
constexpr int kSumBlock = 256;
constexpr int kSumChunk = 16 * kSumBlock;

__global__ __launch_bounds__(kSumBlock) void beam_sums_partial(int np, qz_beam_soa P, Real gref, Real xiref,
                                                               Real* __restrict__ part) {
    __shared__ Real sh[17][kSumBlock];
    const int t = threadIdx.x;
    const long lo = static_cast<long>(blockIdx.x) * kSumChunk;
    const long hi = min(static_cast<long>(np), lo + kSumChunk);
    Real acc[17];
#pragma unroll
    for (int c = 0; c < 17; ++c) acc[c] = Real(0);
    for (long i = lo + t; i < hi; i += kSumBlock) {
        const Real w = P.w[i];
        if (w == Real(0)) continue;
        acc[16] += Real(1);
        const Real PX = P.px[i], PY = P.py[i], PZ = P.pz[i];
        const Real g = dsqrt(Real(1) + PX * PX + PY * PY + PZ * PZ);
        const Real X = P.x[i], Y = P.y[i];
        const Real G = g - gref, XI = P.xi[i] - xiref;
        const Real v[15] = {G, G * G, XI, XI * XI, X * X + Y * Y, X, Y, X * X, Y * Y, PX, PY, PX * PX, PY * PY,
                            X * PX, Y * PY};
        acc[0] += w;
#pragma unroll
        for (int c = 0; c < 15; ++c) acc[c + 1] += w * v[c];
    }
#pragma unroll
    for (int c = 0; c < 17; ++c) sh[c][t] = acc[c];
    __syncthreads();
    for (int h = kSumBlock / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int c = 0; c < 17; ++c) sh[c][t] += sh[c][t + h];
        __syncthreads();
    }
    if (t < 17) part[static_cast<size_t>(blockIdx.x) * 17 + t] = sh[t][0];
}

__global__ void beam_sums_final(int nb, const Real* __restrict__ part, Real* __restrict__ out) {
    const int c = threadIdx.x;
    if (c >= 17) return;
    Real s = Real(0);
    for (int b = 0; b < nb; ++b) s += part[static_cast<size_t>(b) * 17 + c];
    out[c] = s;
}

extern "C" size_t qz_beam_sums_workspace_bytes(int np) {
    const long nb = (static_cast<long>(np) + kSumChunk - 1) / kSumChunk;
    return static_cast<size_t>(nb > 0 ? nb : 1) * 17 * sizeof(Real);
}

extern "C" int qz_beam_sums_device(int np, const qz_beam_soa* p, double gref, double xiref, double* out17, void* work,
                                   void* stream) {
    if (!p || !out17 || !work || np < 0) return qz_fail(QZ_EINVAL, "qz_beam_sums_device: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nb = static_cast<int>((static_cast<long>(np) + kSumChunk - 1) / kSumChunk);
    Real* part = static_cast<Real*>(work);
    if (nb > 0) {
        hipLaunchKernelGGL(beam_sums_partial, dim3(nb), dim3(kSumBlock), 0, st, np, *p, gref, xiref, part);
        QZ_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(beam_sums_final, dim3(1), dim3(64), 0, st, nb, part, out17);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

extern "C" int qz_beam_sums(int np, const qz_beam_soa* p, double gref, double xiref, double* out, void* stream) {
    if (!p || !out || np < 0) return qz_fail(QZ_EINVAL, "qz_beam_sums: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::vector<double> h[7];
    double* src[7] = {p->x, p->y, p->px, p->py, p->pz, p->xi, p->w};
    for (int c = 0; c < 7; ++c) {
        h[c].resize(static_cast<size_t>(np));
        if (np) QZ_HIP(hipMemcpyAsync(h[c].data(), src[c], 8 * static_cast<size_t>(np), hipMemcpyDeviceToHost, st));
    }
    QZ_HIP(hipStreamSynchronize(st));
    for (int c = 0; c < 17; ++c) out[c] = 0;
    for (int i = 0; i < np; ++i) {
        const double w = h[6][i];
        if (w == 0) continue;
        out[16] += 1;
        const double PX = h[2][i], PY = h[3][i], PZ = h[4][i];
        const double g = std::sqrt(1 + PX * PX + PY * PY + PZ * PZ);
        const double X = h[0][i], Y = h[1][i];
        const double G = g - gref, XI = h[5][i] - xiref;
        const double v[15] = {G, G * G, XI, XI * XI, X * X + Y * Y, X, Y, X * X, Y * Y, PX, PY, PX * PX, PY * PY, X * PX, Y * PY};
        out[0] += w;
        for (int c = 0; c < 15; ++c) out[c + 1] += w * v[c];
    }
    return QZ_OK;
}
