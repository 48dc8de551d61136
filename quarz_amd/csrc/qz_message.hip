// The state a sweep hands to the next slice block: the xi-pipeline message of the reference
// (Simulation::make_message, src/Simulation.cpp:851-870; PlasmaSpecies::pack_state / unpack_state,
// src/PlasmaSpecies.cpp:317-364), packed and unpacked in device memory.  The reference copies every array to the
// host and appends it to a std::vector; here the message is one device array written by device copies and two
// small kernels, so it can go straight to a peer GPU (xGMI) or to a checkpoint file without a host round trip.
//
// Layout (doubles), the reference's order:
//   [step][dt][last]  [0 beam particles per beam (one run on one GPU keeps its beams)]
//   per mobile species: [Np][x][y][px][py][Delta][w] (Np each) [fresh] (Np) [history (order, 5) slabs] (Np each)
//                       [pushes][xi of the state][xi of every history slot] (order)
//   [B+ r0 i0 r1 i1 r2 i2] (M each)   [beam sources of the block's last slice] (M x 15, component fastest)
#include <hip/hip_runtime.h>

#include <vector>

#include "qz_device.hpp"
#include "qz_internal.hpp"
#include "qz_sweep_common.hpp"

namespace qz {
namespace {

// scalars of the message, written by a kernel that takes them by value (capturable; no host memory read)
struct Scalars {
    int n;
    long long at[24];
    double v[24];
};

__global__ void put_scalars(Scalars s, double* __restrict__ msg) {
    const int i = threadIdx.x;
    if (i < s.n) msg[s.at[i]] = s.v[i];
}

__global__ void fill_value(double* __restrict__ p, long long n, double v) {
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        p[i] = v;
}

// the beam sources of one slice row, M nodes x nc components, padded to 15 components per node
__global__ void guard_row(const double* __restrict__ row, int M, int nc, double* __restrict__ out) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < M * B_NC1; e += gridDim.x * blockDim.x) {
        const int j = e / B_NC1, c = e - j * B_NC1;
        out[e] = c < nc ? row[static_cast<size_t>(j) * nc + c] : 0.0;
    }
}

// the B+ arrays in the message's order r0 i0 r1 i1 r2 i2 against the workspace order r1 i1 r0 i0 r2 i2
constexpr int kBpFromWs[6] = {W_BPR0, W_BPI0, W_BPR1, W_BPI1, W_BPR2, W_BPI2};

size_t species_doubles(int np, int order) { return 1 + 7 * static_cast<size_t>(np) + 5 * static_cast<size_t>(order) * np + 2 + order; }
}  // namespace
}  // namespace qz

using namespace qz;

extern "C" size_t qz_message_doubles(const qz_sweep_args* a, int nbeams) {
    if (!a) return 0;
    size_t n = 3 + static_cast<size_t>(nbeams < 0 ? 0 : nbeams);
    for (int s = 0; s < a->nspecies; ++s) n += species_doubles(a->sp[s].np, a->ab_order);
    const size_t M = static_cast<size_t>(a->grid.N) + 1;
    return n + 6 * M + M * B_NC1;
}

extern "C" int qz_pack_message(const qz_sweep_args* a, int nbeams, int step, double dt, int last, int k_next,
                               double* msg, void* stream) {
    if (!a || !msg || nbeams < 0) return qz_fail(QZ_EINVAL, "qz_pack_message: null argument");
    if (a->groups > 1 || !a->ws || !a->bsrc) return qz_fail(QZ_EINVAL, "qz_pack_message: needs ws and bsrc of the run");
    if (k_next < 0 || k_next > a->nslices) return qz_fail(QZ_EINVAL, "qz_pack_message: slice outside the box");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int M = a->grid.N + 1, ord = a->ab_order;
    Scalars sc{};
    auto put = [&](long long at, double v) {
        if (sc.n == 24) {   // a full set goes out; errors are read after the last launch
            hipLaunchKernelGGL(put_scalars, dim3(1), dim3(64), 0, st, sc, msg);
            sc.n = 0;
        }
        sc.at[sc.n] = at;
        sc.v[sc.n] = v;
        ++sc.n;
    };
    long long off = 0;
    put(off++, step);
    put(off++, dt);
    put(off++, last ? 1.0 : 0.0);
    for (int b = 0; b < nbeams; ++b) put(off++, 0.0);   // no beam particles change hands on one GPU
    const double xi_state = a->dxi * k_next;   // internal xi of the slice the state enters (box head at 0)
    for (int s = 0; s < a->nspecies; ++s) {
        const qz_species_state& sp = a->sp[s];
        const size_t np = static_cast<size_t>(sp.np);
        put(off++, static_cast<double>(sp.np));
        const double* arr[6] = {sp.x, sp.y, sp.px, sp.py, sp.dl, sp.w};
        for (int v = 0; v < 6; ++v, off += np)
            if (np) QZ_HIP(hipMemcpyAsync(msg + off, arr[v], 8 * np, hipMemcpyDeviceToDevice, st));
        if (np) {   // particles not yet pushed since their load: all of them before the first slice, none after
            hipLaunchKernelGGL(fill_value, dim3(static_cast<unsigned>((np + 255) / 256)), dim3(256), 0, st, msg + off,
                               static_cast<long long>(np), k_next == 0 ? 1.0 : 0.0);
            QZ_HIP(hipGetLastError());
        }
        off += np;
        for (int sl = 0; sl < ord; ++sl)
            for (int v = 0; v < 5; ++v, off += np)
                if (np)
                    QZ_HIP(hipMemcpyAsync(msg + off, sp.hist + (static_cast<size_t>(sl) * 5 + v) * sp.cap, 8 * np,
                                          hipMemcpyDeviceToDevice, st));
        // pushes done (one after every slice before k_next), xi of the state, xi at which each slot was written
        put(off++, static_cast<double>(k_next));
        put(off++, xi_state);
        for (int sl = 0; sl < ord; ++sl) {
            int kw = k_next - 1;   // the last slice before k_next whose push wrote slot sl (k % order)
            while (kw >= 0 && kw % ord != sl) --kw;
            put(off++, kw >= 0 ? a->dxi * kw : 0.0);
        }
    }
    for (int f = 0; f < 6; ++f, off += M) {
        const bool present = a->modes == 1 || kBpFromWs[f] == W_BPR1 || kBpFromWs[f] == W_BPI1;
        if (present)
            QZ_HIP(hipMemcpyAsync(msg + off, a->ws + static_cast<size_t>(kBpFromWs[f]) * M, 8 * static_cast<size_t>(M),
                                  hipMemcpyDeviceToDevice, st));
        else {
            hipLaunchKernelGGL(fill_value, dim3(static_cast<unsigned>((M + 255) / 256)), dim3(256), 0, st, msg + off,
                               static_cast<long long>(M), 0.0);
            QZ_HIP(hipGetLastError());
        }
    }
    const int krow = k_next > 0 ? k_next - 1 : 0;   // the block's last slice
    hipLaunchKernelGGL(guard_row, dim3(static_cast<unsigned>((M * B_NC1 + 255) / 256)), dim3(256), 0, st,
                       a->bsrc + static_cast<size_t>(krow) * M * a->ncb, M, a->ncb, msg + off);
    QZ_HIP(hipGetLastError());
    if (sc.n) {
        hipLaunchKernelGGL(put_scalars, dim3(1), dim3(64), 0, st, sc, msg);
        QZ_HIP(hipGetLastError());
    }
    return QZ_OK;
}

extern "C" int qz_unpack_message(qz_sweep_args* a, int nbeams, const double* msg, double* guard, void* stream) {
    if (!a || !msg || nbeams < 0) return qz_fail(QZ_EINVAL, "qz_unpack_message: null argument");
    if (a->groups > 1 || !a->ws) return qz_fail(QZ_EINVAL, "qz_unpack_message: needs ws of the run");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int M = a->grid.N + 1, ord = a->ab_order;
    size_t off = 3 + static_cast<size_t>(nbeams);
    for (int s = 0; s < a->nspecies; ++s) {
        qz_species_state& sp = a->sp[s];
        double npd = 0;
        QZ_HIP(hipMemcpyAsync(&npd, msg + off, 8, hipMemcpyDeviceToHost, st));
        QZ_HIP(hipStreamSynchronize(st));
        const int np = static_cast<int>(npd);
        if (np < 0 || np > sp.cap) return qz_fail(QZ_ERANGE, "qz_unpack_message: more particles than the species holds");
        sp.np = np;
        ++off;
        double* arr[6] = {sp.x, sp.y, sp.px, sp.py, sp.dl, sp.w};
        for (int v = 0; v < 6; ++v, off += np)
            if (np) QZ_HIP(hipMemcpyAsync(arr[v], msg + off, 8 * static_cast<size_t>(np), hipMemcpyDeviceToDevice, st));
        off += np;   // fresh flags: the sweep keeps no per-particle start-up state
        for (int sl = 0; sl < ord; ++sl)
            for (int v = 0; v < 5; ++v, off += np)
                if (np)
                    QZ_HIP(hipMemcpyAsync(sp.hist + (static_cast<size_t>(sl) * 5 + v) * sp.cap, msg + off,
                                          8 * static_cast<size_t>(np), hipMemcpyDeviceToDevice, st));
        off += 2 + ord;
    }
    for (int f = 0; f < 6; ++f, off += M)
        if (a->modes == 1 || kBpFromWs[f] == W_BPR1 || kBpFromWs[f] == W_BPI1)
            QZ_HIP(hipMemcpyAsync(a->ws + static_cast<size_t>(kBpFromWs[f]) * M, msg + off, 8 * static_cast<size_t>(M),
                                  hipMemcpyDeviceToDevice, st));
    if (guard)
        QZ_HIP(hipMemcpyAsync(guard, msg + off, 8 * static_cast<size_t>(M) * B_NC1, hipMemcpyDeviceToDevice, st));
    a->load = 0;   // the arrays now hold the state: a launch from slice 0 must not load rings over it
    return QZ_OK;
}
