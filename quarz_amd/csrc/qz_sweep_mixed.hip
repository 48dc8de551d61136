// One persistent launch for a batched step whose runs take one or several workgroups each.  Blocks [0, slots)
// are the multi-workgroup runs' workgroups, placed by the host's block map (run -1: an empty slot that works the
// queue, -2: one that ends at once); blocks after them take the one-workgroup runs from a device counter, in the
// host's order (longest first).
// Every block of the launch holds a whole compute unit's LDS and the launch has at most one block per compute
// unit, so all its blocks are resident together; blocks are dispatched in order, so the multi-workgroup runs'
// workgroups never wait behind the queue's.  A one-workgroup run is swept with the fewest particles per thread
// it needs (2, 4 or 8: the three instantiations of sweep_run), not the batch's largest.  Each run's arithmetic
// is that of its own kernel (sweep_run, sweep_multi_run), so every run equals its plain launch bit for bit.
#include "qz_sweep_single.hpp"
#include "qz_sweep_multi_body.hpp"

#include <string>

namespace qz {

static_assert(kBS == mBS && kLdsBytes == mLds, "both sweep bodies run in blocks of one shape");

template <bool M1>
__global__ __launch_bounds__(kBS) void sweep_mixed_kernel(const qz_sweep_args* __restrict__ single, int nsingle,
                                                          unsigned int* __restrict__ queue,
                                                          const qz_sweep_args* __restrict__ multi,
                                                          const int* __restrict__ block_map, int slots) {
    const int b = static_cast<int>(blockIdx.x);
    if (b < slots) {
        const int run = block_map[2 * b];
        if (run == -2) return;   // an empty slot with no queue work
        if (run >= 0) {
            sweep_multi_run<M1, 2>(multi, block_map, b);
            return;
        }
    }
    extern __shared__ double smem_d[];
    int* s_next = reinterpret_cast<int*>(smem_d);   // free between runs
    auto take = [&]() {
        __syncthreads();
        if (threadIdx.x == 0) *s_next = static_cast<int>(atomicAdd(queue, 1u));
        __syncthreads();
        const int r = *s_next;
        __syncthreads();
        return r;
    };
    for (int r = take(); r < nsingle; r = take()) {   // each run with the fewest particles per thread it needs
        const qz_sweep_args& A = single[r];
        int np = 0;
        for (int sp = 0; sp < A.nspecies; ++sp) np += A.sp[sp].np;
        if (np <= 2 * kBS) sweep_run<M1, 2>(A);
        else if (np <= 4 * kBS) sweep_run<M1, 4>(A);
        else sweep_run<M1, 8>(A);
    }
}

}  // namespace qz

using namespace qz;

int qz_sweep_mixed_launch(const qz_sweep_args* dev_single, const qz_sweep_args* host_single, int nsingle,
                          unsigned int* queue, const qz_sweep_args* dev_multi, const int* dev_map,
                          const qz_sweep_args* host_multi, int nmulti, int slots, int workers, hipStream_t stream) {
    if (nsingle < 0 || nmulti < 0 || slots < 0 || workers < 0 || (nsingle > 0 && !queue))
        return qz_fail(QZ_EINVAL, "mixed sweep: bad counts");
    const qz_sweep_args* first = nsingle > 0 ? host_single : host_multi;
    if (!first) return qz_fail(QZ_EINVAL, "mixed sweep: no runs");
    const bool m1 = first->modes == 1;
    int maxNm = 0, need_slots = 0;
    for (int i = 0; i < nsingle; ++i) {
        const qz_sweep_args& a = host_single[i];
        if ((a.modes == 1) != m1) return qz_fail(QZ_EINVAL, "mixed sweep: all runs need the same modes");
        int np = 0;
        for (int s = 0; s < a.nspecies; ++s) np += a.sp[s].np;
        if (np > kBS * 8)
            return qz_fail(QZ_ERANGE, "mixed sweep: more than 4096 plasma particles in a one-workgroup run");
        const int pt = np <= 2 * kBS ? 2 : (np <= 4 * kBS ? 4 : 8);   // as the kernel picks it
        if (sweep_lds_needed(m1, a.grid.N, pt, np) > kLdsBytes)
            return qz_fail(QZ_ERANGE, "mixed sweep: a one-workgroup run does not fit the LDS of one workgroup");
    }
    for (int i = 0; i < nmulti; ++i) {
        const qz_sweep_args& a = host_multi[i];
        if ((a.modes == 1) != m1) return qz_fail(QZ_EINVAL, "mixed sweep: all runs need the same modes");
        if (a.groups < 2 || !a.mws) return qz_fail(QZ_EINVAL, "mixed sweep: a multi-workgroup run needs groups and mws");
        int np = 0;
        for (int s = 0; s < a.nspecies; ++s) np += a.sp[s].np;
        if (qz_sweep_multi_pt(np, a.groups) != 2)
            return qz_fail(QZ_ERANGE, "mixed sweep: a multi-workgroup run with more than 1024 plasma particles per "
                                      "workgroup; give it more workgroups");
        maxNm = std::max(maxNm, a.grid.N);
        need_slots += a.groups;
    }
    if (need_slots > slots) return qz_fail(QZ_EINVAL, "mixed sweep: fewer slots than multi-workgroup workgroups");
    if (nsingle > 0 && workers + slots - need_slots < 1) return qz_fail(QZ_EINVAL, "mixed sweep: no worker");
    if (nmulti > 0 && sweep_multi_lds_needed(m1, maxNm, 2) > mLds)
        return qz_fail(QZ_ERANGE, "mixed sweep: a grid does not fit the LDS of the multi-workgroup sweep");
    auto launch = [&](auto kern) -> int {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
        if (e != hipSuccess) return qz_fail_hip(e, "hipFuncSetAttribute");
        int per_cu = 0, dev = 0, ncu = 0;
        QZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), kBS,
                                                            kLdsBytes));
        QZ_HIP(hipGetDevice(&dev));
        QZ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int nb = slots + workers;
        if (nmulti > 0 && nb > per_cu * ncu)
            return qz_fail(QZ_ERANGE, "mixed sweep: " + std::to_string(nb) + " workgroups cannot be resident together (" +
                                          std::to_string(per_cu * ncu) + " fit)");
        if (nsingle > 0) {
            const int rc = qz_zero_async(queue, sizeof(unsigned int), stream);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(kern, dim3(nb), dim3(kBS), kLdsBytes, stream, dev_single, nsingle, queue, dev_multi, dev_map,
                           slots);
        QZ_HIP(hipGetLastError());
        return QZ_OK;
    };
    return m1 ? launch(sweep_mixed_kernel<true>) : launch(sweep_mixed_kernel<false>);
}
