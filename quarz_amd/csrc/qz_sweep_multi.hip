// Persistent gfx950 kernel of one quasi-static sweep run by G cooperating workgroups; the body
// (sweep_multi_run) and its description are in qz_sweep_multi_body.hpp, shared with the mixed batch kernel
// (qz_sweep_mixed.hip).
#include "qz_sweep_multi_body.hpp"

#include <algorithm>
#include <string>
#include <vector>

namespace qz {

template <bool M1, int PT>
__global__ __launch_bounds__(mBS) void sweep_multi_kernel(const qz_sweep_args* __restrict__ runs,
                                                          const int* __restrict__ block_map) {
    sweep_multi_run<M1, PT>(runs, block_map, static_cast<int>(blockIdx.x));
}

// LDS bytes of the multi-workgroup kernel for N cells and PT particles per thread
long sweep_multi_lds_needed(bool m1, int N, int pt) {
    const MLds plan(N, pt);
    const long nreg = m1 ? MSlots<true>::NREGION : MSlots<false>::NREGION;
    return plan.region + nreg * plan.Mp * 8;
}

}  // namespace qz

// ============================================================================ host launch
using namespace qz;

extern "C" size_t qz_sweep_multi_workspace_doubles(int M, int np, int groups, int modes) {
    if (M < 3 || np < 0 || groups < 1) return 0;
    return MwsLayout(M, M - 1, np, groups, modes == 1).total;
}

std::vector<int> qz_sweep_block_map(const qz_sweep_args* runs, int nruns) {
    std::vector<int> map;
    for (int r = 0; r < nruns; ++r)
        for (int g = 0; g < std::max(runs[r].groups, 1); ++g) {
            map.push_back(r);
            map.push_back(g);
        }
    return map;
}

int qz_sweep_multi_pt(int np, int groups) {
    const int per = (np + groups - 1) / std::max(groups, 1);
    return per <= 2 * mBS ? 2 : (per <= 4 * mBS ? 4 : 8);
}

// An ordinary launch whose workgroups all fit the device at once: the launch's own, or `resident` workgroups when
// the caller keeps several such launches in flight on its streams (R <= the workgroups the device holds with this
// kernel's LDS, one per compute unit).  Every workgroup then gets its compute unit as the ones before it finish,
// and launches on different streams overlap (cooperative launches run one at a time on this runtime, and could
// not be issued once many streams held the hardware queues).  A barrier that is not reached in time ends the
// launch with the run's error flag set, never a hang.
int qz_sweep_multi_launch(const qz_sweep_args* dev_runs, const int* dev_map, const qz_sweep_args* host_runs, int nruns,
                          hipStream_t stream, int resident) {
    int nblocks = 0, maxN = 0, pt = 2;
    const bool m1 = host_runs[0].modes == 1;
    for (int i = 0; i < nruns; ++i) {
        const qz_sweep_args& a = host_runs[i];
        if ((a.modes == 1) != m1) return qz_fail(QZ_EINVAL, "sweep: all runs of a launch need the same modes");
        if (a.groups < 1 || !a.mws) return qz_fail(QZ_EINVAL, "sweep: a multi-workgroup run needs groups and mws");
        int np = 0;
        for (int s = 0; s < a.nspecies; ++s) np += a.sp[s].np;
        if (np > a.groups * mBS * 8)
            return qz_fail(QZ_ERANGE, "sweep: more than 4096 plasma particles per workgroup; give the run more groups");
        pt = std::max(pt, qz_sweep_multi_pt(np, a.groups));
        maxN = std::max(maxN, a.grid.N);
        nblocks += a.groups;
    }
    if (sweep_multi_lds_needed(m1, maxN, pt) > mLds)
        return qz_fail(QZ_ERANGE, "sweep: the grid does not fit the LDS of the multi-workgroup sweep");
    auto launch = [&](auto kern) -> int {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, mLds);
        if (e != hipSuccess) return qz_fail_hip(e, "hipFuncSetAttribute");
        int per_cu = 0, dev = 0, ncu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), mBS, mLds);
        if (e != hipSuccess) return qz_fail_hip(e, "occupancy");
        QZ_HIP(hipGetDevice(&dev));
        QZ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int need = std::max(resident, nblocks);
        if (need > per_cu * ncu)
            return qz_fail(QZ_ERANGE, "sweep: " + std::to_string(need) + " workgroups cannot be resident together (" +
                                          std::to_string(per_cu * ncu) + " fit)");
        hipLaunchKernelGGL(kern, dim3(nblocks), dim3(mBS), mLds, stream, dev_runs, dev_map);
        QZ_HIP(hipGetLastError());
        return QZ_OK;
    };
    if (pt == 2) return m1 ? launch(sweep_multi_kernel<true, 2>) : launch(sweep_multi_kernel<false, 2>);
    if (pt == 4) return m1 ? launch(sweep_multi_kernel<true, 4>) : launch(sweep_multi_kernel<false, 4>);
    return m1 ? launch(sweep_multi_kernel<true, 8>) : launch(sweep_multi_kernel<false, 8>);
}
