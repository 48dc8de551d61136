// Persistent gfx950 kernel of one quasi-static sweep in one workgroup per run; the body (sweep_run) and its
// description are in qz_sweep_single.hpp, shared with the mixed batch kernel (qz_sweep_mixed.hip).
#include "qz_sweep_single.hpp"

#include <algorithm>

namespace qz {

// The sweep kernel: workgroup b sweeps run b
template <bool M1, int PT>
__global__ __launch_bounds__(kBS) void sweep_kernel(const qz_sweep_args* __restrict__ runs) {
    sweep_run<M1, PT>(runs[blockIdx.x]);
}

// LDS bytes the kernel needs for N cells, PT particles per thread and L live particles
long sweep_lds_needed(bool m1, int N, int pt, int L) {
    const LdsPlan plan(N, pt);
    const long Mp = plan.Mp;
    const long nmax = m1 ? Lay<true>::NMAX : Lay<false>::NMAX;
    const long nsdep = m1 ? Lay<true>::NSDEP : Lay<false>::NSDEP;
    long need = nmax * Mp;
    need = std::max(need, nsdep * Mp + 2L * L);   // staged fields + one S+ contribution list
    need = std::max(need, 2L * L);                // one deposit contribution list
    return plan.region + need * 8;
}

}  // namespace qz

// ============================================================================ host launch
using namespace qz;

static_assert(W_NARR == QZ_SWEEP_WS_ARRAYS, "the workspace size the no-allocation sweep reserves");
size_t qz_sweep_workspace_doubles_impl(int M) { return static_cast<size_t>(W_NARR) * M; }

bool qz_sweep_fits_one_group(bool m1, int N, int np) {
    if (np > kBS * 8) return false;
    const int pt = np <= 2 * kBS ? 2 : (np <= 4 * kBS ? 4 : 8);
    return sweep_lds_needed(m1, N, pt, np) <= kLdsBytes;
}

int qz_sweep_launch(const qz_sweep_args* dev_runs, const qz_sweep_args* host_runs, int nruns, hipStream_t stream) {
    int maxnp = 0, maxN = 0;
    const bool m1 = host_runs[0].modes == 1;
    for (int i = 0; i < nruns; ++i) {
        int np = 0;
        for (int s = 0; s < host_runs[i].nspecies; ++s) np += host_runs[i].sp[s].np;
        maxnp = std::max(maxnp, np);
        maxN = std::max(maxN, host_runs[i].grid.N);
        if ((host_runs[i].modes == 1) != m1) return qz_fail(QZ_EINVAL, "qz_sweep_batched: all runs need the same modes");
    }
    if (maxnp > kBS * 8)
        return qz_fail(QZ_ERANGE, "sweep: more than 4096 plasma particles per slice (one workgroup per run)");
    const int pt = maxnp <= 2 * kBS ? 2 : (maxnp <= 4 * kBS ? 4 : 8);
    if (sweep_lds_needed(m1, maxN, pt, maxnp) > kLdsBytes)
        return qz_fail(QZ_ERANGE, "sweep: the grid and the particle count do not fit the LDS of one workgroup");
    auto launch = [&](auto kern) -> int {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
        if (e != hipSuccess) return qz_fail_hip(e, "hipFuncSetAttribute");
        hipLaunchKernelGGL(kern, dim3(nruns), dim3(kBS), kLdsBytes, stream, dev_runs);
        e = hipGetLastError();
        if (e != hipSuccess) return qz_fail_hip(e, "sweep launch");
        return QZ_OK;
    };
    if (pt == 2) return m1 ? launch(sweep_kernel<true, 2>) : launch(sweep_kernel<false, 2>);
    if (pt == 4) return m1 ? launch(sweep_kernel<true, 4>) : launch(sweep_kernel<false, 4>);
    return m1 ? launch(sweep_kernel<true, 8>) : launch(sweep_kernel<false, 8>);
}
